"""Python binding of libafmm.so (include/afmm.h): the MI355X adaptive FMM
force solver behind the reference's ForceSolver3D seam.

    solver = AdaptiveFMM3D(G=1.0, eps=0.01, p=8, threshold=64)
    acc = solver.computeAccelerations(pos, mass)     # (n,3) float64, input order

mirrors galaxy3d::AdaptiveFMM3D(KernelParams3D{G, eps}, p, threshold)
(src3d/AdaptiveFMM3D.hpp:22-23, :29-110) and its ForceSolver3D interface
(src3d/ForceSolver3D.hpp:20-28). Errors raise AfmmError (the reference has no
error channel; its CUDA twin calls exit). The library must be built
(`make -C on-galaxy_amd`); there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafmm.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)

PRECISION_FP64 = 0
PRECISION_MIXED = 1


class AfmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"afmm error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("G", C.c_double), ("eps", C.c_double), ("p", C.c_int32), ("leaf_cap", C.c_int32),
                ("precision", C.c_int32), ("device", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n_bodies", C.c_int64), ("n_boxes", C.c_int64), ("n_leaves", C.c_int64),
                ("n_colleague", C.c_int64), ("n_u", C.c_int64), ("n_v", C.c_int64), ("n_w", C.c_int64),
                ("n_x", C.c_int64), ("p2p_pairs", C.c_int64), ("max_level", C.c_int32),
                ("max_leaf_occ", C.c_int32), ("ms_tree", C.c_double), ("ms_lists", C.c_double),
                ("ms_upward", C.c_double), ("ms_m2l", C.c_double), ("ms_downward", C.c_double),
                ("ms_leaf", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tree(C.Structure):
    _fields_ = [("parent", _ip), ("child", _ip), ("level", _ip), ("center", _dp), ("half", _dp),
                ("leaf", _bp), ("pfirst", _ip), ("pcount", _ip), ("px", _dp), ("py", _dp), ("pz", _dp),
                ("mass", _dp), ("orig_index", _ip), ("leaves", _ip),
                ("colleague_off", _ip), ("colleague_idx", _ip), ("u_off", _ip), ("u_idx", _ip),
                ("v_off", _ip), ("v_idx", _ip), ("w_off", _ip), ("w_idx", _ip), ("x_off", _ip),
                ("x_idx", _ip), ("tol", C.c_double)]


_lib = None


def library():
    """Load libafmm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: build it with `make -C on-galaxy_amd`")
        L = C.CDLL(LIB_PATH)
        L.afmm_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.afmm_destroy.argtypes = [C.c_void_p]
        L.afmm_strerror.argtypes = [C.c_int]
        L.afmm_strerror.restype = C.c_char_p
        L.afmm_last_error.argtypes = [C.c_void_p]
        L.afmm_last_error.restype = C.c_char_p
        L.afmm_eval.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp]
        L.afmm_eval_device.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.afmm_synchronize.argtypes = [C.c_void_p]
        L.afmm_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.afmm_tree_export.argtypes = [C.c_void_p, C.POINTER(Tree)]
        _lib = L
    return _lib


def _p(a):
    t = {np.float64: _dp, np.int32: _ip, np.int8: _bp}[a.dtype.type]
    return a.ctypes.data_as(t)


class AdaptiveFMM3D:
    """ForceSolver3D drop-in: computeAccelerations(pos, mass) -> (n,3)."""

    def __init__(self, G=1.0, eps=0.05, p=8, threshold=64, precision=PRECISION_FP64, device=0):
        self.lib = library()
        self.params = Params(G, eps, p, threshold, precision, device)
        h = C.c_void_p()
        rc = self.lib.afmm_create(C.byref(self.params), C.byref(h))
        if rc:
            raise AfmmError(rc, self.lib.afmm_strerror(rc).decode())
        self.ctx = h

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.afmm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise AfmmError(rc, self.lib.afmm_last_error(self.ctx).decode() or
                            self.lib.afmm_strerror(rc).decode())

    def name(self):
        return "AdaptiveFMM3D(MI355X HIP, FP64)"

    def computeAccelerations(self, pos, mass):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        if len(mass) != len(pos):
            raise ValueError("pos and mass lengths differ")
        acc = np.zeros_like(pos)
        self._check(self.lib.afmm_eval(self.ctx, len(mass), _p(pos), _p(mass), _p(acc)))
        return acc

    def compute_device(self, pos_ptr, mass_ptr, acc_ptr, n):
        """Device-resident evaluation on caller-owned device pointers."""
        self._check(self.lib.afmm_eval_device(self.ctx, n, C.c_void_p(pos_ptr), C.c_void_p(mass_ptr),
                                               C.c_void_p(acc_ptr)))

    def synchronize(self):
        self._check(self.lib.afmm_synchronize(self.ctx))

    def stats(self):
        st = Stats()
        self._check(self.lib.afmm_get_stats(self.ctx, C.byref(st)))
        return st.as_dict()

    def tree(self):
        """Export of the last evaluation's octree in the reference numbering."""
        st = self.stats()
        n, nb, nl = st["n_bodies"], st["n_boxes"], st["n_leaves"]
        counts = {"colleague": st["n_colleague"], "u": st["n_u"], "v": st["n_v"], "w": st["n_w"],
                  "x": st["n_x"]}
        a = {"parent": np.empty(nb, np.int32), "child": np.empty(8 * nb, np.int32),
             "level": np.empty(nb, np.int32), "center": np.empty(3 * nb), "half": np.empty(nb),
             "leaf": np.empty(nb, np.int8), "pfirst": np.empty(nb, np.int32),
             "pcount": np.empty(nb, np.int32), "px": np.empty(n), "py": np.empty(n), "pz": np.empty(n),
             "mass": np.empty(n), "orig_index": np.empty(n, np.int32), "leaves": np.empty(nl, np.int32)}
        for k, cnt in counts.items():
            a[k + "_off"] = np.empty(nb + 1, np.int32)
            a[k + "_idx"] = np.empty(max(cnt, 1), np.int32)
        t = Tree()
        for k, v in a.items():
            setattr(t, k, _p(v))
        self._check(self.lib.afmm_tree_export(self.ctx, C.byref(t)))
        for k, cnt in counts.items():
            a[k + "_idx"] = a[k + "_idx"][:cnt]
        a["tol"] = t.tol
        return a
