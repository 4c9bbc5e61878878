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
PRECISION_MIXED_BF16 = 2
# one reduced-precision phase alone (error attribution): bf16 M2L with FP64
# P2P, and FP64 M2L with FP32 P2P
PRECISION_M2L_BF16 = 3
PRECISION_P2P_FP32 = 4
# afmm_set_overlap modes
OVERLAP_OFF, OVERLAP_AUTO, OVERLAP_BESIDE_M2L, OVERLAP_IN_ORDER = 0, 1, 2, 3

AFMM_ERR_DEVICE = -5
AFMM_ERR_NO_MEMORY = -6


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
                ("ms_leaf", C.c_double), ("ms_total", C.c_double),
                ("ms_p2p", C.c_double), ("graph_instantiations", C.c_int64), ("graph_updates", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Shapes(C.Structure):
    _fields_ = [("m2l_waves", C.c_int32), ("p2p_pair_bodies", C.c_int32), ("p2p_min_group", C.c_int32),
                ("p2p_schedule", C.c_int32)]


class Tree(C.Structure):
    _fields_ = [("parent", _ip), ("child", _ip), ("level", _ip), ("center", _dp), ("half", _dp),
                ("leaf", _bp), ("pfirst", _ip), ("pcount", _ip), ("px", _dp), ("py", _dp), ("pz", _dp),
                ("mass", _dp), ("orig_index", _ip), ("leaves", _ip),
                ("colleague_off", _ip), ("colleague_idx", _ip), ("u_off", _ip), ("u_idx", _ip),
                ("v_off", _ip), ("v_idx", _ip), ("w_off", _ip), ("w_idx", _ip), ("x_off", _ip),
                ("x_idx", _ip), ("tol", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("leaf_cap", C.c_int32), ("precision", C.c_int32),
                ("device", C.c_int32), ("max_batch", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("n", C.c_int64), ("pos_xyz", _dp), ("vel_xyz", _dp), ("mass", _dp), ("G", C.c_double),
                ("eps", C.c_double), ("dt", C.c_double), ("p", C.c_int32), ("steps", C.c_int32)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("hbm_bytes", C.c_int64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("steps_done", C.c_int32),
                ("batches", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3

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
        L.afmm_eval_potential.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp]
        L.afmm_eval_device_potential.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]
        L.afmm_synchronize.argtypes = [C.c_void_p]
        L.afmm_eval_batch.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), _dp, _dp, _dp, _dp, _dp, _dp]
        L.afmm_eval_batch_device.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), _dp, _dp, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p]
        L.afmm_state_create.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), _dp, _dp, _dp, _dp, _dp,
                                        C.POINTER(C.c_void_p)]
        L.afmm_state_step.argtypes = [C.c_void_p, C.c_double, C.c_int32]
        L.afmm_state_download.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.afmm_state_diagnostics.argtypes = [C.c_void_p, _dp]
        L.afmm_state_destroy.argtypes = [C.c_void_p]
        L.afmm_job_bytes.argtypes = [C.c_int64, C.c_int32]
        L.afmm_job_bytes.restype = C.c_int64
        L.afmm_server_create.argtypes = [C.POINTER(ServerParams), C.POINTER(C.c_void_p)]
        L.afmm_server_warm.argtypes = [C.c_void_p, C.c_int32]
        L.afmm_server_submit.argtypes = [C.c_void_p, C.POINTER(JobDesc), C.POINTER(C.c_int64)]
        L.afmm_server_wait.argtypes = [C.c_void_p, C.c_int64]
        L.afmm_server_drain.argtypes = [C.c_void_p]
        L.afmm_server_info.argtypes = [C.c_void_p, C.c_int64, C.POINTER(JobInfo)]
        L.afmm_server_result.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp]
        L.afmm_server_hbm_in_use.argtypes = [C.c_void_p]
        L.afmm_server_hbm_in_use.restype = C.c_int64
        L.afmm_server_destroy.argtypes = [C.c_void_p]
        L.afmm_direct.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, C.c_int64, C.POINTER(C.c_int64), _dp, _dp]
        L.afmm_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.afmm_set_graph.argtypes = [C.c_void_p, C.c_int]
        L.afmm_set_overlap.argtypes = [C.c_void_p, C.c_int]
        L.afmm_set_shapes.argtypes = [C.c_void_p, C.POINTER(Shapes)]
        L.afmm_get_shapes.argtypes = [C.c_void_p, C.POINTER(Shapes)]
        L.afmm_direct_device_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.afmm_tree_export.argtypes = [C.c_void_p, C.POINTER(Tree)]
        L.afmm_make_plummer.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint32, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    t = {np.float64: _dp, np.int32: _ip, np.int8: _bp}[a.dtype.type]
    return a.ctypes.data_as(t)


def _offsets(sys_off, n):
    """System offsets as int64, checked against the n bodies they index: the C
    side reads and writes 3 * sys_off[-1] doubles of the position and result
    arrays, so offsets that do not start at 0, are not strictly increasing or
    do not end at n are rejected here rather than overrunning those buffers."""
    off = np.ascontiguousarray(sys_off, np.int64).reshape(-1)
    if off.size < 2:
        raise ValueError("sys_off needs at least two entries (one system)")
    if off[0] != 0:
        raise ValueError("sys_off[0] must be 0")
    if np.any(np.diff(off) <= 0):
        raise ValueError("sys_off must be strictly increasing (every system needs a body)")
    if off[-1] != n:
        raise ValueError(f"sys_off[-1] = {int(off[-1])} does not match the {n} bodies")
    return off


def make_plummer(n, a=1.0, total_mass=1.0, G=1.0, seed=7):
    """makePlummerSphere (InitialConditions3D.hpp:100-152) on libstdc++'s
    mt19937 stream, generated on the host by the library: (pos (n,3), vel (n,3),
    mass (n,)) identical to the reference's bodies for the same seed."""
    n = int(n)
    pos, vel, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    rc = library().afmm_make_plummer(n, a, total_mass, G, int(seed) & 0xFFFFFFFF, _p(pos), _p(vel), _p(mass))
    if rc:
        raise AfmmError(rc, library().afmm_strerror(rc).decode())
    return pos, vel, mass


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
        prec = {PRECISION_MIXED: "mixed", PRECISION_MIXED_BF16: "mixed-bf16", PRECISION_M2L_BF16: "bf16-M2L",
                PRECISION_P2P_FP32: "FP32-P2P"}.get(self.params.precision, "FP64")
        return f"AdaptiveFMM3D(MI355X HIP, {prec})"

    def computeAccelerations(self, pos, mass):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        if len(mass) != len(pos):
            raise ValueError("pos and mass lengths differ")
        acc = np.zeros_like(pos)
        self._check(self.lib.afmm_eval(self.ctx, len(mass), _p(pos), _p(mass), _p(acc)))
        return acc

    def computeAccelerationsAndPotential(self, pos, mass):
        """(acc (n,3), phi (n,)): phi_i = -G sum_{j != i} m_j / sqrt(|d|^2 + eps^2)
        through the same expansions; 0.5 * mass @ phi is the potential energy."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        if len(mass) != len(pos):
            raise ValueError("pos and mass lengths differ")
        acc = np.zeros_like(pos)
        phi = np.zeros(len(mass))
        self._check(self.lib.afmm_eval_potential(self.ctx, len(mass), _p(pos), _p(mass), _p(acc), _p(phi)))
        return acc, phi

    def compute_device(self, pos_ptr, mass_ptr, acc_ptr, n, phi_ptr=None):
        """Device-resident evaluation on caller-owned device pointers."""
        if phi_ptr is None:
            self._check(self.lib.afmm_eval_device(self.ctx, n, C.c_void_p(pos_ptr), C.c_void_p(mass_ptr),
                                                   C.c_void_p(acc_ptr)))
        else:
            self._check(self.lib.afmm_eval_device_potential(self.ctx, n, C.c_void_p(pos_ptr),
                                                             C.c_void_p(mass_ptr), C.c_void_p(acc_ptr),
                                                             C.c_void_p(phi_ptr)))

    def computeBatch(self, pos, mass, sys_off, G=None, eps=None, potential=False):
        """Batched evaluation of independent systems (one launch per phase):
        system q = rows sys_off[q]:sys_off[q+1] of pos/mass; G, eps per system
        (None: the solver's). Returns acc (and phi when potential=True)."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        if len(mass) != len(pos):
            raise ValueError("pos and mass lengths differ")
        off = _offsets(sys_off, len(mass))
        nsys = len(off) - 1
        g = None if G is None else np.ascontiguousarray(np.broadcast_to(G, (nsys,)), np.float64)
        e = None if eps is None else np.ascontiguousarray(np.broadcast_to(eps, (nsys,)), np.float64)
        acc = np.zeros_like(pos)
        phi = np.zeros(len(mass)) if potential else None
        self._check(self.lib.afmm_eval_batch(self.ctx, nsys, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                             None if g is None else _p(g), None if e is None else _p(e),
                                             _p(pos), _p(mass), _p(acc), None if phi is None else _p(phi)))
        return (acc, phi) if potential else acc

    def compute_batch_device(self, sys_off, pos_ptr, mass_ptr, acc_ptr, G=None, eps=None, phi_ptr=None):
        """Batched evaluation on device pointers (sys_off, G, eps on the host);
        the device arrays must hold sys_off[-1] bodies."""
        off = np.ascontiguousarray(sys_off, np.int64).reshape(-1)
        off = _offsets(off, int(off[-1]) if off.size else 0)
        nsys = len(off) - 1
        g = None if G is None else np.ascontiguousarray(np.broadcast_to(G, (nsys,)), np.float64)
        e = None if eps is None else np.ascontiguousarray(np.broadcast_to(eps, (nsys,)), np.float64)
        self._check(self.lib.afmm_eval_batch_device(
            self.ctx, nsys, off.ctypes.data_as(C.POINTER(C.c_int64)), None if g is None else _p(g),
            None if e is None else _p(e), C.c_void_p(pos_ptr), C.c_void_p(mass_ptr), C.c_void_p(acc_ptr),
            C.c_void_p(phi_ptr) if phi_ptr else None))

    def direct(self, pos, mass, targets, potential=False):
        """FP64 direct sum (DirectSolver3D semantics, Kernel3D.hpp:31-42) of all
        bodies onto the bodies `targets`: (ntgt,3) accelerations, plus the
        softened potentials when potential=True."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        tgt = np.ascontiguousarray(targets, np.int64)
        acc = np.zeros((len(tgt), 3))
        phi = np.zeros(len(tgt))
        self._check(self.lib.afmm_direct(self.ctx, len(mass), _p(pos), _p(mass), len(tgt),
                                         tgt.ctypes.data_as(C.POINTER(C.c_int64)), _p(acc), _p(phi)))
        return (acc, phi) if potential else acc

    def synchronize(self):
        self._check(self.lib.afmm_synchronize(self.ctx))

    def set_graph(self, enable=True):
        """Graph mode: the FMM passes of each evaluation are captured into a
        HIP graph and launched from it (results unchanged)."""
        self._check(self.lib.afmm_set_graph(self.ctx, 1 if enable else 0))

    def set_shapes(self, m2l_waves=0, p2p_pair_bodies=0, p2p_min_group=0, p2p_schedule=0):
        """Kernel shapes (afmm_set_shapes; 0 keeps a default). p2p_schedule 1 is
        the persistent work-binned P2P, 2 the per-leaf grid in tree order."""
        sh = Shapes(m2l_waves, p2p_pair_bodies, p2p_min_group, p2p_schedule)
        self._check(self.lib.afmm_set_shapes(self.ctx, C.byref(sh)))

    def shapes(self):
        sh = Shapes()
        self._check(self.lib.afmm_get_shapes(self.ctx, C.byref(sh)))
        return {k: getattr(sh, k) for k, _ in Shapes._fields_}

    def set_overlap(self, enable=True):
        """Two-stream overlap (afmm_set_overlap): True or OVERLAP_AUTO (the
        default: the upward pass beside the list build, P2P beside the
        far-field passes, and with the bf16 or split-FP32 M2L, or FP64 up to
        p = 8, from p = 14 and at p = 9-13 for up to 500,000 bodies, the
        lists, P2L, M2P and P2P beside M2L),
        OVERLAP_BESIDE_M2L,
        OVERLAP_IN_ORDER, or
        False / OVERLAP_OFF (one stream, in order). Results are the same bits
        in every mode."""
        mode = OVERLAP_AUTO if enable is True else OVERLAP_OFF if enable is False else int(enable)
        self._check(self.lib.afmm_set_overlap(self.ctx, mode))

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


class LeapfrogState:
    """Device-resident velocity Verlet of one or more systems
    (VelocityVerlet3D, Integrator3D.hpp:19-40; afmm_state_* in afmm.h).

        st = LeapfrogState(solver, pos, vel, mass)          # initialize
        st.step(dt, nsteps); pos, vel, acc = st.download()
        diag = st.diagnostics()   # per system: kinetic, potential, total, momentum, angular
    """

    def __init__(self, solver, pos, vel, mass, sys_off=None, G=None, eps=None):
        self.solver = solver
        lib = solver.lib
        self.lib = lib
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        if not (len(mass) == len(pos) == len(vel)):
            raise ValueError("pos, vel and mass lengths differ")
        off = _offsets([0, len(mass)] if sys_off is None else sys_off, len(mass))
        self.nsys = len(off) - 1
        self.n = len(mass)
        g = None if G is None else np.ascontiguousarray(np.broadcast_to(G, (self.nsys,)), np.float64)
        e = None if eps is None else np.ascontiguousarray(np.broadcast_to(eps, (self.nsys,)), np.float64)
        h = C.c_void_p()
        solver._check(lib.afmm_state_create(solver.ctx, self.nsys, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                            None if g is None else _p(g), None if e is None else _p(e),
                                            _p(pos), _p(vel), _p(mass), C.byref(h)))
        self.h = h

    def step(self, dt, nsteps=1):
        self.solver._check(self.lib.afmm_state_step(self.h, dt, nsteps))

    def download(self):
        pos = np.empty((self.n, 3))
        vel = np.empty((self.n, 3))
        acc = np.empty((self.n, 3))
        self.solver._check(self.lib.afmm_state_download(self.h, _p(pos), _p(vel), _p(acc)))
        return pos, vel, acc

    def diagnostics(self):
        out = np.empty(9 * self.nsys)
        self.solver._check(self.lib.afmm_state_diagnostics(self.h, _p(out)))
        out = out.reshape(self.nsys, 9)
        return [{"kinetic": r[0], "potential": r[1], "total": r[2], "momentum": r[3:6], "angular": r[6:9]}
                for r in out]

    def close(self):
        if getattr(self, "h", None):
            self.lib.afmm_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server with continuous batching (afmm_server_* in afmm.h).

        sv = Server(hbm_budget=64 << 30)
        jid = sv.submit(pos, vel, mass, G=1, eps=0.01, dt=0.01, p=8, steps=10)
        sv.wait(jid); pos, vel, acc = sv.result(jid)
    """

    def __init__(self, hbm_budget=0, leaf_cap=64, precision=PRECISION_FP64, device=0, max_batch=0):
        self.lib = library()
        self.prm = ServerParams(hbm_budget, leaf_cap, precision, device, max_batch)
        h = C.c_void_p()
        rc = self.lib.afmm_server_create(C.byref(self.prm), C.byref(h))
        if rc:
            raise AfmmError(rc, self.lib.afmm_strerror(rc).decode())
        self.h = h
        self.sizes = {}

    def _check(self, rc):
        if rc:
            raise AfmmError(rc, self.lib.afmm_strerror(rc).decode())

    def warm(self, p):
        self._check(self.lib.afmm_server_warm(self.h, p))

    def submit(self, pos, vel, mass, G=1.0, eps=0.01, dt=0.01, p=8, steps=0):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        d = JobDesc(len(mass), _p(pos), None if vel is None else _p(vel), _p(mass), G, eps, dt, p, steps)
        jid = C.c_int64()
        self._check(self.lib.afmm_server_submit(self.h, C.byref(d), C.byref(jid)))
        self.sizes[jid.value] = len(mass)
        return jid.value

    def wait(self, jid):
        self._check(self.lib.afmm_server_wait(self.h, jid))

    def drain(self):
        self._check(self.lib.afmm_server_drain(self.h))

    def info(self, jid):
        i = JobInfo()
        self._check(self.lib.afmm_server_info(self.h, jid, C.byref(i)))
        return i.as_dict()

    def result(self, jid):
        n = self.sizes.pop(jid)
        pos, vel, acc = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3))
        self._check(self.lib.afmm_server_result(self.h, jid, _p(pos), _p(vel), _p(acc)))
        return pos, vel, acc

    def hbm_in_use(self):
        return self.lib.afmm_server_hbm_in_use(self.h)

    @staticmethod
    def job_bytes(n, p):
        return library().afmm_job_bytes(n, p)

    def close(self):
        if getattr(self, "h", None):
            self.lib.afmm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """An array (float64 or int64) in device memory, allocated through the HIP runtime that
    libafmm.so itself links (libamdhip64.so.7), so that device-pointer entry
    points (compute_device, compute_batch_device) can be driven without a
    second runtime in the process. `ptr` is the device address."""
    _hip = None

    @classmethod
    def _runtime(cls):
        if cls._hip is None:
            library()  # loads libamdhip64.so.7 as libafmm's dependency
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, shape, host=None, dtype=np.float64):
        self.dtype = np.dtype(dtype if host is None else np.asarray(host).dtype)
        if self.dtype not in (np.dtype(np.float64), np.dtype(np.int64)):
            raise ValueError("DeviceArray holds float64 or int64")
        self.shape = tuple(np.atleast_1d(shape)) if host is None else np.shape(host)
        self.nbytes = int(np.prod(self.shape)) * 8
        h = self._runtime()
        p = C.c_void_p()
        rc = h.hipMalloc(C.byref(p), max(self.nbytes, 8))
        if rc:
            raise AfmmError(AFMM_ERR_NO_MEMORY, f"hipMalloc({self.nbytes}) failed: {rc}")
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host):
        a = np.ascontiguousarray(host, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        rc = self._runtime().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, 1)
        if rc:
            raise AfmmError(AFMM_ERR_DEVICE, f"hipMemcpy H2D failed: {rc}")

    def copy_from(self, other):
        """Device-to-device copy of another DeviceArray of the same size."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        rc = self._runtime().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), self.nbytes, 3)
        if rc:
            raise AfmmError(AFMM_ERR_DEVICE, f"hipMemcpy D2D failed: {rc}")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        rc = self._runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2)
        if rc:
            raise AfmmError(AFMM_ERR_DEVICE, f"hipMemcpy D2H failed: {rc}")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            self._runtime().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
