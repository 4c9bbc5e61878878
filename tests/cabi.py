"""ctypes plumbing shared by the tests: the afmm_tree export struct and thin
wrappers over the oracle restatement (liboracle.so, prefix orc_) and the
compiled reference (oracle/_ref/libref.so, prefix ref_). Test infrastructure
only; the product's own binding lives in on-galaxy_amd/afmm.py."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)


class AfmmTree(C.Structure):
    """Mirror of afmm_tree in include/afmm.h."""
    _fields_ = [
        ("parent", _ip), ("child", _ip), ("level", _ip), ("center", _dp),
        ("half", _dp), ("leaf", _bp), ("pfirst", _ip), ("pcount", _ip),
        ("px", _dp), ("py", _dp), ("pz", _dp), ("mass", _dp),
        ("orig_index", _ip), ("leaves", _ip),
        ("colleague_off", _ip), ("colleague_idx", _ip),
        ("u_off", _ip), ("u_idx", _ip), ("v_off", _ip), ("v_idx", _ip),
        ("w_off", _ip), ("w_idx", _ip), ("x_off", _ip), ("x_idx", _ip),
        ("tol", C.c_double),
    ]


LISTS = ("colleague", "u", "v", "w", "x")


def ptr(a):
    if a is None:
        return None
    t = {np.float64: _dp, np.int32: _ip, np.int8: _bp}[a.dtype.type]
    return a.ctypes.data_as(t)


def alloc_tree(sizes):
    """Allocate numpy buffers for an export; sizes = (n_bodies, n_boxes,
    n_leaves, n_colleague, n_u, n_v, n_w, n_x)."""
    n, nb, nl, nc, nu, nv, nw, nx = (int(s) for s in sizes)
    a = {
        "parent": np.empty(nb, np.int32), "child": np.empty(8 * nb, np.int32),
        "level": np.empty(nb, np.int32), "center": np.empty(3 * nb, np.float64),
        "half": np.empty(nb, np.float64), "leaf": np.empty(nb, np.int8),
        "pfirst": np.empty(nb, np.int32), "pcount": np.empty(nb, np.int32),
        "px": np.empty(n, np.float64), "py": np.empty(n, np.float64),
        "pz": np.empty(n, np.float64), "mass": np.empty(n, np.float64),
        "orig_index": np.empty(n, np.int32), "leaves": np.empty(nl, np.int32),
    }
    for name, cnt in zip(LISTS, (nc, nu, nv, nw, nx)):
        a[name + "_off"] = np.empty(nb + 1, np.int32)
        a[name + "_idx"] = np.empty(max(cnt, 1), np.int32)[:cnt]
    t = AfmmTree()
    for k, v in a.items():
        setattr(t, k, ptr(v) if v.size else ptr(np.empty(1, v.dtype)))
    return t, a


class CpuFmm:
    """One of the two CPU implementations (prefix 'orc' or 'ref')."""

    def __init__(self, path, prefix):
        self.lib = C.CDLL(path)
        self.prefix = prefix
        L = self.lib
        f = self._f
        f("fmm").argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp, _ip]
        f("direct").argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, _dp]
        f("tree_new").argtypes = [C.c_int64, _dp, _dp, C.c_int, C.c_double]
        f("tree_new").restype = C.c_void_p
        f("tree_free").argtypes = [C.c_void_p]
        f("tree_sizes").argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        f("tree_export").argtypes = [C.c_void_p, C.POINTER(AfmmTree)]
        f("harmonics").argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        f("harmonics").restype = C.c_double
        f("acoef").argtypes = [C.c_int, C.c_int]
        f("acoef").restype = C.c_double
        f("p2m").argtypes = [_dp, C.c_int, _dp, _dp, C.c_int, _dp]
        for op in ("m2m", "m2l", "l2l"):
            f(op).argtypes = [_dp, C.c_int, _dp, _dp, _dp]
        for op in ("local_eval", "multipole_eval"):
            f(op).argtypes = [_dp, C.c_int, _dp, _dp, _dp, _dp]
        f("plummer").argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint, _dp, _dp, _dp]
        f("diagnostics").argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, _dp]
        f("leapfrog").argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                  C.c_double, C.c_int, _dp]
        del L

    def _f(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    def set_threads(self, n):
        """Threads for the oracle's FMM passes (OpenMP over independent boxes,
        the same bits at any count); 0 = the starting count (OMP_NUM_THREADS,
        else all host cores). The compiled
        reference has no threads: returns 1 there."""
        fn = getattr(self.lib, f"{self.prefix}_set_threads", None)
        if fn is None:
            return 1
        fn.argtypes = [C.c_int]
        return int(fn(int(n)))

    # ---- solvers ---------------------------------------------------------
    def fmm(self, pos, mass, G=1.0, eps=0.0, p=8, cap=64):
        pos = np.ascontiguousarray(pos, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        acc = np.empty_like(pos)
        st = np.zeros(4, np.int32)
        rc = self._f("fmm")(len(mass), ptr(pos), ptr(mass), G, eps, p, cap, ptr(acc), ptr(st))
        assert rc == 0
        return acc, st

    def direct(self, pos, mass, G=1.0, eps=0.0):
        pos = np.ascontiguousarray(pos, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        acc = np.empty_like(pos)
        assert self._f("direct")(len(mass), ptr(pos), ptr(mass), G, eps, ptr(acc)) == 0
        return acc

    def diagnostics(self, pos, vel, mass, G=1.0, eps=0.0):
        """computeDiagnostics (Diagnostics3D.hpp:20-38): dict of kinetic,
        potential, total, momentum (3), angular (3)."""
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        out = np.empty(9)
        assert self._f("diagnostics")(len(mass), ptr(pos), ptr(vel), ptr(mass), G, eps, ptr(out)) == 0
        return {"kinetic": out[0], "potential": out[1], "total": out[2], "momentum": out[3:6],
                "angular": out[6:9]}

    def leapfrog(self, pos, vel, mass, G, eps, p, cap, dt, nsteps):
        """VelocityVerlet3D over the CPU FMM: (pos, vel, acc) after nsteps."""
        pos = np.array(pos, np.float64, order="C")
        vel = np.array(vel, np.float64, order="C")
        mass = np.ascontiguousarray(mass, np.float64)
        acc = np.empty_like(pos)
        assert self._f("leapfrog")(len(mass), ptr(pos), ptr(vel), ptr(mass), G, eps, p, cap, dt, nsteps,
                                   ptr(acc)) == 0
        return pos, vel, acc

    def tree(self, pos, mass, cap=64, eps_floor=0.0):
        pos = np.ascontiguousarray(pos, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        h = self._f("tree_new")(len(mass), ptr(pos), ptr(mass), cap, eps_floor)
        try:
            sizes = (C.c_int64 * 8)()
            self._f("tree_sizes")(h, sizes)
            t, arrays = alloc_tree(list(sizes))
            self._f("tree_export")(h, C.byref(t))
            arrays["tol"] = t.tol
            return arrays
        finally:
            self._f("tree_free")(h)

    # ---- operators -------------------------------------------------------
    def harmonics(self, v, L):
        out = np.empty(2 * (L + 1) ** 2)
        r = self._f("harmonics")(float(v[0]), float(v[1]), float(v[2]), L, ptr(out))
        return r, out[0::2] + 1j * out[1::2]

    def acoef(self, n, m):
        return self._f("acoef")(n, m)

    @staticmethod
    def _pack(c):
        c = np.asarray(c, np.complex128)
        out = np.empty(2 * c.size)
        out[0::2], out[1::2] = c.real, c.imag
        return out

    def p2m(self, c, spos, q, p):
        c = np.ascontiguousarray(c, np.float64)
        spos = np.ascontiguousarray(spos, np.float64)
        q = np.ascontiguousarray(q, np.float64)
        out = np.empty(2 * (p + 1) ** 2)
        self._f("p2m")(ptr(c), len(q), ptr(spos), ptr(q), p, ptr(out))
        return out[0::2] + 1j * out[1::2]

    def translate(self, op, coeffs, p, a, b):
        inp = self._pack(coeffs)
        out = np.empty_like(inp)
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        self._f(op)(ptr(inp), p, ptr(a), ptr(b), ptr(out))
        return out[0::2] + 1j * out[1::2]

    def evaluate(self, kind, coeffs, p, c, t):
        inp = self._pack(coeffs)
        pot = np.empty(2)
        grad = np.empty(3)
        self._f(kind + "_eval")(ptr(inp), p, ptr(np.ascontiguousarray(c, np.float64)),
                                ptr(np.ascontiguousarray(t, np.float64)), ptr(pot), ptr(grad))
        return pot[0] + 1j * pot[1], grad

    def plummer(self, n, a=1.0, M=1.0, G=1.0, seed=7):
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        mass = np.empty(n)
        self._f("plummer")(n, a, M, G, seed, ptr(pos), ptr(vel), ptr(mass))
        return pos, vel, mass


_cache = {}


def oracle():
    if "orc" not in _cache:
        _cache["orc"] = CpuFmm(ORACLE_SO, "orc")
    return _cache["orc"]


def reference():
    """The compiled reference, or None where oracle/_ref was not built."""
    if not os.path.exists(REF_SO):
        return None
    if "ref" not in _cache:
        _cache["ref"] = CpuFmm(REF_SO, "ref")
    return _cache["ref"]


def direct_targets(pos, mass, idx, G=1.0, eps=0.0):
    """FP64 direct sum on a subset of targets (oracle restatement)."""
    lib = oracle().lib
    lib.orc_direct_targets.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_int64,
                                       C.POINTER(C.c_int64), _dp]
    pos = np.ascontiguousarray(pos, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    idx = np.ascontiguousarray(idx, np.int64)
    acc = np.empty((len(idx), 3))
    lib.orc_direct_targets(len(mass), ptr(pos), ptr(mass), G, eps, len(idx),
                           idx.ctypes.data_as(C.POINTER(C.c_int64)), ptr(acc))
    return acc


def operator_convergence(dir_first):
    lib = oracle().lib
    lib.orc_operator_convergence.argtypes = [C.c_int, _dp]
    out = np.empty(40)
    lib.orc_operator_convergence(int(dir_first), ptr(out))
    return out.reshape(4, 5, 2)


def direct_potential_targets(pos, mass, idx, G=1.0, eps=0.0):
    """Oracle softened potential (Kernel3D.hpp:39-42) at bodies idx."""
    lib = oracle().lib
    lib.orc_direct_potential_targets.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_int64,
                                                 C.POINTER(C.c_int64), _dp]
    pos = np.ascontiguousarray(pos, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    idx = np.ascontiguousarray(idx, np.int64)
    phi = np.empty(len(idx))
    lib.orc_direct_potential_targets(len(mass), ptr(pos), ptr(mass), G, eps, len(idx),
                                     idx.ctypes.data_as(C.POINTER(C.c_int64)), ptr(phi))
    return phi


def rel_err(ref, test):
    """relErr of adaptive_tests.cpp:40-50 / adaptive_gpu_harness.cu:36-46:
    max_i |d_i| / max(|a_i|, 1e-30 + 1e-6 max|a|) and sqrt(sum|d|^2/sum|a|^2)."""
    ref = np.asarray(ref, np.float64).reshape(-1, 3)
    test = np.asarray(test, np.float64).reshape(-1, 3)
    an = np.linalg.norm(ref, axis=1)
    floor = 1e-30 + 1e-6 * an.max()
    d = np.linalg.norm(test - ref, axis=1)
    return float((d / np.maximum(an, floor)).max()), float(np.sqrt((d ** 2).sum() / (an ** 2).sum()))


def uniform_cube(n, seed):
    """Deterministic cube [-1,1]^3 with masses in [0.5,1.5] (numpy stream)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, 3)), rng.uniform(0.5, 1.5, n)


def product():
    """The product binding (on-galaxy_amd/afmm.py) loaded by path."""
    import importlib.util
    if "product" not in _cache:
        spec = importlib.util.spec_from_file_location(
            "afmm_product", os.path.join(ROOT, "on-galaxy_amd", "afmm.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _cache["product"] = mod
    return _cache["product"]


def model_fmm(pos, mass, G, eps, p, cap):
    """Host model of the HIP arithmetic (tests/model/libmodel.so)."""
    if "model" not in _cache:
        lib = C.CDLL(os.path.join(ROOT, "tests", "model", "libmodel.so"))
        lib.model_fmm.argtypes = [C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        _cache["model"] = lib
    pos = np.ascontiguousarray(pos, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    acc = np.empty_like(pos)
    assert _cache["model"].model_fmm(len(mass), ptr(pos), ptr(mass), G, eps, p, cap, ptr(acc)) == 0
    return acc
