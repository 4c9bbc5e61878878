"""CPU checks of the HIP library's arithmetic and boundary, no GPU needed:

  * the host model (tests/model/model.cpp: the oracle's tree + the product's
    scaled real-basis operators, exact centres and translation tables)
    reproduces the reference's AdaptiveFMM3D accelerations to <= 1e-12;
  * libafmm.so loads and exports every function include/afmm.h declares.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import cabi

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FORCE_CASES = sorted(os.path.basename(f)[6:-4] for f in glob.glob(os.path.join(GOLD, "force_*.npz")))


@pytest.mark.parametrize("case", FORCE_CASES)
def test_model_matches_reference_fixture(case):
    g = np.load(os.path.join(GOLD, f"force_{case}.npz"))
    for p in g["ps"]:
        acc = cabi.model_fmm(g["pos"], g["mass"], 1.0, float(g["eps"]), int(p), int(g["cap"]))
        ref = g[f"acc_p{p}"]
        if len(ref) == 1:
            assert np.array_equal(acc, ref)
            continue
        mx, _ = cabi.rel_err(ref, acc)
        assert mx <= 1e-12, (case, p, mx)


def test_model_all_orders_against_oracle():
    orc = cabi.oracle()
    pos, _, m = orc.plummer(1200, seed=23)
    for p in range(1, 13):
        ref, _ = orc.fmm(pos, m, 1.0, 0.003, p, 20)
        acc = cabi.model_fmm(pos, m, 1.0, 0.003, p, 20)
        mx, _ = cabi.rel_err(ref, acc)
        assert mx <= 1e-12, (p, mx)


def test_model_orders_above_12_against_oracle():
    """Orders 13-16 (S up to 289, three coefficient windows in P2M/P2L) on
    a smaller sample: the oracle's O(p^4) M2L is slow on the CPU."""
    orc = cabi.oracle()
    pos, _, m = orc.plummer(600, seed=24)
    for p in (13, 16):
        ref, _ = orc.fmm(pos, m, 1.0, 0.003, p, 24)
        acc = cabi.model_fmm(pos, m, 1.0, 0.003, p, 24)
        mx, _ = cabi.rel_err(ref, acc)
        assert mx <= 1e-12, (p, mx)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(cabi.ROOT, "include", "afmm.h")).read()
    names = set(re.findall(r"^\s*(?:int|int64_t|void|const char\*)\s+(afmm_\w+)\s*\(", hdr, re.M))
    assert {"afmm_create", "afmm_eval", "afmm_eval_device", "afmm_tree_export"} <= names
    lib = C.CDLL(os.path.join(cabi.ROOT, "on-galaxy_amd", "libafmm.so"))
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_parameter_validation_without_device():
    """Argument errors are reported before any device work."""
    mod = cabi.product()
    for kw, code in (({"p": 0}, -2), ({"p": 17}, -2), ({"threshold": 0}, -3), ({"eps": -1.0}, -1),
                     ({"precision": 7}, -9)):
        with pytest.raises(mod.AfmmError) as e:
            mod.AdaptiveFMM3D(**kw)
        assert e.value.code == code


def test_context_switches_reject_null_without_device():
    """The mode switches check their context before touching the device."""
    lib = C.CDLL(os.path.join(cabi.ROOT, "on-galaxy_amd", "libafmm.so"))
    for name in ("afmm_set_overlap", "afmm_set_graph"):
        f = getattr(lib, name)
        f.argtypes = [C.c_void_p, C.c_int]
        f.restype = C.c_int
        assert f(None, 1) == -1 and f(None, 0) == -1, name
    for name in ("afmm_set_shapes", "afmm_get_shapes"):
        f = getattr(lib, name)
        f.argtypes = [C.c_void_p, C.c_void_p]
        f.restype = C.c_int
        assert f(None, None) == -1, name


def test_system_offsets_rejected_before_device_work():
    """The binding's batch and leapfrog wrappers check sys_off against the
    arrays it indexes (the C side copies 3 * sys_off[-1] doubles from and into
    them), so a short, non-increasing or shifted offset array raises
    ValueError instead of overrunning a numpy buffer."""
    mod = cabi.product()
    pos = np.zeros((10, 3))
    mass = np.ones(10)
    good = mod._offsets([0, 4, 10], 10)
    assert good.dtype == np.int64 and list(good) == [0, 4, 10]
    for bad in ([0, 4, 12], [0, 4, 9], [1, 4, 10], [0, 4, 4, 10], [0, 6, 4, 10], [0], []):
        with pytest.raises(ValueError):
            mod._offsets(bad, 10)
    solver = mod.AdaptiveFMM3D.__new__(mod.AdaptiveFMM3D)  # no context: nothing may reach the device
    solver.lib = None
    solver.ctx = None
    for bad in ([0, 4, 12], [0, 11], [2, 10]):
        with pytest.raises(ValueError):
            solver.computeBatch(pos, mass, bad)
        with pytest.raises(ValueError):
            mod.LeapfrogState(solver, pos, pos, mass, sys_off=bad)
    with pytest.raises(ValueError):
        solver.computeBatch(pos, mass[:9], [0, 9])
