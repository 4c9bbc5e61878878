"""afmm_direct_device_noalloc: the allocation-free direct sum (partial sums in
a static device array) against afmm_direct, the entry point it varies, through
the C ABI. Same bodies and targets give the same bits; two calls in a row, the
second smaller, so that the static scratch's leftover contents from the first
call cannot leak into the second; the size limit and out-of-range indices are
refused or flagged without touching memory outside the arrays."""
import ctypes as C

import numpy as np
import pytest

import cabi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afmm():
    return cabi.product()


def _noalloc(s, afmm, pos, mass, tgt, potential=True):
    d_pos, d_mass = afmm.DeviceArray(None, pos), afmm.DeviceArray(None, mass)
    d_tgt = afmm.DeviceArray(None, np.ascontiguousarray(tgt, np.int64))
    d_acc = afmm.DeviceArray((len(tgt), 3))
    d_phi = afmm.DeviceArray((len(tgt),)) if potential else None
    rc = s.lib.afmm_direct_device_noalloc(s.ctx, len(mass), d_pos.ptr, d_mass.ptr, len(tgt), d_tgt.ptr, d_acc.ptr,
                                          d_phi.ptr if d_phi else None, None)
    assert rc == 0, s.lib.afmm_last_error(s.ctx)
    assert s.lib.afmm_synchronize(s.ctx) == 0
    out = d_acc.numpy(), (d_phi.numpy() if d_phi else None)
    for a in (d_pos, d_mass, d_tgt, d_acc, d_phi):
        if a is not None:
            a.close()
    return out


@pytest.mark.parametrize("G,eps", [(1.0, 0.01), (1.7, 0.0)])
def test_noalloc_equals_direct_twice(afmm, G, eps):
    s = afmm.AdaptiveFMM3D(G=G, eps=eps, p=4)
    rng = np.random.default_rng(3)
    # first call: 20000 bodies, 5000 targets; second: 3000 bodies, 700 targets
    for n, nt in ((20000, 5000), (3000, 700)):
        print(f"noalloc direct n={n} ntgt={nt} G={G} eps={eps}", flush=True)
        pos = rng.standard_normal((n, 3))
        mass = rng.uniform(0.5, 1.5, n) / n
        tgt = rng.choice(n, nt, replace=False).astype(np.int64)
        ref_acc, ref_phi = s.direct(pos, mass, tgt, potential=True)
        acc, phi = _noalloc(s, afmm, pos, mass, tgt)
        assert np.array_equal(acc, ref_acc)
        assert np.array_equal(phi, ref_phi)
        acc2, _ = _noalloc(s, afmm, pos, mass, tgt, potential=False)
        assert np.array_equal(acc2, ref_acc)
    s.close()


def test_noalloc_limits(afmm):
    s = afmm.AdaptiveFMM3D(p=4)
    pos = np.random.default_rng(4).standard_normal((100, 3))
    mass = np.full(100, 0.01)
    # more targets than the static scratch holds: refused before any launch
    rc = s.lib.afmm_direct_device_noalloc(s.ctx, 100, C.c_void_p(8), C.c_void_p(8), 65537, C.c_void_p(8),
                                          C.c_void_p(8), None, None)
    assert rc == -8
    # an index outside [0, n) reads nothing and gives NaN for that target only
    tgt = np.array([3, 100, -1, 7], np.int64)
    acc, phi = _noalloc(s, afmm, pos, mass, tgt)
    assert np.isnan(acc[1]).all() and np.isnan(acc[2]).all() and np.isnan(phi[1:3]).all()
    ref = s.direct(pos, mass, np.array([3, 7]))
    assert np.array_equal(acc[[0, 3]], ref)
    s.close()
