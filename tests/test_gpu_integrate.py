"""GPU checks of the device-resident velocity Verlet (afmm_state_*,
VelocityVerlet3D of Integrator3D.hpp:19-40) and its diagnostics
(Diagnostics3D.hpp:20-38).

  * bit-identical to the same update formulas applied on the host to the
    library's own accelerations (the integrator adds no rounding of its own);
  * within 1e-11 of the oracle's leapfrog over the CPU adaptive FMM after
    three steps (the forces agree to 1e-12);
  * several systems integrated together equal each one integrated alone;
  * kinetic energy, momentum and angular momentum equal the oracle's to
    1e-12 and the FMM potential energy is within the FMM error of the
    reference's O(N^2) pair sum.
"""
import numpy as np
import pytest

import cabi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afmm():
    return cabi.product()


def test_state_steps_equal_host_update_of_gpu_forces(afmm):
    orc = cabi.oracle()
    pos, vel, m = orc.plummer(8000, seed=12)
    dt, G, eps = 0.01, 1.0, 0.01
    print("leapfrog 8000 p=6 vs host update", flush=True)
    s = afmm.AdaptiveFMM3D(G=G, eps=eps, p=6, threshold=48)
    st = afmm.LeapfrogState(s, pos, vel, m)
    st.step(dt, 3)
    gp, gv, ga = st.download()
    x, v = pos.copy(), vel.copy()
    a = s.computeAccelerations(x, m)
    for _ in range(3):
        x = x + (v * dt + a * (0.5 * dt * dt))
        an = s.computeAccelerations(x, m)
        v = v + (a + an) * (0.5 * dt)
        a = an
    assert np.array_equal(gp, x) and np.array_equal(gv, v) and np.array_equal(ga, a)


def test_state_matches_oracle_leapfrog(afmm):
    orc = cabi.oracle()
    pos, vel, m = orc.plummer(5000, seed=13)
    print("leapfrog 5000 p=8 vs oracle", flush=True)
    op, ov, oa = orc.leapfrog(pos, vel, m, 1.0, 0.01, 8, 64, 0.01, 3)
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.01, p=8, threshold=64)
    st = afmm.LeapfrogState(s, pos, vel, m)
    st.step(0.01, 3)
    gp, gv, ga = st.download()
    assert np.abs(gp - op).max() <= 1e-11 * np.abs(op).max()
    assert np.abs(gv - ov).max() <= 1e-11 * np.abs(ov).max()
    mx, _ = cabi.rel_err(oa, ga)
    assert mx <= 1e-11


def test_multi_system_state_equals_single(afmm):
    orc = cabi.oracle()
    parts = [orc.plummer(n, seed=40 + n) for n in (3000, 700, 64, 5000)]
    off = np.concatenate([[0], np.cumsum([len(p[2]) for p in parts])])
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    m = np.concatenate([p[2] for p in parts])
    eps = np.array([0.01, 0.02, 0.0, 0.005])
    print("multi-system leapfrog", flush=True)
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.01, p=5, threshold=32, precision=afmm.PRECISION_MIXED)
    st = afmm.LeapfrogState(s, pos, vel, m, sys_off=off, eps=eps)
    st.step(0.005, 4)
    bp, bv, ba = st.download()
    for q, (p0, v0, m0) in enumerate(parts):
        s1 = afmm.AdaptiveFMM3D(G=1.0, eps=eps[q], p=5, threshold=32, precision=afmm.PRECISION_MIXED)
        st1 = afmm.LeapfrogState(s1, p0, v0, m0)
        st1.step(0.005, 4)
        p1, v1, a1 = st1.download()
        sl = slice(off[q], off[q + 1])
        assert np.array_equal(bp[sl], p1) and np.array_equal(bv[sl], v1) and np.array_equal(ba[sl], a1), q


def test_diagnostics(afmm):
    """eps = 0, where the FMM potential differs from the reference's pair sum
    by truncation only (the far field is never softened)."""
    orc = cabi.oracle()
    pos, vel, m = orc.plummer(12000, seed=14)
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.0, p=8, threshold=64)
    st = afmm.LeapfrogState(s, pos, vel, m)
    st.step(0.001, 2)
    gp, gv, _ = st.download()
    d = st.diagnostics()[0]
    ref = orc.diagnostics(gp, gv, m, 1.0, 0.0)
    print(f"diag: KE {d['kinetic']:.12e} vs {ref['kinetic']:.12e}; "
          f"PE {d['potential']:.9e} vs {ref['potential']:.9e}", flush=True)
    assert d["kinetic"] == pytest.approx(ref["kinetic"], rel=1e-12)
    assert np.allclose(d["momentum"], ref["momentum"], rtol=0, atol=1e-12)
    assert np.allclose(d["angular"], ref["angular"], rtol=0, atol=1e-12)
    assert d["potential"] == pytest.approx(ref["potential"], rel=2e-6)
    assert d["total"] == d["kinetic"] + d["potential"]
    # the diagnostics evaluation leaves the state unchanged
    p2, v2, _ = st.download()
    assert np.array_equal(p2, gp) and np.array_equal(v2, gv)
