"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the
checker of the HIP library. Each test names the reference behaviour it pins
(file:line in the reference tree) and is chosen so that a plausible mistake --
a dropped term, a wrong sign or index, a transposed operand -- fails it."""
import csv
import glob
import os

import numpy as np
import pytest

import cabi

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TREE_CASES = sorted(os.path.basename(f)[5:-4] for f in glob.glob(os.path.join(GOLD, "tree_*.npz")))
FORCE_CASES = sorted(os.path.basename(f)[6:-4] for f in glob.glob(os.path.join(GOLD, "force_*.npz")))
TREE_KEYS = ("parent", "child", "level", "center", "half", "leaf", "pfirst", "pcount", "px", "py",
             "pz", "mass", "orig_index", "leaves") + tuple(
                 f"{k}_{s}" for k in cabi.LISTS for s in ("off", "idx"))


@pytest.fixture(scope="module")
def orc():
    return cabi.oracle()


# ---- golden fixtures written by the compiled reference ----------------------
@pytest.mark.parametrize("case", TREE_CASES)
def test_tree_matches_reference_fixture(orc, case):
    """AdaptiveOctree::build + buildLists (AdaptiveOctree.hpp:65-189): every box
    field, centre bit pattern, permutation and list order must be identical."""
    g = np.load(os.path.join(GOLD, f"tree_{case}.npz"))
    t = orc.tree(g["in_pos"], g["in_mass"], int(g["cap"]), float(g["eps_floor"]))
    assert t["tol"] == float(g["tol"])
    for k in TREE_KEYS:
        assert np.array_equal(t[k], g[k]), k


@pytest.mark.parametrize("case", FORCE_CASES)
def test_forces_match_reference_fixture(orc, case):
    """AdaptiveFMM3D::computeAccelerations (AdaptiveFMM3D.hpp:29-110) and
    DirectSolver3D (DirectSolver3D.hpp:15-34): same expression order and
    flags, so the accelerations are bit-identical."""
    g = np.load(os.path.join(GOLD, f"force_{case}.npz"))
    eps, cap = float(g["eps"]), int(g["cap"])
    assert np.array_equal(orc.direct(g["pos"], g["mass"], 1.0, eps), g["direct"])
    # One thread, then several: the OpenMP passes split independent boxes
    # only, each box summing in the reference's order, so the bits must not
    # depend on the thread count.
    try:
        for nt in (1, 4):
            assert orc.set_threads(nt) == nt
            for p in g["ps"]:
                acc, _ = orc.fmm(g["pos"], g["mass"], 1.0, eps, int(p), cap)
                assert np.array_equal(acc, g[f"acc_p{p}"]), (nt, p)
    finally:
        orc.set_threads(0)


@pytest.mark.parametrize("case", FORCE_CASES)
def test_direct_targets_match_reference_fixture(case):
    """orc_direct_targets, the checker of the GPU yardstick afmm_direct, over
    every target against DirectSolver3D's accelerations written by the compiled
    reference (DirectSolver3D.hpp:15-34). The reference sums pairs i < j once
    and the target sum runs j in order, so the bits differ by summation order
    only: max relative difference <= 1e-13 (measured <= 2e-14)."""
    g = np.load(os.path.join(GOLD, f"force_{case}.npz"))
    n = len(g["mass"])
    acc = cabi.direct_targets(g["pos"], g["mass"], np.arange(n), 1.0, float(g["eps"]))
    ref = g["direct"]
    mx, _ = cabi.rel_err(ref, acc)
    assert mx <= 1e-13, mx
    # a subset of targets in a scrambled order returns the same rows
    sel = np.random.default_rng(5).permutation(n)[: max(1, n // 3)]
    assert np.array_equal(cabi.direct_targets(g["pos"], g["mass"], sel, 1.0, float(g["eps"])), acc[sel])


def test_direct_targets_closed_forms():
    """The softened pair law with G != 1 and eps > 0 (Kernel3D.hpp:31-35,
    DirectSolver3D.hpp:25-30): two bodies, a_1 = G m_2 d / (|d|^2 + eps^2)^{3/2}
    with d = x_2 - x_1 and the opposite sign on body 2; three bodies are the sum
    of their pair terms, and the self pair is skipped. A wrong G or eps
    placement, a sign flip or a self term fails one of these."""
    G, eps = 1.7, 0.02
    pos = np.array([[0.1, -0.2, 0.3], [0.4, 0.25, -0.9], [-1.1, 0.6, 0.05]])
    m = np.array([0.7, 1.9, 0.4])

    def pair(i, j):
        d = pos[j] - pos[i]
        return G * m[j] * d / (d @ d + eps * eps) ** 1.5

    two = cabi.direct_targets(pos[:2], m[:2], [0, 1], G, eps)
    assert np.allclose(two[0], pair(0, 1), rtol=1e-15, atol=0)
    assert np.allclose(two[1], pair(1, 0), rtol=1e-15, atol=0)
    three = cabi.direct_targets(pos, m, [2, 0, 1], G, eps)
    for row, i in zip(three, (2, 0, 1)):
        want = sum(pair(i, j) for j in range(3) if j != i)
        assert np.allclose(row, want, rtol=1e-14, atol=0)
    # G enters as a factor, eps inside the distance: doubling G doubles the
    # result bit for bit; eps -> 0 gives the plain inverse square
    assert np.array_equal(cabi.direct_targets(pos, m, [1], 2 * G, eps), 2 * cabi.direct_targets(pos, m, [1], G, eps))
    d = pos[1] - pos[0]
    plain = cabi.direct_targets(pos[:2], m[:2], [0], 1.0, 0.0)[0]
    assert np.allclose(plain, m[1] * d / np.linalg.norm(d) ** 3, rtol=1e-15, atol=0)
    ref = cabi.reference()
    if ref is not None:  # the compiled DirectSolver3D at the same G, eps
        pos4, m4 = cabi.uniform_cube(300, 8)
        mx, _ = cabi.rel_err(ref.direct(pos4, m4, G, eps), cabi.direct_targets(pos4, m4, np.arange(300), G, eps))
        assert mx <= 1e-13, mx


@pytest.mark.parametrize("p", [4, 8, 12])
def test_operators_match_reference_fixture(orc, p):
    """P2M/M2M/M2L/L2L/localField/multipoleField/computeY
    (FmmExpansions3D.hpp:75-330) on fixed inputs."""
    g = np.load(os.path.join(GOLD, "ops.npz"))
    k = lambda s: g[f"p{p}_{s}"]  # noqa: E731
    M = orc.p2m(k("c0"), k("src"), k("q"), p)
    assert np.array_equal(M, k("p2m"))
    assert np.array_equal(orc.translate("m2m", M, p, k("c0"), k("c1")), k("m2m"))
    L = orc.translate("m2l", M, p, k("c0"), k("c2"))
    assert np.array_equal(L, k("m2l"))
    assert np.array_equal(orc.translate("l2l", L, p, k("c2"), k("c3")), k("l2l"))
    pot, grad = orc.evaluate("local", L, p, k("c2"), k("tnear"))
    assert pot == k("local_pot")[0] and np.array_equal(grad, k("local_grad"))
    pot, grad = orc.evaluate("multipole", M, p, k("c0"), k("tfar"))
    assert pot == k("mult_pot")[0] and np.array_equal(grad, k("mult_grad"))
    _, Y = orc.harmonics(k("hv"), 2 * p)
    assert np.array_equal(Y, k("Y"))


def test_plummer_stream_matches_reference(orc):
    """makePlummerSphere (InitialConditions3D.hpp:100-152) with libstdc++'s
    mt19937 stream; pinned by a digest of the C0 input (N=16384, seed 42)."""
    import hashlib
    pos, vel, m = orc.plummer(16384, seed=42)
    h = hashlib.sha256(pos.tobytes() + vel.tobytes() + m.tobytes()).hexdigest()
    ref = cabi.reference()
    if ref is not None:
        rp, rv, rm = ref.plummer(16384, seed=42)
        assert hashlib.sha256(rp.tobytes() + rv.tobytes() + rm.tobytes()).hexdigest() == h
    # digest written from oracle/_ref/libref.so in the build container
    assert h == "6bb78e6d31f40b2c491a99e471c02b9a0c9f600b93c35e7859564854653dc5d5", h


# ---- known-answer tests held by the reference's own tests -------------------
def test_harmonics_closed_forms(orc):
    """fmm3d_tests.cpp:134-157: Y_n^m for n<=2 at theta=1, phi=0.7 against the
    closed forms, |err| < 1e-13, including Y_n^{-m} = conj(Y_n^m)."""
    th, ph = 1.0, 0.7
    v = [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]
    _, Y = orc.harmonics(v, 2)
    c, s = np.cos(th), np.sin(th)
    e = lambda m: np.exp(1j * m * ph)  # noqa: E731
    ref = {(0, 0): 1.0, (1, 0): c, (1, 1): np.sqrt(0.5) * s * e(1), (2, 0): (3 * c * c - 1) / 2,
           (2, 1): np.sqrt(1 / 6) * 3 * c * s * e(1), (2, 2): np.sqrt(1 / 24) * 3 * s * s * e(2)}
    for (n, m), val in ref.items():
        assert abs(Y[n * n + n + m] - val) < 1e-13
        assert abs(Y[n * n + n - m] - np.conj(val)) < 1e-13


def test_gradients_match_finite_differences(orc):
    """fmm3d_tests.cpp:159-188: localField / multipoleField against central
    differences (h=1e-5) of the potentials, random p=6 expansions."""
    rng = np.random.default_rng(123)
    p = 6
    S = (p + 1) ** 2
    Mc = rng.normal(size=S) + 1j * rng.normal(size=S)
    Lc = rng.normal(size=S) + 1j * rng.normal(size=S)
    o = np.zeros(3)
    h = 1e-5
    worst = 0.0
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        for kind, coeffs, t in (("multipole", Mc, 5.0 * d), ("local", Lc, 0.4 * d)):
            _, g = orc.evaluate(kind, coeffs, p, o, t)
            fd = np.empty(3)
            for a in range(3):
                dt = np.zeros(3)
                dt[a] = h
                fd[a] = (orc.evaluate(kind, coeffs, p, o, t + dt)[0].real -
                         orc.evaluate(kind, coeffs, p, o, t - dt)[0].real) / (2 * h)
            worst = max(worst, np.linalg.norm(g - fd) / np.linalg.norm(g))
    assert worst < 1e-6


def test_operator_convergence_table_matches_reference_outputs():
    """The reference's published operator error table
    (outputs/fmm3d_operator_convergence.csv, written by fmm3d_tests.cpp:202-293)
    is reproduced to its printed 6 digits: P2M, M2L (worst of 4 geometries),
    M2M, L2L at p = 2..10, potential and acceleration."""
    ops = ["P2M", "M2L", "M2M", "L2L"]
    ps = [2, 4, 6, 8, 10]
    gold = np.zeros((4, 5, 2))
    with open(os.path.join(GOLD, "fmm3d_operator_convergence.csv")) as fh:
        for r in csv.DictReader(fh):
            gold[ops.index(r["operator"]), ps.index(int(r["p"])),
                 0 if r["quantity"] == "potential" else 1] = float(r["rel_err"])
    ours = cabi.operator_convergence(dir_first=True)
    assert np.abs(ours / gold - 1).max() < 2e-5
    # errors fall strictly with p for every operator (fmm3d_tests.cpp:110-127)
    assert (np.diff(ours, axis=1) < 0).all()


# ---- mathematical identities ---------------------------------------------------
def test_m2m_of_p2m_equals_p2m_at_parent(orc):
    """M2M is exact for n <= p (FmmExpansions3D.hpp:237-241): shifting a child
    multipole equals P2M about the parent. A wrong sign in J or a swapped
    A-coefficient breaks this at roundoff level."""
    rng = np.random.default_rng(5)
    src = rng.uniform(-0.3, 0.3, (25, 3)) + np.array([0.25, -0.25, 0.25])
    q = rng.uniform(0.5, 1.5, 25)
    for p in (3, 8):
        child = np.array([0.25, -0.25, 0.25])
        parent = np.zeros(3)
        shifted = orc.translate("m2m", orc.p2m(child, src, q, p), p, child, parent)
        direct = orc.p2m(parent, src, q, p)
        assert np.abs(shifted - direct).max() < 1e-13 * np.abs(direct).max()


def test_l2l_is_exact_for_polynomial_locals(orc):
    """L2L re-expands a degree-p polynomial exactly (FmmExpansions3D.hpp:301-306):
    potential and gradient at a target agree before and after the shift."""
    rng = np.random.default_rng(8)
    p = 7
    S = (p + 1) ** 2
    # a real field: L_n^{-m} = conj(L_n^m)
    L = np.zeros(S, complex)
    for n in range(p + 1):
        L[n * n + n] = rng.normal()
        for m in range(1, n + 1):
            z = rng.normal() + 1j * rng.normal()
            L[n * n + n + m], L[n * n + n - m] = z, np.conj(z)
    c0, c1 = np.zeros(3), np.array([0.3, -0.2, 0.1])
    L1 = orc.translate("l2l", L, p, c0, c1)
    t = np.array([0.35, -0.1, 0.2])
    pa, ga = orc.evaluate("local", L, p, c0, t)
    pb, gb = orc.evaluate("local", L1, p, c1, t)
    assert abs(pa - pb) < 1e-12 * abs(pa)
    assert np.linalg.norm(ga - gb) < 1e-12 * np.linalg.norm(ga)


def test_m2l_monopole_reduces_to_potential(orc):
    """L_0^0 = Phi(z_L) for a unit point mass (FmmExpansions3D.hpp:271-272)."""
    p = 6
    M = orc.p2m(np.zeros(3), np.array([[0.1, 0.2, -0.1]]), np.array([1.0]), p)
    zL = np.array([3.0, -2.0, 4.0])
    L = orc.translate("m2l", M, p, np.zeros(3), zL)
    exact = 1.0 / np.linalg.norm(zL - np.array([0.1, 0.2, -0.1]))
    assert abs(L[0].real - exact) < 1e-6 * exact


# ---- whole solver --------------------------------------------------------------
def test_single_leaf_is_direct_summation(orc):
    """N <= cap gives one leaf, so the FMM is pure P2P and equals
    DirectSolver3D up to summation order."""
    pos, m = cabi.uniform_cube(60, 3)
    for eps in (0.0, 0.05):
        acc, st = orc.fmm(pos, m, 1.0, eps, 4, 64)
        assert st[1] == 1
        ref = orc.direct(pos, m, 1.0, eps)
        assert np.abs(acc - ref).max() < 1e-13 * np.abs(ref).max()


def test_error_falls_with_p_clustered():
    """adaptive_tests.cpp:107-126 (Gate B): Plummer N=8000-like clustered set
    (smaller here for time), eps=0: rms error falls monotonically in p and is
    at truncation level at p=8."""
    orc = cabi.oracle()
    pos, _, m = orc.plummer(3000, seed=2)
    ref = orc.direct(pos, m, 1.0, 0.0)
    prev = np.inf
    for p in (2, 4, 6, 8):
        _, rms = cabi.rel_err(ref, orc.fmm(pos, m, 1.0, 0.0, p, 64)[0])
        assert rms < prev
        prev = rms
    assert prev < 5e-5


# ---- tree invariants and brute-force list definitions ------------------------
def _brute_lists(t):
    """Lists straight from their geometric definitions (AdaptiveOctree.hpp:14-26),
    O(B^2), used to check the traversal-built ones on small trees."""
    nb = len(t["parent"])
    c = t["center"].reshape(-1, 3)
    h = t["half"]
    tol = t["tol"]

    def adj(a, b):
        return bool(np.all(np.abs(c[a] - c[b]) <= h[a] + h[b] + tol))

    leaf = t["leaf"].astype(bool)
    lev = t["level"]
    par = t["parent"]
    coll = [[0]] + [[] for _ in range(nb - 1)]
    V = [[] for _ in range(nb)]
    for b in range(1, nb):
        coll[b] = [x for x in range(nb) if lev[x] == lev[b] and adj(b, x)]
        pc = set(coll[par[b]])
        V[b] = [x for x in range(nb) if par[x] in pc and lev[x] == lev[b] and not adj(b, x)]
    anc = []
    for b in range(nb):
        a, s = b, set()
        while a >= 0:
            s.add(a)
            a = par[a]
        anc.append(s)
    U = [[] for _ in range(nb)]
    W = [[] for _ in range(nb)]
    for b in t["leaves"]:
        U[b] = sorted((x for x in t["leaves"] if adj(b, x)), key=lambda x: t["pfirst"][x])
        # W: boxes not adjacent to b whose parent is adjacent to b and is a
        # strict descendant of one of b's non-leaf colleagues
        cands = []
        for x in range(nb):
            if x == 0 or adj(b, x) or not adj(b, par[x]):
                continue
            owners = [cc for cc in coll[b] if cc != b and not leaf[cc] and cc in anc[x] and cc != x]
            if owners:
                cands.append(x)
        W[b] = sorted(cands, key=lambda x: (t["pfirst"][x], lev[x]))
    X = [[] for _ in range(nb)]
    for cl in t["leaves"]:
        for w in W[cl]:
            X[w].append(cl)
    return {"colleague": coll, "v": V, "u": U, "w": W, "x": X}


def _unpack(t, name, b):
    o = t[name + "_off"]
    return list(t[name + "_idx"][o[b]:o[b + 1]])


@pytest.mark.parametrize("case", ["clumps", "twoclusters", "two_bodies", "epsfloor"])
def test_tree_invariants_and_brute_force_lists(case):
    g = np.load(os.path.join(GOLD, f"tree_{case}.npz"))
    t = {k: g[k] for k in g.files}
    t["tol"] = float(g["tol"])
    nb = len(t["parent"])
    cap, floor = int(g["cap"]), float(g["eps_floor"])
    leaf = t["leaf"].astype(bool)
    # parent id < child id; no empty boxes; child ranges tile the parent
    for b in range(nb):
        assert t["pcount"][b] > 0
        kids = [k for k in t["child"][8 * b:8 * b + 8] if k >= 0]
        assert leaf[b] == (len(kids) == 0)
        if kids:
            assert all(k > b for k in kids)
            assert sum(t["pcount"][k] for k in kids) == t["pcount"][b]
            assert t["pcount"][b] > cap and 2 * t["half"][b] > floor and t["level"][b] < 24
        else:
            assert t["pcount"][b] <= cap or 2 * t["half"][b] <= floor or t["level"][b] == 24
    # permutation: leaves hold their bodies in ascending input order
    for b in t["leaves"]:
        seg = t["orig_index"][t["pfirst"][b]:t["pfirst"][b] + t["pcount"][b]]
        assert np.all(np.diff(seg) > 0)
    assert np.array_equal(np.sort(t["orig_index"]), np.arange(len(t["orig_index"])))
    if nb > 400:
        return
    brute = _brute_lists(t)
    for name in cabi.LISTS:
        for b in range(nb):
            got = _unpack(t, name, b)
            if name in ("colleague", "v", "x"):
                assert got == sorted(got)
            assert got == list(brute[name][b]), (name, b)


def test_diagnostics_and_potential(orc):
    """Energy and momenta (Diagnostics3D.hpp:20-38) and the per-target softened
    potential (Kernel3D.hpp:39-42): a two-body closed form, the pair-sum
    identity PE = 1/2 sum_i m_i phi_i, a translation/Galilean check of the
    momenta, and bit-equality with the compiled reference where it is built."""
    G, eps = 1.3, 0.05
    # two bodies: PE = -G m1 m2 / sqrt(d^2 + eps^2), phi_i = -G m_j / sqrt(d^2 + eps^2)
    pos = np.array([[0.1, -0.2, 0.3], [0.4, 0.2, -0.9]])
    vel = np.array([[0.0, 1.0, 0.0], [2.0, 0.0, -1.0]])
    m = np.array([0.7, 1.9])
    d2 = np.sum((pos[1] - pos[0]) ** 2) + eps * eps
    dg = orc.diagnostics(pos, vel, m, G, eps)
    assert dg["potential"] == pytest.approx(-G * m[0] * m[1] / np.sqrt(d2), rel=1e-15)
    assert dg["kinetic"] == pytest.approx(0.5 * (0.7 * 1.0 + 1.9 * 5.0), rel=1e-15)
    assert np.allclose(dg["momentum"], [3.8, 0.7, -1.9], rtol=0, atol=1e-15)
    assert np.allclose(dg["angular"], 0.7 * np.cross(pos[0], vel[0]) + 1.9 * np.cross(pos[1], vel[1]),
                       rtol=0, atol=1e-15)
    phi = cabi.direct_potential_targets(pos, m, [0, 1], G, eps)
    assert phi == pytest.approx([-G * m[1] / np.sqrt(d2), -G * m[0] / np.sqrt(d2)], rel=1e-15)
    # pair-sum identity and the reference's own diagnostics on a Plummer sphere
    pos, vel, m = orc.plummer(3000, seed=11)
    dg = orc.diagnostics(pos, vel, m, 1.0, 0.01)
    phi = cabi.direct_potential_targets(pos, m, np.arange(3000), 1.0, 0.01)
    assert 0.5 * np.dot(m, phi) == pytest.approx(dg["potential"], rel=1e-12)
    assert dg["total"] == dg["kinetic"] + dg["potential"]
    assert np.abs(dg["momentum"]).max() < 1e-12  # the generator removes the COM motion
    ref = cabi.reference()
    if ref is not None:
        dr = ref.diagnostics(pos, vel, m, 1.0, 0.01)
        for k in ("kinetic", "potential", "total"):
            assert dg[k] == dr[k], k
        assert np.array_equal(dg["momentum"], dr["momentum"])
        assert np.array_equal(dg["angular"], dr["angular"])


def test_leapfrog_matches_reference_integrator(orc):
    """The oracle's velocity Verlet over its FMM equals the reference's
    VelocityVerlet3D over AdaptiveFMM3D (Integrator3D.hpp:19-40) bit for bit
    where the reference is built; without it, a direct check of one step's
    update formulas against the oracle's own forces."""
    pos, vel, m = orc.plummer(1500, seed=4)
    p1, v1, a1 = orc.leapfrog(pos, vel, m, 1.0, 0.02, 5, 32, 0.01, 3)
    ref = cabi.reference()
    if ref is not None:
        p2, v2, a2 = ref.leapfrog(pos, vel, m, 1.0, 0.02, 5, 32, 0.01, 3)
        assert np.array_equal(p1, p2) and np.array_equal(v1, v2) and np.array_equal(a1, a2)
    a0, _ = orc.fmm(pos, m, 1.0, 0.02, 5, 32)
    x1 = pos + (vel * 0.01 + a0 * (0.5 * 0.01 * 0.01))
    an, _ = orc.fmm(x1, m, 1.0, 0.02, 5, 32)
    p3, v3, a3 = orc.leapfrog(pos, vel, m, 1.0, 0.02, 5, 32, 0.01, 1)
    assert np.array_equal(p3, x1)
    assert np.array_equal(v3, vel + (a0 + an) * (0.5 * 0.01))
    assert np.array_equal(a3, an)
