"""GPU checks of batched evaluation (afmm_eval_batch): many independent
systems in one forest, one launch per FMM phase. Every system's result must
be bit-identical to evaluating it alone (its tree, lists and arithmetic are
the same; MFMA tiles and P2P waves never mix a target's sums with another
system's), in both precisions, with per-system G and eps."""
import numpy as np
import pytest

import cabi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afmm():
    return cabi.product()


def systems(seed, count, nmin=1, nmax=6000):
    """Uneven systems: Plummer spheres and offset uniform cubes, with a few
    tiny ones (1 and 2 bodies, and one at or below the leaf capacity)."""
    orc = cabi.oracle()
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = int(np.exp(rng.uniform(np.log(max(nmin, 2)), np.log(nmax))))
        if q == 1:
            n = 1
        elif q == 2:
            n = 2
        elif q == 3:
            n = 40
        if q % 2 == 0 and n > 2:
            pos, _, m = orc.plummer(n, seed=int(rng.integers(1 << 30)))
        else:
            pos = rng.uniform(-1, 1, (n, 3)) * rng.uniform(0.1, 10) + rng.uniform(-50, 50, 3)
            m = rng.uniform(0.5, 1.5, n) / n
        out.append((np.ascontiguousarray(pos), np.ascontiguousarray(m)))
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed_bf16"])
def test_batch_equals_separate_runs(afmm, prec):
    precision = {"fp64": afmm.PRECISION_FP64, "mixed": afmm.PRECISION_MIXED,
                 "mixed_bf16": afmm.PRECISION_MIXED_BF16}[prec]
    sys_ = systems(7, 24)
    rng = np.random.default_rng(3)
    G = rng.uniform(0.5, 2.0, len(sys_))
    eps = rng.choice([0.0, 0.005, 0.01, 0.05], len(sys_))
    off = np.concatenate([[0], np.cumsum([len(m) for _, m in sys_])])
    pos = np.concatenate([p for p, _ in sys_])
    mass = np.concatenate([m for _, m in sys_])
    print(f"batch {prec}: {len(sys_)} systems, {off[-1]} bodies", flush=True)
    s = afmm.AdaptiveFMM3D(p=6, threshold=32, precision=precision)
    acc, phi = s.computeBatch(pos, mass, off, G, eps, potential=True)
    st = s.stats()
    for q, (p, m) in enumerate(sys_):
        one = afmm.AdaptiveFMM3D(G=G[q], eps=eps[q], p=6, threshold=32, precision=precision)
        a1, f1 = one.computeAccelerationsAndPotential(p, m)
        assert np.array_equal(acc[off[q]:off[q + 1]], a1), q
        assert np.array_equal(phi[off[q]:off[q + 1]], f1), q
    assert st["n_bodies"] == off[-1]


def test_batch_trees_match_oracle_and_stats(afmm):
    """Summed box/leaf counts of the forest equal those of the oracle's
    per-system trees, and each system's forces match the oracle to 1e-12."""
    sys_ = systems(11, 10, nmax=3000)
    off = np.concatenate([[0], np.cumsum([len(m) for _, m in sys_])])
    pos = np.concatenate([p for p, _ in sys_])
    mass = np.concatenate([m for _, m in sys_])
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.01, p=5, threshold=16)
    acc = s.computeBatch(pos, mass, off)
    st = s.stats()
    orc = cabi.oracle()
    nb = nl = 0
    for q, (p, m) in enumerate(sys_):
        t = orc.tree(p, m, 16, 0.01)
        nb += len(t["parent"])
        nl += len(t["leaves"])
        if len(m) > 1:
            ref, _ = orc.fmm(p, m, 1.0, 0.01, 5, 16)
            mx, _ = cabi.rel_err(ref, acc[off[q]:off[q + 1]])
            assert mx <= 1e-12, (q, mx)
    assert (st["n_boxes"], st["n_leaves"]) == (nb, nl)


def test_batch_argument_errors(afmm):
    s = afmm.AdaptiveFMM3D(p=4)
    pos = np.zeros((4, 3))
    m = np.ones(4)
    import ctypes as C
    for off in ([1, 4], [0, 2, 2, 4], [0, 3, 2, 4]):
        # the wrapper refuses offsets that would overrun its arrays ...
        with pytest.raises(ValueError):
            s.computeBatch(pos, m, off)
        # ... and the C ABI rejects them itself (AFMM_ERR_ARG)
        o = np.ascontiguousarray(off, np.int64)
        acc = np.zeros_like(pos)
        dp = C.POINTER(C.c_double)
        rc = s.lib.afmm_eval_batch(s.ctx, len(o) - 1, o.ctypes.data_as(C.POINTER(C.c_int64)), None, None,
                                   pos.ctypes.data_as(dp), m.ctypes.data_as(dp), acc.ctypes.data_as(dp), None)
        assert rc == -1, (off, rc)
    with pytest.raises(afmm.AfmmError):
        s.computeBatch(pos, m, [0, 2, 4], eps=[0.1, -1.0])


@pytest.mark.parametrize("prec", ["fp64", "mixed_bf16"])
def test_graph_mode_replays_equal_stream_launches(afmm, prec):
    """Graph mode (FMM passes captured into a HIP graph, the executable graph
    updated in place or re-instantiated per evaluation) returns bit-identical
    results to plain stream launches: three replays of the same batch with
    eager evaluations and synchronisations between them, then a batch of a
    different shape and a single system with potentials."""
    precision = {"fp64": afmm.PRECISION_FP64, "mixed_bf16": afmm.PRECISION_MIXED_BF16}[prec]
    sys_ = systems(11, 12)
    rng = np.random.default_rng(5)
    G = rng.uniform(0.5, 2.0, len(sys_))
    eps = rng.choice([0.0, 0.01], len(sys_))
    off = np.concatenate([[0], np.cumsum([len(m) for _, m in sys_])])
    pos = np.concatenate([p for p, _ in sys_])
    mass = np.concatenate([m for _, m in sys_])
    print(f"graph {prec}: {len(sys_)} systems, {off[-1]} bodies", flush=True)
    eager = afmm.AdaptiveFMM3D(p=6, threshold=32, precision=precision)
    graph = afmm.AdaptiveFMM3D(p=6, threshold=32, precision=precision)
    graph.set_graph(True)
    ref_acc, ref_phi = eager.computeBatch(pos, mass, off, G, eps, potential=True)
    for rep in range(3):
        acc, phi = graph.computeBatch(pos, mass, off, G, eps, potential=True)
        assert np.array_equal(acc, ref_acc), rep
        assert np.array_equal(phi, ref_phi), rep
        a2, f2 = eager.computeBatch(pos, mass, off, G, eps, potential=True)
        assert np.array_equal(a2, ref_acc)
    st = graph.stats()
    assert st["graph_instantiations"] >= 1
    assert st["graph_instantiations"] + st["graph_updates"] == 3
    # another shape: fewer systems
    sys2 = systems(12, 5)
    off2 = np.concatenate([[0], np.cumsum([len(m) for _, m in sys2])])
    pos2 = np.concatenate([p for p, _ in sys2])
    mass2 = np.concatenate([m for _, m in sys2])
    G2, eps2 = np.ones(len(sys2)), np.full(len(sys2), 0.01)
    assert np.array_equal(graph.computeBatch(pos2, mass2, off2, G2, eps2),
                          eager.computeBatch(pos2, mass2, off2, G2, eps2))
    p1, m1 = sys_[0]
    a_g, f_g = graph.computeAccelerationsAndPotential(p1, m1)
    a_e, f_e = eager.computeAccelerationsAndPotential(p1, m1)
    assert np.array_equal(a_g, a_e) and np.array_equal(f_g, f_e)
    assert graph.stats()["graph_instantiations"] + graph.stats()["graph_updates"] == 5


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed_bf16"])
def test_two_stream_overlap_equals_one_stream(afmm, prec):
    """The two-stream schedule (upward pass beside the list build, P2P near
    field beside the far-field passes, then a separate far + near combine)
    returns the same bits as every pass in order on one stream, for a batch
    with eps = 0 and eps > 0 systems (masked and unmasked self pairs) and a
    single system with and without potentials."""
    precision = {"fp64": afmm.PRECISION_FP64, "mixed_bf16": afmm.PRECISION_MIXED_BF16}[prec]
    sys_ = systems(21, 9)
    rng = np.random.default_rng(8)
    G = rng.uniform(0.5, 2.0, len(sys_))
    eps = rng.choice([0.0, 0.01], len(sys_))
    off = np.concatenate([[0], np.cumsum([len(m) for _, m in sys_])])
    pos = np.concatenate([p for p, _ in sys_])
    mass = np.concatenate([m for _, m in sys_])
    print(f"overlap {prec}: {len(sys_)} systems, {off[-1]} bodies", flush=True)
    two = afmm.AdaptiveFMM3D(p=6, threshold=32, precision=precision)
    one = afmm.AdaptiveFMM3D(p=6, threshold=32, precision=precision)
    one.set_overlap(False)
    for pot in (False, True):
        r2 = two.computeBatch(pos, mass, off, G, eps, potential=pot)
        r1 = one.computeBatch(pos, mass, off, G, eps, potential=pot)
        if pot:
            assert np.array_equal(r2[0], r1[0]) and np.array_equal(r2[1], r1[1])
        else:
            assert np.array_equal(r2, r1)
    p1, m1 = sys_[0]
    for s_ in (two, one):
        s_.close()
    two = afmm.AdaptiveFMM3D(eps=0.01, p=8, threshold=64, precision=precision)
    one = afmm.AdaptiveFMM3D(eps=0.01, p=8, threshold=64, precision=precision)
    one.set_overlap(False)
    assert np.array_equal(two.computeAccelerations(p1, m1), one.computeAccelerations(p1, m1))
    a2, f2 = two.computeAccelerationsAndPotential(p1, m1)
    a1, f1 = one.computeAccelerationsAndPotential(p1, m1)
    assert np.array_equal(a2, a1) and np.array_equal(f2, f1)
    assert two.stats()["ms_p2p"] > 0 and one.stats()["ms_p2p"] > 0
