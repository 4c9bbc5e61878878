"""Kernel shapes (afmm_set_shapes) and the precision-attribution modes.

  * waves per M2L workgroup and the persistent work-binned P2P schedule leave
    every result bit-identical (each target's sums keep their order);
  * the P2P lane grouping changes only the order of each near-field sum:
    within 1e-13 of the default, same tree and statistics;
  * a batch is still bit-identical to its systems alone under any shapes;
  * shapes the kernels cannot run are rejected;
  * the autotuner returns shapes the library accepts and its timings;
  * M2L_BF16 (bf16 M2L, FP64 P2P) and P2P_FP32 (FP64 M2L, FP32 P2P) each
    differ from FP64 by no more than the mode that combines them
    (MIXED_BF16), and each stays within the mixed-precision bound.
"""
import numpy as np
import pytest

import cabi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afmm():
    return cabi.product()


def bodies(n=20000, seed=3):
    orc = cabi.oracle()
    pos, _, m = orc.plummer(n, seed=seed)
    # a dense clump: large leaves at the eps floor for the work bins
    rng = np.random.default_rng(seed)
    k = n // 10
    pos[:k] = 0.3 + 1e-3 * rng.standard_normal((k, 3))
    return np.ascontiguousarray(pos), np.ascontiguousarray(m)


@pytest.mark.parametrize("prec_name", ["fp64", "bf16split"])
def test_shapes_keep_results(afmm, prec_name):
    prec = {"fp64": afmm.PRECISION_FP64, "bf16split": afmm.PRECISION_MIXED_BF16}[prec_name]
    pos, m = bodies()
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.002, p=8, threshold=48, precision=prec)
    a0 = s.computeAccelerations(pos, m)
    st0 = s.stats()
    print(f"shapes {prec_name}: default {s.shapes()}", flush=True)
    for sh in ({"m2l_waves": 8}, {"p2p_schedule": 1}, {"m2l_waves": 8, "p2p_schedule": 1}, {"p2p_schedule": 2}):
        s.set_shapes(**sh)
        print(f"  bit-identical under {s.shapes()}", flush=True)
        a = s.computeAccelerations(pos, m)
        assert np.array_equal(a, a0), sh
    for sh in ({"p2p_pair_bodies": 16}, {"p2p_pair_bodies": 64}, {"p2p_min_group": 4}, {"p2p_min_group": 16},
               {"p2p_min_group": 64, "p2p_schedule": 1}):
        s.set_shapes(**sh)
        print(f"  reordered near sums under {s.shapes()}", flush=True)
        a = s.computeAccelerations(pos, m)
        mx, _ = cabi.rel_err(a0, a)
        assert mx <= 1e-13, (sh, mx)
        st = s.stats()
        for k in ("n_boxes", "n_leaves", "n_u", "n_v", "n_w", "n_x", "p2p_pairs"):
            assert st[k] == st0[k], k
    s.close()


def test_batch_equals_alone_under_shapes(afmm):
    orc = cabi.oracle()
    sys_ = [orc.plummer(n, seed=50 + i)[::2] for i, n in enumerate((3000, 800, 12000, 5000))]
    sys_ = [(np.ascontiguousarray(p), np.ascontiguousarray(m)) for p, m in sys_]
    off = np.concatenate([[0], np.cumsum([len(m) for _, m in sys_])])
    pos = np.concatenate([p for p, _ in sys_])
    mass = np.concatenate([m for _, m in sys_])
    sh = {"m2l_waves": 8, "p2p_schedule": 1, "p2p_pair_bodies": 16, "p2p_min_group": 4}
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.01, p=6, threshold=32)
    s.set_shapes(**sh)
    acc = s.computeBatch(pos, mass, off)
    for q, (pq, mq) in enumerate(sys_):
        assert np.array_equal(acc[off[q]:off[q + 1]], s.computeAccelerations(pq, mq)), q
    s.close()


def test_invalid_shapes_rejected(afmm):
    s = afmm.AdaptiveFMM3D(p=10)
    for sh in ({"m2l_waves": 12}, {"m2l_waves": 4}, {"p2p_min_group": 3}, {"p2p_min_group": 128},
               {"p2p_schedule": 3}, {"p2p_pair_bodies": -1}):
        with pytest.raises(afmm.AfmmError):
            s.set_shapes(**sh)
    assert s.shapes()["m2l_waves"] == 8
    s.set_shapes(m2l_waves=16)  # 16 waves at any order (two per 16 targets above p = 8)
    s.close()
    for p, w in ((8, 16), (12, 8), (14, 16)):
        s = afmm.AdaptiveFMM3D(p=p)
        assert s.shapes()["m2l_waves"] == w, p
        s.close()
    b = afmm.AdaptiveFMM3D(p=10, precision=afmm.PRECISION_MIXED_BF16)
    b.set_shapes(m2l_waves=16)  # the bf16 M2L runs 16 waves at any order
    b.close()


@pytest.mark.parametrize("prec_name,p", [("fp64", 10), ("fp64", 14), ("split", 8)])
def test_row_split_m2l_bit_identical(afmm, prec_name, p):
    """16-wave M2L workgroups in which two waves share 16 targets (half the
    row blocks each) give the bits of the 8-wave workgroups: every row keeps
    its accumulation order."""
    prec = {"fp64": afmm.PRECISION_FP64, "split": afmm.PRECISION_MIXED}[prec_name]
    pos, m = bodies(30000, seed=9)
    out = []
    for w in (8, 16):
        s = afmm.AdaptiveFMM3D(G=1.0, eps=0.002, p=p, threshold=48, precision=prec)
        s.set_shapes(m2l_waves=w)
        out.append(s.computeAccelerations(pos, m))
        s.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("p", [4, 8])
def test_small_tile_row_split_bit_identical(afmm, p):
    """Small trees cut M2L into tiles of 16 or 32 targets, whose 16-target
    blocks the 16-wave FP64 workgroups split by rows over four or two waves
    (64 and more targets: one wave per block): the bits of the 8-wave
    workgroups (tiles of 64 or more targets, one wave per block) at every
    size."""
    for n in (3000, 80000, 200000):
        pos, m = bodies(n, seed=13)
        print(f"row split by tile size: n={n} p={p}", flush=True)
        out = []
        for w in (8, 16):
            s = afmm.AdaptiveFMM3D(G=1.0, eps=0.002, p=p, threshold=64)
            s.set_shapes(m2l_waves=w)
            out.append(s.computeAccelerations(pos, m))
            s.close()
        assert np.array_equal(out[0], out[1]), n


def test_autotuner_returns_accepted_shapes(afmm):
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location("tune_t", os.path.join(cabi.ROOT, "on-galaxy_amd", "tune.py"))
    tune = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tune)
    pos, m = bodies(8000, seed=4)
    off = np.array([0, len(m)], np.int64)
    d_pos, d_mass, d_acc = afmm.DeviceArray(None, pos), afmm.DeviceArray(None, m), afmm.DeviceArray(pos.shape)
    cfg, log = tune.tune(afmm, off, d_pos.ptr, d_mass.ptr, d_acc.ptr, 6, afmm.PRECISION_FP64, caps=(32, 64))
    print(f"autotuned: {cfg}", flush=True)
    assert cfg["leaf_cap"] in (32, 64) and len(log["shapes"]) >= 6 and len(log["overlap"]) == 2
    s = afmm.AdaptiveFMM3D(p=6, threshold=cfg["leaf_cap"])
    s.set_shapes(**cfg["shapes"])
    s.set_overlap(cfg["overlap"])
    assert s.shapes() == cfg["shapes"]
    s.close()


@pytest.mark.parametrize("prec_name,p", [("fp64", 8), ("fp64", 10), ("bf16split", 8), ("split", 6)])
def test_overlap_modes_keep_results(afmm, prec_name, p):
    """Every schedule (one stream; two streams with the list passes beside M2L
    or in order before it; the automatic choice) gives the same bits, with
    16- and 8-wave M2L workgroups."""
    prec = {"fp64": afmm.PRECISION_FP64, "bf16split": afmm.PRECISION_MIXED_BF16,
            "split": afmm.PRECISION_MIXED}[prec_name]
    pos, m = bodies(30000, seed=8)
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.002, p=p, threshold=48, precision=prec)
    a0 = s.computeAccelerations(pos, m)
    for waves in sorted({s.shapes()["m2l_waves"], 8}):
        s.set_shapes(m2l_waves=waves)
        for mode in (afmm.OVERLAP_OFF, afmm.OVERLAP_BESIDE_M2L, afmm.OVERLAP_IN_ORDER, afmm.OVERLAP_AUTO):
            s.set_overlap(mode)
            print(f"  {prec_name} p={p} waves={waves} overlap={mode}", flush=True)
            assert np.array_equal(s.computeAccelerations(pos, m), a0), (waves, mode)
    for bad in (-1, 4):
        with pytest.raises(afmm.AfmmError):
            s.set_overlap(bad)
    s.close()


def test_attribution_modes(afmm):
    pos, m = bodies(60000, seed=5)
    s = afmm.AdaptiveFMM3D(G=1.0, eps=0.002, p=8, threshold=64)
    a64 = s.computeAccelerations(pos, m)
    s.close()
    err = {}
    for name, prec in (("m2l_bf16", afmm.PRECISION_M2L_BF16), ("p2p_fp32", afmm.PRECISION_P2P_FP32),
                       ("mixed_bf16", afmm.PRECISION_MIXED_BF16)):
        s = afmm.AdaptiveFMM3D(G=1.0, eps=0.002, p=8, threshold=64, precision=prec)
        a = s.computeAccelerations(pos, m)
        s.close()
        err[name] = cabi.rel_err(a64, a)[1]
        print(f"{name} vs fp64: rms={err[name]:.3e}", flush=True)
    assert err["m2l_bf16"] > 0 and err["p2p_fp32"] > 0
    assert max(err["m2l_bf16"], err["p2p_fp32"]) <= 1.5 * err["mixed_bf16"] + 1e-12
    assert err["mixed_bf16"] <= 1e-6
