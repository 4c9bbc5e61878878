"""Autotuner for the adaptive FMM on one GPU: leaf capacity and kernel shapes.

The leaf capacity trades near field (P2P pairs grow with it) against far
field (boxes, M2L pairs and tree depth shrink with it); the kernel shapes
(afmm_shapes: waves per M2L workgroup, the P2P lane grouping and the P2P
schedule) trade occupancy, operator reuse and load balance. The best values
depend on N, the clustering, p and the precision, so they are measured on the
actual workload: each candidate is timed by the library's own hipEvents
around tree build and all passes (afmm_stats.ms_total) after one untimed
evaluation that sizes its buffers.

tune_leaf_cap() searches the capacity alone (default shapes). tune() searches
the capacity first, then each shape in turn at the best values found so far
(coordinate descent; every shape leaves the tree, the lists and the accuracy
unchanged), then the two-stream schedule (afmm_set_overlap: the list passes
beside M2L or in order before it; the same bits either way). Results are
cached per (p, precision, size bucket, systems) for the life of the process.
"""
import math

CAPS = (16, 32, 48, 64, 96, 128, 192, 256)
# candidate values per shape (0 in a field keeps the library default)
SHAPES = (("m2l_waves", (16, 8)), ("p2p_schedule", (2, 1)), ("p2p_pair_bodies", (32, 16, 48)),
          ("p2p_min_group", (8, 4, 16)))
_cache = {}


def _time(afmm, cap, shapes, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G, eps, reps, overlap=None):
    s = afmm.AdaptiveFMM3D(p=p, threshold=cap, precision=precision)
    try:
        s.set_shapes(**shapes)
        if overlap is not None:
            s.set_overlap(overlap)
        s.compute_batch_device(sys_off, pos_ptr, mass_ptr, acc_ptr, G=G, eps=eps)
        best = float("inf")
        for _ in range(reps):
            s.compute_batch_device(sys_off, pos_ptr, mass_ptr, acc_ptr, G=G, eps=eps)
            best = min(best, s.stats()["ms_total"])
        return best
    finally:
        s.close()


def tune_leaf_cap(afmm, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G=None, eps=None, caps=CAPS,
                  reps=1):
    """Returns (best_cap, {cap: ms}) for a batch on device pointers (pos AoS
    n x 3, mass, acc scratch of n x 3 doubles); sys_off as for computeBatch."""
    n = int(sys_off[-1])
    key = ("cap", p, precision, int(round(math.log2(max(n, 1)))), len(sys_off) - 1)
    if key in _cache:
        return _cache[key]
    times = {cap: _time(afmm, cap, {}, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G, eps, reps)
             for cap in caps}
    cap = min(times, key=times.get)
    _cache[key] = (cap, times)
    return cap, times


def tune(afmm, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G=None, eps=None, caps=CAPS, reps=1):
    """Leaf capacity, then kernel shapes, by coordinate descent, then the
    overlap mode. Returns (config, log): config = {"leaf_cap": c, "shapes":
    {...}, "overlap": mode} to pass to AdaptiveFMM3D(threshold=c),
    set_shapes(**shapes) and set_overlap(mode); log holds every timing (ms)
    in search order."""
    n = int(sys_off[-1])
    key = ("full", p, precision, int(round(math.log2(max(n, 1)))), len(sys_off) - 1)
    if key in _cache:
        return _cache[key]
    cap, cap_ms = tune_leaf_cap(afmm, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G, eps, caps, reps)
    log = {"leaf_cap": cap_ms, "shapes": []}
    best = {}
    best_ms = cap_ms[cap]
    probe = afmm.AdaptiveFMM3D(p=p, threshold=cap, precision=precision)
    for name, values in SHAPES:
        for v in values:
            trial = dict(best, **{name: v})
            try:
                probe.set_shapes(**trial)  # rejects shapes this order or precision cannot run
            except afmm.AfmmError:
                continue
            ms = _time(afmm, cap, trial, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G, eps, reps)
            log["shapes"].append({"shapes": trial, "ms": ms})
            if ms < best_ms:
                best, best_ms = trial, ms
    probe.set_shapes(**best)
    resolved = probe.shapes()
    probe.close()
    overlap = afmm.OVERLAP_AUTO
    log["overlap"] = []
    for mode in (afmm.OVERLAP_BESIDE_M2L, afmm.OVERLAP_IN_ORDER):
        ms = _time(afmm, cap, best, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G, eps, reps, mode)
        log["overlap"].append({"overlap": mode, "ms": ms})
        if ms < best_ms:
            overlap, best_ms = mode, ms
    out = ({"leaf_cap": cap, "shapes": resolved, "overlap": overlap, "ms": best_ms}, log)
    _cache[key] = out
    return out
