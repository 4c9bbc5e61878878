"""Leaf-capacity autotuner for the adaptive FMM on one GPU.

The leaf capacity trades near field (P2P pairs grow with it) against far
field (boxes, M2L pairs and tree depth shrink with it); the best value depends
on N, the clustering, p and the precision. tune_leaf_cap() measures the
device time (the library's own hipEvents around tree build and all passes,
afmm_stats.ms_total) of one evaluation of the actual workload per candidate,
after one untimed evaluation that sizes the buffers, and returns the fastest.
Results are cached per (p, precision, size bucket, systems) for the life of
the process.

The P2P and M2L tile shapes are not tuned here: the P2P kernel picks its
lane-group width per leaf at run time from the leaf's body count, and the M2L
tile (16-target column blocks per wave) is fixed per order at compile time.
"""
import math

CAPS = (16, 32, 48, 64, 96, 128, 192, 256)
_cache = {}


def tune_leaf_cap(afmm, sys_off, pos_ptr, mass_ptr, acc_ptr, p, precision, G=None, eps=None, caps=CAPS,
                  reps=1):
    """Returns (best_cap, {cap: ms}) for a batch on device pointers (pos AoS
    n x 3, mass, acc scratch of n x 3 doubles); sys_off as for computeBatch."""
    n = int(sys_off[-1])
    key = (p, precision, int(round(math.log2(max(n, 1)))), len(sys_off) - 1)
    if key in _cache:
        return _cache[key]
    times = {}
    for cap in caps:
        s = afmm.AdaptiveFMM3D(p=p, threshold=cap, precision=precision)
        s.compute_batch_device(sys_off, pos_ptr, mass_ptr, acc_ptr, G=G, eps=eps)
        best = float("inf")
        for _ in range(reps):
            s.compute_batch_device(sys_off, pos_ptr, mass_ptr, acc_ptr, G=G, eps=eps)
            best = min(best, s.stats()["ms_total"])
        times[cap] = best
        s.close()
    cap = min(times, key=times.get)
    _cache[key] = (cap, times)
    return cap, times
