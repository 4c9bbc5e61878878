#!/usr/bin/env python3
"""Headline benchmark: seconds per adaptive-FMM force evaluation of a
1,000,000-body two-galaxy merger (2 x 500k, disk + Hernquist bulge) at p=8,
leaf cap 64, eps=0.005, FP64, on one MI355X.

A step is one force evaluation (octree + lists + all FMM passes) of the
device-resident bodies through the C ABI (afmm_eval_device); inputs are in
HBM when timing starts. Prints one JSON line. Extra fields:
  roofline      the dominant kernel (M2L: 2 S^2 |V| flops per launch) timed
                with HIP events on the library's stream over the timed steps
  cpu_baseline  the reference CPU adaptive FMM (oracle/_ref, single thread)
                or, without it, the oracle restatement, on a random subsample
                of the same bodies, scaled linearly in N to the full workload
  phases_ms     per-phase device time (hipEvents) averaged over timed steps
  host_io_s_per_eval  afmm_eval with host arrays in and out (PCIe-inclusive)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "1M-body adaptive FMM s/eval at p=8 (FP64 vs mixed); p50/p99 request latency"
PEAK_FP64_TFLOPS = 78.6


def load_pkg():
    import importlib.util
    out = {}
    for name in ("afmm", "ics"):
        spec = importlib.util.spec_from_file_location(f"on_galaxy_amd_{name}",
                                                      os.path.join(ROOT, "on-galaxy_amd", f"{name}.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        out[name] = mod
    return out["afmm"], out["ics"]


def cpu_baseline(pos, mass, args):
    """Times the CPU adaptive FMM on a bounded random subsample (single
    thread) and scales it to s/eval at the full N (the method is O(N))."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cabi
    ref = cabi.reference()
    impl, kind = (ref, "reference") if ref is not None else (cabi.oracle(), "port")
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(len(mass), args.cpu_sample, replace=False))
    sp, sm = np.ascontiguousarray(pos[idx]), np.ascontiguousarray(mass[idx])
    t0 = time.perf_counter()
    impl.fmm(sp, sm, 1.0, args.eps, args.p, args.cap)
    dt = time.perf_counter() - t0
    return {"value": dt * len(mass) / args.cpu_sample, "unit": "s/eval", "cores": 1, "kind": kind,
            "sample": f"one {('AdaptiveFMM3D (reference, compiled from its sources)' if kind == 'reference' else 'oracle restatement')} "
                      f"evaluation of a random {args.cpu_sample}-body subsample of the same merger "
                      f"(p={args.p}, cap {args.cap}, eps {args.eps}) took {dt:.2f} s on 1 host core; "
                      f"value = that x N/{args.cpu_sample}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-per-galaxy", type=int, default=500_000)
    ap.add_argument("--p", type=int, default=8)
    ap.add_argument("--cap", type=int, default=64)
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards as described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    afmm, ics = load_pkg()
    pos, _, mass, _ = ics.merger(args.n_per_galaxy, seed=args.seed)
    n = len(mass)
    dev = torch.device("cuda:0")
    d_pos = torch.from_numpy(np.ascontiguousarray(pos)).to(dev)
    d_mass = torch.from_numpy(np.ascontiguousarray(mass)).to(dev)
    d_acc = torch.empty_like(d_pos)
    solver = afmm.AdaptiveFMM3D(G=1.0, eps=args.eps, p=args.p, threshold=args.cap)

    def step():
        solver.compute_device(d_pos.data_ptr(), d_mass.data_ptr(), d_acc.data_ptr(), n)

    torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    phases = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        phases.append(solver.stats())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s_per_eval = dt / args.steps

    st = phases[-1]
    keys = ("ms_tree", "ms_upward", "ms_m2l", "ms_downward", "ms_leaf", "ms_total")
    avg = {k[3:]: float(np.mean([ph[k] for ph in phases])) for k in keys}
    S = (args.p + 1) ** 2
    m2l_flops = 2.0 * S * S * st["n_v"]
    achieved = m2l_flops / (avg["m2l"] * 1e-3) / 1e12
    roofline = {"kernel": "M2L (k_m2l)", "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                "algorithmic_per_launch": f"2*S^2*|V| = 2*{S}^2*{st['n_v']} = {m2l_flops:.4e} flop"}

    # PCIe-inclusive drop-in semantics (host arrays in and out), reported aside
    t1 = time.perf_counter()
    solver.computeAccelerations(pos, mass)
    host_io = time.perf_counter() - t1

    out = {"metric": METRIC, "value": s_per_eval, "unit": "s/eval", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": s_per_eval * 1e3, "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"two-galaxy merger 2x{args.n_per_galaxy} (80% exp. disk + 20% Hernquist "
                                  f"bulge per galaxy), adaptive FMM p={args.p}, leaf cap {args.cap}, "
                                  f"eps {args.eps}, FP64, device-resident",
                      "n_bodies": n, "p": args.p, "leaf_cap": args.cap, "eps": args.eps,
                      "precision": "fp64", "parallelism": "single GPU"},
           "roofline": roofline,
           "phases_ms": avg,
           "tree": {"boxes": st["n_boxes"], "leaves": st["n_leaves"], "max_level": st["max_level"],
                    "max_leaf_occ": st["max_leaf_occ"], "v_pairs": st["n_v"], "p2p_pairs": st["p2p_pairs"],
                    "u_entries": st["n_u"], "w_entries": st["n_w"]},
           "host_io_s_per_eval": host_io}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pos, mass, args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
