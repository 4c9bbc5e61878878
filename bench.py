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
                of the same bodies, scaled linearly in N to the full workload;
                all_cores: the oracle restatement on every host core (OpenMP)
  phases_ms     per-phase device time (hipEvents) averaged over timed steps
  host_io_s_per_eval  afmm_eval with host arrays in and out (PCIe-inclusive)
  mixed         the same evaluation in mixed precision (FP32 P2P with FP64
                accumulation; M2L as split-FP32 MFMA or as three-piece bf16
                MFMA): s/eval, speedup, phases of both forms
  accuracy      relative RMS and max relErr (floor 1e-6 max|a|) of both forms
                against FP64 direct summation at random targets (GPU)
  large_4m      the same at N = 4M (2 x 2M): s/eval and eps = 0 error per form
  energy_drift  100 leapfrog steps of the merger per form: |E - E0|/|E0|
  batched       256 systems in one batched evaluation vs 256 sequential solves
                (aggregate body-interactions/s), leaf capacity autotuned
  serving       a Poisson job stream through the continuous-batching server vs
                one job at a time: p50/p99 time-to-result, bodies*steps/s
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))


def progress(msg):
    """One progress line per leg on stderr (the JSON result stays the last stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
METRIC = "1M-body adaptive FMM s/eval at p=8 (FP64 vs mixed); p50/p99 request latency"
PEAK_FP64_TFLOPS = 78.6


def load_pkg():
    import importlib.util
    out = {}
    for name in ("afmm", "ics", "tune"):
        spec = importlib.util.spec_from_file_location(f"on_galaxy_amd_{name}",
                                                      os.path.join(ROOT, "on-galaxy_amd", f"{name}.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        out[name] = mod
    return out["afmm"], out["ics"], out["tune"]


def measured_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed counter summary
    (profiles/m2l_traffic_merger1M_p8.md, lines `traffic <kernel> <bytes>`):
    FETCH_SIZE and WRITE_SIZE of two rocprofv3 --pmc runs, scaled by what the
    same counters read for a plain 1 GiB device copy. None if absent."""
    path = os.path.join(ROOT, "profiles", "m2l_traffic_merger1M_p8.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.strip().strip("`").split()
        if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel_key:
            return float(parts[2])
    return None


def host_cpu():
    """Core count and CPU model of the host the CPU figures were taken on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline(pos, mass, args):
    """Times the CPU adaptive FMM (single thread, as the reference has no
    threads) on two bounded random subsamples of the workload and scales each
    to s/eval at the full N (the method is O(N)); the value is the larger
    sample's figure, the smaller one shows how far the linear scaling holds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cabi
    ref = cabi.reference()
    impl, kind = (ref, "reference") if ref is not None else (cabi.oracle(), "port")
    what = ("AdaptiveFMM3D (reference, compiled from its sources)" if kind == "reference"
            else "oracle restatement")
    rng = np.random.default_rng(1)
    runs, samples = {}, {}
    orc = cabi.oracle()
    orc.set_threads(1)
    for ns in (args.cpu_sample, 4 * args.cpu_sample):
        idx = np.sort(rng.choice(len(mass), ns, replace=False))
        sp, sm = np.ascontiguousarray(pos[idx]), np.ascontiguousarray(mass[idx])
        t0 = time.perf_counter()
        acc, _ = impl.fmm(sp, sm, 1.0, args.eps, args.p, args.cap)
        dt = time.perf_counter() - t0
        runs[ns] = {"sample_bodies": ns, "seconds": dt, "scaled_s_per_eval": dt * len(mass) / ns}
        samples[ns] = (sp, sm, acc)
    big = runs[4 * args.cpu_sample]
    small = runs[args.cpu_sample]
    # The same method on every host core: the oracle restatement, whose
    # OpenMP passes split the boxes of one level (and the leaves) and keep each
    # box's sums in the reference's order. Its one-thread time on the smaller
    # sample calibrates it against the figure above; its all-core result on
    # the larger sample must be the same bits.
    sp, sm, _ = samples[args.cpu_sample]
    t0 = time.perf_counter()
    orc.fmm(sp, sm, 1.0, args.eps, args.p, args.cap)
    orc1 = time.perf_counter() - t0
    nt = orc.set_threads(0)
    sp, sm, acc_big = samples[4 * args.cpu_sample]
    t0 = time.perf_counter()
    acc_all, _ = orc.fmm(sp, sm, 1.0, args.eps, args.p, args.cap)
    tall = time.perf_counter() - t0
    all_cores = {"kind": "port", "threads": nt, "sample_bodies": 4 * args.cpu_sample, "seconds": tall,
                 "scaled_s_per_eval": tall * len(mass) / (4 * args.cpu_sample),
                 "speedup_vs_1_core": big["seconds"] / tall,
                 "oracle_1_thread_vs_1_core_figure": orc1 / small["seconds"],
                 "same_bits_as_1_core": bool(np.array_equal(acc_all, acc_big))}
    return {"value": big["scaled_s_per_eval"], "unit": "s/eval", "cores": 1, "kind": kind,
            "host": host_cpu(),
            "sample": f"one {what} evaluation of a random {big['sample_bodies']}-body subsample of the same "
                      f"merger (p={args.p}, cap {args.cap}, eps {args.eps}) took {big['seconds']:.2f} s on 1 "
                      f"host core; value = that x N/{big['sample_bodies']}",
            "subsamples": list(runs.values()),
            "scaling_check": big["scaled_s_per_eval"] / small["scaled_s_per_eval"],
            "all_cores": all_cores}


def c0_leg(afmm, args):
    """The reference's CPU configuration: N = 16384 Plummer sphere
    (makePlummerSphere, a = 1, M = 1, G = 1, seed 42), eps 0.01, p = 8, leaf
    capacity 64, VelocityVerlet3D initialize + one step of dt 0.01 (two force
    evaluations), host arrays in and out. GPU: FP64 and both mixed forms on the
    resident leapfrog (upload, two evaluations, download), median of 5 after a
    warm-up; CPU: the reference's own integrator over its AdaptiveFMM3D on one
    core (oracle/_ref), best of 3. Also the GPU FP64 state against the CPU
    state after the step, and the initial accelerations against FP64 direct
    all-pairs (softened, and at eps = 0 for the truncation error)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cabi
    n, eps, p, cap, dt = 16384, 0.01, 8, 64, 0.01
    pos, vel, mass = afmm.make_plummer(n, seed=42)
    out = {"n": n, "p": p, "eps": eps, "leaf_cap": cap, "dt": dt, "work": "initialize + 1 step (2 evaluations)"}
    states = {}
    for name, prec in (("fp64", afmm.PRECISION_FP64), ("mixed_fp32_split", afmm.PRECISION_MIXED),
                       ("mixed_bf16_split", afmm.PRECISION_MIXED_BF16)):
        s = afmm.AdaptiveFMM3D(eps=eps, p=p, threshold=cap, precision=prec)
        ts = []
        for rep in range(6):
            t0 = time.perf_counter()
            st = afmm.LeapfrogState(s, pos, vel, mass)
            st.step(dt, 1)
            res = st.download()
            t = time.perf_counter() - t0
            st.close()
            if rep:
                ts.append(t)
        states[name] = res
        out[name] = {"s_median": float(np.median(ts)), "s_min": float(min(ts))}
        s.close()
    sd = afmm.AdaptiveFMM3D(eps=eps, p=p, threshold=cap)
    a0 = sd.computeAccelerations(pos, mass)
    allt = np.arange(n)
    out["fp64_vs_direct_softened"] = rel_errors(sd.direct(pos, mass, allt), a0)
    sd.close()
    s0 = afmm.AdaptiveFMM3D(eps=0.0, p=p, threshold=cap)
    out["fp64_vs_direct_eps0"] = rel_errors(s0.direct(pos, mass, allt), s0.computeAccelerations(pos, mass))
    s0.close()
    ref = cabi.reference()
    if ref is not None:
        best, cpu_state = float("inf"), None
        for _ in range(3):
            t0 = time.perf_counter()
            cpu_state = ref.leapfrog(pos, vel, mass, 1.0, eps, p, cap, dt, 1)
            best = min(best, time.perf_counter() - t0)
        out["cpu_reference"] = {"s": best, "cores": 1, "host": host_cpu(),
                                "what": "VelocityVerlet3D over AdaptiveFMM3D, compiled from the reference's sources"}
        out["speedup_fp64_vs_cpu"] = best / out["fp64"]["s_median"]
        g = states["fp64"]
        out["fp64_state_vs_cpu"] = {k: float(np.abs(g[i] - cpu_state[i]).max() / np.abs(cpu_state[i]).max())
                                    for i, k in enumerate(("pos", "vel", "acc"))}
    return out


def drift_leg(afmm, ics, args):
    """C1/C2 energy drift: the merger (with velocities from the solver's forces)
    integrated for --drift-steps leapfrog steps of dt = 0.01 on the resident
    state, in FP64 and both mixed forms at p = 8, and in FP64 and bf16 mixed at
    p = 4 and 12; |E - E0|/|E0| every 10 steps with the FMM potential
    (afmm_state_diagnostics)."""
    ic = afmm.AdaptiveFMM3D(eps=args.eps, p=4, threshold=args.cap)
    pos, vel, mass, _ = ics.merger(args.drift_n // 2, seed=args.seed, solver=ic)
    ic.close()
    out = {"n": int(len(mass)), "steps": args.drift_steps, "dt": 0.01, "p": args.p, "eps": args.eps}
    forms = [(name, prec, args.p) for name, prec in (("fp64", afmm.PRECISION_FP64),
                                                       ("mixed_fp32_split", afmm.PRECISION_MIXED),
                                                       ("mixed_bf16_split", afmm.PRECISION_MIXED_BF16))]
    for p in (4, 12):
        forms += [(f"p{p}_fp64", afmm.PRECISION_FP64, p), (f"p{p}_mixed_bf16_split", afmm.PRECISION_MIXED_BF16, p)]
    for name, prec, p in forms:
        s = afmm.AdaptiveFMM3D(eps=args.eps, p=p, threshold=args.cap, precision=prec)
        st = afmm.LeapfrogState(s, pos, vel, mass)
        e0 = st.diagnostics()[0]["total"]
        drift = []
        t_step = 0.0
        for _ in range(max(args.drift_steps // 10, 1)):
            t0 = time.perf_counter()
            st.step(0.01, 10)
            t_step += time.perf_counter() - t0
            drift.append(abs(st.diagnostics()[0]["total"] - e0) / abs(e0))
        out[name] = {"p": p, "E0": e0, "max_rel_drift": float(max(drift)), "final_rel_drift": float(drift[-1]),
                     "s_per_step": t_step / (10 * len(drift))}
        st.close()
        s.close()
    out["mixed_minus_fp64_final_drift"] = {k: out[k]["final_rel_drift"] - out["fp64"]["final_rel_drift"]
                                           for k in ("mixed_fp32_split", "mixed_bf16_split")}
    return out


def rel_errors(ref, a):
    """RMS and max-rel (floor 1e-30 + 1e-6 max|ref|, adaptive_gpu_harness.cu:36-46)."""
    rn = np.linalg.norm(ref, axis=1)
    floor = 1e-30 + 1e-6 * rn.max()
    d = np.linalg.norm(a - ref, axis=1)
    return {"rms": float(np.sqrt((d ** 2).sum() / (rn ** 2).sum())), "max_rel": float((d / np.maximum(rn, floor)).max())}


def large_leg(afmm, ics, args):
    """C2 at N = 4M (2 x 2M merger): s/eval of FP64 and both mixed forms at
    p = 8, FP64 and bf16 mixed at p = 4 and 12, the eps = 0 truncation error
    of each against FP64 direct on 4096 targets, and the 100-step energy drift
    of the FP64 and bf16 forms at p = 8."""
    ic = afmm.AdaptiveFMM3D(eps=args.eps, p=4, threshold=args.cap)  # disk velocities from its forces
    pos, vel, mass, _ = ics.merger(args.large_n // 2, seed=args.seed + 9, solver=ic)
    ic.close()
    n = len(mass)
    d_pos, d_mass, d_acc = afmm.DeviceArray(None, pos), afmm.DeviceArray(None, mass), afmm.DeviceArray(pos.shape)
    idx = np.random.default_rng(args.seed + 10).choice(n, 4096, replace=False)
    out = {"n": n}
    ref0 = None
    forms = [(name, prec, args.p) for name, prec in (("fp64", afmm.PRECISION_FP64),
                                                       ("mixed_fp32_split", afmm.PRECISION_MIXED),
                                                       ("mixed_bf16_split", afmm.PRECISION_MIXED_BF16))]
    for p in (4, 12):
        forms += [(f"p{p}_fp64", afmm.PRECISION_FP64, p), (f"p{p}_mixed_bf16_split", afmm.PRECISION_MIXED_BF16, p)]
    for name, prec, p in forms:
        s = afmm.AdaptiveFMM3D(eps=args.eps, p=p, threshold=args.cap, precision=prec)
        s.compute_device(d_pos.ptr, d_mass.ptr, d_acc.ptr, n)
        t0 = time.perf_counter()
        for _ in range(3):
            s.compute_device(d_pos.ptr, d_mass.ptr, d_acc.ptr, n)
        t = (time.perf_counter() - t0) / 3
        s0 = afmm.AdaptiveFMM3D(eps=0.0, p=p, threshold=args.cap, precision=prec)
        s0.compute_device(d_pos.ptr, d_mass.ptr, d_acc.ptr, n)
        a0 = d_acc.numpy()[idx]
        if ref0 is None:
            ref0 = s0.direct(pos, mass, idx)
        e = rel_errors(ref0, a0)
        out[name] = {"p": p, "s_per_eval": t, "eps0_rms_vs_direct": e["rms"], "eps0_max_rel_vs_direct": e["max_rel"]}
        s.close()
        s0.close()
    for k, ref in (("mixed_fp32_split", "fp64"), ("mixed_bf16_split", "fp64"),
                   ("p4_mixed_bf16_split", "p4_fp64"), ("p12_mixed_bf16_split", "p12_fp64")):
        out[k]["speedup_vs_fp64"] = out[ref]["s_per_eval"] / out[k]["s_per_eval"]
    # energy drift over --drift-steps leapfrog steps (dt = 0.01) at 4M, p = 8,
    # from the generator's velocities, with the FMM potential every 10 steps
    for name, prec in (("fp64", afmm.PRECISION_FP64), ("mixed_bf16_split", afmm.PRECISION_MIXED_BF16)):
        s = afmm.AdaptiveFMM3D(eps=args.eps, p=args.p, threshold=args.cap, precision=prec)
        st = afmm.LeapfrogState(s, pos, vel, mass)
        e0 = st.diagnostics()[0]["total"]
        drift, t_step = [], 0.0
        for _ in range(max(args.drift_steps // 10, 1)):
            t0 = time.perf_counter()
            st.step(0.01, 10)
            t_step += time.perf_counter() - t0
            drift.append(abs(st.diagnostics()[0]["total"] - e0) / abs(e0))
        out[name]["drift_100_steps"] = {"steps": 10 * len(drift), "max_rel_drift": float(max(drift)),
                                        "final_rel_drift": float(drift[-1]), "s_per_step": t_step / (10 * len(drift))}
        st.close()
        s.close()
    return out


def batched_leg(afmm, ics, tune, args):
    """256 independent systems (N log-uniform over 2k-200k; even: Plummer a=1,
    odd: Hernquist a=0.5 truncated at 20a; total mass 1 each), p=8,
    eps=0.01, mixed precision (bf16-split M2L, FP32 P2P), leaf capacity from
    the autotuner, FMM passes replayed as one HIP graph per step: one batched
    evaluation (one launch per FMM phase for all systems) against 256
    sequential plain (FP64, leaf capacity 64) solves of the same systems. The
    same step with stream launches is reported beside it."""
    parts = ics.batch_systems(args.batch_systems, seed=args.seed + 3)
    ns = np.array([len(mq) for _, mq in parts], np.int64)
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    d_pos = afmm.DeviceArray(None, np.concatenate([p for p, _ in parts]))
    d_mass = afmm.DeviceArray(None, np.concatenate([m for _, m in parts]))
    d_acc = afmm.DeviceArray(d_pos.shape)
    P, E = d_pos.ptr, d_mass.ptr
    A = d_acc.ptr
    cfg, tune_log = tune.tune(afmm, off, P, E, A, 8, afmm.PRECISION_MIXED_BF16, eps=0.01)
    cap, shapes, overlap = cfg["leaf_cap"], cfg["shapes"], cfg["overlap"]

    def timed_batch(precision, leaf_cap, reps=3, graph=False, sh=None, ov=None):
        s = afmm.AdaptiveFMM3D(eps=0.01, p=8, threshold=leaf_cap, precision=precision)
        s.set_graph(graph)
        if sh:
            s.set_shapes(**sh)
        if ov is not None:
            s.set_overlap(ov)
        s.compute_batch_device(off, P, E, A)
        t0 = time.perf_counter()
        for _ in range(reps):
            s.compute_batch_device(off, P, E, A)
        return (time.perf_counter() - t0) / reps, s.stats()

    t_mix, st = timed_batch(afmm.PRECISION_MIXED_BF16, cap, graph=True, sh=shapes)
    t_mix_stream, st_stream = timed_batch(afmm.PRECISION_MIXED_BF16, cap, sh=shapes, ov=overlap)
    t_mix_default, _ = timed_batch(afmm.PRECISION_MIXED_BF16, cap)
    t_b64, _ = timed_batch(afmm.PRECISION_FP64, 64)
    # 256 sequential plain solves (FP64, cap 64), device-resident slices
    s = afmm.AdaptiveFMM3D(eps=0.01, p=8, threshold=64)
    s.compute_device(P, E, A, int(ns[0]))
    t0 = time.perf_counter()
    for q in range(len(ns)):
        o = int(off[q])
        s.compute_device(P + 24 * o, E + 8 * o, A + 24 * o, int(ns[q]))
    t_seq = time.perf_counter() - t0
    inter = float(np.sum(ns.astype(np.float64) * (ns - 1)))
    return {"systems": int(len(ns)), "bodies": int(off[-1]), "p": 8, "eps": 0.01,
            "leaf_cap": int(cap), "shapes": shapes, "overlap_stream_launch": int(overlap),
            "autotune_ms_total_by_cap": {str(k): round(v, 3) for k, v in tune_log["leaf_cap"].items()},
            "autotune_shape_trials": tune_log["shapes"], "autotune_overlap_trials": tune_log["overlap"],
            "batched_mixed_default_shapes_s_per_step": t_mix_default,
            "phases_ms_stream_launch": {k[3:]: st_stream[k] for k in ("ms_tree", "ms_lists", "ms_upward", "ms_m2l",
                                                                      "ms_downward", "ms_leaf", "ms_p2p", "ms_total")},
            "precision": "mixed_bf16", "graph": {"instantiations": int(st["graph_instantiations"]),
                                                 "in_place_updates": int(st["graph_updates"])},
            "batched_mixed_s_per_step": t_mix, "batched_mixed_stream_launch_s_per_step": t_mix_stream,
            "batched_fp64_cap64_s_per_step": t_b64,
            "sequential_plain_s_per_step": t_seq,
            "body_interactions_per_s": inter / t_mix,
            "sequential_body_interactions_per_s": inter / t_seq,
            "speedup_vs_sequential": t_seq / t_mix,
            "speedup_fp64_batched_vs_sequential_fp64": t_seq / t_b64,
            "actual_p2p_pairs": int(st["p2p_pairs"]), "v_pairs": int(st["n_v"]),
            "metric_note": "body-interactions/s = sum_i N_i (N_i - 1) / t_step (direct-equivalent)"}


def serving_leg(afmm, ics, args):
    """A seeded Poisson stream (mean 20 jobs/s) of Plummer or two-galaxy jobs,
    N log-uniform over 10k-2M, p in {4, 8, 12}, 1-50 leapfrog steps of 0.01,
    through the continuous-batching server in FP64 and in bf16 mixed
    precision, and the same jobs one at a time (FP64, each on its own resident
    leapfrog, in arrival order). Also the trace submitted all at once (the
    server's sustainable job rate for this mix), and for a sample of the mixed
    server's jobs the error of their final accelerations against the FP64 FMM
    at the same positions and against FP64 direct summation on 1024 targets."""
    import threading
    ic_solver = afmm.AdaptiveFMM3D(eps=0.005, p=4, threshold=64)
    jobs = ics.serving_jobs(args.serve_jobs, seed=args.seed + 4, rate=20.0, solver=ic_solver)
    ic_solver.close()
    work = float(sum(len(j["mass"]) * j["steps"] for j in jobs))

    def summarize(lat, makespan):
        lat = np.asarray(lat)
        return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "makespan_s": makespan, "bodies_steps_per_s": work / makespan}

    def served(precision, at_once=False):
        sv = afmm.Server(precision=precision)
        for p in (4, 8, 12):
            sv.warm(p)
        ids = [None] * len(jobs)
        t0 = time.perf_counter()

        def submitter():
            for k, j in enumerate(jobs):
                dt = t0 + j["t"] - time.perf_counter()
                if dt > 0 and not at_once:
                    time.sleep(dt)
                ids[k] = sv.submit(j["pos"], j["vel"], j["mass"], 1.0, j["eps"], 0.01, j["p"], j["steps"])

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        sv.drain()
        infos = [sv.info(i) for i in ids]
        results = [sv.result(i) for i in ids]
        sv.close()
        lat = [i["t_done"] - i["t_submit"] for i in infos]
        span = max(i["t_done"] for i in infos) - min(i["t_submit"] for i in infos)
        out = summarize(lat, span)
        out["most_batched_evaluations_of_a_job"] = int(max(i["batches"] for i in infos))
        return out, results

    fp64, _ = served(afmm.PRECISION_FP64)
    bf16, res_bf = served(afmm.PRECISION_MIXED_BF16)
    burst, _ = served(afmm.PRECISION_FP64, at_once=True)
    burst["sustainable_jobs_per_s"] = len(jobs) / burst["makespan_s"]
    # accuracy of sampled mixed-precision results at their final positions
    rng = np.random.default_rng(args.seed + 5)
    acc_rows = []
    for k in sorted(rng.choice(len(jobs), min(4, len(jobs)), replace=False).tolist()):
        j = jobs[k]
        pos, _, acc = res_bf[k]
        s = afmm.AdaptiveFMM3D(eps=j["eps"], p=j["p"], threshold=64)
        a64 = s.computeAccelerations(pos, j["mass"])
        idx = rng.choice(len(j["mass"]), min(1024, len(j["mass"])), replace=False)
        ref = s.direct(pos, j["mass"], idx)
        s.close()
        acc_rows.append({"job": k, "kind": j["kind"], "n": int(len(j["mass"])), "p": j["p"], "steps": j["steps"],
                         "mixed_vs_fp64_fmm": rel_errors(a64, acc), "mixed_vs_fp64_direct": rel_errors(ref, acc[idx]),
                         "fp64_fmm_vs_fp64_direct": rel_errors(ref, a64[idx])})
    # plain one-job-at-a-time path
    lat2 = []
    t0 = time.perf_counter()
    for j in jobs:
        dt = t0 + j["t"] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        s = afmm.AdaptiveFMM3D(eps=j["eps"], p=j["p"], threshold=64)
        st = afmm.LeapfrogState(s, j["pos"], j["vel"], j["mass"])
        st.step(0.01, j["steps"])
        st.download()
        st.close()
        s.close()
        lat2.append(time.perf_counter() - (t0 + j["t"]))
    span2 = time.perf_counter() - t0
    return {"jobs": len(jobs), "arrival_rate_per_s": 20.0, "total_bodies_steps": work,
            "n_range": [int(min(len(j["mass"]) for j in jobs)), int(max(len(j["mass"]) for j in jobs))],
            "orders": sorted({j["p"] for j in jobs}),
            "server": fp64, "server_mixed_bf16": bf16, "server_all_at_once_fp64": burst,
            "mixed_accuracy_sampled_jobs": acc_rows,
            "accuracy_note": "eps > 0: the FMM softens the near field only (reference semantics), so against "
                             "softened direct summation the far field differs by ~eps^2/r^2; mixed_vs_fp64_fmm "
                             "is the error the reduced precision adds",
            "one_at_a_time": summarize(lat2, span2)}


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
    ap.add_argument("--targets", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=100)
    ap.add_argument("--drift-n", type=int, default=1_000_000)
    ap.add_argument("--large-n", type=int, default=4_000_000)
    ap.add_argument("--no-drift", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-orders", action="store_true")
    ap.add_argument("--no-c0", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (scale for HBM counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        afmm, _, _ = load_pkg()
        a = afmm.DeviceArray(None, np.ones(args.copy_calibration << 17))
        b = afmm.DeviceArray(a.shape)
        for _ in range(5):
            b.copy_from(a)
        print(json.dumps({"copy_bytes_per_launch": a.nbytes, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: "
              "collectives and sharding are not built (independent systems would be bin-packed "
              "across GPUs, one large system partitioned by Morton key range)", file=sys.stderr)
        sys.exit(2)

    afmm, ics, tune = load_pkg()
    pos, _, mass, _ = ics.merger(args.n_per_galaxy, seed=args.seed)
    n = len(mass)
    d_pos, d_mass, d_acc = afmm.DeviceArray(None, pos), afmm.DeviceArray(None, mass), afmm.DeviceArray(pos.shape)
    def timed(precision, p=None, steps=None, overlap=True):
        """W warmup + K timed evaluations; returns (s/eval, averaged phase ms, last stats, solver).
        overlap=False runs every pass in order on one stream."""
        p = args.p if p is None else p
        nsteps = args.steps if steps is None else steps
        solver = afmm.AdaptiveFMM3D(G=1.0, eps=args.eps, p=p, threshold=args.cap, precision=precision)
        solver.set_overlap(overlap)

        def step():
            solver.compute_device(d_pos.ptr, d_mass.ptr, d_acc.ptr, n)

        for _ in range(args.warmup):
            step()
        phases = []
        t0 = time.perf_counter()
        for _ in range(nsteps):
            step()
            phases.append(solver.stats())
        dt = time.perf_counter() - t0
        keys = ("ms_tree", "ms_lists", "ms_upward", "ms_m2l", "ms_downward", "ms_leaf", "ms_total", "ms_p2p")
        avg = {k[3:]: float(np.mean([ph[k] for ph in phases])) for k in keys}
        return dt / nsteps, avg, phases[-1], solver

    progress(f"FP64 timing, n={n}")
    s_per_eval, avg, st, solver = timed(afmm.PRECISION_FP64)
    acc64 = d_acc.numpy()
    S = (args.p + 1) ** 2
    m2l_flops = 2.0 * S * S * st["n_v"]
    achieved = m2l_flops / (avg["m2l"] * 1e-3) / 1e12
    roofline = {"kernel": "M2L (k_m2l_mfma, FP64 MFMA)", "bound": "mfma", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": measured_traffic("k_m2l_mfma_fp64"),
                "traffic_source": "profiles/m2l_traffic_merger1M_p8.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
                "algorithmic_per_launch": f"2*S^2*|V| = 2*{S}^2*{st['n_v']} = {m2l_flops:.4e} flop"}

    # the same evaluation with every pass in order on one stream: P2P timed
    # alone (on two streams it shares the GPU with M2L), and the overlap's gain
    progress("FP64 timing, one stream")
    seq_s, seq_avg, _, seq_solver = timed(afmm.PRECISION_FP64, steps=max(3, args.steps // 2), overlap=False)
    seq_solver.close()
    # the same launch with nothing beside it: in the timed (two-stream) run P2P
    # starts beside M2L and the two share the CUs for part of M2L's duration
    alone = m2l_flops / (seq_avg["m2l"] * 1e-3) / 1e12
    roofline["alone"] = {"ms": seq_avg["m2l"], "achieved": alone, "frac": alone / PEAK_FP64_TFLOPS,
                         "note": "k_m2l_mfma in the one-stream run (afmm_set_overlap 0), no other kernel beside it"}

    # PCIe-inclusive drop-in semantics (host arrays in and out), reported aside:
    # median of 5 after a warm-up
    solver.computeAccelerations(pos, mass)
    host_runs = []
    for _ in range(5):
        t1 = time.perf_counter()
        solver.computeAccelerations(pos, mass)
        host_runs.append(time.perf_counter() - t1)
    host_io = float(np.median(host_runs))

    progress("mixed-precision timing")
    # mixed precision on the same bodies: split-FP32 and three-piece bf16 M2L
    mix_s, mix_avg, _, _ = timed(afmm.PRECISION_MIXED)
    accmx = d_acc.numpy()
    mix_achieved = m2l_flops / (mix_avg["m2l"] * 1e-3) / 1e12
    bf_s, bf_avg, _, _ = timed(afmm.PRECISION_MIXED_BF16)
    accbf = d_acc.numpy()
    bf_achieved = m2l_flops / (bf_avg["m2l"] * 1e-3) / 1e12

    # accuracy: FP64 direct summation (on the GPU) at random targets. Like the
    # reference, the FMM softens only the near field (P2P), so at eps > 0 its
    # far field differs from a fully softened direct sum by ~1.5 eps^2/r^2,
    # which in the dense cores of this merger reaches 1e-2; the truncation
    # error is measured at eps = 0 on the same bodies (the reference's own
    # accuracy gates use eps = 0, adaptive_tests.cpp:107-126).
    idx = np.random.default_rng(args.seed + 1).choice(n, args.targets, replace=False)

    def errs(ref, a):
        floor = 1e-30 + 1e-6 * np.linalg.norm(ref, axis=1).max()
        d = np.linalg.norm(a - ref, axis=1)
        return {"rms": float(np.sqrt((d ** 2).sum() / (np.linalg.norm(ref, axis=1) ** 2).sum())),
                "max_rel": float((d / np.maximum(np.linalg.norm(ref, axis=1), floor)).max())}

    progress("accuracy against FP64 direct")
    ref_soft = solver.direct(pos, mass, idx)
    accuracy = {"targets": args.targets,
                "eps0": {}, "softened": {"eps": args.eps, "fp64_vs_direct": errs(ref_soft, acc64[idx]),
                                         "mixed_vs_direct": errs(ref_soft, accmx[idx]),
                                         "note": "far field unsoftened by design (reference semantics)"},
                "mixed_vs_fp64": errs(acc64, accmx), "mixed_bf16_vs_fp64": errs(acc64, accbf)}
    ref0 = None
    for name, prec in (("fp64", afmm.PRECISION_FP64), ("mixed", afmm.PRECISION_MIXED),
                       ("mixed_bf16", afmm.PRECISION_MIXED_BF16)):
        s0 = afmm.AdaptiveFMM3D(G=1.0, eps=0.0, p=args.p, threshold=args.cap, precision=prec)
        s0.compute_device(d_pos.ptr, d_mass.ptr, d_acc.ptr, n)
        a0 = d_acc.numpy()
        if ref0 is None:
            ref0 = s0.direct(pos, mass, idx)
        accuracy["eps0"][name + "_vs_direct"] = errs(ref0, a0[idx])
        s0.close()
    # the other orders of the same workload: FP64 and bf16-split timing (softened,
    # as the headline) and accuracy against FP64 direct at eps = 0
    orders = {}
    if not args.no_orders:
        progress("orders 4, 12, 14 and 16")
        for po in (4, 12, 14, 16):
            if po == args.p:
                continue
            row = {}
            for name, prec in (("fp64", afmm.PRECISION_FP64), ("mixed_bf16", afmm.PRECISION_MIXED_BF16)):
                t_o, ph_o, _, so = timed(prec, p=po, steps=3)
                so.close()
                s0 = afmm.AdaptiveFMM3D(G=1.0, eps=0.0, p=po, threshold=args.cap, precision=prec)
                s0.compute_device(d_pos.ptr, d_mass.ptr, d_acc.ptr, n)
                a0 = d_acc.numpy()
                s0.close()
                row[name] = {"s_per_eval": t_o, "m2l_ms": ph_o["m2l"], "eps0_vs_direct": errs(ref0, a0[idx])}
            row["mixed_bf16_speedup_vs_fp64"] = row["fp64"]["s_per_eval"] / row["mixed_bf16"]["s_per_eval"]
            orders[f"p{po}"] = row
    best = min((mix_s, "fp32_split"), (bf_s, "bf16_split"))
    mixed = {"s_per_eval": best[0], "speedup_vs_fp64": s_per_eval / best[0], "form": best[1],
             "fp32_split": {"s_per_eval": mix_s, "speedup_vs_fp64": s_per_eval / mix_s, "phases_ms": mix_avg,
                            "m2l_tflops_algorithmic": mix_achieved,
                            "m2l_form": "split FP32 MFMA (hi/lo operands, 3 products)"},
             "bf16_split": {"s_per_eval": bf_s, "speedup_vs_fp64": s_per_eval / bf_s, "phases_ms": bf_avg,
                            "m2l_tflops_algorithmic": bf_achieved,
                            "m2l_form": "bf16 MFMA 16x16x32 (3-piece operands, 6 products)"},
             "p2p_form": "FP32, FP64 accumulation"}

    out = {"metric": METRIC, "value": s_per_eval, "unit": "s/eval", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": s_per_eval * 1e3, "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"two-galaxy merger 2x{args.n_per_galaxy} (80% exp. disk + 20% Hernquist "
                                  f"bulge per galaxy), adaptive FMM p={args.p}, leaf cap {args.cap}, "
                                  f"eps {args.eps}, FP64, device-resident",
                      "n_bodies": n, "p": args.p, "leaf_cap": args.cap, "eps": args.eps,
                      "precision": "fp64", "parallelism": "single GPU"},
           "roofline": roofline,
           # the near field on the VALU, at the conventional 20 flop per softened
           # pair interaction, against the FP64 peak
           "p2p_valu": {"kernel": "k_p2p (FP64 VALU), timed alone in the one-stream run",
                        "pairs": st["p2p_pairs"], "ms": seq_avg["p2p"],
                        "achieved_tflops": 20.0 * st["p2p_pairs"] / (seq_avg["p2p"] * 1e-3) / 1e12,
                        "peak_tflops": PEAK_FP64_TFLOPS,
                        "frac": 20.0 * st["p2p_pairs"] / (seq_avg["p2p"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
           "phases_ms": avg,
           "one_stream": {"s_per_eval": seq_s, "phases_ms": seq_avg,
                          "note": "every pass in order on one stream (afmm_set_overlap 0)"},
           "mixed": mixed,
           "accuracy": accuracy,
           "orders": orders,
           "tree": {"boxes": st["n_boxes"], "leaves": st["n_leaves"], "max_level": st["max_level"],
                    "max_leaf_occ": st["max_leaf_occ"], "v_pairs": st["n_v"], "p2p_pairs": st["p2p_pairs"],
                    "colleague_entries": st["n_colleague"], "u_entries": st["n_u"], "w_entries": st["n_w"],
                    "x_entries": st["n_x"],
                    # algorithmic bytes of the build (68 N + 32 N + 80 B + 4 E, E = all list entries)
                    # over the one-stream tree phase
                    "build_bytes": 100.0 * n + 80.0 * st["n_boxes"] + 4.0 * (st["n_colleague"] + st["n_u"] + st["n_v"]
                                                                           + st["n_w"] + st["n_x"]),
                    "build_gbps_one_stream": (100.0 * n + 80.0 * st["n_boxes"] + 4.0 * (
                        st["n_colleague"] + st["n_u"] + st["n_v"] + st["n_w"] + st["n_x"]))
                    / (seq_avg["tree"] * 1e-3) / 1e9},
           "host_io_s_per_eval": host_io}
    if not args.no_large:
        progress("4M leg")
        out["large_4m"] = large_leg(afmm, ics, args)
    if not args.no_drift:
        progress("energy drift leg")
        out["energy_drift"] = drift_leg(afmm, ics, args)
    if not args.no_batch:
        progress("batched leg")
        out["batched"] = batched_leg(afmm, ics, tune, args)
    if not args.no_serve:
        progress("serving leg")
        out["serving"] = serving_leg(afmm, ics, args)
        out["latency_p50_s"] = out["serving"]["server"]["p50_s"]
        out["latency_p99_s"] = out["serving"]["server"]["p99_s"]
    if not args.no_c0:
        progress("C0 leg (16384-body Plummer, initialize + one step)")
        out["c0"] = c0_leg(afmm, args)
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(pos, mass, args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
