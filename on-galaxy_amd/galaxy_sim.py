#!/usr/bin/env python3
"""Galaxy simulation driver on the device-resident integrator: the MI355X
counterpart of the reference's gpu/galaxy_sim.cu loop (:178-283).

    python on-galaxy_amd/galaxy_sim.py --scenario merger --n-per-galaxy 50000 \\
        --steps 100 --dt 0.01 --p 8 --eps 0.005 --out run/merger

Bodies, velocities and accelerations stay in HBM (afmm_state); positions come
back to the host only for frames and the separation record. Writes the
reference's files (<out>_frames.bin, _species.bin, _sep.csv, _energy.csv; see
simio.py). The energy record uses the FMM potential of the current positions,
so unlike the reference (n <= 60000, O(N^2) pair sum, :196) it is kept at any N.
"""
import argparse
import importlib.util
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    spec = importlib.util.spec_from_file_location(f"on_galaxy_amd_{name}", os.path.join(_HERE, f"{name}.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def run(scenario="merger", n_per_galaxy=50000, steps=100, dt=0.01, p=8, eps=0.005, cap=64, frame_stride=10,
        out="galaxy", precision="fp64", seed=42, energy=True, log=print):
    afmm, ics, simio = _load("afmm"), _load("ics"), _load("simio")
    prec = afmm.PRECISION_MIXED if precision == "mixed" else afmm.PRECISION_FP64
    solver = afmm.AdaptiveFMM3D(G=1.0, eps=eps, p=p, threshold=cap, precision=prec)
    if scenario == "merger":
        pos, vel, mass, species = ics.merger(n_per_galaxy, seed=seed, solver=solver)
    else:
        pos, vel, mass = ics.plummer(n_per_galaxy, seed=seed)
        species = np.zeros(len(mass), np.int8)
    n = len(mass)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    n_frames = steps // frame_stride + 1
    frames = simio.FrameWriter(out, n, n_frames)
    simio.write_species(out, species)
    st = afmm.LeapfrogState(solver, pos, vel, mass)
    frames.write(pos)
    merger = scenario == "merger"
    sep_csv = simio.CsvWriter(out + "_sep.csv", "step,time,sep") if merger else None
    min_sep, peri = np.inf, 0
    if merger:
        s0 = simio.com_separation(pos, mass, species)
        sep_csv.row(0, 0.0, s0)
        min_sep = s0
    en_csv, e0, max_drift = None, None, 0.0
    stride_e = max(1, steps // 50)
    if energy:
        e0 = st.diagnostics()[0]["total"]
        en_csv = simio.CsvWriter(out + "_energy.csv", "step,time,total,relDrift")
        en_csv.row(0, 0.0, e0, 0.0)
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        st.step(dt, 1)
        need_pos = step % frame_stride == 0 or merger
        if need_pos:
            cur, _, _ = st.download()
            if step % frame_stride == 0:
                frames.write(cur)
            if merger:
                sep = simio.com_separation(cur, mass, species)
                if step % 5 == 0 or step == steps:
                    sep_csv.row(step, step * dt, sep)
                if sep < min_sep:
                    min_sep, peri = sep, step
        if energy and (step % stride_e == 0 or step == steps):
            e = st.diagnostics()[0]["total"]
            rel = abs(e - e0) / abs(e0)
            max_drift = max(max_drift, rel)
            en_csv.row(step, step * dt, e, rel)
        if step % 100 == 0:
            log(f"  step {step}/{steps}  t={step * dt:.2f}  sep={min_sep if merger else 0:.2f}")
    sec = (time.perf_counter() - t0) / max(steps, 1)
    frames.close()
    for c in (sep_csv, en_csv):
        if c:
            c.close()
    summary = {"n": n, "steps": steps, "frames": n_frames, "s_per_step": sec, "max_rel_energy_drift": max_drift,
               "min_sep": float(min_sep) if merger else None, "peri_step": peri if merger else None}
    log(f"frames={n_frames}  per-step={sec:.4f} s  total-sim={sec * steps:.1f} s")
    if energy:
        log(f"E0={e0:.6e}  max|dE|/|E0|={max_drift:.3e}")
    log(f"wrote {out}_frames.bin ({n_frames} frames, N={n})")
    st.close()
    solver.close()
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scenario", choices=["merger", "plummer"], default="merger")
    ap.add_argument("--n-per-galaxy", type=int, default=50000)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--p", type=int, default=8)
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--cap", type=int, default=64)
    ap.add_argument("--frame-stride", type=int, default=10)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--out", default="galaxy")
    a = ap.parse_args(argv)
    run(a.scenario, a.n_per_galaxy, a.steps, a.dt, a.p, a.eps, a.cap, a.frame_stride, a.out, a.precision, a.seed,
        not a.no_energy)


if __name__ == "__main__":
    sys.exit(main())
