"""Write the golden fixtures of the reference adaptive FMM.

Runs in the build container only: it needs oracle/_ref/libref.so, the
reference's own header-only CPU path (src3d/AdaptiveOctree.hpp,
AdaptiveFMM3D.hpp, FmmExpansions3D.hpp, DirectSolver3D.hpp,
InitialConditions3D.hpp) compiled by oracle/Makefile. Every fixture stores its
inputs next to the reference outputs, so the tests never regenerate inputs:

  tree_<case>.npz   inputs (in_pos, in_mass) + full AdaptiveOctree export (box arrays, permuted
                    bodies, origIndex, leaves, colleague/U/V/W/X CSR, tol)
  force_<case>.npz  inputs + AdaptiveFMM3D accelerations at several p, plus
                    DirectSolver3D accelerations
  ops.npz           operator input/output pairs (P2M, M2M, M2L, L2L,
                    local/multipole evaluation, harmonics) at p = 4, 8, 12
  MANIFEST.json     cases, seeds, compiler and flags, sha256 of every file

Usage: python tests/golden/make_golden.py   (after `make -C oracle`)
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cabi  # noqa: E402


def clumps(seed):
    """Three clumps of 80 bodies within 1e-9 of a point (drives the tree to
    the depth-24 cap with cap 16) plus 300 background bodies."""
    rng = np.random.default_rng(seed)
    parts = [rng.uniform(-1, 1, (300, 3))]
    for c in ((0.3, -0.2, 0.1), (-0.5, 0.5, 0.5), (0.70001, 0.7, -0.7)):
        parts.append(np.array(c) + 1e-9 * rng.uniform(-1, 1, (80, 3)))
    pos = np.concatenate(parts)
    return pos, rng.uniform(0.5, 1.5, len(pos))


def two_clusters(ref):
    """A dense Plummer clump beside a diffuse one: adjacent leaves at very
    different levels, so W and X are non-empty."""
    p1, _, m1 = ref.plummer(1500, a=0.05, seed=11)
    p2, _, m2 = ref.plummer(900, a=1.0, seed=12)
    p1 = p1 + np.array([0.4, 0.1, -0.2])
    return np.concatenate([p1, p2]), np.concatenate([m1, m2])


def duplicates(seed):
    """Exactly coincident bodies (tree only: forces are singular at eps=0)."""
    rng = np.random.default_rng(seed)
    base = rng.uniform(-1, 1, (200, 3))
    pos = np.concatenate([base, np.repeat(base[:3], 40, axis=0), np.zeros((20, 3))])
    return pos, np.ones(len(pos))


def main():
    ref = cabi.reference()
    if ref is None:
        sys.exit("oracle/_ref/libref.so missing: run `make -C oracle` with the reference present")
    cases = {}

    pos, _, m = ref.plummer(2000, seed=5)
    cases["plummer2k"] = dict(pos=pos, mass=m, cap=64, eps=0.01, ps=(4, 8, 12))
    pos, m = cabi.uniform_cube(1000, 21)
    cases["cube1k"] = dict(pos=pos, mass=m, cap=32, eps=0.0, ps=(8,))
    pos, m = clumps(31)
    cases["clumps"] = dict(pos=pos, mass=m, cap=16, eps=0.0, ps=(6,))
    pos, _, m = ref.plummer(1500, seed=41)
    cases["epsfloor"] = dict(pos=pos, mass=m, cap=8, eps=0.05, ps=(8,))
    pos, m = two_clusters(ref)
    cases["twoclusters"] = dict(pos=pos, mass=m, cap=24, eps=0.001, ps=(8,))
    pos, m = cabi.uniform_cube(40, 51)
    cases["single_leaf"] = dict(pos=pos, mass=m, cap=64, eps=0.0, ps=(4,))
    cases["one_body"] = dict(pos=np.array([[0.5, -0.25, 2.0]]), mass=np.array([3.0]), cap=64,
                             eps=0.0, ps=(4,))
    cases["two_bodies"] = dict(pos=np.array([[0.0, 0.0, 0.0], [1.0, 2.0, -1.0]]),
                               mass=np.array([1.0, 2.0]), cap=1, eps=0.0, ps=(4,))
    pos, m = duplicates(61)
    cases["duplicates"] = dict(pos=pos, mass=m, cap=16, eps=0.0, ps=())

    manifest = {"generator": "tests/golden/make_golden.py",
                "reference": ["src3d/AdaptiveOctree.hpp", "src3d/AdaptiveFMM3D.hpp",
                              "src3d/FmmExpansions3D.hpp", "src3d/DirectSolver3D.hpp",
                              "src3d/InitialConditions3D.hpp"],
                "compiler": subprocess.run(["g++", "--version"], capture_output=True,
                                           text=True).stdout.splitlines()[0],
                "flags": "-O3 -std=c++17 -ffp-contract=off (oracle/Makefile)",
                "cases": {}, "files": {}}
    for name, c in cases.items():
        pos, m = np.ascontiguousarray(c["pos"]), np.ascontiguousarray(c["mass"])
        t = ref.tree(pos, m, c["cap"], c["eps"])
        tol = t.pop("tol")
        np.savez_compressed(os.path.join(HERE, f"tree_{name}.npz"), in_pos=pos, in_mass=m,
                            cap=c["cap"], eps_floor=c["eps"], tol=tol, **t)
        info = {"n": len(m), "cap": c["cap"], "eps": c["eps"], "boxes": int(len(t["parent"])),
                "max_level": int(t["level"].max()),
                "lists": {k: int(len(t[k + "_idx"])) for k in cabi.LISTS}}
        if c["ps"]:
            out = {"pos": pos, "mass": m, "cap": c["cap"], "eps": c["eps"], "G": 1.0,
                   "ps": np.array(c["ps"]), "direct": ref.direct(pos, m, 1.0, c["eps"])}
            for p in c["ps"]:
                out[f"acc_p{p}"], _ = ref.fmm(pos, m, 1.0, c["eps"], p, c["cap"])
            np.savez_compressed(os.path.join(HERE, f"force_{name}.npz"), **out)
            info["ps"] = list(c["ps"])
        manifest["cases"][name] = info
        print(name, info)

    # operator pairs
    rng = np.random.default_rng(77)
    ops = {}
    for p in (4, 8, 12):
        S = (p + 1) ** 2
        src = rng.uniform(-0.5, 0.5, (30, 3)) + np.array([0.1, -0.2, 0.3])
        q = rng.uniform(0.5, 1.5, 30)
        c0 = np.array([0.1, -0.2, 0.3])
        Mc = ref.p2m(c0, src, q, p)
        ops[f"p{p}_src"], ops[f"p{p}_q"], ops[f"p{p}_c0"], ops[f"p{p}_p2m"] = src, q, c0, Mc
        c1 = c0 + np.array([0.5, 0.5, -0.5])
        ops[f"p{p}_c1"] = c1
        ops[f"p{p}_m2m"] = ref.translate("m2m", Mc, p, c0, c1)
        c2 = c0 + np.array([2.0, -4.0, 2.0])
        ops[f"p{p}_c2"] = c2
        Lc = ref.translate("m2l", Mc, p, c0, c2)
        ops[f"p{p}_m2l"] = Lc
        c3 = c2 + np.array([-0.25, 0.25, 0.25])
        ops[f"p{p}_c3"] = c3
        ops[f"p{p}_l2l"] = ref.translate("l2l", Lc, p, c2, c3)
        t_near = c2 + np.array([0.2, -0.1, 0.15])
        t_far = c0 + np.array([3.0, 2.0, -2.5])
        ops[f"p{p}_tnear"], ops[f"p{p}_tfar"] = t_near, t_far
        pot, grad = ref.evaluate("local", Lc, p, c2, t_near)
        ops[f"p{p}_local_pot"], ops[f"p{p}_local_grad"] = np.array([pot]), grad
        pot, grad = ref.evaluate("multipole", Mc, p, c0, t_far)
        ops[f"p{p}_mult_pot"], ops[f"p{p}_mult_grad"] = np.array([pot]), grad
        v = np.array([0.3, -0.7, 0.4])
        r, Y = ref.harmonics(v, 2 * p)
        ops[f"p{p}_hv"], ops[f"p{p}_Y"] = v, Y
        assert len(Mc) == S
    np.savez_compressed(os.path.join(HERE, "ops.npz"), **ops)

    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            with open(os.path.join(HERE, f), "rb") as fh:
                manifest["files"][f] = hashlib.sha256(fh.read()).hexdigest()
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
