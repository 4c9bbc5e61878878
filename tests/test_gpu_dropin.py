"""The C++ drop-in (include/galaxy3d/AdaptiveFMM3D_HIP.hpp) used exactly like
the reference's solver: a kick-drift-kick leapfrog through the ForceSolver3D
interface (Integrator3D.hpp:19-40) on a clustered Plummer sphere. The energy
check mirrors adaptive_tests.cpp Gate D (max |dE|/|E0| < 1e-2)."""
import json
import os
import subprocess

import pytest

import cabi

pytestmark = pytest.mark.gpu


def test_cpp_dropin_leapfrog_energy():
    exe = os.path.join(cabi.ROOT, "on-galaxy_amd", "cpp", "drop_in_demo")
    assert os.path.exists(exe), "build with `make -C on-galaxy_amd`"
    print("drop_in_demo 4000 bodies 20 steps p=6", flush=True)
    out = subprocess.run([exe, "4000", "20", "6"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    assert rec["n"] == 4000 and rec["steps"] == 20
    assert rec["dE_rel"] < 1e-2
    assert rec["max_leaf_occ"] <= 64 and rec["boxes"] > 1
