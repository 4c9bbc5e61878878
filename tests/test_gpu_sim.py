"""The device-resident simulation driver (galaxy_sim.py) writes the
reference's files with the right sizes and keeps the energy bounded."""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_merger_run_outputs(tmp_path):
    spec = importlib.util.spec_from_file_location("gsim", os.path.join(ROOT, "on-galaxy_amd", "galaxy_sim.py"))
    g = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(g)
    base = str(tmp_path / "m")
    print("galaxy_sim merger 2x3000, 20 steps", flush=True)
    out = g.run("merger", 3000, steps=20, dt=0.01, p=6, eps=0.005, frame_stride=10, out=base, log=lambda *_: None)
    n = 6000
    assert os.path.getsize(base + "_frames.bin") == 8 + 3 * n * 3 * 4
    assert os.path.getsize(base + "_species.bin") == n
    sep = np.loadtxt(base + "_sep.csv", delimiter=",", skiprows=1)
    assert sep.shape == (5, 3) and list(sep[:, 0]) == [0, 5, 10, 15, 20]
    assert sep[-1, 2] < sep[0, 2]  # the galaxies approach
    en = np.loadtxt(base + "_energy.csv", delimiter=",", skiprows=1)
    assert en[0, 3] == 0.0 and np.all(en[:, 3] < 1e-2)
    assert out["max_rel_energy_drift"] < 1e-2
