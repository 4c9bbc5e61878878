"""The reference's output formats (gpu/galaxy_sim.cu:180-225; readers
python/core_sep.py:17-29, python/render_galaxy.py:33-52): frame file layout,
species file, CSV number formatting as C++ ostream prints doubles, and the
COM separation."""
import importlib.util
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def simio():
    spec = importlib.util.spec_from_file_location("simio", os.path.join(ROOT, "on-galaxy_amd", "simio.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_frames_and_species_layout(tmp_path):
    s = simio()
    base = str(tmp_path / "run")
    rng = np.random.default_rng(0)
    frames = [rng.normal(size=(37, 3)) for _ in range(3)]
    w = s.FrameWriter(base, 37, 3)
    for f in frames:
        w.write(f)
    w.close()
    s.write_species(base, np.arange(37) % 2)
    raw = open(base + "_frames.bin", "rb").read()
    assert len(raw) == 8 + 3 * 37 * 3 * 4
    assert struct.unpack("<ii", raw[:8]) == (37, 3)
    # read the way core_sep.py does: frame by frame, N*3 float32 each
    for k, f in enumerate(frames):
        got = np.frombuffer(raw[8 + k * 37 * 12: 8 + (k + 1) * 37 * 12], dtype=np.float32).reshape(37, 3)
        assert np.array_equal(got, f.astype(np.float32))
    sp = np.frombuffer(open(base + "_species.bin", "rb").read(), dtype=np.int8)
    assert np.array_equal(sp, np.arange(37) % 2)
    n, nf, fr = s.read_frames(base)
    assert (n, nf) == (37, 3) and np.array_equal(fr[2], frames[2].astype(np.float32))


def test_csv_numbers_match_cpp_ostream(tmp_path):
    vals = [0.0, 0.01, 0.5, 1.0, 100.0, 1234567.0, -0.2924234708532302, 3.0e-9, 7.071067811865476, 1e300]
    src = tmp_path / "p.cpp"
    src.write_text("#include <iostream>\nint main(){double v[]={" + ",".join(repr(v) for v in vals) +
                   "};for(double x: v) std::cout << x << '\\n';}\n")
    exe = tmp_path / "p"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    ref = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    s = simio()
    assert [s.fmt(v) for v in vals] == ref
    assert s.fmt(5) == "5"
    w = s.CsvWriter(str(tmp_path / "x_sep.csv"), "step,time,sep")
    w.row(5, 5 * 0.01, 7.25)
    w.close()
    assert open(tmp_path / "x_sep.csv").read() == "step,time,sep\n5,0.05,7.25\n"


def test_com_separation():
    s = simio()
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0], [10.0, 1, 0], [12.0, 1, 0]])
    m = np.array([1.0, 3.0, 2.0, 2.0])
    sp = np.array([0, 0, 1, 1], np.int8)
    # COM0 = (1.5, 0, 0), COM1 = (11, 1, 0)
    assert s.com_separation(pos, m, sp) == pytest.approx(np.hypot(9.5, 1.0), rel=1e-15)
    assert s.com_separation(pos, m, np.zeros(4, np.int8)) == 0.0
