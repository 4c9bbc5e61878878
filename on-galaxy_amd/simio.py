"""Simulation outputs in the reference's file formats (gpu/galaxy_sim.cu:180-225),
read by its renderer and merger gates (python/render_galaxy.py:33-52,
python/core_sep.py:17-29):

  <base>_frames.bin   int32 N, int32 nFrames, then nFrames * N * 3 float32 (x, y, z per body)
  <base>_species.bin  N int8 (galaxy index of each body)
  <base>_sep.csv      "step,time,sep": COM separation of species 0 and 1
  <base>_energy.csv   "step,time,total,relDrift"

Numbers in the CSV files use C++ ostream's default formatting (6 significant
digits, like printf %g), as the reference writes them.
"""
import struct

import numpy as np


def fmt(x):
    """A number as `std::ostream << double` prints it by default."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    return "%g" % float(x)


class FrameWriter:
    """<base>_frames.bin: header (N, nFrames) then one float32 frame per call."""

    def __init__(self, base, n, n_frames):
        self.f = open(base + "_frames.bin", "wb")
        self.n, self.n_frames, self.written = int(n), int(n_frames), 0
        self.f.write(struct.pack("<ii", self.n, self.n_frames))

    def write(self, pos):
        pos = np.asarray(pos, np.float64).reshape(self.n, 3)
        self.f.write(pos.astype(np.float32).tobytes())
        self.written += 1

    def close(self):
        if self.f:
            self.f.close()
            self.f = None


def write_species(base, species):
    np.asarray(species, np.int8).tofile(base + "_species.bin")


def read_frames(base):
    """(N, nFrames, frames[nFrames, N, 3] float32) as the reference's readers parse them."""
    with open(base + "_frames.bin", "rb") as f:
        n, nf = struct.unpack("<ii", f.read(8))
        data = np.frombuffer(f.read(), dtype=np.float32)
    return n, nf, data.reshape(-1, n, 3)


def com_separation(pos, mass, species):
    """|COM(species 0) - COM(species 1)| (galaxy_sim.cu:209-217); 0 if a species is empty."""
    pos = np.asarray(pos)
    mass = np.asarray(mass)
    s0, s1 = species == 0, species != 0
    m0, m1 = mass[s0].sum(), mass[s1].sum()
    if m0 <= 0 or m1 <= 0:
        return 0.0
    c0 = (mass[s0, None] * pos[s0]).sum(0) / m0
    c1 = (mass[s1, None] * pos[s1]).sum(0) / m1
    return float(np.linalg.norm(c0 - c1))


class CsvWriter:
    def __init__(self, path, header):
        self.f = open(path, "w")
        self.f.write(header + "\n")

    def row(self, *vals):
        self.f.write(",".join(fmt(v) for v in vals) + "\n")

    def close(self):
        if self.f:
            self.f.close()
            self.f = None
