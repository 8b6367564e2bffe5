"""Writes c4_sphere_markers.npy: the 1,810 Fibonacci markers of the 24-cell sphere at
(96, 64, 64) with seeded N(0, 0.05 cell) noise, as luma_amd.cases.c4_markers generates them."""
from pathlib import Path

import numpy as np

from luma_amd.cases import c4_markers

if __name__ == "__main__":
    np.save(Path(__file__).resolve().parent / "c4_sphere_markers.npy", c4_markers())
