"""The marker generators of LUMA's prefab bodies (Body.h:144-255, restated in luma_amd/setups.py and
in the geometry reader, which the GPU tests compare with them) against what their geometry fixes:
counts, distances, spacings and angles.  CPU only."""
import math

import numpy as np
import pytest

from luma_amd import setups


@pytest.mark.parametrize("r,dh", [(0.5, 0.1), (4.0, 1.0), (0.3, 0.0125)])
def test_circle_markers(r, dh):
    c = (1.3, -0.7, 0.25)
    n = int(math.floor(2.0 * math.pi * r / dh))
    pos = setups.circle_markers(c, r, dh)
    # LUMA drops the first of n equally spaced angles (theta = 0): n - 1 markers
    assert len(pos) == n - 1
    d = pos - np.array(c)
    assert np.allclose(np.hypot(d[:, 0], d[:, 1]), r, rtol=0, atol=1e-12)
    assert np.all(d[:, 2] == 0.0)
    theta = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2.0 * math.pi)
    assert np.allclose(theta, 2.0 * math.pi * np.arange(1, n) / n, atol=1e-12)
    full = setups.circle_markers(c, r, dh, luma_skip_first=False)
    assert len(full) == n and np.allclose(full[0], (c[0] + r, c[1], c[2]), atol=1e-15)
    assert len(setups.circle_markers(c, r, dh, n=193)) == 192


@pytest.mark.parametrize("r,dh", [(2.5, 1.0), (12.0, 1.0), (0.2, 0.05)])
def test_sphere_markers(r, dh):
    c = (8.3, 8.1, 7.9)
    n = int(math.floor(4.0 * math.pi * r * r / (dh * dh)))
    pos = setups.sphere_markers(c, r, dh)
    assert len(pos) == n
    d = pos - np.array(c)
    assert np.allclose(np.linalg.norm(d, axis=1), r, rtol=0, atol=1e-12 * max(r, 1.0))
    # Fibonacci lattice: heights equally spaced in (-r, r), azimuth advancing by the golden angle
    off = 2.0 / n
    assert np.allclose(d[:, 1] / r, -1.0 + off / 2.0 + off * np.arange(n), atol=1e-6)
    phi = np.arctan2(d[:, 2], d[:, 0])
    golden = math.pi * (3.0 - math.sqrt(5.0))
    step = np.mod(np.diff(phi) - golden + math.pi, 2.0 * math.pi) - math.pi
    assert np.allclose(step, 0.0, atol=1e-9)
    # the lattice is balanced: its centroid is the centre to O(1/n)
    assert np.linalg.norm(d.mean(axis=0)) < 2.0 * r / n


@pytest.mark.parametrize("length,angle,dh", [(1.0, 30.0, 0.1), (2.0, 0.0, 0.05), (0.4, -60.0, 0.01)])
def test_filament_markers(length, angle, dh):
    start = (4.0, 1.0, 0.0)
    pos = setups.filament_markers(start, length, angle, dh)
    n = int(math.floor(length / dh)) + 1
    assert len(pos) == n
    a = math.radians(angle)
    assert np.allclose(pos[0], start, atol=1e-15)
    assert np.allclose(pos[-1], (start[0] + length * math.cos(a), start[1] + length * math.sin(a), 0.0), atol=1e-12)
    seg = np.linalg.norm(np.diff(pos, axis=0), axis=1)
    assert np.allclose(seg, length / (n - 1), rtol=0, atol=1e-12)
