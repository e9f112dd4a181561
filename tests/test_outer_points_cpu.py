"""The order in which identification takes a marker's outer points
(Identification.cpp:1192-1210: sorted by atan2 angle about the ellipse
centre, then every step-th): the library's radix sort of the angles against
the oracle's sort of (angle, index) pairs, on point sets with many equal
angles (points on rays from the centre), points on the axes through the
centre, dense level-0 rings, and sizes from 1 point to beyond the
requested count."""
import ctypes

import numpy as np
import pytest

from tests import oracle_lib


def mismatch(x, y, cx, cy, requested=150):
    o = oracle_lib.load()
    f = o.oracle_outer_points_mismatch
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                  ctypes.c_int]
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return f(x.ctypes.data, y.ctypes.data, len(x), cx, cy, requested)


def ring(rng, n, cx, cy, a, b, noise):
    t = rng.uniform(0, 2 * np.pi, n)
    return (np.round(cx + a * np.cos(t) + rng.normal(0, noise, n)),
            np.round(cy + b * np.sin(t) + rng.normal(0, noise, n)))


@pytest.mark.parametrize("n", [0, 1, 2, 5, 149, 150, 151, 400, 3000, 9000])
def test_rings(n):
    rng = np.random.Generator(np.random.PCG64(n))
    x, y = ring(rng, n, 640.0, 360.0, 180.0, 120.0, 1.5)
    assert mismatch(x, y, 640.25, 359.75) == 0
    assert mismatch(x, y, 640.0, 360.0, requested=37) == 0


def test_equal_angles_and_axes():
    """Points on a few rays from the centre (equal angles, ties broken by
    index) and on the axes through it (angles 0, pi/2, pi, -pi/2)."""
    rng = np.random.Generator(np.random.PCG64(4))
    cx, cy = 300.0, 200.0
    rays = rng.integers(-5, 6, size=(12, 2))
    rays = rays[np.any(rays != 0, axis=1)]
    pts = []
    for _ in range(800):
        dx, dy = rays[rng.integers(len(rays))]
        k = rng.integers(1, 20)
        pts.append((cx + k * dx, cy + k * dy))
    for k in range(1, 30):
        pts += [(cx + k, cy), (cx - k, cy), (cx, cy + k), (cx, cy - k)]
    x, y = np.array(pts, np.float64).T
    perm = rng.permutation(len(x))
    assert mismatch(x[perm], y[perm], cx, cy) == 0
    assert mismatch(x[perm], y[perm], cx, cy, requested=len(x) + 5) == 0
    # the input reversed: equal angles now meet in the other index order
    assert mismatch(x[perm][::-1].copy(), y[perm][::-1].copy(), cx, cy,
                    requested=len(x)) == 0
