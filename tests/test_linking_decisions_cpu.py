"""The candidate stage's edge linking takes its two angle decisions -- the
first neighbour direction (a sector of the gradient angle) and the convexity
test on the sine and cosine of an angle difference -- from gradient cross and
dot products, and evaluates the reference's float expressions on
fmod(atan2(gy, gx) + 2pi, 2pi) (Vote.cpp:296-340) only within 1e-4 rad of a
decision boundary.  Checked against those expressions on random integer
gradients of every magnitude and on gradients placed next to every boundary
(sector edges (2k+1)pi/8, angle differences 0, pi, +-pi/2 and those whose
sine is -0.707)."""
import numpy as np

from tests import oracle_lib


def _near(rng, angles, n):
    """Integer gradients whose direction is within ~2e-4 rad of `angles`."""
    a = rng.choice(angles, n) + rng.uniform(-2e-4, 2e-4, n)
    mag = rng.integers(1, 32768, n)
    return np.rint(mag * np.cos(a)), np.rint(mag * np.sin(a))


def test_linking_decisions_match_float_angle_expressions():
    rng = np.random.default_rng(2024)
    n = 400000
    cases = []
    for lim in (3, 40, 1000, 32767):  # random gradients, zero included
        g = rng.integers(-lim, lim + 1, (n, 4)).astype(np.float64)
        cases.append(np.column_stack([g, rng.choice([-1.0, 1.0], n)]))
    # front next to a sector boundary
    fx, fy = _near(rng, (2 * np.arange(8) + 1) * np.pi / 8, n)
    g = rng.integers(-32767, 32768, (n, 2)).astype(np.float64)
    cases.append(np.column_stack([fx, fy, g, rng.choice([-1.0, 1.0], n)]))
    # back next to front turned by a boundary of the convexity test
    f = rng.integers(-32767, 32768, (n, 2)).astype(np.float64)
    base = np.arctan2(f[:, 1], f[:, 0])
    s = np.arcsin(0.707)
    offs = np.array([0.0, np.pi, -np.pi, np.pi / 2, -np.pi / 2, s, -s, np.pi - s, s - np.pi])
    a = base + rng.choice(offs, n) + rng.uniform(-2e-4, 2e-4, n)
    mag = rng.integers(1, 32768, n)
    cases.append(np.column_stack([f, np.rint(mag * np.cos(a)), np.rint(mag * np.sin(a)),
                                  rng.choice([-1.0, 1.0], n)]))
    cases = np.concatenate(cases).astype(np.float32)
    assert oracle_lib.linking_mismatches(cases) == 0
