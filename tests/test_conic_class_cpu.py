"""The candidate stage's RANSAC trials (Vote.cpp:486-503) classify each fitted
conic from its invariants instead of deriving the ellipse in float: every
outcome the fast test decides must be the float derivation's, near the
ratio bounds, near a zero centre value and on ordinary five-point fits."""
import numpy as np

from tests import oracle_lib


def _five_point_fits(rng, n):
    out = []
    while len(out) < n:
        cx, cy = rng.uniform(100, 1800), rng.uniform(100, 1000)
        a, b = rng.uniform(5, 300), rng.uniform(5, 300)
        phi = rng.uniform(0, np.pi)
        th = rng.uniform(0, 2 * np.pi, 5)
        x = cx + a * np.cos(th) * np.cos(phi) - b * np.sin(th) * np.sin(phi)
        y = cy + a * np.cos(th) * np.sin(phi) + b * np.sin(th) * np.cos(phi)
        x = np.round(x + rng.normal(0, rng.choice([0.0, 0.7, 5.0]), 5))
        y = np.round(y + rng.normal(0, rng.choice([0.0, 0.7, 5.0]), 5))
        A = np.stack([x * x, 2 * x * y, y * y, 2 * x, 2 * y], 1)
        try:
            t = np.linalg.solve(A, -np.ones(5)).astype(np.float32)
        except np.linalg.LinAlgError:
            continue
        if np.float32(t[0] * t[2]) - np.float32(t[1] * t[1]) > 0:
            out.append(t)
    return np.array(out, np.float32)


def _conics(rng, lam1, lam2, phi, c):
    """x^T M x + 2 u^T x + 1 with M = R diag(lam1, lam2) R^T, centre c."""
    co, si = np.cos(phi), np.sin(phi)
    m00 = lam1 * co * co + lam2 * si * si
    m11 = lam1 * si * si + lam2 * co * co
    m01 = (lam1 - lam2) * co * si
    u0 = -(m00 * c[:, 0] + m01 * c[:, 1])
    u1 = -(m01 * c[:, 0] + m11 * c[:, 1])
    t = np.stack([m00, m01, m11, u0, u1], 1).astype(np.float32)
    keep = t[:, 0] * t[:, 2] - t[:, 1] * t[:, 1] > 0
    return t[keep]


def _near_bounds(rng, n, lo, hi):
    phi = rng.uniform(0, np.pi, n)
    c = np.stack([rng.uniform(10, 1900, n), rng.uniform(10, 1000, n)], 1)
    rel = rng.choice([-1, 1], n) * 10.0 ** rng.uniform(-9, -1, n)
    sign = rng.choice([-1.0, 1.0], n)
    # ratio^2 = lam_small / lam_large at lo^2 (1 + rel), or hi^2 for negative M
    r = np.where(sign > 0, lo * lo, 1.0 / (hi * hi)) * (1 + rel)
    big = 10.0 ** rng.uniform(-7, -3, n)
    lam1, lam2 = sign * big, sign * big * r
    # scale so that the centre value 1 - c^T M c sits near its sign change
    cmc = lam1 * (c[:, 0] * np.cos(phi) + c[:, 1] * np.sin(phi)) ** 2 + \
        lam2 * (-c[:, 0] * np.sin(phi) + c[:, 1] * np.cos(phi)) ** 2
    s = (1 + rng.choice([-1, 1], n) * 10.0 ** rng.uniform(-9, 0, n)) / np.abs(cmc)
    return _conics(rng, lam1 * s, lam2 * s, phi, c)


def test_conic_class_matches_float_derivation():
    rng = np.random.default_rng(7)
    fits = _five_point_fits(rng, 20000)
    bad, open_ = oracle_lib.conic_class_mismatches(fits)
    assert bad == 0
    # about a fifth are left to the float derivation: ellipses far from the
    # origin, whose centre value is small beside the constant term 1, so that
    # the float path's own cancellation may decide its sign
    assert open_ < 0.25 * len(fits)
    edge = _near_bounds(rng, 200000, 0.04, 25.0)
    bad, open_ = oracle_lib.conic_class_mismatches(edge)
    assert bad == 0
    # the outer-ellipse test's bounds of the segment merge (Vote.cpp:640)
    edge = _near_bounds(rng, 50000, 0.12, 8.0)
    bad, _ = oracle_lib.conic_class_mismatches(edge, 0.12, 8.0)
    assert bad == 0
