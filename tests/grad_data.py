"""Refinement test markers built from the synthetic frames' ground truth:
the outer ellipse (image of the unit circle), cut stops on it and the true
imaged centre, all in detector coordinates (pixel centres at integers)."""
from __future__ import annotations

import math

import numpy as np

SHIFT = np.array([[1.0, 0.0, -0.5], [0.0, 1.0, -0.5], [0.0, 0.0, 1.0]])


def ellipse_of(H: np.ndarray):
    """(cx, cy, a, b, angle) of the image of the unit circle under H."""
    Hi = np.linalg.inv(H)
    C = Hi.T @ np.diag([1.0, 1.0, -1.0]) @ Hi
    A = C[:2, :2]
    ctr = np.linalg.solve(A, -C[:2, 2])
    f = C[2, 2] + C[:2, 2] @ ctr
    lam, vec = np.linalg.eigh(A)
    if f > 0:
        lam, f = -lam, -f
    ax = np.sqrt(-f / lam)
    i = int(np.argmax(ax))
    angle = math.atan2(vec[1, i], vec[0, i])
    return float(ctr[0]), float(ctr[1]), float(ax[i]), float(ax[1 - i]), angle


def marker_from_gt(H_synth: np.ndarray, n_cuts: int = 22, phase: float = 0.1):
    H = SHIFT @ np.asarray(H_synth, np.float64)
    th = phase + 2 * np.pi * np.arange(n_cuts) / n_cuts
    pts = H @ np.stack([np.cos(th), np.sin(th), np.ones_like(th)])
    stops = (pts[:2] / pts[2]).T
    e = ellipse_of(H)
    c = H @ np.array([0.0, 0.0, 1.0])
    return {"ellipse": e, "center0": (e[0], e[1]), "stops": stops,
            "center_gt": (c[0] / c[2], c[1] / c[2]), "H": H / H[2, 2]}


def markers_of(gts, w, h, n_cuts: int = 22):
    out = []
    for g in gts:
        m = marker_from_gt(g.H, n_cuts)
        if all(0 <= x < w - 1 and 0 <= y < h - 1 for x, y in m["stops"]):
            out.append(m)
    return out
