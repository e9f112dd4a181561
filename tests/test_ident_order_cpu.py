"""The host identification prepares every marker's refinement before its
cut selection has ended, with the conditioner scale (a+b)/2 of the frame's
first marker that has cuts to select; the reference takes it from the first
marker refined (conditioner.hpp:22-46).  When that first marker's selection
turns out empty, the frame is refined again with the right marker's scale,
so a marker's result cannot depend on a failing marker before it."""
import numpy as np

from cctag_amd import synth
from tests import frames, oracle_lib


def _outer_points(e, n=150, mag=100.0):
    """n points on the ellipse (cx, cy, a, b, angle) with outward normals."""
    cx, cy, a, b, ang = e
    t = 2 * np.pi * np.arange(n) / n
    ca, sa = np.cos(ang), np.sin(ang)
    x = cx + a * np.cos(t) * ca - b * np.sin(t) * sa
    y = cy + a * np.cos(t) * sa + b * np.sin(t) * ca
    u, v = (x - cx) * ca + (y - cy) * sa, -(x - cx) * sa + (y - cy) * ca
    gu, gv = u / a ** 2, v / b ** 2
    gx, gy = gu * ca - gv * sa, gu * sa + gv * ca
    nrm = np.hypot(gx, gy)
    return np.column_stack([x, y, mag * gx / nrm, mag * gy / nrm])


def test_marker_results_do_not_depend_on_a_failing_first_marker():
    img, gts = frames.config_frame(1)
    real = []
    for g in gts:
        m = synth.refine_marker(g.H, 22)
        real.append({"ellipse": m["ellipse"], "points": _outer_points(m["ellipse"]),
                     "center0": m["center0"]})
    # cut stops 2 px from the left border with gradients along x: every cut
    # is collected, every refinement cut across the stop leaves the image,
    # so nothing is selected
    ys = np.linspace(280.0, 320.0, 12)
    bad = {"ellipse": (6.0, 300.0, 40.0, 40.0, 0.0),
           "points": np.column_stack([np.full(12, 3.0), ys, np.full(12, -100.0), np.zeros(12)]),
           "center0": (6.0, 300.0)}
    alone = oracle_lib.identify_given(img, real)
    after_bad = oracle_lib.identify_given(img, [bad] + real)
    before_bad = oracle_lib.identify_given(img, real + [bad])
    assert after_bad[0][0] == -2 and before_bad[0][-1] == -2   # no selected cuts
    assert (alone[0] == 1).sum() >= 3                          # the real markers are identified
    for got in ((after_bad[0][1:], after_bad[1][1:], after_bad[2][1:]),
                (before_bad[0][:-1], before_bad[1][:-1], before_bad[2][:-1])):
        np.testing.assert_array_equal(got[0], alone[0])
        np.testing.assert_array_equal(got[1], alone[1])
        np.testing.assert_array_equal(got[2], alone[2])
