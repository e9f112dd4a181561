"""The grid search conditions every marker with the (a+b)/2 of the first
ellipse conditionerFromEllipse sees (conditioner.hpp:22-46): the first
marker refined, so a marker whose cut selection fails does not set it, and
the other markers' results do not depend on whether it comes first."""
import numpy as np

from cctag_amd import synth
from tests import frames, oracle_lib


def test_marker_results_do_not_depend_on_a_failing_first_marker():
    img, gts = frames.config_frame(1)
    real = []
    for g in gts:
        m = synth.refine_marker(g.H, 22)
        real.append({"ellipse": m["ellipse"], "points": frames.outer_points(m["ellipse"]),
                     "center0": m["center0"]})
    bad = frames.failing_marker()
    alone = oracle_lib.identify_given(img, real)
    after_bad = oracle_lib.identify_given(img, [bad] + real)
    before_bad = oracle_lib.identify_given(img, real + [bad])
    assert after_bad[0]["status"] == -2 and before_bad[-1]["status"] == -2  # no selected cuts
    assert sum(r["status"] == 1 for r in alone) >= 3   # the real markers are identified
    for got in (after_bad[1:], before_bad[:-1]):
        for g, a in zip(got, alone):
            assert (g["status"], g["id"], g["x"], g["y"]) == (a["status"], a["id"], a["x"], a["y"])
