"""Refinement test markers built from the synthetic frames' ground truth
(cctag_amd.synth.refine_marker): the outer ellipse, cut stops on it, the
ellipse centre as the start and the true imaged centre."""
from __future__ import annotations

from cctag_amd import synth


def markers_of(gts, w, h, n_cuts: int = 22):
    out = []
    for g in gts:
        m = synth.refine_marker(g.H, n_cuts)
        if all(0 <= x < w - 1 and 0 <= y < h - 1 for x, y in m["stops"]):
            out.append(m)
    return out
