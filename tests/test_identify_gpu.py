"""Identification of given markers through the C ABI (cctagx_identify: GPU
cut selection, grid search and read-out, host control) against the oracle's
own serial identification on the same inputs: identical statuses and IDs,
the same selected cuts with stops within 1e-5 px, imaged centres within
0.05 px and the normalised residual within 1e-4 relative.  Also the
drop-in's checks: frames of another size, custom marker banks."""
import math

import numpy as np
import pytest

from cctag_amd import _lib
from cctag_amd.detector import Detector
from tests import frames, golden_tables, oracle_lib

CENTER_TOL_PX = 0.05   # imaged centre, GPU vs oracle
STOP_TOL_PX = 1e-5     # selected cut stops (identify_step_1)
RESIDUAL_RTOL = 1e-4   # normalised residual = 1 / quality, at the same grid point
# The grid search compares residuals that the reference sums in float and
# the kernels in double (the two sums differ by <= 1e-5 relative); where
# two grid points of an iteration tie to that precision the two searches
# can end on neighbouring points of the last 5x5 grid (<= 0.02 px apart),
# whose residuals differ by up to a few 1e-4.
RESIDUAL_RTOL_NEIGHBOUR = 2e-3
SAME_POINT_PX = 1e-3


def compare(got, ref, where):
    assert len(got) == len(ref)
    for k, (g, r) in enumerate(zip(got, ref)):
        w = (where, k)
        assert g["status"] == r["status"] and g["id"] == r["id"], (w, g["status"], r["status"],
                                                                   g["id"], r["id"])
        assert g["stops"].shape == r["stops"].shape, (w, g["stops"].shape, r["stops"].shape)
        if len(r["stops"]):
            err = np.abs(g["stops"] - r["stops"]).max()
            assert err <= STOP_TOL_PX, (w, "stops", err)
        if r["status"] in (1, -4, -5):
            d = math.hypot(g["x"] - r["x"], g["y"] - r["y"])
            assert d <= CENTER_TOL_PX, (w, "centre", d)
            rel = abs(1 / g["quality"] - 1 / r["quality"]) * r["quality"]
            tol = RESIDUAL_RTOL if d < SAME_POINT_PX else RESIDUAL_RTOL_NEIGHBOUR
            assert rel <= tol, (w, "residual", rel, "centres", d, "px apart")


def identify_both(img, markers, where):
    print("identify", where, img.shape, len(markers), "markers", flush=True)
    ref = oracle_lib.identify_given(img, markers)
    h, w = img.shape
    with Detector(w, h) as det:
        det.front_end(img)
        got = det.identify(markers)
    compare(got, ref, where)
    return got, ref


@pytest.mark.gpu
@pytest.mark.parametrize("config,k", [(1, 0), (1, 1), (2, 0), (2, 2)])
def test_identify_ground_truth_markers(config, k):
    """Every marker of a configuration frame, from its exact outer ellipse."""
    img, gts = frames.config_frame(config, k)
    markers = [frames.ident_input(g.H) for g in gts]
    got, _ = identify_both(img, markers, "config %d frame %d" % (config, k))
    ok = [(g, m) for g, m in zip(got, gts) if g["status"] == 1]
    assert len(ok) >= len(gts) - 2
    assert all(g["id"] == m.id for g, m in ok)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [7000, 7003, 7005])
def test_identify_border_markers(seed):
    """Markers cut by the frame border: cuts leaving the image are left out
    of the vote, never the whole marker (Identification.cpp:78)."""
    img, gts = frames.border_frame(seed)
    markers = [frames.ident_input(g.H) for g in gts]
    got, _ = identify_both(img, markers, "border %d" % seed)
    for g, m in zip(got, gts):
        assert g["status"] != 1 or g["id"] == m.id, (g["status"], g["id"], m.id)


@pytest.mark.gpu
def test_identify_clean_and_faded_markers():
    for k in range(4):
        img, _, m = frames.clean_marker(100 + k)
        identify_both(img, [m], "clean %d" % k)
    img, mid, m = frames.clean_marker(200, radius=(70, 70), tilt=10.0)
    got, _ = identify_both(frames.faded_sector(img, m["ellipse"]), [m], "faded sector")
    assert got[0]["status"] == 1 and got[0]["id"] == mid and len(got[0]["stops"]) == 22


@pytest.mark.gpu
def test_identify_does_not_depend_on_a_failing_first_marker():
    """The conditioning scale comes from the first marker refined: a marker
    whose selection fails, put first, changes nothing for the others."""
    img, gts = frames.config_frame(1)
    real = [frames.ident_input(g.H) for g in gts]
    bad = frames.failing_marker()
    alone, _ = identify_both(img, real, "alone")
    first, _ = identify_both(img, [bad] + real, "failing first")
    assert first[0]["status"] == -2
    for a, b in zip(alone, first[1:]):
        assert (a["status"], a["id"], a["x"], a["y"]) == (b["status"], b["id"], b["x"], b["y"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [7000, 7003])
def test_detect_border_frames_match_oracle(seed):
    from tests.test_detect_gpu import match
    img, gts = frames.border_frame(seed)
    print("detect border frame", seed, flush=True)
    ref = oracle_lib.detect(img, threads=8)
    h, w = img.shape
    with Detector(w, h) as det:
        got = det.detect(img)
    match(got, ref)
    for m in got:
        assert m["status"] != 1 or 0 <= m["id"] < 32, m
    assert sum(m["status"] == 1 for m in got) >= len(gts) - 1


@pytest.mark.gpu
def test_frame_of_another_size_is_rejected():
    """The reference stops on a resolution change (Detection.cpp:753-757);
    the drop-in returns CCTAGX_ERR_ARG and the handle stays usable."""
    img, _ = frames.small_marker_frame()
    h, w = img.shape
    with Detector(w, h) as det:
        a = det.detect(img)
        for bad in (img[:, :-2], img[:-2, :], np.zeros((h * 2, w * 2), np.uint8)):
            with pytest.raises(_lib.CCTagError) as e:
                det.detect(bad)
            assert e.value.rc == -1
            with pytest.raises(_lib.CCTagError):
                det.front_end(bad)
        assert det.detect(img) == a


@pytest.mark.gpu
def test_custom_bank():
    """A bank given in place of the built-in one (cctagDetection's pBank):
    the rows reversed give every marker the mirrored ID; invalid banks are
    rejected and leave the handle's bank as it was."""
    img, gts = frames.config_frame(1)
    h, w = img.shape
    bank = np.asarray(golden_tables.tables()["bank3"], np.float32)
    with Detector(w, h) as det:
        a = det.detect(img)
        det.set_bank(bank[::-1])
        b = det.detect(img)
        for bad, rc in ((np.ones((129, 5), np.float32), -6), (np.ones((32, 9), np.float32), -6),
                        (-bank, -1), (np.full((32, 5), np.nan, np.float32), -1)):
            with pytest.raises(_lib.CCTagError) as e:
                det.set_bank(bad)
            assert e.value.rc == rc
        c = det.detect(img)
        det.set_bank(bank)
        d = det.detect(img)
    assert d == a and c == b
    ka = sorted((m["status"], m["id"], round(m["x"], 3)) for m in a)
    kb = sorted((m["status"], 31 - m["id"] if m["id"] >= 0 else m["id"], round(m["x"], 3))
                for m in b)
    assert ka == kb
    assert sum(m["status"] == 1 for m in a) >= 3
