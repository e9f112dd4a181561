"""The reference's file formats through the C ABI (host code, no device):
the marker bank text (CCTagMarkersBank.cpp:60-78, pinned by the reference's
own bank files markersToPrint/generators/{3,4}Crowns/ids.txt, which must give
the built-in tables of CCTagMarkersBank.cpp:120-282 bit for bit), the
Parameters XML archive (Params.hpp:210-248) and the regression log with its
comparison rules (regression/TestLog.h, Regression.cpp:114-170)."""
import os

import numpy as np
import pytest

from cctag_amd import _lib, regression
from tests import golden_tables

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("n", [3, 4])
def test_reference_bank_files_give_the_built_in_banks(n):
    got = _lib.bank_from_text(path=os.path.join(HERE, "golden", "bank%d_ids.txt" % n))
    ref = np.asarray(golden_tables.tables()["bank%d" % n], np.float32)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))


def test_bank_line_grammar():
    """Items are "p/q" ratios (white space allowed around the slash) or
    decimals; empty lines are skipped; a line ends at the first item that is
    neither, keeping what came before."""
    b = _lib.bank_from_text("100/50 1.5 3/4\n\n  2 / 4 0.25e1 7\n")
    assert b.tolist() == [[2.0, 1.5, 0.75], [0.5, 2.5, 7.0]]
    assert _lib.bank_from_text("1.5 2.5 x 3\n4 5").tolist() == [[1.5, 2.5], [4.0, 5.0]]
    with pytest.raises(_lib.CCTagError) as e:
        _lib.bank_from_text("1.5 2.5 x 3\n4 5 6")
    assert e.value.rc == -6   # rows of 2 and 3 ratios


def test_bank_errors():
    with pytest.raises(_lib.CCTagError) as e:
        _lib.bank_from_text(path="/nonexistent/bank.txt")
    assert e.value.rc == -1 and "Unable to open the bank file" in str(e.value)
    with pytest.raises(_lib.CCTagError) as e:
        _lib.bank_from_text("\n  \n")
    assert e.value.rc == -1


BOOST_ARCHIVE = """<?xml version="1.0" encoding="UTF-8" standalone="yes" ?>
<!DOCTYPE boost_serialization>
<boost_serialization signature="serialization::archive" version="17">
<CCTagsParams class_id="0" tracking_level="0" version="0">
\t<_cannyThrLow>0.0099999998</_cannyThrLow>
\t<_cannyThrHigh>0.039999999</_cannyThrHigh>
\t<_distSearch>30</_distSearch>
\t<_thrGradientMagInVote>2500</_thrGradientMagInVote>
\t<_angleVoting>0</_angleVoting>
\t<_ratioVoting>4</_ratioVoting>
\t<_averageVoteMin>0</_averageVoteMin>
\t<_thrMedianDistanceEllipse>3</_thrMedianDistanceEllipse>
\t<_maximumNbSeeds>500</_maximumNbSeeds>
\t<_maximumNbCandidatesLoopTwo>40</_maximumNbCandidatesLoopTwo>
\t<_nCrowns>4</_nCrowns>
\t<_minPointsSegmentCandidate>10</_minPointsSegmentCandidate>
\t<_minVotesToSelectCandidate>3</_minVotesToSelectCandidate>
\t<_threshRobustEstimationOfOuterEllipse>30</_threshRobustEstimationOfOuterEllipse>
\t<_ellipseGrowingEllipticHullWidth>2.29999995</_ellipseGrowingEllipticHullWidth>
\t<_windowSizeOnInnerEllipticSegment>20</_windowSizeOnInnerEllipticSegment>
\t<_numberOfMultiresLayers>4</_numberOfMultiresLayers>
\t<_numberOfProcessedMultiresLayers>3</_numberOfProcessedMultiresLayers>
\t<_nSamplesOuterEllipse>150</_nSamplesOuterEllipse>
\t<_numCutsInIdentStep>22</_numCutsInIdentStep>
\t<_numSamplesOuterEdgePointsRefinement>20</_numSamplesOuterEdgePointsRefinement>
\t<_cutsSelectionTrials>500</_cutsSelectionTrials>
\t<_sampleCutLength>100</_sampleCutLength>
\t<_imagedCenterNGridSample>5</_imagedCenterNGridSample>
\t<_imagedCenterNeighbourSize>0.200000003</_imagedCenterNeighbourSize>
\t<_minIdentProba>9.99999997e-07</_minIdentProba>
\t<_useLMDif>1</_useLMDif>
\t<_searchForAnotherSegment>0</_searchForAnotherSegment>
\t<_writeOutput>0</_writeOutput>
\t<_doIdentification>1</_doIdentification>
\t<_maxEdges>20000</_maxEdges>
\t<_useCuda>1</_useCuda>
\t<_pinnedCounters>100</_pinnedCounters>
\t<_pinnedNearbyPoints>7000</_pinnedNearbyPoints>
</CCTagsParams>
</boost_serialization>
"""


def test_parameter_archive():
    p = _lib.params_from_xml(BOOST_ARCHIVE)
    d = _lib.default_params(4)
    d.number_of_processed_multires_layers = 3
    d.search_for_another_segment = 0
    for f, _ in _lib.Params._fields_:
        assert getattr(p, f) == getattr(d, f), f
    # write then read back, every field
    q = _lib.default_params(3)
    q.canny_thr_low, q.ratio_voting, q.maximum_nb_seeds = 0.0123, 3.5, 321
    back = _lib.params_from_xml(_lib.params_to_xml(q))
    for f, _ in _lib.Params._fields_:
        assert getattr(back, f) == getattr(q, f), f


def test_parameter_archive_errors(tmp_path):
    with pytest.raises(_lib.CCTagError) as e:
        _lib.params_from_xml(BOOST_ARCHIVE.replace("\t<_distSearch>30</_distSearch>\n", ""))
    assert "_distSearch" in str(e.value)
    with pytest.raises(_lib.CCTagError) as e:
        _lib.params_from_xml(BOOST_ARCHIVE.replace(">2500<", ">abc<"))
    assert "_thrGradientMagInVote" in str(e.value)
    with pytest.raises(_lib.CCTagError) as e:
        _lib.params_from_xml(path=str(tmp_path / "missing.xml"))
    assert "is missing" in str(e.value)
    f = tmp_path / "p.xml"
    f.write_text(BOOST_ARCHIVE)
    assert _lib.params_from_xml(path=str(f)).n_crowns == 4


def _log(tags, name="frame.png", frames=1):
    return regression.FileLog(name, _lib.default_params(3),
                              [regression.FrameLog(k, 0.01 * (k + 1),
                                                   [regression.DetectedTag(*t) for t in tags])
                               for k in range(frames)])


def test_regression_log_round_trip_and_rules(tmp_path):
    tags = [(3, 1, 100.25, 50.5, 2.5), (7, 1, 10.0, 20.0, 3.0), (-1, -2, 5.0, 6.0, 0.0)]
    ref = _log(tags, frames=2)
    path = tmp_path / "ref.xml"
    ref.save(str(path))
    back = regression.FileLog.load(str(path))
    assert back.filename == ref.filename and len(back.frame_logs) == 2
    assert [(t.id, t.status, t.x, t.y, t.quality) for t in back.frame_logs[1].tags] == tags
    assert back.parameters.n_crowns == 3
    out = regression.compare(ref, back)
    assert len(out["quality_diff"]) == 4
    # tags in another order, a small shift: still equal
    regression.compare(ref, _log([tags[1], (3, 1, 100.6, 50.5, 2.4), tags[2]], frames=2))
    for bad, msg in (
            (_log([(3, 1, 101.0, 50.5, 2.5)] + tags[1:], frames=2), "different positions"),
            (_log([(4, 1, 100.25, 50.5, 2.5)] + tags[1:], frames=2), "different IDs"),
            (_log([(3, -4, 100.25, 50.5, 2.5)] + tags[1:], frames=2), "different status"),
            (_log(tags[:2], frames=2), "different # of tags"),
            (_log(tags, frames=1), "frame counts"),
            (_log(tags, name="other.png", frames=2), "filenames"),
            (_log([(7, 1, 100.25, 50.5, 2.5)] + tags[1:], frames=2), "duplicate IDs")):
        with pytest.raises(regression.CheckError) as e:
            regression.compare(ref, bad)
        assert msg in str(e.value), (msg, str(e.value))
