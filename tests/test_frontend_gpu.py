"""Front-end parity: HIP kernels through the C ABI against the oracle, bit
for bit, at every pyramid level (pyramid, derivatives, Canny, thinning, edge
list, descent links, votes, voter lists, flow lengths, seeds)."""
import numpy as np
import pytest

from cctag_amd.detector import Detector
from tests import frames, oracle_lib

KEYS_EXACT = ["img", "dx", "dy", "edges", "edge_map", "edge_x", "edge_y", "edge_dx",
              "edge_dy", "before", "after", "chosen", "voters_index", "voters_list",
              "is_max", "seeds"]
KEYS_FLOAT = ["total_dist", "flow_length"]


def compare_levels(got, ref, where):
    assert got["width"] == ref["width"] and got["height"] == ref["height"], where
    assert got["shape"] == ref["shape"], where
    assert got["n_edges"] == ref["n_edges"], (where, got["n_edges"], ref["n_edges"])
    for k in KEYS_EXACT:
        a, b = got[k], ref[k]
        assert a.shape == b.shape, (where, k, a.shape, b.shape)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError("%s %s differs at %d places, first %s: got %s ref %s" % (
                where, k, len(bad), bad[0], a[tuple(bad[0])], b[tuple(bad[0])]))
    for k in KEYS_FLOAT:  # bit-exact floats
        assert np.array_equal(got[k].view(np.int32), ref[k].view(np.int32)), (where, k)


def run_parity(img, n_levels=4):
    ref = oracle_lib.front_end(img, n_levels)
    h, w = img.shape
    with Detector(w, h) as det:
        det.front_end(img)
        for l in range(n_levels):
            compare_levels(det.level(l), ref[l], "level %d" % l)
    return ref


@pytest.mark.gpu
@pytest.mark.parametrize("config,k", [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1)])
def test_front_end_parity_configs(config, k):
    img, _ = frames.config_frame(config, k)
    print("front-end parity config", config, "frame", k, img.shape, flush=True)
    ref = run_parity(img)
    assert ref[0]["n_edges"] > 1000 and len(ref[0]["seeds"]) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("w,h", frames.ODD_SIZES)
def test_front_end_parity_floor_halved_pyramid(w, h):
    """Sizes whose levels are floor halves (ImagePyramid.cpp:23-33), not exact
    halves: every level bit for bit, resampling included."""
    img, _ = frames.odd_size_frame(w, h)
    print("front-end parity floor-halved", img.shape, flush=True)
    ref = run_parity(img)
    assert [r["width"] for r in ref] == [w >> l for l in range(4)]
    assert [r["height"] for r in ref] == [h >> l for l in range(4)]
    assert ref[0]["n_edges"] > 1000


@pytest.mark.gpu
def test_front_end_parity_small_markers():
    img, _ = frames.small_marker_frame()
    run_parity(img)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["blank", "noise", "checker", "rings", "tiny", "spiral", "stripes"])
def test_front_end_parity_edge_cases(name):
    img = frames.edge_case(name)
    print("front-end parity edge case", name, img.shape, flush=True)
    n_levels = 1 if name == "tiny" else 4
    if name == "tiny":
        n_levels = 1
    ref = oracle_lib.front_end(img, n_levels)
    h, w = img.shape
    from cctag_amd import _lib
    prm = _lib.default_params(3)
    prm.number_of_processed_multires_layers = n_levels
    with Detector(w, h, params=prm) as det:
        det.front_end(img)
        for l in range(n_levels):
            compare_levels(det.level(l), ref[l], "%s level %d" % (name, l))
    if name == "blank":
        assert ref[0]["n_edges"] == 0


@pytest.mark.gpu
def test_front_end_repeatable():
    """Two runs on one handle give identical results (union-find and atomics
    leave no trace in the outputs)."""
    img, _ = frames.config_frame(1)
    h, w = img.shape
    with Detector(w, h) as det:
        det.front_end(img)
        a = [det.level(l) for l in range(4)]
        det.front_end(img)
        for l in range(4):
            compare_levels(det.level(l), a[l], "rerun level %d" % l)


@pytest.mark.gpu
def test_front_end_parity_four_crowns():
    """The vote walks four crowns (nCrowns 4) on a four-ring marker frame."""
    from cctag_amd import _lib, synth
    img, _ = synth.four_crown_frame(0)
    prm = _lib.default_params(4)
    ref = oracle_lib.front_end_params(img, prm)
    h, w = img.shape
    with Detector(w, h, params=prm) as det:
        det.front_end(img)
        for l in range(4):
            compare_levels(det.level(l), ref[l], "four crowns level %d" % l)
    assert len(ref[0]["seeds"]) > 0
