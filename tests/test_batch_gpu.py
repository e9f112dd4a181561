"""Batched detection through the C ABI: every frame of an uneven-size batch
gets exactly the markers the plain path finds for it alone, under every
work partition the autotuner can choose, and graph replays repeat."""
import pytest

from cctag_amd.detector import BatchDetector, Detector
from tests import frames, oracle_lib
from tests.test_detect_gpu import match


def plain(imgs):
    out = []
    for img in imgs:
        h, w = img.shape
        with Detector(w, h) as det:
            out.append(det.detect(img))
    return out


@pytest.mark.gpu
def test_batch_matches_plain_per_frame():
    imgs = frames.small_batch()
    ref = plain(imgs)
    print("batch of", [im.shape for im in imgs], flush=True)
    with BatchDetector([(im.shape[1], im.shape[0]) for im in imgs]) as b:
        for f, im in enumerate(imgs):
            b.load(f, im)
        got = b.detect()
        st = b.stats()
        print("stats", st, flush=True)
        assert st["graph"] == 1 and st["tune_us"] > 0
        again = b.detect()
        assert b.stats()["tune_us"] == 0
    assert got == ref
    assert again == got
    assert sum(len(r) for r in ref) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("fv,eb", [(0, 64), (1, 512), (2, 1024), (2, 128), (1, 2048)])
def test_batch_partitions_agree(fv, eb):
    imgs = frames.small_batch()
    ref = plain(imgs)
    with BatchDetector([(im.shape[1], im.shape[0]) for im in imgs]) as b:
        for f, im in enumerate(imgs):
            b.load(f, im)
        b.set_partition(fv, eb)
        print("partition", fv, eb, flush=True)
        assert b.detect() == ref


@pytest.mark.gpu
def test_batch_gradient_refinement_matches_plain_gradient_refinement():
    """Batched detection with the gradient refinement gives every frame what
    the plain path gives it alone with the same refinement."""
    imgs = frames.small_batch()
    ref = []
    for img in imgs:
        h, w = img.shape
        with Detector(w, h) as det:
            det.set_refinement(20)
            ref.append(det.detect(img))
    with BatchDetector([(im.shape[1], im.shape[0]) for im in imgs]) as b:
        b.set_refinement(20)
        for f, im in enumerate(imgs):
            b.load(f, im)
        got = b.detect()
    assert got == ref


# frames of the 64-frame uneven-size batch: 640x480 (2, 3), 1280x720 (6),
# 1920x1080 (7, 8) and 3840x2160 (0, 5)
UNEVEN = [2, 7, 0, 6, 8, 3, 5]


@pytest.mark.gpu
def test_uneven_batch_matches_oracle_per_frame():
    """Frames of the uneven-size batch, up to 3840x2160, detected in one
    batch: each frame's markers against the oracle's for that frame alone."""
    imgs = [frames.batch_frame(k)[0] for k in UNEVEN]
    print("uneven batch", [im.shape for im in imgs], flush=True)
    with BatchDetector([(im.shape[1], im.shape[0]) for im in imgs]) as b:
        for f, im in enumerate(imgs):
            b.load(f, im)
        got = b.detect()
    for f, (k, im) in enumerate(zip(UNEVEN, imgs)):
        print("oracle on batch frame", k, im.shape, flush=True)
        match(got[f], oracle_lib.detect(im, threads=8))
    assert sum(m["status"] == 1 for ms in got for m in ms) >= 20


@pytest.mark.gpu
def test_floor_halved_batch_matches_oracle_per_frame():
    """A batch mixing frames whose pyramids halve exactly with frames whose
    levels are floor halves: the level-by-level resampling runs inside the
    captured graph for the whole plane set."""
    imgs = [frames.odd_size_frame(w, h)[0] for (w, h) in frames.ODD_SIZES]
    imgs.insert(1, frames.small_marker_frame()[0])
    print("floor-halved batch", [im.shape for im in imgs], flush=True)
    with BatchDetector([(im.shape[1], im.shape[0]) for im in imgs]) as b:
        for f, im in enumerate(imgs):
            b.load(f, im)
        got = b.detect()
        again = b.detect()
    assert again == got
    for f, im in enumerate(imgs):
        match(got[f], oracle_lib.detect(im, threads=8))
