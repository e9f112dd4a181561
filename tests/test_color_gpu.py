"""Color input through the C ABI: BGR / BGRA frames converted on the device
as the reference's applications convert them (cv::cvtColor BGR2GRAY,
applications/detection/main.cpp:357,427,515), bit for bit against the
oracle's conversion, then detected exactly as the converted gray frame."""
import numpy as np
import pytest

from cctag_amd import _lib
from cctag_amd.detector import Detector
from tests import frames, oracle_lib
from tests.test_detect_gpu import match


def color_frame(gray: np.ndarray, seed: int = 31) -> np.ndarray:
    """A BGR frame around a gray one: channels shifted and scaled apart, plus
    per-channel noise, so that the conversion's weights all matter."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = gray.astype(np.int32)
    b = g * 7 // 8 + rng.integers(-12, 13, size=g.shape)
    r = g * 9 // 8 - 10 + rng.integers(-12, 13, size=g.shape)
    gg = g + rng.integers(-6, 7, size=g.shape)
    return np.clip(np.stack([b, gg, r], -1), 0, 255).astype(np.uint8)


@pytest.mark.gpu
@pytest.mark.parametrize("config", [1, 2])
def test_detect_bgr_matches_converted_gray(config):
    gray0, _ = frames.config_frame(config)
    bgr = color_frame(gray0)
    gray = oracle_lib.bgr_to_gray(bgr)
    print("color frame", bgr.shape, flush=True)
    h, w = gray.shape
    with Detector(w, h) as det:
        got = det.detect_bgr(bgr)
        assert np.array_equal(det.level(0)["img"], gray)
        want = det.detect(gray)
        alpha = np.concatenate([bgr, np.full((h, w, 1), 7, np.uint8)], -1)
        assert det.detect_bgr(alpha) == got
        padded = np.zeros((h, w + 5, 3), np.uint8)
        padded[:, :w] = bgr
        assert det.detect_bgr(padded[:, :w]) == got
        det.load_bgr(bgr)
        assert det.detect_loaded() == got
    assert got == want
    match(got, oracle_lib.detect(gray, threads=8))
    assert sum(m["status"] == 1 for m in got) >= 4


@pytest.mark.gpu
def test_color_conversion_odd_width_bit_exact():
    """A width that is not a multiple of the 4 pixels a thread converts."""
    rng = np.random.Generator(np.random.PCG64(5))
    bgr = rng.integers(0, 256, size=(67, 1023, 3), dtype=np.uint8)
    bgra = rng.integers(0, 256, size=(67, 1023, 4), dtype=np.uint8)
    prm = _lib.default_params(3)
    prm.number_of_processed_multires_layers = 1
    with Detector(1023, 67, params=prm) as det:
        for img in (bgr, bgra):
            det.load_bgr(img)
            det.detect_loaded()
            assert np.array_equal(det.level(0)["img"], oracle_lib.bgr_to_gray(img))


@pytest.mark.gpu
def test_color_argument_errors():
    with Detector(64, 48) as det:
        with pytest.raises(ValueError):
            det.detect_bgr(np.zeros((48, 64, 2), np.uint8))
        with pytest.raises(_lib.CCTagError) as e:
            det.detect_bgr(np.zeros((48, 63, 3), np.uint8))
        assert e.value.rc == -1
        lib = det._lib
        img = np.zeros((48, 64, 3), np.uint8)
        assert lib.cctagx_load_bgr(det._h, img.ctypes.data, 64, 48, 64 * 3 - 1, 3) == -1
        assert lib.cctagx_load_bgr(det._h, img.ctypes.data, 64, 48, 64 * 3, 2) == -1
