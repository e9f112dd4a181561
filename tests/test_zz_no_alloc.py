"""The allocation-free grid search (cctagx_refine_grid_static, its buffers
in a fixed device array) against the allocating one (cctagx_refiner_grid)
on the same frames and markers: identical centres and ok flags, on two
calls in a row, the second on a smaller frame with fewer markers; inputs
beyond the array's capacity are rejected."""
import ctypes

import numpy as np
import pytest

from cctag_amd import _lib, synth
from cctag_amd.detector import Refiner
from tests import frames


def _markers(gts):
    return [synth.refine_marker(g.H) for g in gts]


def _static(img, markers):
    lib = _lib.load()
    p = _lib.default_params(3)
    arr = (_lib.RefineMarker * max(len(markers), 1))()
    for i, m in enumerate(markers):
        arr[i].ellipse[:] = [float(v) for v in m["ellipse"]]
        arr[i].center0[:] = [float(v) for v in m["center0"]]
        st = np.asarray(m["stops"], np.float32)
        arr[i].n_cuts = len(st)
        for k, (x, y) in enumerate(st):
            arr[i].stops[k][0] = float(x)
            arr[i].stops[k][1] = float(y)
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    c = np.zeros((len(markers), 2), np.float32)
    ok = np.zeros(len(markers), np.int32)
    rc = lib.cctagx_refine_grid_static(ctypes.byref(p), img.ctypes.data, w, h, img.strides[0], arr,
                                       len(markers), c.ctypes.data, ok.ctypes.data)
    return rc, c, ok.astype(bool)


def _allocating(img, markers):
    with Refiner(len(markers)) as r:
        r.set_image(img)
        r.set_markers(markers)
        return r.grid()


@pytest.mark.gpu
def test_static_grid_search_matches_the_allocating_one():
    calls = [frames.config_frame(2, 0), frames.config_frame(1, 0)]  # 1920x1080 x16, 1280x720 x4
    for img, gts in calls:
        markers = _markers(gts)
        print("static grid search", img.shape, len(markers), "markers", flush=True)
        rc, c, ok = _static(img, markers)
        assert rc == 0, _lib.load().cctagx_last_error()
        c_ref, ok_ref = _allocating(img, markers)
        assert np.array_equal(ok, ok_ref)
        assert np.array_equal(c.view(np.int32), c_ref.view(np.int32))
        assert ok.sum() >= len(markers) - 1
        gt = np.array([m["center_gt"] for m in markers])
        assert np.abs(c[ok] - gt[ok]).max() < 0.5


@pytest.mark.gpu
def test_static_grid_search_capacity():
    img, gts = frames.config_frame(1, 0)
    markers = _markers(gts) * 17   # 68 > 64
    rc, _, _ = _static(img, markers)
    assert rc == -3
    big = np.zeros((1090, 1920), np.uint8)
    rc, _, _ = _static(big, _markers(gts))
    assert rc == -3
    rc, _, ok = _static(img, _markers(gts))   # still works afterwards
    assert rc == 0 and ok.sum() >= 3
