"""Differentiable centre refinement through the C ABI: the closed-form
homography's defining properties, the forward score and the adjoint
gradient against the float64 oracle and float64 central finite
differences (relative error <= 1e-4), and the batched descent against the
ground truth and the reference grid search on the same markers."""
import math

import numpy as np
import pytest

from cctag_amd import synth
from cctag_amd.detector import Refiner
from tests import frames, grad_data, oracle_lib
from tests.test_grad_cpu import fd_grad

GRAD_REL_TOL = 1e-4     # adjoint (fp64 sampling) vs FP64 central differences
FLOAT_REL_TOL = 2e-3    # fp32 sampling vs fp64 sampling


def setup(img, gts):
    ms = grad_data.markers_of(gts, img.shape[1], img.shape[0])
    r = Refiner(len(ms))
    r.set_image(img)
    r.set_markers(ms)
    return r, ms


@pytest.mark.gpu
def test_closed_form_homography_properties():
    img, gts = frames.config_frame(1)
    r, ms = setup(img, gts)
    with r:
        c = np.array([m["center_gt"] for m in ms], np.float32) + np.float32(0.7)
        H = r.homography(c).astype(np.float64)
        for Hm, cm, m in zip(H, c, ms):
            o = Hm @ np.array([0.0, 0.0, 1.0])
            assert np.hypot(o[0] / o[2] - cm[0], o[1] / o[2] - cm[1]) < 1e-3
            # the unit circle lands on the outer ellipse
            cx, cy, a, b, ang = m["ellipse"]
            th = np.linspace(0, 2 * np.pi, 16, endpoint=False)
            p = Hm @ np.stack([np.cos(th), np.sin(th), np.ones_like(th)])
            x, y = p[0] / p[2] - cx, p[1] / p[2] - cy
            u = x * math.cos(ang) + y * math.sin(ang)
            v = -x * math.sin(ang) + y * math.cos(ang)
            assert np.allclose((u / a) ** 2 + (v / b) ** 2, 1.0, atol=1e-3)


@pytest.mark.gpu
def test_score_and_gradient_against_oracle_and_finite_differences():
    img, gts = frames.config_frame(1)
    r, ms = setup(img, gts)
    with r:
        offs = np.array([[0.01, -0.008], [0.0, 0.0], [-0.012, 0.005], [0.007, 0.011]])
        Hs = []
        for m, (dx, dy) in zip(ms, offs):
            H = m["H"] @ np.array([[1, 0, dx], [0, 1, dy], [0, 0, 1.0]])
            Hs.append(H / H[2, 2])
        Hs = np.array(Hs, np.float32).astype(np.float64)   # what the device receives
        cost64, grad64 = r.cost_grad(Hs, fp64=True)
        cost32, grad32 = r.cost_grad(Hs, fp64=False)
        for i, m in enumerate(ms):
            print("marker", i, cost64[i], flush=True)
            c_or, g_or, nv = oracle_lib.grad_cost(img, Hs[i], m["stops"])
            assert nv == len(m["stops"])
            assert abs(cost64[i] - c_or) <= 1e-9 * c_or
            g = grad64[i, :8]
            assert np.linalg.norm(g - g_or.ravel()[:8]) <= 1e-7 * np.linalg.norm(g_or.ravel()[:8])
            gfd = fd_grad(img, Hs[i], m["stops"]).ravel()[:8]
            assert np.linalg.norm(g - gfd) <= GRAD_REL_TOL * np.linalg.norm(gfd)
            # centre offsets: d/dx, d/dy of H [1 0 x; 0 1 y; 0 0 1]
            for j, (ex, ey) in enumerate(((1e-7, 0), (0, 1e-7))):
                T = np.array([[1, 0, ex], [0, 1, ey], [0, 0, 1.0]])
                Ti = np.array([[1, 0, -ex], [0, 1, -ey], [0, 0, 1.0]])
                cp, _, _ = oracle_lib.grad_cost(img, Hs[i] @ T, m["stops"], grad=False)
                cm, _, _ = oracle_lib.grad_cost(img, Hs[i] @ Ti, m["stops"], grad=False)
                fd = (cp - cm) / 2e-7
                assert abs(grad64[i, 8 + j] - fd) <= GRAD_REL_TOL * max(abs(fd), 1e-3 * np.abs(g).max())
            assert abs(cost32[i] - cost64[i]) <= FLOAT_REL_TOL * cost64[i]
            assert np.linalg.norm(grad32[i] - grad64[i]) <= 1e-2 * np.linalg.norm(grad64[i])


@pytest.mark.gpu
@pytest.mark.parametrize("config", [1, 2])
def test_descent_against_ground_truth_and_grid_search(config):
    img, gts = frames.config_frame(config)
    r, ms = setup(img, gts)
    with r:
        c_grad, cost = r.descend(20)
        c_grid, ok = r.grid()
        t = r.times()
    gt = np.array([m["center_gt"] for m in ms])
    c0 = np.array([m["center0"] for m in ms])
    e0 = np.hypot(*(c0 - gt).T)
    eg = np.hypot(*(c_grad - gt).T)
    er = np.hypot(*(c_grid - gt).T)
    print("initial", e0.mean(), "descent", eg.mean(), eg.max(), "grid", er[ok].mean(), er[ok].max(), t,
          flush=True)
    assert ok.all() and np.isfinite(cost).all()
    assert eg.mean() < e0.mean()
    assert eg.mean() <= er.mean() + 0.05
    assert eg.max() < 0.5


@pytest.mark.gpu
def test_coarse_start_refinement_on_a_mosaic():
    """The grid search's first iteration followed by 20 descent steps reaches
    the grid search's centre accuracy on a mosaic of strongly tilted markers,
    including those whose starting centre is pixels off, where plain descent
    can settle in another basin."""
    img, ms = synth.refine_mosaic(512)
    gt = np.array([m["center_gt"] for m in ms])
    r = Refiner(len(ms))
    with r:
        r.set_image(img)
        r.set_markers(ms)
        c_ref, cost = r.refine(20)
        c_grid, ok = r.grid()
    e_ref = np.hypot(*(np.asarray(c_ref, np.float64) - gt).T)
    e_grid = np.hypot(*(np.asarray(c_grid, np.float64) - gt).T)
    print("coarse+descent", e_ref.mean(), e_ref.max(), "grid", e_grid[ok].mean(), e_grid[ok].max(),
          flush=True)
    assert ok.all() and np.isfinite(cost).all()
    assert e_ref.mean() <= e_grid.mean() * 1.05
    assert e_ref.max() < 0.5
