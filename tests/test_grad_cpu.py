"""Pins the float64 restatement of the refinement score's gradient (forward
mode, tests/oracle_lib.grad_cost) against float64 central finite
differences on synthetic markers: relative error <= 1e-4."""
import numpy as np
import pytest

from tests import frames, grad_data, oracle_lib

FD_REL_TOL = 1e-4


# Steps that move the samples by ~1e-6 px: entries of the last row act
# through the projective division, with leverage ~ image coordinates.
FD_STEPS = np.array([1e-6] * 6 + [1e-9] * 3)


def fd_grad(img, H, stops, steps=FD_STEPS):
    g = np.zeros(9)
    for e in range(9):
        d = np.zeros(9)
        d[e] = steps[e]
        cp, _, _ = oracle_lib.grad_cost(img, H.reshape(9) + d, stops, grad=False)
        cm, _, _ = oracle_lib.grad_cost(img, H.reshape(9) - d, stops, grad=False)
        g[e] = (cp - cm) / (2 * d[e])
    return g.reshape(3, 3)


@pytest.mark.parametrize("offset", [(0.0, 0.0), (1.3, -0.7)])
def test_oracle_gradient_matches_finite_differences(offset):
    img, gts = frames.config_frame(1)
    ms = grad_data.markers_of(gts, img.shape[1], img.shape[0])
    assert len(ms) == 4
    for m in ms:
        H = m["H"].copy()
        # move the centre off the optimum so the gradient is not ~0
        T = np.array([[1, 0, offset[0] * 0.01], [0, 1, offset[1] * 0.01], [0, 0, 1.0]])
        H = H @ T
        c, g, nv = oracle_lib.grad_cost(img, H, m["stops"])
        assert nv == len(m["stops"]) and np.isfinite(c)
        gfd = fd_grad(img, H, m["stops"])
        err = np.linalg.norm(gfd - g) / np.linalg.norm(g)
        assert err <= FD_REL_TOL, (err, g, gfd)


def test_score_is_lowest_near_the_true_centre():
    """The true homography scores lower than ones whose centre is moved by
    a few pixels in the marker plane."""
    img, gts = frames.config_frame(1)
    for m in grad_data.markers_of(gts, img.shape[1], img.shape[0]):
        H = m["H"]
        c0, _, _ = oracle_lib.grad_cost(img, H, m["stops"], grad=False)
        for dx, dy in ((0.05, 0), (0, 0.05), (-0.05, 0.03)):
            T = np.array([[1, 0, dx], [0, 1, dy], [0, 0, 1.0]])
            c1, _, _ = oracle_lib.grad_cost(img, H @ T, m["stops"], grad=False)
            assert c1 > c0
