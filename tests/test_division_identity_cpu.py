"""The statistics kernels divide a float by a small integer n as
float(double(x) * double(1/n)) instead of an IEEE float division: the
double product is within 2^-52 of x/n, and x/n is never a float midpoint
and lies more than 2^-33 (relative) from every one, so both round the same.
Checked here on many normal-range floats for every n the kernels use."""
import numpy as np


def test_reciprocal_multiply_rounds_like_division():
    rng = np.random.default_rng(7)
    for n in range(1, 256):
        inv = 1.0 / np.float64(n)
        bits = rng.integers(0, 2**32, 40000, dtype=np.uint64).astype(np.uint32)
        expo = rng.integers(27, 227, bits.size).astype(np.uint32)
        x = ((bits & np.uint32(0x807FFFFF)) | (expo << np.uint32(23))).view(np.float32)
        halves = (rng.integers(0, 100000, 20000) * 0.5 + rng.integers(0, 7, 20000) * 0.125)
        x = np.concatenate([x, halves.astype(np.float32)])
        want = x / np.float32(n)
        got = (x.astype(np.float64) * inv).astype(np.float32)
        assert np.array_equal(want.view(np.uint32), got.view(np.uint32)), n
