"""Closed-form checks of the learned tuner's block statistics as restated in
oracle/features_ref.py (the GPU kernel is compared with this restatement in
tests/test_gpu_learned.py): blocks whose statistics the mathematics fixes.
Each check targets a plausible mistake: a wrong stride or axis, a missing
factor 2 in the Laplacian, a sign error in the correlation, a wrong count."""
import numpy as np
import pytest

from oracle.features_ref import block_features

S1, S2, S4 = 4, 13, 22  # bases of the stride-1, -2, -4 groups


def _grid(shape):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")


@pytest.mark.parametrize("shape", [(33, 33, 33), (17, 17, 17), (65, 65)])
def test_constant_block(shape):
    f = block_features(np.full(shape, 3.0, np.float32), 1.0, 4.0, 1e-3)
    assert f[0] == pytest.approx(0.5) and f[1] == 0 and f[2] == f[3] == np.float32(0.5)
    assert np.all(f[4:31] == 0)            # no differences, no Laplacian
    assert np.all(f[31:37] == 0)           # correlation undefined -> 0
    assert f[37] == 0 and f[38] == 0
    assert f[39] == 1.0                    # every difference below the bound


def test_linear_ramp_exact_differences_zero_laplacian():
    shape = (33, 33, 33)
    i, j, k = _grid(shape)
    # dyadic slopes keep every value and difference exact in float32
    v = (0.25 * i + 0.125 * j + 0.0625 * k).astype(np.float32)
    f = block_features(v, 0.0, 1.0, 0.1)
    for si, s in enumerate((1, 2, 4)):
        base = 4 + 9 * si
        assert np.allclose(f[base:base + 3], [0.25 * s, 0.125 * s, 0.0625 * s], rtol=0, atol=0)
        assert np.allclose(f[base + 3:base + 6], [0.25 * s, 0.125 * s, 0.0625 * s], rtol=0, atol=0)
        assert f[base + 6] == 0 and f[base + 7] == 0 and f[base + 8] == 0   # Laplacian of a plane
    assert f[39] == pytest.approx(1.0 / 3.0)  # only the 0.0625 differences lie below 0.1


def test_quadratic_constant_laplacian():
    shape = (33, 33, 33)
    i, _, _ = _grid(shape)
    v = (i * i / 1024.0).astype(np.float32)  # exact in float32
    f = block_features(v, 0.0, 1.0, 1e-6)
    for si, s in enumerate((1, 2, 4)):
        L = 2.0 * s * s / 1024.0              # d^2/di^2 of i^2, sampled at stride s
        base = 4 + 9 * si
        assert f[base + 6] == pytest.approx(L, rel=1e-6)   # mean |L|
        assert f[base + 7] == pytest.approx(L, rel=1e-6)   # RMS L
        assert f[base + 8] == pytest.approx(L, rel=1e-6)   # max |L|
        assert f[base + 1] == 0 and f[base + 2] == 0        # no change along axes 1, 2
    assert f[37] == pytest.approx(1.0, rel=1e-6)  # kurtosis of a constant
    assert f[38] == 0


def test_checkerboard_correlations():
    shape = (32, 32, 32)
    i, j, k = _grid(shape)
    v = ((i + j + k) % 2).astype(np.float32)
    f = block_features(v, 0.0, 1.0, 0.5)
    assert np.allclose(f[31:34], -1.0, atol=1e-6)   # lag 1: perfect anti-correlation
    assert np.allclose(f[34:37], 1.0, atol=1e-6)    # lag 2: perfect correlation
    assert np.allclose(f[4:7], 1.0) and np.allclose(f[13:16], 0.0)
    assert f[10] == pytest.approx(6.0)              # |L1| = 6 at every interior point
    assert f[39] == 0.0


def test_ar1_lag_correlations():
    """x_i = rho x_(i-1) + e along axis 1 (stationary start): the lag-1 and
    lag-2 correlations estimate rho and rho^2; the other axes are independent."""
    r = np.random.default_rng(5)
    rho = 0.8
    shape = (64, 257, 64)
    x = np.empty(shape)
    x[:, 0, :] = r.standard_normal((64, 64)) / np.sqrt(1 - rho * rho)
    for t in range(1, shape[1]):
        x[:, t, :] = rho * x[:, t - 1, :] + r.standard_normal((64, 64))
    f = block_features(x.astype(np.float32), float(x.min()), float(x.max() - x.min()), 1e-3)
    assert f[32] == pytest.approx(rho, abs=0.02)
    assert f[35] == pytest.approx(rho * rho, abs=0.03)
    assert abs(f[31]) < 0.03 and abs(f[33]) < 0.03


def test_two_dimensional_block_leaves_axis_2_empty():
    r = np.random.default_rng(6)
    v = r.standard_normal((65, 65)).astype(np.float32)
    f = block_features(v, float(v.min()), float(v.max() - v.min()), 1e-2)
    for si in range(3):
        assert f[4 + 9 * si + 2] == 0 and f[7 + 9 * si + 2] == 0
    assert f[33] == 0 and f[36] == 0
    assert f[1] > 0 and 0 < f[39] < 1
