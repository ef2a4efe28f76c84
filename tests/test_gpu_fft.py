"""GPU FFT passes against numpy.fft (a library routine) through the C ABI.

Shapes include extents of 256 along z or y, where a z or y pass block moves 16 pencils of 256
elements and its global loads run in two chunks of 8 per thread."""
import numpy as np
import pytest

import srx_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 16, 16), (4, 6, 10), (12, 20, 9), (32, 16, 64), (1, 8, 8),
                                   (8, 1, 5), (96, 8, 4), (4, 32, 256), (16, 256, 8)])
@pytest.mark.parametrize("ncomp", [1, 6, 9])
def test_forward_matches_numpy_fp64(shape, ncomp):
    rng = np.random.default_rng(hash((shape, ncomp)) % 2**32)
    x = rng.standard_normal((ncomp,) + shape)
    got = srx_amd.fft_forward(x)
    ref = np.fft.rfftn(x, axes=(1, 2, 3))
    assert got.shape == ref.shape
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 16, 16), (12, 20, 9), (32, 16, 64), (96, 8, 4), (4, 32, 256),
                                   (16, 256, 8)])
def test_inverse_matches_numpy_fp64(shape):
    rng = np.random.default_rng(7)
    x = rng.standard_normal((9,) + shape)
    X = np.fft.rfftn(x, axes=(1, 2, 3))
    n = shape[0] * shape[1] * shape[2]
    got = srx_amd.fft_inverse(X, shape[2], scale=1.0 / n)
    assert np.abs(got - x).max() < 1e-13 * np.abs(x).max() * 10


def test_mixed_precision_forward_error_bound():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 32, 32, 32))
    got = srx_amd.fft_forward(x, precision=srx_amd.PREC_MIXED)
    ref = np.fft.rfftn(x, axes=(1, 2, 3))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 5e-6, err
