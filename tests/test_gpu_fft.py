"""3D R2C / C2R of the HIP library against numpy's pocketfft (FP64).

Tolerances: FP32 transforms max|dX| / max|X| < 2e-6 (a few ulps times log N);
FP64 < 1e-13.  Shapes cover every radix (2,3,4,5,8, and primes 7..31 as
direct DFT stages), odd and unit axes and the half-length R2C path."""
import numpy as np
import pytest
import torch

import ffcz_amd

pytestmark = pytest.mark.gpu

SHAPES = [
    (1, 16, 16), (2, 3, 4), (6, 10, 12), (5, 9, 15), (4, 6, 1), (1, 1, 8), (3, 1, 2), (12, 10, 8),
    (8, 8, 8), (16, 32, 64), (60, 72, 96), (1, 256, 256), (25, 27, 30), (7 * 0 + 30, 36, 50),
    (64, 64, 64), (125, 24, 18), (1, 1, 1), (2, 2, 2), (128, 16, 600),
]
TOL = {np.float32: 2e-6, np.float64: 1e-13}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_matches_numpy(gpu, shape, dtype):
    rng = np.random.default_rng(hash(shape) % 2**32)
    x = rng.standard_normal(shape).astype(dtype)
    X = ffcz_amd.fft_r2c(torch.from_numpy(x).to(gpu)).cpu().numpy()
    ref = np.fft.rfftn(x.astype(np.float64))
    assert X.shape == ref.shape
    err = np.abs(X - ref).max() / max(np.abs(ref).max(), 1e-300)
    assert err < TOL[dtype], err


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", SHAPES)
def test_c2r_matches_numpy(gpu, shape, dtype):
    rng = np.random.default_rng(1 + hash(shape) % 2**32)
    y = rng.standard_normal(shape)
    Y = np.fft.rfftn(y)  # Hermitian-consistent half spectrum
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    out = ffcz_amd.fft_c2r(torch.from_numpy(Y.astype(cdt)).to(gpu), shape[2]).cpu().numpy()
    ref = np.fft.irfftn(Y, s=shape, axes=(0, 1, 2), norm="forward")  # unnormalised
    err = np.abs(out - ref).max() / max(np.abs(ref).max(), 1e-300)
    assert err < TOL[dtype] * 4, err


# axis lengths with prime factors 7..31 (direct DFT stages of the run-time
# planned kernels): lone primes, prime powers, the half-length z path over an
# odd prime, and lengths whose prime stage spans 64-128 butterflies
PRIME_SHAPES = [(7, 11, 13), (14, 22, 26), (23, 1, 46), (1, 69, 115), (2, 98, 62), (29, 17, 19), (49, 8, 14),
                (448, 2, 3), (1, 3, 896), (31, 31, 31)]
PRIME_TOL = {np.float32: 1e-5, np.float64: 1e-13}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", PRIME_SHAPES)
def test_prime_factor_lengths_match_numpy(gpu, shape, dtype):
    rng = np.random.default_rng(7 + hash(shape) % 2**32)
    x = rng.standard_normal(shape).astype(dtype)
    X = ffcz_amd.fft_r2c(torch.from_numpy(x).to(gpu)).cpu().numpy()
    ref = np.fft.rfftn(x.astype(np.float64))
    err = np.abs(X - ref).max() / np.abs(ref).max()
    assert err < PRIME_TOL[dtype], err
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    out = ffcz_amd.fft_c2r(torch.from_numpy(ref.astype(cdt)).to(gpu), shape[2]).cpu().numpy()
    back = np.fft.irfftn(ref, s=shape, axes=(0, 1, 2), norm="forward")
    err = np.abs(out - back).max() / np.abs(back).max()
    assert err < PRIME_TOL[dtype] * 4, err


def test_prime_factor_above_31_is_rejected(gpu):
    with pytest.raises(ffcz_amd.FFCzError, match="prime factor above 31"):
        ffcz_amd.fft_r2c(torch.zeros((2, 37, 4), dtype=torch.float32, device=gpu))


def test_impulse_and_cosine_closed_forms(gpu):
    shape = (6, 10, 16)
    x = np.zeros(shape, np.float64)
    x[0, 0, 0] = 1.0
    X = ffcz_amd.fft_r2c(torch.from_numpy(x).to(gpu)).cpu().numpy()
    assert np.allclose(X, 1.0, atol=1e-14)
    n = np.arange(16)
    c = np.broadcast_to(np.cos(2 * np.pi * 3 * n / 16), shape).copy()
    C = ffcz_amd.fft_r2c(torch.from_numpy(c).to(gpu)).cpu().numpy()
    N = c.size
    expect = np.zeros_like(C)
    expect[0, 0, 3] = N / 2
    assert np.abs(C - expect).max() < 1e-9 * N


@pytest.mark.parametrize("shape", [(4, 6, 8), (5, 9, 15), (1, 16, 16)])
def test_c2r_of_non_hermitian_input_takes_hermitian_part(gpu, shape):
    """Per-bin bounds can leave the half spectrum non-Hermitian on the
    self-conjugate planes; the C2R then follows pocketfft (imaginary parts of
    each row's DC/Nyquist bins ignored after the x/y passes)."""
    rng = np.random.default_rng(7)
    H = rng.standard_normal((shape[0], shape[1], shape[2] // 2 + 1)) + 1j * rng.standard_normal(
        (shape[0], shape[1], shape[2] // 2 + 1))
    out = ffcz_amd.fft_c2r(torch.from_numpy(H).to(gpu), shape[2]).cpu().numpy()
    ref = np.fft.irfftn(H, s=shape, axes=(0, 1, 2), norm="forward")
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
