"""MFMA split-precision FFT variants against numpy's FP64 pocketfft, and the
projection loop run on them against the vector-ALU FP32 loop.

The variants run every radix-16 Stockham stage as a DFT-matrix GEMM on the
matrix cores (v_mfma_f32_16x16x32) with FP32 accumulation:
  mfma-bf16x3  x = hi + lo and W = Whi + Wlo in bfloat16, three products;
               each operand keeps ~16 significant bits (|x - hi - lo| <= 2^-18 |x|)
  mfma-fp16x3  the same split in float16 after a power-of-two scaling of
               each butterfly's inputs into [2^14, 2^15); ~22 bits per operand
Tolerances (relative L2 norm / max|dX| over max|X|, against FP64):
  bf16x3 3e-5 / 1e-4, fp16x3 3e-6 / 1e-5 -- an operand error of a few 2^-18
  (2^-22) per product over at most two GEMM stages plus the FP32 VALU stages.
Lengths without a radix-16 factor run the vector-ALU kernels inside the same
variant; those shapes are here too.
"""
import numpy as np
import pytest
import torch

import ffcz_amd
from ffcz_amd import FFT_BF16X3, FFT_FP16X3, FFT_VALU, MODE_ABS, MODE_DIRECT, ProjectionSolver
from tests._cases import binding_case, direct_case

pytestmark = pytest.mark.gpu

# y/x lengths 32..512 and z half-lengths 16..256 with a radix-16 factor, plus fallbacks
SHAPES = [
    (32, 64, 32), (96, 128, 64), (64, 32, 96), (128, 96, 128), (8, 256, 256), (4, 512, 512),
    (512, 32, 512), (60, 72, 96), (6, 10, 12), (1, 64, 64),
]
TOL = {FFT_BF16X3: (3e-5, 1e-4), FFT_FP16X3: (3e-6, 1e-5), FFT_VALU: (1e-6, 2e-6)}
NAMES = {FFT_VALU: "valu-fp32", FFT_BF16X3: "mfma-bf16x3", FFT_FP16X3: "mfma-fp16x3"}


def errs(a, ref):
    d = np.asarray(a, np.complex128) - ref
    return (np.linalg.norm(d) / np.linalg.norm(ref), np.abs(d).max() / np.abs(ref).max())


@pytest.mark.parametrize("variant", [FFT_VALU, FFT_BF16X3, FFT_FP16X3])
@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_variant_error(gpu, shape, variant):
    rng = np.random.default_rng(7 + sum(shape))
    x = rng.standard_normal(shape).astype(np.float32)
    X = ffcz_amd.fft_r2c(torch.from_numpy(x).to(gpu), variant).cpu().numpy()
    ref = np.fft.rfftn(x.astype(np.float64))
    l2, mx = errs(X, ref)
    print(f"r2c {NAMES[variant]} {shape}: rel-L2 {l2:.2e} max {mx:.2e}")
    assert l2 < TOL[variant][0] and mx < TOL[variant][1], (l2, mx)


@pytest.mark.parametrize("variant", [FFT_BF16X3, FFT_FP16X3])
@pytest.mark.parametrize("shape", SHAPES)
def test_c2r_variant_error(gpu, shape, variant):
    rng = np.random.default_rng(11 + sum(shape))
    Y = np.fft.rfftn(rng.standard_normal(shape))
    out = ffcz_amd.fft_c2r(torch.from_numpy(Y.astype(np.complex64)).to(gpu), shape[2], variant).cpu().numpy()
    ref = np.fft.irfftn(Y, s=shape, axes=(0, 1, 2), norm="forward")
    l2, mx = errs(out, ref)
    print(f"c2r {NAMES[variant]} {shape}: rel-L2 {l2:.2e} max {mx:.2e}")
    assert l2 < TOL[variant][0] and mx < 4 * TOL[variant][1], (l2, mx)


@pytest.mark.parametrize("scale", [1e-30, 1e-8, 1e8, 1e30])
def test_fp16_scaling_covers_float_range(gpu, scale):
    """The per-butterfly power-of-two scaling keeps float16 in range: the
    relative error does not depend on the magnitude of the data."""
    shape = (16, 64, 256)
    x = np.random.default_rng(3).standard_normal(shape).astype(np.float32)
    ref = np.fft.rfftn(x.astype(np.float64))
    X = ffcz_amd.fft_r2c(torch.from_numpy((x * np.float32(scale)).astype(np.float32)).to(gpu), FFT_FP16X3)
    l2, mx = errs(X.cpu().numpy().astype(np.complex128) / scale, ref)
    assert np.isfinite(l2) and l2 < TOL[FFT_FP16X3][0] and mx < TOL[FFT_FP16X3][1], (scale, l2, mx)


def test_impulse_closed_form_on_mfma(gpu):
    """An impulse transforms to the all-ones spectrum exactly enough to see a
    dropped or transposed DFT-matrix entry (every bin would move)."""
    shape = (32, 64, 512)
    x = np.zeros(shape, np.float32)
    x[0, 0, 0] = 1.0
    for v in (FFT_BF16X3, FFT_FP16X3):
        X = ffcz_amd.fft_r2c(torch.from_numpy(x).to(gpu), v).cpu().numpy()
        assert np.abs(X - 1.0).max() < 1e-5, v
    # a single cosine: two bins of N/2 (one in the half spectrum)
    n = np.arange(shape[2])
    c = np.broadcast_to(np.cos(2 * np.pi * 37 * n / shape[2]), shape).astype(np.float32)
    for v in (FFT_BF16X3, FFT_FP16X3):
        C = ffcz_amd.fft_r2c(torch.from_numpy(np.ascontiguousarray(c)).to(gpu), v).cpu().numpy()
        expect = np.zeros(C.shape, np.complex128)
        expect[0, 0, 37] = c.size / 2
        assert np.abs(C - expect).max() < 1e-4 * c.size, v


def _solve(gpu, shape, variant, case, mode):
    with ProjectionSolver(*shape, dtype=np.float32) as s:
        if mode == MODE_ABS:
            x, e0, eps, delta = case
            s.initialize_all_abs(x, eps, delta)
        else:
            x, e0, eps, idx, dk = case
            s.initialize(x, idx, dk, eps, len(idx))
        s.set_fft_precision(variant)
        assert s.fft_precision() == variant
        s.set_margins(0.125, 0.125)
        # the reference driver's tolerance (projection_example.cu:173-246); it sits
        # above every variant's round-trip noise (~1e-5 relative of |e| ~ eps)
        s.solve_loop(mode, e0, 400, 1e-7)
        it, viol, conv = s.status()
        cert = s.certify(False)
        return it, conv, cert, s.get_results()


@pytest.mark.parametrize("mode", [MODE_ABS, MODE_DIRECT])
@pytest.mark.parametrize("shape", [(16, 64, 128), (32, 32, 96), (6, 10, 12)])
def test_loop_on_mfma_variants_certifies(gpu, shape, mode):
    """Both bounds certified in FP64 after the loop on each variant; iteration
    count within +-1 of the vector-ALU loop and the iterate close to it."""
    if mode == MODE_ABS:
        x, e0, eps, delta = binding_case(shape, seed=5, dtype=np.float32, spread=1.3, dfrac=0.95)
        case = (x, e0, eps, delta)
    else:
        x, e0, eps, idx, dk = direct_case(shape, seed=5, dtype=np.float32, frac=0.5)
        case = (x, e0, eps, idx, (dk * np.float32(3.0)).astype(np.float32))
    it0, conv0, cert0, e_valu = _solve(gpu, shape, FFT_VALU, case, mode)
    assert conv0 and cert0.ok, (it0, cert0)
    for v in (FFT_BF16X3, FFT_FP16X3):
        it, conv, cert, e = _solve(gpu, shape, v, case, mode)
        d = np.linalg.norm(e.astype(np.float64) - e_valu) / np.linalg.norm(e_valu)
        print(f"{NAMES[v]} {shape} mode {mode}: iterations {it} (valu {it0}), iterate rel diff {d:.2e}")
        assert conv and cert.ok, (v, it, cert)
        assert abs(it - it0) <= 1, (v, it, it0)
        assert d < 1e-3, (v, d)
