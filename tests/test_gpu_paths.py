"""The length-specialised FFT kernels, the fused iteration form and the
slab-interleaved five-pass iteration give the same results as the generic
run-time-planned kernels and each other (FP64: to rounding; the two paths
use different radix schedules only where noted)."""
import numpy as np
import pytest
import torch

import ffcz_amd
from ffcz_amd import ProjectionSolver
from ffcz_amd._lib import set_tuning
from tests._cases import binding_case

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_tuning():
    prev = set_tuning()
    yield
    set_tuning(*prev)


def run(shape, dtype, force_generic, slab_mb):
    set_tuning(force_generic, slab_mb)
    x, e0, eps, delta = binding_case(shape, seed=12, dtype=dtype)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 25, 1e-12 if dtype == np.float64 else 1e-7)
        return s.iteration_count, s.get_results(), s.get_freq_edits_full(), s.get_spatial_edits_full()


@pytest.mark.parametrize("shape", [(32, 64, 128), (24, 60, 72), (64, 16, 256)])
def test_fast_matches_generic_f64(gpu, restore_tuning, shape):
    a = run(shape, np.float64, 0, 0)
    b = run(shape, np.float64, 1, 0)
    assert a[0] == b[0]
    for u, v in zip(a[1:], b[1:]):
        assert np.linalg.norm(u - v) <= 1e-11 * max(np.linalg.norm(v), 1e-300)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(32, 64, 128), (16, 32, 256),
                                   # rows (nx*ny) not a multiple of the fused z tile's rows: the last
                                   # tile is partial (half lengths 64, 256 and 75)
                                   (1, 60, 128), (1, 72, 512), (2, 60, 150)])
def test_fused_form_matches_five_pass_form(gpu, restore_tuning, shape, dtype):
    """The whole-field loop runs the fused form (inverse z, spatial clamp and
    forward z in one pass, speculative clamps undone after convergence); a
    slab setting (1 MiB slabs: several slabs at these sizes) runs the five
    passes slab by slab.  The same operations in the same order, compiled in
    different kernels (the compiler's contraction of multiply-adds may
    differ): iteration count equal, edits and iterate to rounding (the fused
    form's iterate is recovered from the spectrum rows after the loop, a z
    transform round trip).  Measured: FP64 1.5e-15, 7e-16, 4.4e-15 (e, df,
    ds); FP32 5.7e-7, 2.5e-7, 1.6e-6, the spatial edits amplifying the
    rounding as against the oracle (test_gpu_workloads)."""
    a = run(shape, dtype, 0, 0)
    b = run(shape, dtype, 0, 1)
    assert a[0] == b[0]
    tol = [1e-6, 1e-6, 5e-6] if dtype == np.float32 else [1e-13] * 3
    errs = [float(np.linalg.norm(u - v) / np.linalg.norm(v)) for u, v in zip(a[1:], b[1:])]
    print("fused vs five-pass (e, df, ds):", errs)
    assert all(e <= t for e, t in zip(errs, tol))


@pytest.mark.parametrize("shape", [(8, 60, 72), (4, 512, 16), (2, 600, 8)])
def test_fast_fft_matches_numpy(gpu, restore_tuning, shape):
    set_tuning(0, 64)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    X = ffcz_amd.fft_r2c(torch.from_numpy(x).to(gpu)).cpu().numpy()
    ref = np.fft.rfftn(x)
    assert np.abs(X - ref).max() < 1e-13 * np.abs(ref).max()


def run_direct(shape, dtype, force_generic, mode="direct"):
    from tests._cases import direct_case

    set_tuning(force_generic, 0)
    x, e0, eps, idx, dk = direct_case(shape, seed=7, dtype=dtype)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        if mode == "direct":
            s.initialize(x, idx, dk, eps)
            s.solve(e0, 10, -1.0)  # a fixed count of iterations
        else:
            s.initialize_all_ptw(x, eps, 0.05)
            s.solve_all_ptw(e0, 10, -1.0)
        return s.iteration_count, s.get_results(), s.get_freq_edits_full(), s.get_spatial_edits_full()


@pytest.mark.parametrize("mode", ["direct", "ptw"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(32, 24, 40), (60, 18, 20), (600, 4, 12)])
def test_direct_clamp_pass_matches_generic(gpu, restore_tuning, shape, dtype, mode):
    # per-bin and ptw bounds: the specialised x pass clamps in a pass of its
    # own over the tile (the 600-point axis: one FP64 tile per CU); the
    # generic kernel clamps in its last stage
    a = run_direct(shape, dtype, 0, mode)
    b = run_direct(shape, dtype, 1, mode)
    assert a[0] == b[0] == 10
    tol = 1e-11 if dtype == np.float64 else 2e-5
    for u, v in zip(a[1:], b[1:]):
        assert np.linalg.norm(u - v) <= tol * max(np.linalg.norm(v), 1e-300)
