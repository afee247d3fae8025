"""Projection loop of the HIP solver against the oracle restatement of the
reference loop (projection_algorithm_impl.cuh:772-858).

FP64 solver vs FP64 oracle: same iteration count, edits/iterate within
1e-9 (norm-wise relative).  FP32 solver vs the FP32 oracle: the clamp is
discontinuous, so FFT rounding can flip a bin at the box edge; parity is
norm-wise (||d||/||ref|| <= 1e-5) plus an active-set agreement >= 99.9 %.
"""
import numpy as np
import pytest
import torch

import ffcz_amd
from ffcz_amd import MODE_ABS, MODE_DIRECT, ProjectionSolver
from oracle import ffcz_oracle as O
from tests._cases import binding_case, direct_case

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


SHAPES = [(1, 16, 16), (6, 10, 12), (12, 10, 8), (5, 9, 15), (1, 64, 64), (16, 20, 30)]


@pytest.mark.parametrize("shape", SHAPES)
def test_abs_loop_f64_matches_oracle(gpu, shape):
    x, e0, eps, delta = binding_case(shape, seed=3)
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=50, tol=1e-12)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 50, 1e-12)
        it, v, conv = s.status()
        assert it == ref.iterations
        assert rel(s.get_results(), ref.e.ravel()) < 1e-9
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-9
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-9
        assert abs(v - ref.violation) <= 1e-12 + 1e-9 * abs(ref.violation)


@pytest.mark.parametrize("shape", SHAPES)
def test_direct_loop_f64_matches_oracle(gpu, shape):
    x, e0, eps, idx, dk = direct_case(shape, seed=5)
    ref = O.solve(x, e0, eps, mode="direct", idx=idx, delta_k=dk, max_iter=40, tol=1e-12)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize(x, idx, dk, eps)
        s.solve(e0, 40, 1e-12)
        assert s.iteration_count == ref.iterations
        assert rel(s.get_results(), ref.e.ravel()) < 1e-9
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-9
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-9


@pytest.mark.parametrize("shape", [(1, 64, 64), (16, 20, 30), (32, 32, 32)])
def test_abs_loop_f32_matches_fp32_oracle(gpu, shape):
    x, e0, eps, delta = binding_case(shape, seed=11, dtype=np.float32)
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=30, tol=1e-7, dtype=np.float32)
    with ProjectionSolver(*shape, dtype=np.float32) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 30, 1e-7)
        assert abs(s.iteration_count - ref.iterations) <= 1
        df = s.get_freq_edits_full()
        ds = s.get_spatial_edits_full()
        assert rel(df, ref.df) < 1e-5
        assert rel(ds, ref.ds) < 1e-5
        assert rel(s.get_results(), ref.e.ravel()) < 1e-5
        agree = np.mean((np.abs(df) > 1e-7) == (np.abs(ref.df) > 1e-7))
        assert agree > 0.999


@pytest.mark.parametrize("shape", [(6, 10, 12), (16, 20, 30), (32, 12, 10)])
def test_ptw_loop_f64_matches_oracle(gpu, shape):
    x, e0, eps, _ = binding_case(shape, seed=21)
    ref = O.solve(x, e0, eps, mode="ptw", delta_ptw=0.05, max_iter=20, tol=1e-12)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_ptw(x, eps, 0.05)
        s.solve_all_ptw(e0, 20, 1e-12)
        assert s.iteration_count == ref.iterations
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-9
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-9


def test_inactive_frequency_bound_closed_form(gpu):
    # delta huge: iteration 1 is the identity, the spatial clamp fires, and
    # iteration 2 converges with ds = clamp(e0) - e0 exactly (bit-exact).
    shape = (4, 6, 8)
    x, e0, eps, _ = binding_case(shape, seed=2, dtype=np.float32)
    with ProjectionSolver(*shape, dtype=np.float32) as s:
        s.initialize_all_abs(x, eps, 1e30)
        s.solve_all_abs(e0, 10, 1e-7)
        assert s.iteration_count == 2
        assert not np.any(s.get_freq_edits_full())
        ds = s.get_spatial_edits_full()
        # the FP32 round trip adds < 1e-9 noise to e; ds is computed from it
        expect = np.clip(e0, -eps, eps).ravel() - e0.ravel()
        assert np.abs(ds - expect).max() < 1e-8


def test_solve_before_initialize_raises(gpu):
    with ProjectionSolver(2, 2, 4) as s:
        with pytest.raises(ffcz_amd.FFCzError, match="Solver not initialized"):
            s.solve_all_abs(np.zeros(16, np.float32), 5, 1e-7)


def test_repeated_solve_accumulates_and_reset(gpu):
    shape = (6, 10, 12)
    x, e0, eps, delta = binding_case(shape, seed=9)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 3, 1e-12)
        d1 = s.get_freq_edits_full().copy()
        s.solve_all_abs(e0, 3, 1e-12)
        d2 = s.get_freq_edits_full()
        assert rel(d2, 2 * d1) < 1e-9  # same e0 again: the same edits accumulate twice
        s.reset_edit_accumulators()
        assert not np.any(s.get_freq_edits_full())


@pytest.mark.parametrize("shape", [(16, 20, 30), (16, 20, 32)])
def test_device_inputs_and_max_iterations_epilogue(gpu, shape):
    # (16, 20, 30): the five-pass form and its epilogue clamp; (16, 20, 32):
    # the fused form, whose last z pass made the final clamp
    x, e0, eps, delta = binding_case(shape, seed=4, dfrac=0.2)
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=3, tol=0.0)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(torch.from_numpy(x).to(gpu), eps, delta)
        s.solve_all_abs(torch.from_numpy(e0).to(gpu), 3, 0.0)
        it, v, conv = s.status()
        assert it == 3 and not conv
        # the last iteration's spatial clamp is applied (impl:851)
        assert rel(s.get_results(), ref.e.ravel()) < 1e-9
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-9


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(16, 20, 32), (8, 64, 64)])
def test_converged_iteration_clamps_are_dropped(gpu, shape, dtype):
    """tol = 2 % of eps: the converged iteration still has elements beyond
    eps by less than tol.  The reference stops before clamping them
    (projection_algorithm_impl.cuh:845); the fused form clamped them
    speculatively and must leave them out of ds, and report the unclamped
    iterate."""
    x, e0, eps, delta = binding_case(shape, seed=17, dtype=dtype, dfrac=0.3)
    tol = 0.02 * eps
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=40, tol=tol, dtype=dtype)
    # the final iterate has elements in (eps, eps + tol): clamps to undo
    over = np.abs(ref.e) - eps
    assert np.count_nonzero((over > 0) & (over < tol)) > 0
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 40, tol)
        it, v, conv = s.status()
        assert conv and it == ref.iterations
        lim = 1e-9 if dtype == np.float64 else 1e-5
        assert rel(s.get_spatial_edits_full(), ref.ds) < lim
        assert rel(s.get_freq_edits_full(), ref.df) < lim
        e = s.get_results()
        assert rel(e, ref.e.ravel()) < lim
        # elements the reference left beyond eps stay beyond it (FP64: no
        # element sits within rounding of eps)
        if dtype == np.float64:
            assert np.count_nonzero(np.abs(e) > eps) == np.count_nonzero(np.abs(ref.e) > eps)


DEGENERATE = [(1, 1, 1), (1, 1, 2), (2, 1, 1), (1, 3, 1), (7, 1, 5), (1, 1, 17), (3, 2, 1), (1, 2, 64), (2, 2, 2)]


@pytest.mark.parametrize("shape", DEGENERATE)
def test_degenerate_axes_match_oracle(gpu, shape):
    """Axes of length 1 and 2 (a lone DC bin, a DC and Nyquist pair, lines and
    planes along every axis): both bound modes against the FP64 oracle, then
    the certification of the margin-shrunk solve against the oracle's."""
    x, e0, eps, delta = binding_case(shape, seed=9)
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=50, tol=1e-12)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 50, 1e-12)
        assert s.iteration_count == ref.iterations
        assert rel(s.get_results(), ref.e.ravel()) < 1e-9
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-9
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-9
        s.reset_edit_accumulators()
        s.set_margins(0.125, 0.125)
        s.solve_all_abs(e0, 50, 1e-12)
        c = s.certify(False)
        r = O.certify(O.reconstruct_edit(shape, e0, s.get_freq_edits_full(), s.get_spatial_edits_full()), eps, delta)
        assert (c.ok, c.spatial_violations, c.freq_violations) == (r["ok"], r["spatial_violations"],
                                                                   r["freq_violations"])
        assert c.ok
    x, e0, eps, idx, dk = direct_case(shape, seed=9)
    ref = O.solve(x, e0, eps, mode="direct", idx=idx, delta_k=dk, max_iter=50, tol=1e-12)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize(x, idx, dk, eps)
        s.solve(e0, 50, 1e-12)
        assert s.iteration_count == ref.iterations
        assert rel(s.get_results(), ref.e.ravel()) < 1e-9
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-9
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-9


PRIME_LOOP_SHAPES = [(7, 11, 13), (1, 69, 115), (2, 98, 14), (23, 6, 46)]


@pytest.mark.parametrize("shape", PRIME_LOOP_SHAPES)
def test_prime_factor_axes_match_oracle(gpu, shape):
    """Axis lengths with prime factors 7..31 (69 = 3·23, 98 = 2·7², 115 = 5·23,
    half-length z over 7 and 23): both bound modes in FP64 against the oracle,
    the abs loop in FP32 against the FP32 oracle."""
    x, e0, eps, delta = binding_case(shape, seed=13)
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=50, tol=1e-12)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 50, 1e-12)
        assert s.iteration_count == ref.iterations
        assert rel(s.get_results(), ref.e.ravel()) < 1e-9
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-9
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-9
    x, e0, eps, idx, dk = direct_case(shape, seed=13)
    ref = O.solve(x, e0, eps, mode="direct", idx=idx, delta_k=dk, max_iter=50, tol=1e-12)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize(x, idx, dk, eps)
        s.solve(e0, 50, 1e-12)
        assert s.iteration_count == ref.iterations
        assert rel(s.get_results(), ref.e.ravel()) < 1e-9
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-9
    x, e0, eps, delta = binding_case(shape, seed=14, dtype=np.float32)
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=30, tol=1e-7, dtype=np.float32)
    with ProjectionSolver(*shape, dtype=np.float32) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 30, 1e-7)
        assert abs(s.iteration_count - ref.iterations) <= 1
        assert rel(s.get_freq_edits_full(), ref.df) < 1e-5
        assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-5
        assert rel(s.get_results(), ref.e.ravel()) < 1e-5
