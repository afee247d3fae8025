"""Pointwise power-spectrum bounds of the HIP solver against the oracle: the
partial ptw projection (projection_algorithm.cu:261-333 over
initialize_partial_ptw's list, projection_algorithm_impl.cuh:303-333,
701-738) and the corrected ptw bound (opt-in, not the reference's
arithmetic), both through the generic and the specialised x pass.

FP64 solver vs FP64 oracle: same iteration count, edits within 1e-9
(norm-wise relative).  FP32 solver vs FP32 oracle: iteration count within one,
edits within 1e-5 norm-wise (a bin at the bound's edge can flip under FFT
rounding).  Measured on MI355X: FP64 at most 6.2e-15 (df, ds, e; equal
iteration counts), FP32 at most 1.7e-6."""
import numpy as np
import pytest

import ffcz_amd
from ffcz_amd import ProjectionSolver
from ffcz_amd._lib import set_tuning
from oracle import ffcz_oracle as O
from tests._cases import binding_case

pytestmark = pytest.mark.gpu

DPTW = 0.05


@pytest.fixture
def restore_tuning():
    prev = set_tuning()
    yield
    set_tuning(*prev)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def partial_list(shape, seed, frac=0.4):
    """A constraint list as a caller might give it: idx[0] a nonzero bin (the
    DC rule takes bin 0 whatever it holds), a repeated bin, and bin 0 again
    later in the list."""
    F = shape[0] * shape[1] * (shape[2] // 2 + 1)
    rng = np.random.default_rng(seed)
    sel = rng.choice(np.arange(1, F), size=max(2, int(frac * F)), replace=False)
    return np.concatenate([[sel[0]], sel, [sel[1], 0]]).astype(np.uint64)


def run_gpu(shape, dtype, x, e0, eps, idx, corrected, max_iter, tol, partial=True):
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.set_ptw_semantics(corrected)
        if partial:
            s.initialize_partial_ptw(x, idx, DPTW, eps)
            s.solve_partial_ptw(e0, max_iter, tol)
        else:
            s.initialize_all_ptw(x, eps, DPTW)
            s.solve_all_ptw(e0, max_iter, tol)
        return s.iteration_count, s.get_freq_edits_full(), s.get_spatial_edits_full(), s.get_results()


def check(shape, dtype, mode, corrected, force_generic, seed):
    set_tuning(force_generic, 0)
    x, e0, eps, _ = binding_case(shape, seed=seed, dtype=dtype)
    idx = partial_list(shape, seed + 1)
    f64 = dtype == np.float64
    max_iter, tol = (20, 1e-12) if f64 else (20, 1e-7)
    ref = O.solve(x, e0, eps, mode=mode, idx=idx if mode == "partial_ptw" else None, delta_ptw=DPTW,
                  max_iter=max_iter, tol=tol, dtype=dtype, ptw_corrected=corrected)
    it, df, ds, e = run_gpu(shape, dtype, x, e0, eps, idx, corrected, max_iter, tol, partial=mode == "partial_ptw")
    # the bound binds: some frequency edit is non-zero
    assert np.any(ref.df != 0)
    errs = [rel(df, ref.df), rel(ds, ref.ds), rel(e, ref.e.ravel())]
    print(f"{mode} corrected={corrected} generic={force_generic} {shape} {np.dtype(dtype).name}: "
          f"iterations {it} vs {ref.iterations}, (df, ds, e) {errs}")
    if f64:
        assert it == ref.iterations
        assert max(errs) < 1e-9
    else:
        assert abs(it - ref.iterations) <= 1
        assert max(errs) < 1e-5


SHAPES = [(6, 10, 12), (16, 20, 30), (32, 32, 32)]


@pytest.mark.parametrize("force_generic", [0, 1])
@pytest.mark.parametrize("shape", SHAPES)
def test_partial_ptw_reference_f64(gpu, restore_tuning, shape, force_generic):
    check(shape, np.float64, "partial_ptw", False, force_generic, seed=31)


@pytest.mark.parametrize("force_generic", [0, 1])
@pytest.mark.parametrize("shape", SHAPES)
def test_partial_ptw_corrected_f64(gpu, restore_tuning, shape, force_generic):
    check(shape, np.float64, "partial_ptw", True, force_generic, seed=41)


@pytest.mark.parametrize("force_generic", [0, 1])
@pytest.mark.parametrize("shape", SHAPES)
def test_all_ptw_corrected_f64(gpu, restore_tuning, shape, force_generic):
    check(shape, np.float64, "ptw", True, force_generic, seed=51)


@pytest.mark.parametrize("mode,corrected", [("partial_ptw", False), ("partial_ptw", True), ("ptw", True)])
def test_ptw_f32(gpu, restore_tuning, mode, corrected):
    check((32, 32, 32), np.float32, mode, corrected, 0, seed=61)


def test_all_ptw_runs_after_partial_initialize(gpu):
    # initialize_partial_ptw computes everything solve_all_ptw needs; the
    # reference's solve_all_ptw runs after it over every bin
    shape = (16, 20, 30)
    x, e0, eps, _ = binding_case(shape, seed=71)
    idx = partial_list(shape, 72)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_partial_ptw(x, idx, DPTW, eps)
        s.solve_all_ptw(e0, 20, 1e-12)
        got = s.get_freq_edits_full()
    ref = run_gpu(shape, np.float64, x, e0, eps, idx, False, 20, 1e-12, partial=False)
    assert np.array_equal(got, ref[1])


def test_semantics_switch_between_solves(gpu):
    # the per-bin rule table follows the semantics of each solve
    shape = (16, 20, 30)
    x, e0, eps, _ = binding_case(shape, seed=81)
    idx = partial_list(shape, 82)
    out = []
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_partial_ptw(x, idx, DPTW, eps)
        for corrected in (True, False, True):
            s.set_ptw_semantics(corrected)
            s.reset_edit_accumulators()
            s.solve_partial_ptw(e0, 20, 1e-12)
            out.append(s.get_freq_edits_full())
    for k, corrected in enumerate((True, False, True)):
        ref = O.solve(x, e0, eps, mode="partial_ptw", idx=idx, delta_ptw=DPTW, max_iter=20, tol=1e-12,
                      ptw_corrected=corrected)
        assert rel(out[k], ref.df) < 1e-9
    assert rel(out[0], out[1]) > 1e-3


def test_partial_ptw_rejects_bad_input(gpu):
    shape = (4, 6, 8)
    x, e0, eps, _ = binding_case(shape, seed=91)
    F = 4 * 6 * 5
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        with pytest.raises(ffcz_amd.FFCzError, match="out of range"):
            s.initialize_partial_ptw(x, np.array([0, F], np.uint64), DPTW, eps)
        s.initialize_all_ptw(x, eps, DPTW)
        with pytest.raises(ffcz_amd.FFCzError, match="does not match"):
            s.solve_partial_ptw(e0, 5, 1e-12)
        with pytest.raises(ffcz_amd.FFCzError, match="ptw semantics"):
            ffcz_amd._lib.call("ffcz_set_ptw_semantics", s._h, 2)


def test_partial_ptw_empty_list_constrains_nothing(gpu):
    # K = 0: no frequency constraint, so the loop is the spatial clamp alone
    shape = (8, 10, 12)
    x, e0, eps, _ = binding_case(shape, seed=95)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_partial_ptw(x, np.zeros(0, np.uint64), DPTW, eps)
        s.solve_partial_ptw(e0, 10, 1e-12)
        assert not np.any(s.get_freq_edits_full())
        assert s.iteration_count == 2
