"""Edit accumulators across resets and solves (projection_algorithm_impl.cuh:60,
133-147: zeroed in the constructor and by reset_edit_accumulators, then
accumulated by every solve).

The library zeroes them lazily: after a reset the first spatial clamp of the
next solve stores ds whole (forward z pass of iteration 1, or the epilogue
after one unconverged iteration), and df is zeroed only when a clamp may have
touched it.  Every path must equal the oracle loop started from zero edits,
and a device-resident initial error (copied by the first iteration) must
give the same results as a host one."""
import numpy as np
import pytest
import torch

from ffcz_amd import MODE_ABS, ProjectionSolver
from oracle import ffcz_oracle as O
from tests._cases import binding_case

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a, np.float64) - b) / (nb if nb > 0 else 1.0)


SHAPE = (12, 10, 16)


@pytest.mark.parametrize("max_iter,tol", [(1, 0.0), (1, 1e30), (2, 0.0), (5, 0.0), (0, 0.0)])
def test_fresh_accumulators_match_oracle(gpu, max_iter, tol):
    """max_iter=1/tol=0: epilogue stores ds; tol=1e30: converged after one
    iteration, ds never written (zeroed); 2 and 5: iteration 1 stores ds."""
    x, e0, eps, delta = binding_case(SHAPE, seed=6, dfrac=0.7)
    ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=max_iter, tol=tol)
    for second in (False, True):
        with ProjectionSolver(*SHAPE, dtype=np.float64) as s:
            s.initialize_all_abs(x, eps, delta)
            if second:
                # dirty accumulators, then a reset: the lazy path from a used solver
                s.solve_all_abs(e0 * 0.7, 4, 0.0)
                s.reset_edit_accumulators()
            s.solve_loop(MODE_ABS, e0, max_iter, tol)
            assert s.status()[0] == ref.iterations
            assert rel(s.get_results(), ref.e.ravel()) < 1e-12
            assert rel(s.get_spatial_edits_full(), ref.ds) < 1e-12
            assert rel(s.get_freq_edits_full(), ref.df) < 1e-12


def test_reset_without_solve_reads_zero(gpu):
    x, e0, eps, delta = binding_case(SHAPE, seed=7, dfrac=0.5)
    with ProjectionSolver(*SHAPE, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 3, 0.0)
        assert np.abs(s.get_spatial_edits_full()).max() > 0
        s.reset_edit_accumulators()
        assert not s.get_spatial_edits_full().any()
        assert not s.get_freq_edits_full().any()


def test_device_initial_error_is_copied_by_the_first_iteration(gpu):
    """The solver keeps its own copy of e0 for certification even when the
    caller's device buffer changes after the solve."""
    x, e0, eps, delta = binding_case(SHAPE, seed=8, dfrac=0.8)
    out = []
    for dev in (False, True):
        with ProjectionSolver(*SHAPE, dtype=np.float64) as s:
            s.initialize_all_abs(x, eps, delta)
            s.set_margins(0.125, 0.125)
            src = torch.from_numpy(e0.copy()).to(gpu) if dev else e0.copy()
            s.solve_loop(MODE_ABS, src, 50, 1e-12)
            if dev:
                src.fill_(1.0)
                torch.cuda.synchronize()
            c = s.certify(False)
            out.append((s.status(), s.get_results(), s.get_spatial_edits_full(), c))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3] and out[1][3].ok
