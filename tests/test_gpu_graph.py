"""Graph replay of the projection loop (ffcz_set_graph).

Each iteration after a solve's first is one HIP graph launch.  The replayed
loop must give the same bits as eager launches: iteration count, violation,
edits and iterate, over repeated solves with eager work and host
synchronisations between them, and after the bounds change (a new capture).
"""
import numpy as np
import pytest
import torch

from ffcz_amd import MODE_ABS, MODE_DIRECT, ProjectionSolver
from tests._cases import binding_case, direct_case

pytestmark = pytest.mark.gpu


def _solve(s, mode, e0, max_iter=40):
    s.reset_edit_accumulators()
    s.solve_loop(mode, e0, max_iter, 1e-7)
    return s.status(), s.get_results(), s.get_spatial_edits_full(), s.get_freq_edits_full()


def _same(a, b):
    assert a[0] == b[0]
    for u, v in zip(a[1:], b[1:]):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("shape,dtype", [((32, 64, 128), np.float32), ((30, 36, 50), np.float64),
                                         ((16, 60, 120), np.float32)])
def test_graph_replay_matches_eager_launches(gpu, shape, dtype):
    x, e0, eps, delta = binding_case(shape, seed=11, dtype=dtype, spread=1.4, dfrac=0.9)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        s.set_margins(0.125, 0.125)
        s.set_graph(False)
        ref = _solve(s, MODE_ABS, e0)
        assert ref[0][0] >= 2  # at least one replayed iteration
        s.set_graph(True)
        for rep in range(3):
            got = _solve(s, MODE_ABS, e0)
            _same(got, ref)
            # the program's own eager device work between the replays
            t = torch.ones(1 << 20, device=gpu) * rep
            torch.cuda.synchronize()
            assert float(t.sum()) == float(rep * (1 << 20))
        # new bounds: the next solve captures again
        s.initialize_all_abs(x, eps * 1.25, delta * 0.8)
        got = _solve(s, MODE_ABS, e0)
        s.set_graph(False)
        ref2 = _solve(s, MODE_ABS, e0)
        _same(got, ref2)
        assert not np.array_equal(ref2[2], ref[2])


def test_graph_replay_per_bin_bounds_and_device_input(gpu):
    shape, dtype = (24, 40, 60), np.float64
    x, e0, eps, idx, dk = direct_case(shape, seed=12, dtype=dtype)
    e0d = torch.from_numpy(e0).to(gpu)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize(x, idx, dk, eps)
        s.set_graph(False)
        ref = _solve(s, MODE_DIRECT, e0d)
        s.set_graph(True)
        for _ in range(2):
            _same(_solve(s, MODE_DIRECT, e0d), ref)
        # another device buffer with the same initial error: only the first,
        # eager iteration reads it, so the replayed iterations are unchanged
        e0d2 = e0d.clone()
        _same(_solve(s, MODE_DIRECT, e0d2), ref)
        del e0d2


def test_pass_timing_runs_eagerly_with_the_same_results(gpu):
    shape, dtype = (32, 64, 128), np.float32
    x, e0, eps, delta = binding_case(shape, seed=13, dtype=dtype, spread=1.4, dfrac=0.9)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        ref = _solve(s, MODE_ABS, e0)
        s.set_pass_timing(True)
        got = _solve(s, MODE_ABS, e0)
        ms, n = s.pass_timing()
        s.set_pass_timing(False)
        _same(got, ref)
        assert n == got[0][0]
