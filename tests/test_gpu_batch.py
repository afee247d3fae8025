"""Batched projection loop: many fields of uneven shapes in one sequence of
launches (one per pass and axis length), per-field convergence flags, the
iteration replayed as a HIP graph.

Parity: each field of a batch runs the same pass kernels on the same data in
the same order as the field solved alone, so iterate, edits and iteration
count are bit-identical to the single-field solver (itself pinned against
the oracle restatement of projection_algorithm_impl.cuh:772-858 in
test_gpu_solver).  Graph replays are checked against eager launches of the
same batch, bit for bit, over several solves with host work in between.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import ffcz_amd
from ffcz_amd import MODE_ABS, MODE_DIRECT, QUANT_STEP, ProjectionBatch, ProjectionSolver
from ffcz_amd._lib import FFCzError, load
from oracle import ffcz_oracle as O
from tests._cases import binding_case

pytestmark = pytest.mark.gpu

# batched lengths (60..120 and their z half-lengths) plus fields that fall back
# to per-field launches (z half-lengths 6 and 15 have no specialised kernel)
SHAPES = [(60, 72, 96), (72, 60, 120), (96, 96, 60), (60, 60, 60), (32, 64, 128), (6, 10, 12), (16, 20, 30)]


def _cases(shapes, dtype, seed0=0):
    out = []
    for i, s in enumerate(shapes):
        # different spreads: the fields converge after different iteration counts
        x, e0, eps, delta = binding_case(s, seed=seed0 + i, dtype=dtype, spread=1.2 + 0.15 * i, dfrac=0.9)
        out.append((x, e0, eps, delta))
    return out


def _solo(shape, case, dtype, max_iter, tol, repeat=1):
    x, e0, eps, delta = case
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        for _ in range(repeat):
            s.solve_loop(MODE_ABS, e0, max_iter, tol)
        it, v, conv = s.status()
        return it, conv, s.get_results(), s.get_spatial_edits_full(), s.get_freq_edits_full()


def _batch_results(b):
    return [(f.get_results(), f.get_spatial_edits_full(), f.get_freq_edits_full()) for f in b.fields]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_field_solves(gpu, use_graph):
    dtype = np.float32
    cases = _cases(SHAPES, dtype)
    with ProjectionBatch(SHAPES, dtype=dtype) as b:
        for f, (x, e0, eps, delta) in zip(b.fields, cases):
            f.initialize_all_abs(x, eps, delta)
        b.solve(MODE_ABS, [c[1] for c in cases], 60, 1e-7, use_graph=use_graph)
        its, viols, convs, launched = b.status()
        res = _batch_results(b)
    # iterations are issued in chunks between host checks; a field past its
    # convergence skips the rest on the device
    assert launched >= max(its)
    assert len(set(its)) > 1, "fields should converge at different iterations"
    for shape, case, it, conv, (e, ds, df) in zip(SHAPES, cases, its, convs, res):
        it1, conv1, e1, ds1, df1 = _solo(shape, case, dtype, 60, 1e-7)
        assert (it, conv) == (it1, conv1), shape
        assert np.array_equal(e, e1) and np.array_equal(ds, ds1) and np.array_equal(df, df1), shape


def test_graph_replays_match_eager_over_repeated_solves(gpu):
    """Two solves in a row (edits accumulate, impl:60) with and without the
    graph, host reads in between: every field identical bit for bit."""
    dtype = np.float32
    shapes = SHAPES[:5]
    cases = _cases(shapes, dtype, seed0=10)
    out = {}
    for g in (False, True):
        with ProjectionBatch(shapes, dtype=dtype) as b:
            for f, (x, e0, eps, delta) in zip(b.fields, cases):
                f.initialize_all_abs(x, eps, delta)
            b.solve(MODE_ABS, [c[1] for c in cases], 40, 1e-7, use_graph=g)
            first = _batch_results(b)
            b.solve(MODE_ABS, [c[1] for c in cases], 40, 1e-7, use_graph=g)
            out[g] = (first, _batch_results(b), b.status())
    assert out[False][2] == out[True][2]
    for k in (0, 1):
        for a, c in zip(out[False][k], out[True][k]):
            for u, v in zip(a, c):
                assert np.array_equal(u, v)
    # the second solve accumulates like the single-field solver
    for shape, case, (e, ds, df) in zip(shapes, cases, out[True][1]):
        _, _, e1, ds1, df1 = _solo(shape, case, dtype, 40, 1e-7, repeat=2)
        assert np.array_equal(e, e1) and np.array_equal(ds, ds1) and np.array_equal(df, df1)


def test_batch_f64_against_oracle(gpu):
    shapes = [(6, 10, 12), (12, 10, 8), (16, 20, 30)]
    cases = _cases(shapes, np.float64, seed0=20)
    with ProjectionBatch(shapes, dtype=np.float64) as b:
        for f, (x, e0, eps, delta) in zip(b.fields, cases):
            f.initialize_all_abs(x, eps, delta)
        b.solve(MODE_ABS, [torch.from_numpy(c[1]).to(gpu) for c in cases], 50, 1e-9)
        its, _, convs, _ = b.status()
        for f, s, (x, e0, eps, delta), it in zip(b.fields, shapes, cases, its):
            ref = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=50, tol=1e-9)
            assert it == ref.iterations
            e = f.get_results()
            assert np.linalg.norm(e - ref.e.ravel()) <= 1e-9 * np.linalg.norm(ref.e)


def test_batch_fields_extract_and_certify(gpu):
    dtype = np.float32
    shapes = SHAPES[:4]
    cases = _cases(shapes, dtype, seed0=30)
    with ProjectionBatch(shapes, dtype=dtype) as b:
        for f, (x, e0, eps, delta) in zip(b.fields, cases):
            f.initialize_all_abs(x, eps, delta)
            f.set_margins(0.125, 0.125)
        b.solve(MODE_ABS, [c[1] for c in cases], 200, 1e-7)
        _, _, convs, _ = b.status()
        assert all(convs)
        for f in b.fields:
            f.extract_compact_representation()
            assert f.certify(False).ok
            q = f.get_active_spatial_quant_edit_distances()
            assert q.dtype == np.uint16


def test_batch_max_iterations_zero_and_mode_mismatch(gpu):
    shapes = [(60, 72, 96), (6, 10, 12)]
    cases = _cases(shapes, np.float32, seed0=40)
    with ProjectionBatch(shapes, dtype=np.float32) as b:
        for f, (x, e0, eps, delta) in zip(b.fields, cases):
            f.initialize_all_abs(x, eps, delta)
        b.solve(MODE_ABS, [c[1] for c in cases], 0, 1e-7)
        for f, c in zip(b.fields, cases):
            assert np.array_equal(f.get_results(), c[1].ravel())
        with pytest.raises(FFCzError):
            b.solve(MODE_DIRECT, [c[1] for c in cases], 5, 1e-7)


def test_batch_field_handles_are_borrowed(gpu):
    lib = load()
    with ProjectionBatch([(8, 8, 8), (4, 6, 8)], dtype=np.float32) as b:
        n = C.c_size_t()
        assert lib.ffcz_batch_size(b._h, C.byref(n)) == 0 and n.value == 2
        assert lib.ffcz_destroy(b.fields[0]._h) != 0  # freed with the batch only
        h = C.c_void_p()
        assert lib.ffcz_batch_field(b._h, 2, C.byref(h)) != 0


def test_batch_extract_and_certify_all_match_per_field(gpu):
    """certify_all runs the fields' FP64 passes as batched launches (the
    fields without specialised kernels one by one): the same counts and
    maximal excesses as each field's own certification."""
    dtype = np.float32
    shapes = SHAPES
    cases = _cases(shapes, dtype, seed0=50)
    res = {}
    for overlapped in (False, True):
        with ProjectionBatch(shapes, dtype=dtype) as b:
            for f, (x, e0, eps, delta) in zip(b.fields, cases):
                f.initialize_all_abs(x, eps, delta)
                f.set_margins(0.125, 0.125)
            b.solve(MODE_ABS, [c[1] for c in cases], 100, 1e-7)
            if overlapped:
                c1 = b.certify_all(False)
                b.extract_all()
                c2 = b.certify_all(True)
            else:
                c1 = [f.certify(False) for f in b.fields]
                for f in b.fields:
                    f.extract_compact_representation()
                c2 = [f.certify(True) for f in b.fields]
            q = [(f.get_active_freq_quant_edit_distances(), f.get_spatial_flags_pack()) for f in b.fields]
            res[overlapped] = (c1, c2, q)
    assert res[False][0] == res[True][0] and res[False][1] == res[True][1]
    assert all(c.ok for c in res[True][0]) and all(c.ok for c in res[True][1])
    for (a1, b1), (a2, b2) in zip(res[False][2], res[True][2]):
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2)


def _extraction(f):
    return (f.get_freq_flags_pack(), f.get_spatial_flags_pack(), f.get_active_freq_edit_distances(),
            f.get_active_spatial_edit_distances(), f.get_active_freq_quant_edit_distances(),
            f.get_active_spatial_quant_edit_distances(), f.get_min_freq_edit(), f.get_max_freq_edit(),
            f.get_min_spatial_edit(), f.get_max_spatial_edit())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batched_extraction_matches_per_field(gpu, dtype):
    """extract_all runs every field's two edit arrays through one launch of
    each extraction kernel (tiles of all arrays under one ticket counter):
    flags, compact values, symbols, counts and ranges equal each field's own
    extraction, with the reference quantiser and the step quantiser mixed in
    one batch, and again on a second call (reused workspace)."""
    from ffcz_amd import QUANT_REF16, QUANT_STEP
    shapes = SHAPES
    cases = _cases(shapes, dtype, seed0=60)
    with ProjectionBatch(shapes, dtype=dtype) as b:
        for k, (f, (x, e0, eps, delta)) in enumerate(zip(b.fields, cases)):
            f.initialize_all_abs(x, eps, delta)
            if k % 2:
                f.set_quantizer(QUANT_STEP, float(delta) / 8, float(eps) / 8)
        b.solve(MODE_ABS, [c[1] for c in cases], 60, 1e-7)
        for rep in range(2):
            b.extract_all()
            got = [_extraction(f) for f in b.fields]
            for f in b.fields:
                f.extract_compact_representation()
            ref = [_extraction(f) for f in b.fields]
            for shape, g, r in zip(shapes, got, ref):
                for u, v in zip(g, r):
                    assert np.array_equal(u, v), (shape, rep)
            assert any(len(g[2]) > 0 for g in got) and any(len(g[3]) > 0 for g in got)


def test_batched_extraction_reports_overflow_per_field(gpu):
    """A step too fine for 16-bit symbols in one field: extract_all raises
    the range error and the other fields keep their results."""
    from ffcz_amd import QUANT_STEP
    shapes = SHAPES[:3]
    cases = _cases(shapes, np.float32, seed0=70)
    with ProjectionBatch(shapes, dtype=np.float32) as b:
        for f, (x, e0, eps, delta) in zip(b.fields, cases):
            f.initialize_all_abs(x, eps, delta)
            f.set_margins(0.125, 0.125)  # room for the 16-bit quantiser
        b.fields[1].set_quantizer(QUANT_STEP, float(cases[1][3]) * 1e-9, float(cases[1][2]) * 1e-9)
        b.solve(MODE_ABS, [c[1] for c in cases], 200, 1e-7)
        with pytest.raises(FFCzError):
            b.extract_all()
        assert b.fields[0].certify(True).ok and b.fields[2].certify(True).ok
        with pytest.raises(FFCzError):
            b.fields[1].certify(True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batched_extraction_of_untouched_frequency_edits(gpu, dtype):
    """Fields whose solve fired no frequency clamp keep df at zero, and the
    batched extraction gives their df no tiles: flags, pack, compact values,
    symbols and range equal the oracle's extraction of the downloaded edits,
    for those fields and for the others of the same batch."""
    shapes = SHAPES
    cases = _cases(shapes, dtype, seed0=80)
    with ProjectionBatch(shapes, dtype=dtype) as b:
        for k, (f, (x, e0, eps, delta)) in enumerate(zip(b.fields, cases)):
            # every other field with a bound no frequency bin reaches
            f.initialize_all_abs(x, eps, delta * (1e6 if k % 2 else 1.0))
        b.solve(MODE_ABS, [c[1] for c in cases], 60, 1e-7)
        b.extract_all()
        untouched = 0
        for f in b.fields:
            df, ds = f.get_freq_edits_full(), f.get_spatial_edits_full()
            untouched += not np.any(df != 0)
            got = _extraction(f)
            rf, rs = O.extract_one(df), O.extract_one(ds)
            for g, r in zip(got, (rf.pack, rs.pack, rf.compact, rs.compact, rf.quant, rs.quant, rf.vmin, rf.vmax,
                                  rs.vmin, rs.vmax)):
                assert np.array_equal(g, r)
        assert 0 < untouched < len(shapes)


def test_batch_borrowed_initial_errors(gpu):
    """Fields that borrow their device initial errors (set_borrow_e0) solve,
    extract and certify exactly like fields that copy them."""
    dtype = np.float32
    shapes = SHAPES[:4] + SHAPES[5:]
    cases = _cases(shapes, dtype, seed0=70)
    e0_dev = [torch.from_numpy(c[1]).cuda() for c in cases]
    res = []
    for borrow in (False, True):
        with ProjectionBatch(shapes, dtype=dtype) as b:
            for f, (x, e0, eps, delta) in zip(b.fields, cases):
                f.set_borrow_e0(borrow)
                f.initialize_all_abs(x, eps, delta)
                f.set_margins(0.125, 0.125)
                f.set_quantizer(QUANT_STEP, delta / 8, eps / 8)
            b.solve(MODE_ABS, e0_dev, 80, 1e-7)
            its = b.status()[0]
            c1 = b.certify_all(False)
            b.extract_all()
            c2 = b.certify_all(True)
            res.append((its, c1, c2, _batch_results(b)))
    assert res[0][:3] == res[1][:3]
    for (e, ds, df), (e1, ds1, df1) in zip(res[0][3], res[1][3]):
        assert np.array_equal(e, e1) and np.array_equal(ds, ds1) and np.array_equal(df, df1)


def test_graph_kept_across_solves_with_new_initial_errors(gpu):
    """The captured iteration is kept across solves whose launches are the same
    and reads each solve's arguments from the plan's tables: three solves on
    alternating device initial errors (borrowed, so the tables point at other
    arrays each time) give the same bits with the graph as with eager launches."""
    dtype = np.float32
    shapes = SHAPES[:4]
    cases = _cases(shapes, dtype, seed0=90)
    e_a = [torch.from_numpy(c[1]).cuda() for c in cases]
    e_b = [torch.from_numpy(np.ascontiguousarray(c[1][::-1])).cuda() for c in cases]
    out = {}
    for use_graph in (False, True):
        with ProjectionBatch(shapes, dtype=dtype) as b:
            for f, (x, e0, eps, delta) in zip(b.fields, cases):
                f.set_borrow_e0(True)
                f.initialize_all_abs(x, eps, delta)
            res = []
            for es in (e_a, e_b, e_a):
                b.solve(MODE_ABS, es, 60, 1e-7, use_graph=use_graph)
                res.append((b.status()[0], _batch_results(b)))
                for f in b.fields:
                    f.reset_edit_accumulators()
            out[use_graph] = res
    assert out[False][0][0] != out[False][1][0] or not all(
        np.array_equal(a[0], b[0]) for a, b in zip(out[False][0][1], out[False][1][1]))
    for (it0, r0), (it1, r1) in zip(out[False], out[True]):
        assert it0 == it1
        for (e, ds, df), (e1, ds1, df1) in zip(r0, r1):
            assert np.array_equal(e, e1) and np.array_equal(ds, ds1) and np.array_equal(df, df1)
