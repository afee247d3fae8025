"""Extraction (flags, min/max, compaction, quantisation, packing) and FP64
certification on the GPU, bit-exact against the oracle on the same
accumulators (projection_algorithm_impl.cuh:371-456,
projection_algorithm.cu:363-550, 615-644)."""
import ctypes as C

import numpy as np
import pytest
import torch

import ffcz_amd
from ffcz_amd import MODE_ABS, ProjectionSolver, QUANT_STEP
from ffcz_amd._lib import FFCzError, call
from oracle import ffcz_oracle as O
from tests._cases import binding_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 16, 16), (6, 10, 12), (16, 20, 30), (1, 64, 64)])
def test_extraction_bit_exact(gpu, shape, dtype):
    x, e0, eps, delta = binding_case(shape, seed=7, dtype=dtype)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 20, 1e-7)
        df = s.get_freq_edits_full()
        ds = s.get_spatial_edits_full()
        for edits, flags, pack, quant, comp, (lo, hi) in (
            (df, s.get_freq_flags(), s.get_freq_flags_pack(), s.get_active_freq_quant_edit_distances(),
             s.get_active_freq_edit_distances(), (s.get_min_freq_edit(), s.get_max_freq_edit())),
            (ds, s.get_spatial_flags(), s.get_spatial_flags_pack(), s.get_active_spatial_quant_edit_distances(),
             s.get_active_spatial_edit_distances(), (s.get_min_spatial_edit(), s.get_max_spatial_edit())),
        ):
            ref = O.extract_one(edits)
            assert np.array_equal(flags, ref.flags)
            assert np.array_equal(pack, ref.pack)
            assert np.array_equal(comp, ref.compact)
            assert lo == ref.vmin and hi == ref.vmax
            assert np.array_equal(quant, ref.quant)
            assert ref.flags.any()


@pytest.mark.parametrize("dtype, dt", [(np.float64, ffcz_amd._lib.FFCZ_F64), (np.float32, ffcz_amd._lib.FFCZ_F32)])
def test_kernel_extract_step_symbols_near_half_steps(gpu, dtype, dt):
    # values at and around (k + 1/2) steps, where the quotient's rounding
    # decides the symbol: the device's reciprocal form must give the
    # division's symbols (llround(v/step), ties away from zero)
    step = 1.25e-5
    rng = np.random.default_rng(3)
    k = rng.integers(0, 30000, 4000).astype(np.float64)
    offs = np.array([0.0, 1e-16, -1e-16, 1e-13, -1e-13, 1e-9, -1e-9, 1e-6, -1e-6, 1e-3])
    near = ((k[:, None] + 0.5) * (1.0 + offs[None, :]) * step).ravel()
    v = np.concatenate([near, -near, rng.standard_normal(50000) * step * 100]).astype(dtype)
    n = v.size
    dv = torch.from_numpy(v).to(gpu)
    pack = torch.zeros(((n + 63) // 64) * 8, dtype=torch.uint8, device=gpu)
    comp = torch.zeros(n, dtype=dv.dtype, device=gpu)
    quant = torch.zeros(n, dtype=torch.int16, device=gpu)
    cnt, lo, hi = C.c_size_t(), C.c_double(), C.c_double()
    torch.cuda.synchronize()
    call("ffcz_k_extract", dt, dv.data_ptr(), n, 2, step, pack.data_ptr(), comp.data_ptr(), quant.data_ptr(),
         C.byref(cnt), C.byref(lo), C.byref(hi), None)
    q = O.quantize_step(v, step)
    act = q != 0
    kk = cnt.value
    assert kk == int(act.sum())
    assert np.array_equal(comp.cpu().numpy()[:kk], v[act])
    assert np.array_equal(quant.cpu().numpy().view(np.uint16)[:kk], O.zigzag(q[act]).astype(np.uint16))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4095, 4096, 4097, 100_000, 1_000_003])
def test_kernel_extract_sizes(gpu, n):
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n).astype(np.float32) * 1e-3
    v[rng.random(n) < 0.7] = 0.0
    v[: min(3, n)] = [1e-8, -2e-7, 5e-8][: min(3, n)]  # around the 1e-7 threshold
    dv = torch.from_numpy(v).to(gpu)
    pack = torch.zeros(((n + 63) // 64) * 8, dtype=torch.uint8, device=gpu)
    comp = torch.zeros(n, dtype=torch.float32, device=gpu)
    quant = torch.zeros(n, dtype=torch.int16, device=gpu)
    cnt, lo, hi = C.c_size_t(), C.c_double(), C.c_double()
    torch.cuda.synchronize()
    call("ffcz_k_extract", 0, dv.data_ptr(), n, 1, 0.0, pack.data_ptr(), comp.data_ptr(), quant.data_ptr(),
         C.byref(cnt), C.byref(lo), C.byref(hi), None)
    ref = O.extract_one(v)
    k = cnt.value
    assert k == ref.compact.size
    assert np.array_equal(pack.cpu().numpy()[: (n + 7) // 8], ref.pack)
    assert np.array_equal(comp.cpu().numpy()[:k], ref.compact)
    assert np.array_equal(quant.cpu().numpy().view(np.uint16)[:k], ref.quant)
    assert lo.value == ref.vmin and hi.value == ref.vmax


def test_step_quantizer_and_certification(gpu):
    shape = (16, 20, 30)
    x, e0, eps, delta = binding_case(shape, seed=8)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.set_margins(0.125, 0.125)
        s.set_quantizer(QUANT_STEP, delta / 8, eps / 8)
        s.solve_all_abs(e0, 200, 1e-12)
        c1 = s.certify(False)
        c2 = s.certify(True)
        df = s.get_freq_edits_full()
        ds = s.get_spatial_edits_full()
        q = s.get_active_spatial_quant_edit_distances()
        rq = O.quantize_step(ds, eps / 8)
        rq = rq[rq != 0]
        assert np.array_equal(q, O.zigzag(rq).astype(np.uint16))
        # oracle certification of the same edits
        d1 = O.reconstruct_edit(shape, e0, df, ds)
        r1 = O.certify(d1, eps, delta)
        assert c1.ok == r1["ok"]
        assert c1.spatial_violations == r1["spatial_violations"]
        assert c1.freq_violations == r1["freq_violations"]
        dq = O.reconstruct_edit(shape, e0, O.dequantize_step(O.quantize_step(df, delta / 8), delta / 8),
                                O.dequantize_step(O.quantize_step(ds, eps / 8), eps / 8))
        r2 = O.certify(dq, eps, delta)
        assert c2.ok == r2["ok"]
        assert c2.spatial_violations == r2["spatial_violations"]
        assert c2.freq_violations == r2["freq_violations"]
        assert c1.ok and c2.ok


def test_certification_detects_violation(gpu):
    # no iterations: e0 violates the spatial bound, certification must say so
    shape = (6, 10, 12)
    x, e0, eps, delta = binding_case(shape, seed=1)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 0, 1e-7)
        c = s.certify(False)
        r = O.certify(O.reconstruct_edit(shape, e0, np.zeros(2 * s.F), np.zeros(s.N)), eps, delta)
        assert not c.ok
        assert c.spatial_violations == r["spatial_violations"] > 0
        assert c.freq_violations == r["freq_violations"] > 0
        assert abs(c.spatial_max_excess - r["spatial_max_excess"]) < 1e-15


def test_quantize_dequantize_kernel(gpu):
    rng = np.random.default_rng(0)
    v = (rng.standard_normal(10000) * 1e-3).astype(np.float32)
    v[::7] = 0
    lo, hi = float(v.min()), float(v.max())
    dv = torch.from_numpy(v).to(gpu)
    out = torch.empty_like(dv)
    torch.cuda.synchronize()
    call("ffcz_k_quantize_dequantize", 0, dv.data_ptr(), out.data_ptr(), lo, hi, v.size, 1e-7, None)
    ref = O.quantize_and_dequantize_ref(v, np.float32(lo), np.float32(hi))
    assert np.array_equal(out.cpu().numpy(), ref)


def test_ref16_quantized_certification_reads_device_range(gpu):
    """certify(quantized) with the reference quantiser: the dequantisers take
    min/max from the extraction on the device (no host round trip); counts
    match the oracle on the dequantised edits (projection_algorithm.cu:615-644)."""
    shape = (16, 20, 30)
    x, e0, eps, delta = binding_case(shape, seed=9)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.solve_all_abs(e0, 30, 1e-12)
        c2 = s.certify(True)
        df = s.get_freq_edits_full()
        ds = s.get_spatial_edits_full()
        fmin, fmax = float(df.min()), float(df.max())
        smin, smax = float(ds.min()), float(ds.max())
        dq = O.reconstruct_edit(shape, e0, O.quantize_and_dequantize_ref(df, fmin, fmax),
                                O.quantize_and_dequantize_ref(ds, smin, smax))
        r2 = O.certify(dq, eps, delta)
        assert c2.spatial_violations == r2["spatial_violations"]
        assert c2.freq_violations == r2["freq_violations"]
        assert c2.ok == r2["ok"]


def test_step_quantizer_overflow_is_reported(gpu):
    """Extraction only queues its kernels; a symbol beyond 16 bits surfaces as
    an error on the next result read (getter or quantised certification)."""
    shape = (6, 10, 12)
    x, e0, eps, delta = binding_case(shape, seed=2)
    with ProjectionSolver(*shape, dtype=np.float64) as s:
        s.initialize_all_abs(x, eps, delta)
        s.set_quantizer(QUANT_STEP, delta * 1e-9, eps * 1e-9)
        s.solve_all_abs(e0, 20, 1e-12)  # extraction runs at the end of solve_all_abs
        assert s.certify(False) is not None
        with pytest.raises(FFCzError):
            s.certify(True)
        with pytest.raises(FFCzError):
            s.get_active_spatial_quant_edit_distances()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(16, 20, 30), (32, 64, 128), (6, 10, 12), (9, 15, 25)])
@pytest.mark.parametrize("quant", ["ref16", "step"])
@pytest.mark.parametrize("paired", [True, False])
def test_certify_both_equals_the_two_certifications(gpu, shape, dtype, quant, paired):
    """The paired sequence (one y launch over both spectra; one z launch for
    both or one each) gives exactly the counts and maximal excesses of
    certify(False) and certify(True); frequency edits present (the C2R of df
    runs)."""
    from ffcz_amd import QUANT_REF16
    x, e0, eps, delta = binding_case(shape, seed=11, dtype=dtype, spread=1.4, dfrac=0.8)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.set_cert_pair(paired)
        s.initialize_all_abs(x, eps, delta)
        s.set_margins(0.125, 0.125)
        if quant == "step":
            s.set_quantizer(QUANT_STEP, delta / 8, eps / 8)
        else:
            s.set_quantizer(QUANT_REF16)
        s.solve_loop(MODE_ABS, e0, 60, 1e-9)
        s.extract_compact_representation()
        # called with the extraction still pending
        pair = s.certify_both()
        got = (s.get_spatial_flags_pack(), s.get_freq_flags_pack(), s.get_active_spatial_quant_edit_distances(),
               s.get_active_freq_quant_edit_distances(), s.get_active_spatial_edit_distances())
        single = (s.certify(False), s.certify(True))
        assert pair == single
        assert any(v.freq_violations or v.spatial_violations or v.ok for v in pair)
        # and that extraction is the one made on its own
        s.extract_compact_representation()
        again = (s.get_spatial_flags_pack(), s.get_freq_flags_pack(), s.get_active_spatial_quant_edit_distances(),
                 s.get_active_freq_quant_edit_distances(), s.get_active_spatial_edit_distances())
        assert all(np.array_equal(a, b) for a, b in zip(got, again))


def _expected_extraction(edits, step):
    """The oracle's extraction of one edit array: the reference quantiser's
    (step None) or the step quantiser's, which flags the values whose symbol
    is non-zero; the range is the full array's in both."""
    if step is None:
        return O.extract_one(edits)
    q = O.quantize_step(edits, step)
    flags = q != 0
    return O.Extraction(flags=flags, pack=O.pack_flags(flags), compact=edits[flags],
                        quant=O.zigzag(q[flags]).astype(np.uint16), vmin=edits.min(), vmax=edits.max())


@pytest.mark.parametrize("quant", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_extraction_of_untouched_frequency_edits(gpu, dtype, quant):
    # a solve in which no frequency clamp fires leaves df at zero, and the
    # extraction writes its results without reading df: they must be what the
    # full extraction of an all-zero df gives, also after an extraction whose
    # flag pack was not zero
    shape = (6, 10, 12)
    x, e0, eps, delta = binding_case(shape, seed=11, dtype=dtype)
    with ProjectionSolver(*shape, dtype=dtype) as s:
        # (bound, reset df first, df non-zero after the solve): the edits of a
        # solve add to those kept from the previous one unless reset
        for d, reset, fires in ((delta, False, True), (delta * 1e6, False, True), (delta * 1e6, True, False)):
            s.initialize_all_abs(x, eps, d)
            if reset:
                s.reset_edit_accumulators()
            if quant:
                s.set_quantizer(QUANT_STEP, d / 8, eps / 8)
            s.solve_all_abs(e0, 20, 1e-7)
            df = s.get_freq_edits_full()
            assert bool(np.any(df != 0)) == fires
            for got, edits, step in (
                ((s.get_freq_flags(), s.get_freq_flags_pack(), s.get_active_freq_edit_distances(),
                  s.get_active_freq_quant_edit_distances(), s.get_min_freq_edit(), s.get_max_freq_edit()),
                 df, d / 8 if quant else None),
                ((s.get_spatial_flags(), s.get_spatial_flags_pack(), s.get_active_spatial_edit_distances(),
                  s.get_active_spatial_quant_edit_distances(), s.get_min_spatial_edit(), s.get_max_spatial_edit()),
                 s.get_spatial_edits_full(), eps / 8 if quant else None),
            ):
                ref = _expected_extraction(edits, step)
                flags, pack, comp, qsym, lo, hi = got
                assert np.array_equal(flags, ref.flags)
                assert np.array_equal(pack, ref.pack)
                assert np.array_equal(comp, ref.compact)
                assert np.array_equal(qsym, ref.quant)
                assert lo == ref.vmin and hi == ref.vmax


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_queued_certification_across_solves(gpu, dtype):
    """certify_both_enqueue / certify_both_results: a field's certification
    queued, the next field's solve queued behind it, then its results read
    (the pipelined order of the benchmark's timed steps).  Each equals the
    synchronous certify_both of the same solve; a second certification before
    the results are read, or a read with nothing queued, is refused."""
    shape = (16, 20, 30)
    cases = [binding_case(shape, seed=s, dtype=dtype, spread=1.4, dfrac=0.8) for s in (21, 22)]
    x, _, eps, delta = cases[0]
    with ProjectionSolver(*shape, dtype=dtype) as s:
        s.initialize_all_abs(x, eps, delta)
        s.set_margins(0.125, 0.125)
        s.set_quantizer(QUANT_STEP, delta / 8, eps / 8)
        expected = []
        for _, e0, _, _ in cases:
            s.reset_edit_accumulators()
            s.solve_loop(MODE_ABS, e0, 60, 1e-9)
            s.extract_compact_representation()
            expected.append(s.certify_both())
        with pytest.raises(FFCzError):
            s.certify_both_results()
        got = []
        for k, (_, e0, _, _) in enumerate(cases):
            s.reset_edit_accumulators()
            s.solve_loop(MODE_ABS, e0, 60, 1e-9)
            if k > 0:
                got.append(s.certify_both_results())
            s.extract_compact_representation()
            s.certify_both_enqueue()
            if k == 0:
                with pytest.raises(FFCzError):
                    s.certify_both_enqueue()
                with pytest.raises(FFCzError):
                    s.certify(False)
        got.append(s.certify_both_results())
        assert got == expected
        assert any(v.freq_violations or v.spatial_violations or v.ok for pair in got for v in pair)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_borrowed_initial_error(gpu, dtype):
    """set_borrow_e0: a device initial error read in place by the
    certifications instead of copied gives the same iterations, edits and
    certifications as the copying default; a host initial error is copied
    either way."""
    shape = (16, 20, 30)
    x, e0, eps, delta = binding_case(shape, seed=31, dtype=dtype, spread=1.4, dfrac=0.8)
    e0_dev = torch.from_numpy(e0).cuda()
    out = []
    for borrow, src in ((False, e0_dev), (True, e0_dev), (True, e0)):
        with ProjectionSolver(*shape, dtype=dtype) as s:
            s.set_borrow_e0(borrow)
            s.initialize_all_abs(x, eps, delta)
            s.set_margins(0.125, 0.125)
            s.set_quantizer(QUANT_STEP, delta / 8, eps / 8)
            s.solve_loop(MODE_ABS, src, 60, 1e-9)
            s.extract_compact_representation()
            out.append((s.iteration_count, s.get_spatial_edits_full(), s.get_freq_edits_full(), s.certify_both()))
    for o in out[1:]:
        assert o[0] == out[0][0]
        assert np.array_equal(o[1], out[0][1]) and np.array_equal(o[2], out[0][2])
        assert o[3] == out[0][3]
    assert any(v.freq_violations or v.spatial_violations or v.ok for v in out[0][3])
