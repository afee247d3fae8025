"""CPU oracle for the FFCz alternating-projection post-processor.

TEST INFRASTRUCTURE ONLY.  This module checks the GPU results: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``ffcz_amd``) never routes through it.

It restates, in numpy, the arithmetic of the reference FFCz solver
(``projection_algorithm.cu`` / ``projection_algorithm_impl.cuh`` /
``HuffmanZSTDCoder.cuh``).  Every function cites the reference lines it follows.
The FFT is numpy/scipy pocketfft: FP64 for the exact-arithmetic semantics, or
FP32 (``dtype=np.float32``) to mimic a single-precision iterate.  cuFFT's own
rounding cannot be reproduced, so the FP32 iterate is compared norm-wise.

Pinning (what checks this oracle independently of its own formulas) lives in
``tests/test_oracle.py``: brute-force DFT-matrix loops on tiny shapes, closed
forms of the loop when one bound is inactive, the hand-derived Huffman
known-answer vector of the reference's own test stub (``HuffmanZSTDCoder.cuh``
lines 472-518), brute-force bit packing and quantizer identities.
"""
from __future__ import annotations

import heapq
from dataclasses import dataclass, field

import numpy as np

try:  # scipy's pocketfft takes ``workers`` (multi-threaded baseline)
    import scipy.fft as _fft

    _HAVE_SCIPY = True
except Exception:  # pragma: no cover
    import numpy.fft as _fft

    _HAVE_SCIPY = False

FLAG_THRESHOLD = 1e-7  # projection_algorithm_impl.cuh:389,398
QMAX = (1 << 16) - 1  # projection_algorithm.cu:450,454


# ---------------------------------------------------------------------------
# FFT helpers (cuFFT conventions: forward e^{-i}, inverse unnormalised)
# ---------------------------------------------------------------------------
def rfft3(e: np.ndarray, workers: int = 1) -> np.ndarray:
    """R2C of a row-major (Nx,Ny,Nz) array -> (Nx,Ny,Nz//2+1); cuFFT R2C/D2Z
    layout used at projection_algorithm_impl.cuh:207-213."""
    if _HAVE_SCIPY:
        return _fft.rfftn(e, axes=(0, 1, 2), workers=workers)
    return _fft.rfftn(e, axes=(0, 1, 2))


def irfft3_unnorm(E: np.ndarray, shape, workers: int = 1) -> np.ndarray:
    """Unnormalised C2R (cuFFT C2R/Z2D, projection_algorithm_impl.cuh:231-237)."""
    if _HAVE_SCIPY:
        return _fft.irfftn(E, s=shape, axes=(0, 1, 2), norm="forward", workers=workers)
    return _fft.irfftn(E, s=shape, axes=(0, 1, 2), norm="forward")


# ---------------------------------------------------------------------------
# Clamp kernels
# ---------------------------------------------------------------------------
def _clamp_component(v: np.ndarray, acc: np.ndarray, lo, hi) -> None:
    """In-place box clamp of ``v`` to [lo, hi]; ``acc += clamped - v`` where it
    fires.  Edit computed first, then accumulated (one writer per address),
    as in projection_algorithm.cu:161-172 and :351-360."""
    over = v > hi
    under = v < lo
    if np.ndim(hi) == 0:
        acc[over] += hi - v[over]
        v[over] = hi
        acc[under] += lo - v[under]
        v[under] = lo
    else:
        acc[over] += hi[over] - v[over]
        v[over] = hi[over]
        acc[under] += lo[under] - v[under]
        v[under] = lo[under]


def project_frequency_all_abs(E: np.ndarray, delta, df: np.ndarray) -> None:
    """projection_algorithm.cu:147-188.  ``E`` complex half spectrum (flat view
    of F bins), ``df`` split accumulator [Re(F) | Im(F)] (cu:164,179)."""
    T = df.dtype.type
    F = E.size
    Ef = E.reshape(-1)
    re = Ef.real.copy()
    im = Ef.imag.copy()
    d = T(delta)
    _clamp_component(re, df[:F], -d, d)
    _clamp_component(im, df[F:], -d, d)
    Ef.real = re
    Ef.imag = im


def dense_bounds(F: int, idx, delta_k, dtype) -> np.ndarray:
    """Per-bin bound array of the direct mode: ``delta_k[j]`` on bin
    ``idx[j]``, +inf on bins no constraint lists, the smallest bound where an
    index repeats.  ``idx=None`` means delta_k already lists every bin in
    order (bin j has bound delta_k[j])."""
    if idx is None:
        dense = np.asarray(delta_k, dtype=dtype).reshape(-1)
        if dense.size != F:
            raise ValueError("a dense bound array needs one bound per bin")
        return dense
    dense = np.full(F, np.inf, dtype=dtype)
    np.minimum.at(dense, np.asarray(idx, dtype=np.int64), np.asarray(delta_k, dtype=dtype))
    return dense


def project_frequency_direct(E: np.ndarray, idx: np.ndarray, delta_k: np.ndarray, df: np.ndarray,
                             dense: np.ndarray | None = None) -> None:
    """projection_algorithm.cu:93-145: per-constraint bound ``delta_k`` on the
    bins listed in ``idx``; other bins are unconstrained.  Duplicate indices are
    resolved to the smallest bound (the fixed point of repeated application).
    ``dense`` (optional) is the per-bin bound array of ``dense_bounds``."""
    T = df.dtype.type
    F = E.size
    if dense is None:
        dense = dense_bounds(F, idx, delta_k, df.dtype)
    Ef = E.reshape(-1)
    re = Ef.real.copy()
    im = Ef.imag.copy()
    _clamp_component(re, df[:F], -dense, dense)
    _clamp_component(im, df[F:], -dense, dense)
    Ef.real = re
    Ef.imag = im
    del T


def project_frequency_all_ptw(E: np.ndarray, orig: np.ndarray, delta_ptw, df: np.ndarray) -> None:
    """projection_algorithm.cu:190-259 ("pointwise relative" power-spectrum mode).

    DC: if E0 is outside [X0*(1/sqrt(1+d)-1), X0*(1/sqrt(1-d)-1)] the DC error
    is zeroed (cu:204-216).  Other bins: ratio |X+E|^2/|X|^2 tested against
    sqrt(1 +- d) (cu:219-226); when outside, Re/Im are clamped to the fixed
    bound min(1-sqrt(1-d), sqrt(1+d)-1)/sqrt(2) (cu:228-258)."""
    T = df.dtype.type
    F = E.size
    Ef = E.reshape(-1)
    Of = orig.reshape(-1)
    dp = T(delta_ptw)
    one = T(1)
    re = Ef.real.copy()
    im = Ef.imag.copy()
    # DC (cu:204-216)
    lo = T(Of[0].real) * (one / np.sqrt(one + dp) - one)
    hi = T(Of[0].real) * (one / np.sqrt(one - dp) - one)
    dc_err = re[0]
    dc_fire = (dc_err > hi) or (dc_err < lo)
    # non-DC (cu:219-258)
    omag = Of.real * Of.real + Of.imag * Of.imag
    cr = Of.real + re
    ci = Of.imag + im
    cmag = cr * cr + ci * ci
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = cmag / omag
    inside = (ratio <= np.sqrt(one + dp)) & (ratio >= np.sqrt(one - dp))
    # sqrt_val(2.0f) deduces float even when T is double (cu:228)
    dl = T(min(one - np.sqrt(one - dp), np.sqrt(one + dp) - one) / T(np.sqrt(np.float32(2.0))))
    act = ~inside
    act[0] = False
    sub_re = re[act]
    sub_im = im[act]
    acc_re = df[:F][act]
    acc_im = df[F:][act]
    _clamp_component(sub_re, acc_re, -dl, dl)
    _clamp_component(sub_im, acc_im, -dl, dl)
    re[act] = sub_re
    im[act] = sub_im
    df[:F][act] = acc_re
    df[F:][act] = acc_im
    if dc_fire:
        df[0] += -dc_err
        re[0] = 0
    Ef.real = re
    Ef.imag = im


def project_frequency_partial_ptw(E: np.ndarray, orig: np.ndarray, idx, delta_ptw, df: np.ndarray) -> None:
    """project_frequency_constraints_partial_ptw, projection_algorithm.cu:261-333,
    which the reference's projection_algorithm_impl.cuh:303-333 launches over
    initialize_partial_ptw's index list (impl:701-738).  Constraint 0 is the
    DC rule of cu:204-216 on bin 0 whatever idx[0] holds (cu:276-289); every
    later constraint j applies the ratio rule of cu:219-258 to bin idx[j].  A
    bin listed twice is projected once; a later entry naming bin 0 is left to
    the DC rule (in the reference both would race on bin 0)."""
    T = df.dtype.type
    F = E.size
    Ef = E.reshape(-1)
    idx = np.asarray(idx, dtype=np.int64).reshape(-1)
    if idx.size == 0:
        return
    sel = np.unique(idx[1:])
    sel = sel[sel != 0]
    # the all-ptw rule on the selected bins, and the DC rule on bin 0
    sub = np.concatenate([[0], sel])
    Es = Ef[sub].copy()
    dfs = np.concatenate([df[:F][sub], df[F:][sub]])
    project_frequency_all_ptw(Es, np.asarray(orig).reshape(-1)[sub], delta_ptw, dfs)
    n = sub.size
    Ef[sub] = Es
    df[:F][sub] = dfs[:n]
    df[F:][sub] = dfs[n:]
    del T


def project_frequency_ptw_corrected(E: np.ndarray, orig: np.ndarray, delta_ptw, df: np.ndarray, mask=None) -> None:
    """The pointwise power-spectrum bound with the defects of cu:199-258 (D11)
    corrected; not the reference's arithmetic.  Every constrained bin,
    DC included, must keep its power ratio |X+E|^2/|X|^2 within [1-d, 1+d]
    (the reference tests the squared ratio against sqrt(1 +- d)).  Outside
    it, Re and Im of E are clamped to +-d'|X| with
    d' = min(1 - sqrt(1-d), sqrt(1+d) - 1) / sqrt(2) (the reference clamps to
    an absolute d'), which puts |X+E'| within [sqrt(1-d), sqrt(1+d)]|X|.
    mask: the constrained bins (None: all); edits accumulate new - old."""
    T = df.dtype.type
    F = E.size
    Ef = E.reshape(-1)
    Of = np.asarray(orig).reshape(-1)
    dp = T(delta_ptw)
    one = T(1)
    re = Ef.real.copy()
    im = Ef.imag.copy()
    omag = (Of.real * Of.real + Of.imag * Of.imag).astype(df.dtype)
    cr = Of.real + re
    ci = Of.imag + im
    cmag = cr * cr + ci * ci
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = cmag / omag
    inside = (ratio <= one + dp) & (ratio >= one - dp)
    dl = T(min(one - np.sqrt(one - dp), np.sqrt(one + dp) - one) / np.sqrt(T(2)))
    lvl = (dl * np.sqrt(omag)).astype(df.dtype)
    act = ~inside
    if mask is not None:
        act &= np.asarray(mask, bool).reshape(-1)
    sub_re, sub_im = re[act], im[act]
    acc_re, acc_im = df[:F][act], df[F:][act]
    lv = lvl[act]
    _clamp_component(sub_re, acc_re, -lv, lv)
    _clamp_component(sub_im, acc_im, -lv, lv)
    re[act], im[act] = sub_re, sub_im
    df[:F][act], df[F:][act] = acc_re, acc_im
    Ef.real = re
    Ef.imag = im


def project_spatial(e: np.ndarray, eps, ds: np.ndarray) -> None:
    """projection_algorithm.cu:335-361 (clamp e to [-eps, eps], ds += delta)."""
    T = ds.dtype.type
    _clamp_component(e.reshape(-1), ds.reshape(-1), -T(eps), T(eps))


def convergence_metric(e: np.ndarray, eps):
    """max_i max(|e_i|-eps, 0): projection_algorithm.cu:586-612 reduced on the
    host as in projection_algorithm_impl.cuh:762-766."""
    T = e.dtype.type
    a = np.abs(e)
    v = np.where(a > T(eps), a - T(eps), T(0))
    return T(v.max()) if v.size else T(0)


# ---------------------------------------------------------------------------
# Solver loop
# ---------------------------------------------------------------------------
@dataclass
class SolveResult:
    e: np.ndarray
    df: np.ndarray
    ds: np.ndarray
    iterations: int
    violation: float
    history: list = field(default_factory=list)


def solve(x: np.ndarray, e0: np.ndarray, eps, *, mode: str = "abs", delta=None,
          idx=None, delta_k=None, delta_ptw=None, max_iter: int = 100, tol=1e-7,
          dtype=np.float64, df=None, ds=None, workers: int = 1, ptw_corrected: bool = False) -> SolveResult:
    """Loop of ``solve_all_abs`` / ``solve`` / ``solve_all_ptw``
    (projection_algorithm_impl.cuh:835-852, 791-808, 879-896):

        for it < max_iter:
            E = R2C(e); clamp E (mode); e = C2R(E) * (T(1)/T(N))   (impl:202-249)
            v = max(|e|-eps, 0); iterations += 1                    (impl:841-843)
            if v < tol: break                                       (impl:845)
            spatial clamp of e                                      (impl:851)

    ``df``/``ds`` continue an existing accumulation when given (a second solve
    on the same solver keeps accumulating: reset only in the ctor, impl:60).
    """
    T = np.dtype(dtype).type
    shape = x.shape
    N = int(np.prod(shape))
    F = shape[0] * shape[1] * (shape[2] // 2 + 1)
    e = np.array(e0, dtype=dtype, copy=True)
    df = np.zeros(2 * F, dtype=dtype) if df is None else df
    ds = np.zeros(N, dtype=dtype) if ds is None else ds
    norm = T(1) / T(N)  # impl:242
    orig = rfft3(x.astype(dtype), workers) if mode in ("ptw", "partial_ptw") else None
    pmask = None
    if mode == "partial_ptw" and ptw_corrected:
        pmask = np.zeros(F, bool)
        pmask[np.asarray(idx, dtype=np.int64)] = True
    dense = dense_bounds(F, idx, delta_k, dtype) if mode == "direct" else None
    it = 0
    v = T(0)
    hist = []
    for _ in range(max_iter):
        E = rfft3(e, workers)
        if mode == "abs":
            project_frequency_all_abs(E, delta, df)
        elif mode == "direct":
            project_frequency_direct(E, idx, delta_k, df, dense=dense)
        elif mode == "ptw" and not ptw_corrected:
            project_frequency_all_ptw(E, orig, delta_ptw, df)
        elif mode == "ptw":
            project_frequency_ptw_corrected(E, orig, delta_ptw, df)
        elif mode == "partial_ptw" and not ptw_corrected:
            project_frequency_partial_ptw(E, orig, idx, delta_ptw, df)
        elif mode == "partial_ptw":
            project_frequency_ptw_corrected(E, orig, delta_ptw, df, mask=pmask)
        else:
            raise ValueError(mode)
        e = (irfft3_unnorm(E, shape, workers) * norm).astype(dtype, copy=False)
        v = convergence_metric(e, eps)
        hist.append(float(v))
        it += 1
        if v < T(tol):
            break
        project_spatial(e, eps, ds)
    return SolveResult(e=e, df=df, ds=ds, iterations=it, violation=float(v), history=hist)


# ---------------------------------------------------------------------------
# Extraction: flags, min/max, compaction, quantisation, bit packing
# ---------------------------------------------------------------------------
def quantize_ref(v: np.ndarray, vmin, vmax) -> np.ndarray:
    """16-bit quantiser of projection_algorithm.cu:447-455 in type T:
    n=(v-min)/(max-min); q = n>0.5 ? ceil(n*65535) : floor(n*65535)."""
    T = np.asarray(v).dtype.type
    n = (v - T(vmin)) / (T(vmax) - T(vmin))
    s = n * T(QMAX)
    q = np.where(n > 0.5, np.ceil(s), np.floor(s))
    return q.astype(np.uint16)


def dequantize_ref(q: np.ndarray, vmin, vmax, dtype) -> np.ndarray:
    """projection_algorithm.cu:638: q/65535*(max-min)+min in type T."""
    T = np.dtype(dtype).type
    return (q.astype(dtype) / T(QMAX) * (T(vmax) - T(vmin)) + T(vmin)).astype(dtype)


def quantize_and_dequantize_ref(v: np.ndarray, vmin, vmax, threshold=FLAG_THRESHOLD) -> np.ndarray:
    """projection_algorithm.cu:615-644 (dense: 0 where |v| <= threshold)."""
    T = v.dtype.type
    out = np.zeros_like(v)
    m = np.abs(v) > T(threshold)
    out[m] = dequantize_ref(quantize_ref(v[m], vmin, vmax), vmin, vmax, v.dtype)
    return out


def quantize_step(v: np.ndarray, step: float) -> np.ndarray:
    """Uniform quantiser with a given step (the edit-quantisation step of the
    configs, e.g. eps_s/8): q = llround(v/step) in FP64, ties away from zero.
    Not in the reference (its quantiser is ``quantize_ref``)."""
    r = np.asarray(v, dtype=np.float64) / np.float64(step)
    return (np.sign(r) * np.floor(np.abs(r) + 0.5)).astype(np.int64)


def dequantize_step(q: np.ndarray, step: float) -> np.ndarray:
    return np.asarray(q, dtype=np.float64) * np.float64(step)


def zigzag(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, dtype=np.int64)
    return ((q << 1) ^ (q >> 63)).astype(np.uint64)


def pack_flags(flags: np.ndarray) -> np.ndarray:
    """LSB-first packing: bit i of byte b is flag[8b+i] (the intent of
    projection_algorithm.cu:381-405; its bound bug D1 is not reproduced)."""
    return np.packbits(np.asarray(flags, dtype=bool), bitorder="little")


@dataclass
class Extraction:
    flags: np.ndarray
    pack: np.ndarray
    compact: np.ndarray
    quant: np.ndarray
    vmin: float
    vmax: float


def extract_one(edits: np.ndarray, threshold=FLAG_THRESHOLD) -> Extraction:
    """extract_active_flags (cu:363-379) + compact_and_quantize_active_elements
    (cu:459-550): min/max over the FULL array (zeros included, cu:482-487),
    stable compaction in index order, reference quantiser, packed flags."""
    T = edits.dtype.type
    flags = np.abs(edits) > T(threshold)
    vmin = edits.min() if edits.size else T(0)
    vmax = edits.max() if edits.size else T(0)
    comp = edits[flags]
    quant = quantize_ref(comp, vmin, vmax) if comp.size else np.zeros(0, np.uint16)
    return Extraction(flags=flags, pack=pack_flags(flags), compact=comp, quant=quant,
                      vmin=vmin, vmax=vmax)


# ---------------------------------------------------------------------------
# Huffman (bit-exact restatement of HuffmanZSTDCoder.cuh:48-357)
# ---------------------------------------------------------------------------
def huffman_codebook(symbols: np.ndarray):
    """Returns (sorted unique symbols, codes uint64, lengths int).

    Leaves are the distinct symbols in ascending order (thrust::sort +
    reduce_by_key, Huffman:58-82).  A min-heap of (freq, node) pairs pops the
    minimum as ``left`` and the next as ``right``; internal nodes are numbered
    n, n+1, ... (Huffman:113-150).  Codes from a DFS, left=0 right=1, MSB first;
    a single symbol gets code 0 / length 1 (Huffman:161-177)."""
    sym, cnt = np.unique(np.asarray(symbols), return_counts=True)
    n = len(sym)
    if n == 0:
        raise ValueError("Empty data")
    if n == 1:
        return sym, np.zeros(1, np.uint64), np.ones(1, np.int64)
    left = [-1] * (2 * n - 1)
    right = [-1] * (2 * n - 1)
    heap = [(int(c), i) for i, c in enumerate(cnt)]
    heapq.heapify(heap)
    nxt = n
    while len(heap) > 1:
        f1, a = heapq.heappop(heap)
        f2, b = heapq.heappop(heap)
        left[nxt] = a
        right[nxt] = b
        heapq.heappush(heap, (f1 + f2, nxt))
        nxt += 1
    root = heap[0][1]
    codes = np.zeros(n, np.uint64)
    lens = np.zeros(n, np.int64)
    stack = [(root, 0, 0)]
    while stack:
        node, code, ln = stack.pop()
        if node < n:
            codes[node] = code
            lens[node] = ln
        else:
            stack.append((left[node], code << 1, ln + 1))
            stack.append((right[node], (code << 1) | 1, ln + 1))
    return sym, codes, lens


def huffman_encode(symbols: np.ndarray):
    """Bitstream of Huffman:307-357: codes in input order, each MSB first,
    concatenated MSB first, last byte zero padded.  Returns (bytes, table)
    where table=(sym, codes, lens).  Raises if a code exceeds 32 bits (the
    reference stores codes in uint32, Huffman:159)."""
    symbols = np.asarray(symbols)
    sym, codes, lens = huffman_codebook(symbols)
    if lens.max() > 32:
        raise ValueError("Huffman code longer than 32 bits")
    pos = np.searchsorted(sym, symbols)
    L = lens[pos]
    C = codes[pos]
    total = int(L.sum())
    nbytes = (total + 7) // 8
    bits = np.zeros(nbytes * 8, np.uint8)
    offs = np.concatenate([[0], np.cumsum(L)[:-1]]).astype(np.int64)
    for b in range(int(lens.max())):
        m = L > b
        # bit b (from the MSB) of each code of length L
        bitval = (C[m] >> (L[m] - 1 - b).astype(np.uint64)) & np.uint64(1)
        bits[offs[m] + b] = bitval.astype(np.uint8)
    return np.packbits(bits, bitorder="big"), (sym, codes, lens)


def huffman_decode(stream: np.ndarray, table, num_elements: int) -> np.ndarray:
    """Prefix decode, as Huffman:404-447 (bit-serial, lengths 1..32)."""
    sym, codes, lens = table
    rev = {(int(l), int(c)): s for s, c, l in zip(sym, codes, lens)}
    bits = np.unpackbits(np.asarray(stream, np.uint8), bitorder="big")
    out = []
    p = 0
    nb = bits.size
    while len(out) < num_elements and p < nb:
        code = 0
        found = False
        for ln in range(1, 33):
            if p + ln > nb:
                break
            code = (code << 1) | int(bits[p + ln - 1])
            s = rev.get((ln, code))
            if s is not None:
                out.append(s)
                p += ln
                found = True
                break
        if not found:
            p += 1
    return np.asarray(out, dtype=np.asarray(sym).dtype)


# ---------------------------------------------------------------------------
# FP64 certification (not in the reference; the nearest is impl:741-770)
# ---------------------------------------------------------------------------
def reconstruct_edit(shape, e0, df, ds, workers: int = 1) -> np.ndarray:
    """x_hat - x = e0 + C2R(df)/N + ds in FP64 (the linearity identity of the
    loop: every iteration adds df-clamp deltas in frequency, ds in space)."""
    N = int(np.prod(shape))
    F = shape[0] * shape[1] * (shape[2] // 2 + 1)
    d = np.asarray(e0, np.float64).reshape(shape).copy()
    df = np.asarray(df, np.float64)
    if np.any(df != 0):
        spec = (df[:F] + 1j * df[F:]).reshape(shape[0], shape[1], shape[2] // 2 + 1)
        d += irfft3_unnorm(spec, shape, workers) / N
    d += np.asarray(ds, np.float64).reshape(shape)
    return d


def certify(d: np.ndarray, eps, delta=None, delta_dense=None, workers: int = 1):
    """Both bounds in FP64: |d_i| <= eps for every element and |Re|,|Im| of
    F(d)_k <= delta(_k) for every half-spectrum bin.  Returns a dict with the
    violation counts and the largest excesses."""
    d = np.asarray(d, np.float64)
    sp_ex = np.abs(d) - float(eps)
    D = rfft3(d, workers).reshape(-1)
    dl = np.full(D.size, float(delta)) if delta_dense is None else np.asarray(delta_dense, np.float64)
    fr_ex = np.maximum(np.abs(D.real), np.abs(D.imag)) - dl
    return dict(
        spatial_violations=int((sp_ex > 0).sum()),
        freq_violations=int((fr_ex > 0).sum()),
        spatial_max_excess=float(max(sp_ex.max(), 0.0)),
        freq_max_excess=float(max(np.nanmax(fr_ex), 0.0)),
        ok=bool((sp_ex <= 0).all() and (fr_ex <= 0).all()),
    )
