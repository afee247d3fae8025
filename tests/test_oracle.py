"""Pins the oracle (oracle/ffcz_oracle.py) with checks that do not retype its
formulas: brute-force DFT matrices, closed forms of the loop when one bound is
inactive, invariants, hand-derived known answers and brute-force bit packing.
The reference ships no tests or fixtures for this path; the only test-like
code is the commented round-trip stub of HuffmanZSTDCoder.cuh:472-518, whose
input is the Huffman known-answer vector below."""
import itertools

import numpy as np
import pytest

from oracle import ffcz_oracle as O
from tests._cases import binding_case, direct_case


# ---------------------------------------------------------------------------
# brute-force DFT (independent of pocketfft)
# ---------------------------------------------------------------------------
def dft_matrix(n, sign=-1):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def brute_rfft3(x):
    nx, ny, nz = x.shape
    X = np.einsum("ai,ijk->ajk", dft_matrix(nx), x.astype(np.complex128))
    X = np.einsum("bj,ajk->abk", dft_matrix(ny), X)
    X = np.einsum("ck,abk->abc", dft_matrix(nz), X)
    return X[:, :, : nz // 2 + 1]


def brute_irfft3_unnorm(H, shape):
    """Unnormalised C2R: Hermitian-extend the half spectrum, full inverse DFT,
    real part (the imaginary parts of self-conjugate bins are dropped, as a
    C2R does for a consistent input)."""
    nx, ny, nz = shape
    full = np.zeros(shape, np.complex128)
    full[:, :, : nz // 2 + 1] = H
    for kz in range(nz // 2 + 1, nz):
        full[:, :, kz] = np.conj(H[(-np.arange(nx)) % nx][:, (-np.arange(ny)) % ny][:, :, nz - kz])
    # symmetrise the self-conjugate planes the way C2R reads them
    for kz in {0, nz // 2} if nz % 2 == 0 else {0}:
        P = full[:, :, kz]
        Pc = np.conj(P[(-np.arange(nx)) % nx][:, (-np.arange(ny)) % ny])
        full[:, :, kz] = 0.5 * (P + Pc)
    y = np.einsum("ai,ijk->ajk", dft_matrix(nx, +1), full)
    y = np.einsum("bj,ajk->abk", dft_matrix(ny, +1), y)
    y = np.einsum("ck,abk->abc", dft_matrix(nz, +1), y)
    return y.real


@pytest.mark.parametrize("shape", [(1, 4, 6), (2, 3, 5), (3, 4, 4), (2, 2, 1), (1, 1, 7)])
def test_fft_wrappers_match_brute_force(shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape)
    assert np.allclose(O.rfft3(x), brute_rfft3(x), atol=1e-12)
    H = O.rfft3(x)
    assert np.allclose(O.irfft3_unnorm(H, shape), brute_irfft3_unnorm(H, shape), atol=1e-11)
    assert np.allclose(O.irfft3_unnorm(H, shape) / x.size, x, atol=1e-13)


def test_fft_closed_forms():
    shape = (4, 6, 8)
    x = np.zeros(shape)
    x[0, 0, 0] = 1
    assert np.allclose(O.rfft3(x), 1.0)
    n = np.arange(8)
    c = np.broadcast_to(np.cos(2 * np.pi * 2 * n / 8), shape).copy()
    C = O.rfft3(c)
    expect = np.zeros_like(C)
    expect[0, 0, 2] = c.size / 2
    assert np.allclose(C, expect, atol=1e-9)
    rng = np.random.default_rng(1)
    y = rng.standard_normal(shape)
    Y = np.fft.fftn(y)  # Parseval on the full spectrum
    assert np.isclose((np.abs(Y) ** 2).sum() / y.size, (y ** 2).sum())


# ---------------------------------------------------------------------------
# the loop
# ---------------------------------------------------------------------------
def brute_solve(x, e0, eps, delta, max_iter, tol, dense=None):
    """Loop of projection_algorithm_impl.cuh:835-852 with DFT matrices and
    element-wise Python clamps."""
    shape = x.shape
    N = x.size
    F = shape[0] * shape[1] * (shape[2] // 2 + 1)
    e = e0.astype(np.float64).copy()
    df = np.zeros(2 * F)
    ds = np.zeros(N)
    it = 0
    for _ in range(max_iter):
        E = brute_rfft3(e).reshape(-1)
        for i in range(F):
            d = delta if dense is None else dense[i]
            re, im = E[i].real, E[i].imag
            if re > d:
                df[i] += d - re
                re = d
            elif re < -d:
                df[i] += -d - re
                re = -d
            if im > d:
                df[i + F] += d - im
                im = d
            elif im < -d:
                df[i + F] += -d - im
                im = -d
            E[i] = complex(re, im)
        e = brute_irfft3_unnorm(E.reshape(shape[0], shape[1], -1), shape) * (1.0 / N)
        v = max(max(abs(t) - eps, 0.0) for t in e.ravel())
        it += 1
        if v < tol:
            break
        flat = e.reshape(-1)
        for i in range(N):
            if flat[i] > eps:
                ds[i] += eps - flat[i]
                flat[i] = eps
            elif flat[i] < -eps:
                ds[i] += -eps - flat[i]
                flat[i] = -eps
    return e, df, ds, it


@pytest.mark.parametrize("shape", [(1, 4, 6), (2, 3, 4), (3, 4, 5)])
def test_loop_matches_brute_force(shape):
    x, e0, eps, delta = binding_case(shape, seed=4, spread=4.0, dfrac=0.9)
    r = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=12, tol=1e-13)
    e, df, ds, it = brute_solve(x, e0, eps, delta, 12, 1e-13)
    assert r.iterations == it
    assert np.allclose(r.e, e, atol=1e-12)
    assert np.allclose(r.df, df, atol=1e-12)
    assert np.allclose(r.ds, ds, atol=1e-12)
    assert np.any(df != 0) and np.any(ds != 0)


def test_direct_loop_matches_brute_force():
    shape = (2, 3, 4)
    x, e0, eps, idx, dk = direct_case(shape, seed=3)
    F = 2 * 3 * 3
    dense = np.full(F, np.inf)
    for i, d in zip(idx, dk):
        dense[int(i)] = min(dense[int(i)], d)
    r = O.solve(x, e0, eps, mode="direct", idx=idx, delta_k=dk, max_iter=10, tol=1e-13)
    e, df, ds, it = brute_solve(x, e0, eps, None, 10, 1e-13, dense=dense)
    assert r.iterations == it
    assert np.allclose(r.df, df, atol=1e-12) and np.allclose(r.ds, ds, atol=1e-12)


def test_loop_inactive_frequency_bound_closed_form():
    shape = (3, 4, 6)
    x, e0, eps, _ = binding_case(shape, seed=1)
    r = O.solve(x, e0, eps, mode="abs", delta=1e30, max_iter=10, tol=1e-12)
    assert r.iterations == 2
    assert not r.df.any()
    assert np.allclose(r.ds, (np.clip(e0, -eps, eps) - e0).ravel(), atol=1e-15)


def test_loop_inactive_spatial_bound_closed_form():
    shape = (3, 4, 6)
    x, e0, _, delta = binding_case(shape, seed=2)
    r = O.solve(x, e0, 1e30, mode="abs", delta=delta, max_iter=10, tol=1e-12)
    assert r.iterations == 1
    E0 = brute_rfft3(e0).reshape(-1)
    clamped = np.clip(E0.real, -delta, delta) + 1j * np.clip(E0.imag, -delta, delta)
    F = E0.size
    assert np.allclose(r.df[:F], clamped.real - E0.real, atol=1e-12)
    assert np.allclose(r.df[F:], clamped.imag - E0.imag, atol=1e-12)
    assert np.allclose(r.e, brute_irfft3_unnorm(clamped.reshape(3, 4, 4), shape) / e0.size, atol=1e-13)
    assert not r.ds.any()


def test_loop_invariants_after_convergence():
    shape = (6, 10, 12)
    x, e0, eps, delta = binding_case(shape, seed=6)
    r = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=500, tol=1e-12)
    assert r.violation < 1e-12
    # linearity identity: e_final = e0 + C2R(df)/N + ds
    d = O.reconstruct_edit(shape, e0, r.df, r.ds)
    assert np.allclose(d, r.e, atol=1e-12)
    # frequency bound holds on the final iterate up to rounding
    E = O.rfft3(r.e)
    assert max(np.abs(E.real).max(), np.abs(E.imag).max()) <= delta * (1 + 1e-9)


def test_planted_sign_error_is_caught():
    """A plausible mistake (edit accumulated with the wrong sign) planted in a
    copy of the oracle's clamp breaks the brute-force comparison."""
    shape = (2, 3, 4)
    x, e0, eps, delta = binding_case(shape, seed=4)
    orig = O._clamp_component

    def wrong(v, acc, lo, hi):
        over, under = v > hi, v < lo
        acc[over] -= hi - v[over]
        v[over] = hi
        acc[under] -= lo - v[under]
        v[under] = lo

    try:
        O._clamp_component = wrong
        r = O.solve(x, e0, eps, mode="abs", delta=delta, max_iter=5, tol=1e-13)
    finally:
        O._clamp_component = orig
    e, df, ds, it = brute_solve(x, e0, eps, delta, 5, 1e-13)
    assert not np.allclose(r.df, df, atol=1e-12)


def test_ptw_clamp_hand_values():
    # projection_algorithm.cu:204-258 on three bins
    orig = np.array([10 + 0j, 3 + 4j, 1 + 0j])
    E = np.array([0.5 + 0j, 0.0 + 0j, 0.9 + 0.9j])
    df = np.zeros(6)
    dp = 0.1
    O.project_frequency_all_ptw(E, orig, dp, df)
    lo, hi = 10 * (1 / np.sqrt(1.1) - 1), 10 * (1 / np.sqrt(0.9) - 1)  # DC window [-0.465, 0.541]
    assert lo < 0.5 < hi and E[0] == 0.5 and df[0] == 0  # inside: untouched
    assert E[1] == 0 and df[1] == 0  # ratio 1: inside
    dl = min(1 - np.sqrt(0.9), np.sqrt(1.1) - 1) / float(np.float32(np.sqrt(np.float32(2))))
    assert np.isclose(E[2].real, dl) and np.isclose(E[2].imag, dl)
    assert np.isclose(df[2], dl - 0.9) and np.isclose(df[5], dl - 0.9)
    E2 = np.array([0.7 + 0j, 0j, 0j])
    df2 = np.zeros(6)
    O.project_frequency_all_ptw(E2, orig, dp, df2)  # DC outside -> zeroed
    assert E2[0] == 0 and np.isclose(df2[0], -0.7)


# ---------------------------------------------------------------------------
# extraction
# ---------------------------------------------------------------------------
def test_pack_flags_brute_force():
    rng = np.random.default_rng(0)
    for n in (1, 7, 8, 9, 64, 100):
        f = rng.random(n) < 0.5
        p = O.pack_flags(f)
        assert p.size == (n + 7) // 8
        for i in range(n):
            assert ((p[i // 8] >> (i % 8)) & 1) == f[i]


def test_reference_quantizer_hand_values():
    v = np.array([-1.0, 1.0, 0.0, 1e-3, -0.5, 0.75], np.float64)
    q = O.quantize_ref(v, -1.0, 1.0)
    # n = (v+1)/2 -> ceil(n*65535) above 0.5, floor at or below
    assert list(q) == [0, 65535, 32767, 32801, 16383, 57344]
    d = O.dequantize_ref(q, -1.0, 1.0, np.float64)
    assert np.all(np.abs(d - v) <= 2 / 65535)


def test_extract_one_semantics():
    v = np.array([0, 2e-7, -1e-7, 5e-8, -3e-3, 0, 1e-3], np.float32)
    r = O.extract_one(v)
    assert list(r.flags) == [False, True, False, False, True, False, True]  # |v| > 1e-7
    assert r.vmin == np.float32(-3e-3) and r.vmax == np.float32(1e-3)  # zeros included
    assert np.array_equal(r.compact, v[[1, 4, 6]])


def test_step_quantizer_and_zigzag():
    q = O.quantize_step(np.array([0.0, 0.49, 0.5, -0.5, 1.49, -2.6]), 1.0)
    assert list(q) == [0, 0, 1, -1, 1, -3]
    assert list(O.zigzag(np.array([0, -1, 1, -2, 2]))) == [0, 1, 2, 3, 4]


# ---------------------------------------------------------------------------
# Huffman
# ---------------------------------------------------------------------------
KAT = np.array([1, 2, 3, 1, 2, 1, 4, 5, 1, 2, 3, 3, 3])


def test_huffman_known_answer():
    # hand-derived from HuffmanZSTDCoder.cuh:86-178 and 307-357:
    # counts 1:4 2:3 3:4 4:1 5:1 -> codes 1:10 2:01 3:11 4:000 5:001, 28 bits
    b, (sym, codes, lens) = O.huffman_encode(KAT)
    assert b.tobytes() == bytes([0x9E, 0x60, 0x67, 0xF0])
    table = {int(s): (int(c), int(l)) for s, c, l in zip(sym, codes, lens)}
    assert table == {1: (0b10, 2), 2: (0b01, 2), 3: (0b11, 2), 4: (0b000, 3), 5: (0b001, 3)}
    assert np.array_equal(O.huffman_decode(b, (sym, codes, lens), KAT.size), KAT)


def test_huffman_tie_break_second_case():
    # counts a:1 b:1 c:2: heap pops (1,a)(1,b)->n3(2); then (2,c) before (2,n3)
    # because index 2 < 3 -> left=c, right=n3: c=0, a=10, b=11
    b, (sym, codes, lens) = O.huffman_encode(np.array([7, 8, 9, 9]))
    assert list(lens) == [2, 2, 1] and list(codes) == [0b10, 0b11, 0b0]


def _optimal_cost(counts):
    import heapq

    h = list(counts)
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a, b = heapq.heappop(h), heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


@pytest.mark.parametrize("seed", range(6))
def test_huffman_prefix_kraft_optimal_roundtrip(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 300))
    data = rng.choice(n, size=int(rng.integers(50, 3000)), p=rng.dirichlet(np.ones(n) * 0.3))
    b, (sym, codes, lens) = O.huffman_encode(data)
    words = [format(int(c), f"0{int(l)}b") for c, l in zip(codes, lens)]
    for u, v in itertools.permutations(words, 2):
        assert not v.startswith(u)
    if len(sym) > 1:
        assert abs(sum(2.0 ** -l for l in lens) - 1.0) < 1e-12
    _, cnt = np.unique(data, return_counts=True)
    total = int((lens[np.searchsorted(sym, data)]).sum())
    assert total == (_optimal_cost(cnt) if len(cnt) > 1 else len(data))
    assert b.size == (total + 7) // 8
    assert np.array_equal(O.huffman_decode(b, (sym, codes, lens), data.size), data)


def test_huffman_single_symbol():
    b, (sym, codes, lens) = O.huffman_encode(np.full(9, 5))
    assert list(lens) == [1] and list(codes) == [0]
    assert b.tobytes() == b"\x00\x00"


# ---------------------------------------------------------------------------
# certification
# ---------------------------------------------------------------------------
def test_certify_brute_force():
    shape = (2, 3, 4)
    rng = np.random.default_rng(5)
    d = rng.uniform(-1.2e-3, 1.2e-3, shape)
    D = brute_rfft3(d).reshape(-1)
    delta = float(np.median(np.abs(D.real)))
    r = O.certify(d, 1e-3, delta)
    assert r["spatial_violations"] == int((np.abs(d) > 1e-3).sum())
    assert r["freq_violations"] == int((np.maximum(np.abs(D.real), np.abs(D.imag)) > delta).sum())
    assert not r["ok"]
    assert O.certify(d * 0.5, 1e-3, 1e9)["ok"]


# ---------------------------------------------------------------------------
# step quantiser: the device's reciprocal form decides only where it must
# agree with the division (restated here in the same IEEE double operations)
# ---------------------------------------------------------------------------
def _quant_step_reciprocal(v, step):
    v = np.asarray(v, dtype=np.float64)
    inv = np.float64(1.0) / np.float64(step)
    r = v * inv
    n = np.rint(r)  # round half to even, as the device's rint; no tie is taken where fast
    with np.errstate(invalid="ignore"):
        fast = (np.abs(r) < 16777216.0) & (np.abs(r - n) < 0.499999)
    return fast, n


@pytest.mark.parametrize("step", [1e-5, 1.25e-5, 3.0e-3 / 7, 0.1, 6716.3 / 8])
def test_step_quantiser_reciprocal_agrees_with_division(step):
    rng = np.random.default_rng(11)
    k = rng.integers(0, 1 << 23, 20000).astype(np.float64)
    # half-integer neighbourhoods at every scale of relative offset, and
    # plain random values of both signs
    offs = np.array([0.0, 1e-16, -1e-16, 1e-13, -1e-13, 1e-10, -1e-10, 1e-7, -1e-7, 2e-6, -2e-6, 1e-3])
    near = ((k[:, None] + 0.5) * (1.0 + offs[None, :]) * step).ravel()
    rand = rng.standard_normal(200000) * step * rng.choice([1.0, 10.0, 1e4, 1e7], 200000)
    v = np.concatenate([near, -near, rand, [0.0, -0.0, 0.49 * step, 0.51 * step]])
    fast, q_fast = _quant_step_reciprocal(v, step)
    q_div = O.quantize_step(v, step).astype(np.float64)
    # the multiply decides almost every ordinary value below 2^24 steps
    moderate = np.abs(rand) < 1e6 * step
    assert fast[2 * near.size: 2 * near.size + rand.size][moderate].mean() > 0.99
    assert np.array_equal(q_fast[fast], q_div[fast])
    # the values the margin sends to the division include the exact half-integers
    assert not fast[: near.size].reshape(k.size, offs.size)[:, 0].any()
