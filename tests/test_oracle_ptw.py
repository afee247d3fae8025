"""Pins the oracle's pointwise-relative ("ptw") frequency projection,
projection_algorithm.cu:190-259, by properties over thousands of random bins
rather than by retyping its formula:

* a bin whose squared-magnitude ratio |X+E|^2/|X|^2 lies well inside
  [sqrt(1-d), sqrt(1+d)] is left untouched (cu:219-226);
* a bin well outside ends with each component within the clamp level d'
  (cu:228-258): a component beyond +-d' becomes exactly +-d', one within is
  kept, and the accumulated edit is new - old;
* d' equals min(1 - sqrt(1-d), sqrt(1+d) - 1) / sqrt(2), evaluated here in
  40-digit decimal arithmetic, with sqrt(2) rounded to float: the reference
  writes ``sqrt_val(2.0f)``, which picks the float specialisation even when T
  is double (cu:58-67, 228).  A double-precision field separates the two: its
  clamp level matches the float sqrt(2) and misses the double one by far
  more than its rounding;
* the DC bin is zeroed exactly when its error leaves
  [X0 (1/sqrt(1+d) - 1), X0 (1/sqrt(1-d) - 1)] (cu:204-216), and is otherwise
  untouched.

Bins within a relative 1e-4 of a window edge, and components within a few
ulps of the clamp level, are left out: there the reference's own rounding in
T decides.
"""
from __future__ import annotations

import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import ffcz_oracle as O

getcontext().prec = 40
SQRT2_F32 = Decimal(float(np.float32(math.sqrt(2.0))))  # sqrtf(2.0f), correctly rounded
SQRT2 = Decimal(2).sqrt()


def clamp_level(d: float, sqrt2: Decimal) -> Decimal:
    D = Decimal(d)
    one = Decimal(1)
    return min(one - (one - D).sqrt(), (one + D).sqrt() - one) / sqrt2


def random_bins(n, d, dtype, seed):
    """Original bins X with |X| log-uniform over six decades and errors E
    placing |X+E|^2/|X|^2 across [1-3d, 1+3d] (inside and outside the
    window), at random phases."""
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-3, 3, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    X = mag * np.exp(1j * ph)
    r = rng.uniform(1 - 3 * d, 1 + 3 * d, n)  # target squared ratio
    th = rng.normal(0, 0.02, n)
    E = X * (np.sqrt(np.maximum(r, 1e-6)) * np.exp(1j * th) - 1)
    X[0] = 5.0 + 0j  # the DC bin gets its own test
    E[0] = 0.0
    ct = np.complex64 if dtype == np.float32 else np.complex128
    return X.astype(ct), E.astype(ct)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [0.02, 0.1, 0.3])
def test_ptw_properties_over_random_bins(dtype, d):
    T = np.dtype(dtype).type
    n = 20000
    X, E = random_bins(n, d, dtype, seed=int(d * 1000) + (7 if dtype == np.float32 else 0))
    E_in = E.copy()
    df = np.zeros(2 * n, dtype)
    O.project_frequency_all_ptw(E, X, T(d), df)

    # exact ratio of the stored T values (the reference forms X+E in T first)
    xr, xi = X.real.astype(np.float64), X.imag.astype(np.float64)
    cr = (X.real + E_in.real).astype(np.float64)  # the sum rounded in T, as cu:221-222
    ci = (X.imag + E_in.imag).astype(np.float64)
    ratio = (cr * cr + ci * ci) / (xr * xr + xi * xi)
    lo, hi = math.sqrt(1 - d), math.sqrt(1 + d)
    inside = (ratio < hi * (1 - 1e-4)) & (ratio > lo * (1 + 1e-4))
    outside = (ratio > hi * (1 + 1e-4)) | (ratio < lo * (1 - 1e-4))
    inside[0] = outside[0] = False
    assert inside.sum() > 2000 and outside.sum() > 2000

    # untouched inside the window
    assert np.array_equal(E[inside], E_in[inside])
    assert not np.any(df[:n][inside]) and not np.any(df[n:][inside])

    dl = float(clamp_level(d, SQRT2_F32))
    # d' is formed in T from 1 - sqrt(1-d) and sqrt(1+d) - 1: the cancellation
    # leaves an absolute error of a few ulps of 1, not of d'
    ulp = 4 * np.finfo(dtype).eps
    dl_T = None
    for new, old, acc in ((E.real, E_in.real, df[:n]), (E.imag, E_in.imag, df[n:])):
        fired = outside & ((old > dl + ulp) | (old < -dl - ulp))
        kept = outside & (np.abs(old) < dl - ulp)
        assert fired.sum() > 100 and kept.sum() > 100
        # beyond the level: exactly +-d' (one value, the level in T)
        mags = np.unique(np.abs(new[fired]))
        assert mags.size == 1
        dl_T = mags[0] if dl_T is None else dl_T
        assert mags[0] == dl_T
        assert abs(float(mags[0]) - dl) <= ulp
        assert np.array_equal(np.sign(new[fired]), np.sign(old[fired]))
        # within the level: kept
        assert np.array_equal(new[kept], old[kept])
        # edit = new - old, accumulated once
        assert np.array_equal(acc[fired], (new[fired] - old[fired]).astype(dtype))
        assert not np.any(acc[kept])
        # every outside bin ends within the level
        assert np.all(np.abs(new[outside]) <= dl_T)


def test_ptw_clamp_level_uses_float_sqrt2_in_double():
    """sqrt_val(2.0f) is float even in the double kernel (cu:228): the
    clamp level of a double field follows the float sqrt(2)."""
    d = 0.1
    X = np.array([5 + 0j, 1000 + 0j, 1000 + 0j], np.complex128)
    E = np.array([0j, 500 + 500j, -500 - 500j], np.complex128)
    df = np.zeros(6)
    O.project_frequency_all_ptw(E, X, d, df)
    with_f32 = float(clamp_level(d, SQRT2_F32))
    with_f64 = float(clamp_level(d, SQRT2))
    assert abs(with_f32 - with_f64) > 1e-9 * with_f64  # the two differ by ~2e-8 relative
    for v in (E[1].real, E[1].imag, -E[2].real, -E[2].imag):
        assert abs(v - with_f32) <= 4 * np.finfo(np.float64).eps  # rounding of the T arithmetic
        assert abs(v - with_f64) > 1e-9 * with_f64


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [0.05, 0.2])
def test_ptw_dc_window(dtype, d):
    T = np.dtype(dtype).type
    ct = np.complex64 if dtype == np.float32 else np.complex128
    rng = np.random.default_rng(3)
    one = Decimal(1)
    for _ in range(300):
        x0 = float(rng.choice([-1, 1]) * 10 ** rng.uniform(-2, 4))
        X0 = Decimal(float(T(x0)))
        a = X0 * (one / (one + Decimal(d)).sqrt() - one)
        b = X0 * (one / (one - Decimal(d)).sqrt() - one)
        lo, hi = float(min(a, b)), float(max(a, b))
        w = hi - lo
        e0 = float(rng.uniform(lo - 0.5 * w, hi + 0.5 * w))
        margin = 1e-4 * max(abs(lo), abs(hi))
        if min(abs(e0 - lo), abs(e0 - hi)) < margin:
            continue
        X = np.array([x0, 3 + 4j, 1 + 0j], ct)
        E = np.array([e0, 0, 0], ct)
        df = np.zeros(6, dtype)
        O.project_frequency_all_ptw(E, X, T(d), df)
        e0T = T(e0)
        # the reference tests e0 against [X0(1/sqrt(1+d)-1), X0(1/sqrt(1-d)-1)]
        # literally (cu:206-209): for X0 < 0 that interval is empty, so every
        # DC error is zeroed
        leaves = not (float(a) <= e0 <= float(b))
        if leaves:
            assert E[0] == 0 and df[0] == -e0T
        else:
            assert E[0] == e0T and df[0] == 0
        assert df[3] == 0  # the imaginary part of DC is never edited


# ---------------------------------------------------------------------------
# corrected semantics (D11, opt-in) and the partial projection
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [0.02, 0.1, 0.3])
def test_ptw_corrected_keeps_every_power_ratio_in_the_band(dtype, d):
    """The corrected rule is a projection onto a box inscribed in the annulus
    sqrt(1-d)|X| <= |X+E| <= sqrt(1+d)|X|: after one application every bin's
    power ratio |X+E|^2/|X|^2 lies in [1-d, 1+d] (checked in exact arithmetic
    of the stored values), bins already inside are untouched, a clamped
    component sits at +-d'|X| with d' = min(1-sqrt(1-d), sqrt(1+d)-1)/sqrt(2)
    (40-digit decimal), and a bin of zero amplitude ends with zero edit."""
    T = np.dtype(dtype).type
    n = 20000
    X, E = random_bins(n, d, dtype, seed=int(d * 1000) + 3)
    X[5] = 0
    E[5] = 0.3 + 0.1j  # zero amplitude: only E = 0 keeps |X+E| = |X|
    E_in = E.copy()
    df = np.zeros(2 * n, dtype)
    O.project_frequency_ptw_corrected(E, X, T(d), df)
    xr, xi = X.real.astype(np.float64), X.imag.astype(np.float64)
    p0 = xr * xr + xi * xi
    r0 = ((xr + E_in.real) ** 2 + (xi + E_in.imag) ** 2) / np.where(p0 > 0, p0, 1)
    r1 = ((xr + E.real.astype(np.float64)) ** 2 + (xi + E.imag.astype(np.float64)) ** 2) / np.where(p0 > 0, p0, 1)
    tol = 16 * np.finfo(dtype).eps
    nz = p0 > 0
    assert np.all(r1[nz] <= 1 + d + tol) and np.all(r1[nz] >= 1 - d - tol)
    assert E[5] == 0 and df[5] == -np.float64(E_in[5].real) and df[n + 5] == -np.float64(E_in[5].imag)
    inside = nz & (r0 < (1 + d) * (1 - 1e-4)) & (r0 > (1 - d) * (1 + 1e-4))
    assert inside.sum() > 2000 and np.array_equal(E[inside], E_in[inside])
    dl = float(clamp_level(d, SQRT2))
    amp = np.sqrt(p0)
    outside = nz & ((r0 > (1 + d) * (1 + 1e-4)) | (r0 < (1 - d) * (1 - 1e-4)))
    assert outside.sum() > 2000
    for new, old, acc in ((E.real, E_in.real, df[:n]), (E.imag, E_in.imag, df[n:])):
        lv = dl * amp
        fired = outside & (np.abs(old) > lv * (1 + 1e-4))
        kept = outside & (np.abs(old) < lv * (1 - 1e-4))
        assert fired.sum() > 100 and kept.sum() > 100
        # d' formed in T loses bits to the cancellation in 1 - sqrt(1-d): a few
        # ulps of 1 absolute, so relative to d' itself
        rtol = 8 * np.finfo(dtype).eps * (1 + 1 / dl)
        assert np.allclose(np.abs(new[fired]), lv[fired], rtol=rtol, atol=0)
        assert np.array_equal(np.sign(new[fired]), np.sign(old[fired]))
        assert np.array_equal(new[kept], old[kept])
        assert np.array_equal(acc[fired], (new[fired] - old[fired]).astype(dtype))


def test_partial_ptw_reference_dc_is_the_first_constraint():
    """cu:276-289: constraint 0 runs the DC rule on bin 0 whatever idx[0]
    names; later constraints run the ratio rule on their bins; unlisted bins
    are untouched.  The listed bins end as the all-bins projection leaves
    them, and bin idx[0] (listed nowhere else) is untouched."""
    rng = np.random.default_rng(11)
    n, d = 4000, 0.1
    X, E = random_bins(n, d, np.float64, seed=17)
    X[0] = 10.0 + 0j
    E[0] = 0.9  # outside the DC window [-0.465, 0.541]: zeroed by the DC rule
    idx = np.concatenate([[1234], rng.choice(np.arange(1, n), 1500, replace=False)])
    idx = idx[idx != 1234]
    idx = np.concatenate([[1234], idx])
    E_in = E.copy()
    df = np.zeros(2 * n)
    O.project_frequency_partial_ptw(E, X, idx, d, df)
    Ea = E_in.copy()
    dfa = np.zeros(2 * n)
    O.project_frequency_all_ptw(Ea, X, d, dfa)
    listed = np.zeros(n, bool)
    listed[idx[1:]] = True
    assert E[0] == 0 and df[0] == -0.9  # the DC rule on bin 0
    assert E[1234] == E_in[1234] and df[1234] == 0 and df[n + 1234] == 0
    assert np.array_equal(E[listed], Ea[listed])
    assert np.array_equal(df[:n][listed], dfa[:n][listed]) and np.array_equal(df[n:][listed], dfa[n:][listed])
    rest = ~listed
    rest[0] = False
    assert np.array_equal(E[rest], E_in[rest]) and not np.any(df[:n][rest]) and not np.any(df[n:][rest])
    assert np.any(E[listed] != E_in[listed])  # the ratio rule fired somewhere


def test_partial_ptw_corrected_constrains_exactly_the_listed_bins():
    d, n = 0.1, 3000
    X, E = random_bins(n, d, np.float64, seed=29)
    idx = np.random.default_rng(2).choice(n, 1000, replace=False)
    E_in = E.copy()
    df = np.zeros(2 * n)
    mask = np.zeros(n, bool)
    mask[idx] = True
    O.project_frequency_ptw_corrected(E, X, d, df, mask=mask)
    assert np.array_equal(E[~mask], E_in[~mask]) and not np.any(df[:n][~mask])
    p0 = np.abs(X) ** 2
    r1 = np.abs(X + E) ** 2 / p0
    assert np.all(r1[mask] <= 1 + d + 1e-12) and np.all(r1[mask] >= 1 - d - 1e-12)
