"""Synthetic fields for the benchmark and the tests.

Gaussian random fields with an isotropic power spectrum P(k) ~ k^slope: complex
normals on the rfft half grid with amplitude sqrt(P(|k|)) (P(0) = 0), inverse
real FFT (Hermitian symmetry is implied), normalised to [0, 1].  The lossy
reconstruction x' is x quantised uniformly with step 2*eps_s, which gives
|x' - x| <= eps_s element-wise (SZ-like uniform error); the rounding of x' to
the storage type is nudged so the bound holds exactly on the stored values.

torch is used for the generator so a GPU can produce the 512^3 inputs in place;
the same seed gives the same field on the same device type.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


def grf(shape, slope: float, seed: int, device="cpu", dtype=torch.float64) -> torch.Tensor:
    nx, ny, nz = shape
    g = torch.Generator(device=device).manual_seed(int(seed))
    kx = torch.fft.fftfreq(nx, d=1.0 / nx, device=device, dtype=dtype)
    ky = torch.fft.fftfreq(ny, d=1.0 / ny, device=device, dtype=dtype)
    kz = torch.fft.rfftfreq(nz, d=1.0 / nz, device=device, dtype=dtype)
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    amp = torch.where(k2 > 0, k2.clamp_min(1e-30) ** (slope / 4.0), torch.zeros((), dtype=dtype, device=device))
    del k2
    re = torch.randn(amp.shape, generator=g, device=device, dtype=dtype)
    im = torch.randn(amp.shape, generator=g, device=device, dtype=dtype)
    spec = torch.complex(re * amp, im * amp)
    del re, im, amp
    f = torch.fft.irfftn(spec, s=shape, dim=(0, 1, 2))
    del spec
    return f


def normalize01(f: torch.Tensor) -> torch.Tensor:
    lo, hi = f.min(), f.max()
    return (f - lo) / (hi - lo)


def uniform_quantize(x: torch.Tensor, eps: float) -> torch.Tensor:
    """x' = round(x / (2 eps)) * 2 eps in FP64, stored in x.dtype with
    |x' - x| <= eps guaranteed on the stored values."""
    x64 = x.to(torch.float64)
    xp = torch.round(x64 / (2.0 * eps)) * (2.0 * eps)
    xs = xp.to(x.dtype)
    bad = (xs.to(torch.float64) - x64).abs() > eps
    if bool(bad.any()):
        toward = torch.nextafter(xs, x)
        xs = torch.where(bad, toward, xs)
    return xs


@dataclass
class Workload:
    name: str
    x: torch.Tensor        # original field
    xp: torch.Tensor       # lossy reconstruction
    e0: torch.Tensor       # x' - x in the field type (exact)
    eps: float             # spatial bound
    delta: float           # uniform frequency bound (abs mode), or
    delta_k: torch.Tensor | None = None  # per-bin bound (direct mode), compact half-spectrum order
    max_iter: int = 100
    seed: int = 0


def make_config(name: str, device="cpu", seed: int = 1234, shape=None) -> Workload:
    """The benchmark configurations (shapes can be overridden for tests).

    c1: 1x256x256 f32, k^-3, eps 1e-3, delta 1e-3 max|F(x)|, <= 100 iterations
    c2: 256^3 f32, k^-11/3, eps 1e-3 range, delta 5e-4 max|F(x)|
    c3: 512^3 f32, k^-3 + 5% i.i.d. noise, eps 1e-4 range, delta 1e-4 max|F(x)|
    c5: 600x600x300 f64, k^-5/3, 0.1% +-10 sigma spikes, eps 1e-5 range,
        delta_k = 1e-6 |F(x)_k| + 1e-9 max|F(x)|, <= 1000 iterations
    """
    name = name.lower()
    if name == "c1":
        shape = shape or (1, 256, 256)
        f = normalize01(grf(shape, -3.0, seed, device))
        x = f.to(torch.float32)
        eps = 1e-3
        X = torch.fft.rfftn(x.to(torch.float64), dim=(0, 1, 2))
        delta = 1e-3 * float(X.abs().max())
        xp = uniform_quantize(x, eps)
        return Workload("c1", x, xp, xp - x, eps, delta, max_iter=100, seed=seed)
    if name == "c2":
        shape = shape or (256, 256, 256)
        x = normalize01(grf(shape, -11.0 / 3.0, seed, device)).to(torch.float32)
        eps = 1e-3 * float(x.max() - x.min())
        X = torch.fft.rfftn(x.to(torch.float64), dim=(0, 1, 2))
        delta = 5e-4 * float(X.abs().max())
        del X
        xp = uniform_quantize(x, eps)
        return Workload("c2", x, xp, xp - x, eps, delta, max_iter=100, seed=seed)
    if name == "c3":
        shape = shape or (512, 512, 512)
        f = grf(shape, -3.0, seed, device)
        g = torch.Generator(device=device).manual_seed(int(seed) + 1)
        f = f + 0.05 * f.std() * torch.randn(f.shape, generator=g, device=device, dtype=f.dtype)
        x = normalize01(f).to(torch.float32)
        del f
        eps = 1e-4 * float(x.max() - x.min())
        X = torch.fft.rfftn(x, dim=(0, 1, 2))
        delta = 1e-4 * float(X.abs().max())
        del X
        xp = uniform_quantize(x, eps)
        return Workload("c3", x, xp, xp - x, eps, delta, max_iter=100, seed=seed)
    if name == "c5":
        shape = shape or (600, 600, 300)
        f = grf(shape, -5.0 / 3.0, seed, device)
        sig = f.std()
        g = torch.Generator(device=device).manual_seed(int(seed) + 2)
        n = f.numel()
        k = max(1, int(round(1e-3 * n)))
        idx = torch.randperm(n, generator=g, device=device)[:k]
        sgn = torch.where(torch.rand(k, generator=g, device=device) < 0.5, -1.0, 1.0).to(f.dtype)
        f.view(-1)[idx] = 10.0 * sig * sgn
        x = normalize01(f)
        del f
        eps = 1e-5 * float(x.max() - x.min())
        X = torch.fft.rfftn(x, dim=(0, 1, 2))
        mx = float(X.abs().max())
        delta_k = (1e-6 * X.abs() + 1e-9 * mx).reshape(-1)
        del X
        xp = uniform_quantize(x, eps)
        return Workload("c5", x, xp, xp - x, eps, 0.0, delta_k=delta_k, max_iter=1000, seed=seed)
    raise ValueError(f"unknown config {name}")


def batch_shapes(seed: int = 7, count: int = 48):
    """Batching configuration: ``count`` fields, each axis drawn from
    {60,72,96,120,150,180,240,250}, slope U[-4,-2], eps U[1e-4,1e-2]*range."""
    rng = np.random.default_rng(seed)
    dims = np.array([60, 72, 96, 120, 150, 180, 240, 250])
    out = []
    for _ in range(count):
        shape = tuple(int(v) for v in rng.choice(dims, 3))
        out.append(dict(shape=shape, slope=float(rng.uniform(-4, -2)), eps_rel=float(rng.uniform(1e-4, 1e-2))))
    return out


def make_batch(device="cpu", seed: int = 7, count: int = 48, specs=None) -> list:
    """c4: the batching configuration as float32 workloads.  Field i uses the
    slope and eps of batch_shapes; its frequency bound scales with its eps,
    delta = 0.5 * (eps / range) * max|F(x)| (the c2 ratio of 5e-4 to 1e-3)."""
    out = []
    for i, sp in enumerate(specs or batch_shapes(seed, count)):
        x = normalize01(grf(sp["shape"], sp["slope"], seed * 1000 + i, device)).to(torch.float32)
        rng_ = float(x.max() - x.min())
        eps = sp["eps_rel"] * rng_
        X = torch.fft.rfftn(x.to(torch.float64), dim=(0, 1, 2))
        delta = 0.5 * sp["eps_rel"] * float(X.abs().max())
        del X
        xp = uniform_quantize(x, eps)
        out.append(Workload(f"c4[{i}]", x, xp, xp - x, eps, delta, max_iter=100, seed=seed))
    return out


__all__ = ["grf", "normalize01", "uniform_quantize", "Workload", "make_config", "batch_shapes", "make_batch"]
del math
