"""Small deterministic problem instances for the parity tests."""
from __future__ import annotations

import numpy as np


def binding_case(shape, seed=0, dtype=np.float64, spread=1.6, dfrac=0.5):
    """Field x, initial error e0 and bounds for which both projections fire.

    e0 is uniform in [-spread*eps, spread*eps] (so the spatial clamp is active)
    and delta is the ``dfrac`` quantile of |Re|,|Im| of F(e0) (so the
    frequency clamp is active)."""
    rng = np.random.default_rng(seed)
    x = rng.random(shape).astype(dtype)
    eps = 1e-3
    e0 = (rng.uniform(-spread, spread, shape) * eps).astype(dtype)
    E = np.fft.rfftn(e0.astype(np.float64))
    comp = np.concatenate([np.abs(E.real).ravel(), np.abs(E.imag).ravel()])
    delta = float(np.quantile(comp, dfrac)) if comp.size else 1.0
    return x, e0, eps, delta


def direct_case(shape, seed=0, dtype=np.float64, frac=0.6):
    """Per-bin bounds on a random subset of bins (direct mode)."""
    x, e0, eps, delta = binding_case(shape, seed, dtype)
    rng = np.random.default_rng(seed + 100)
    F = shape[0] * shape[1] * (shape[2] // 2 + 1)
    K = max(1, int(frac * F))
    idx = rng.choice(F, size=K, replace=False).astype(np.uint64)
    dk = (delta * rng.uniform(0.5, 1.5, K)).astype(dtype)
    return x, e0, eps, idx, dk
