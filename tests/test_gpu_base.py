"""The GPU base compressor (dual-quantisation Lorenzo) and the whole
pipeline: x' is round(x/(2 eps)) * 2 eps exactly as a numpy restatement
forms it; the coded symbols are the zigzag of numpy's mixed third difference
of the integers (bit-exact through the oracle's Huffman decoder); decoding
rebuilds x' bit for bit, outliers included; base x' -> solver -> container ->
decode meets both bounds (checked by the oracle in FP64)."""
import struct

import numpy as np
import pytest
import torch

from ffcz_amd import container
from oracle import ffcz_oracle as O
from tests import _container as K

pytestmark = pytest.mark.gpu

BHDR = struct.Struct("<4sIII3QdQ")


def smooth_field(shape, seed, dtype):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    for ax in range(3):
        x = np.cumsum(x, axis=ax)
    x = (x - x.min()) / (x.max() - x.min())
    return x.astype(dtype)


def reference(x, eps, dtype):
    q = np.rint(x.astype(np.float64) * (1.0 / (2.0 * eps))).astype(np.int64)
    xp = (q.astype(np.float64) * (2.0 * eps)).astype(dtype)
    r = q
    for ax in range(3):
        r = np.diff(r, axis=ax, prepend=0)
    z = ((r << 1) ^ (r >> 63)).astype(np.uint64).ravel()
    sym = np.where(z < 65535, z, 65535).astype(np.uint16)
    out = np.flatnonzero(z >= 65535)
    return xp, sym, out, r.ravel()[out]


def parse_base(buf):
    h = BHDR.unpack_from(buf, 0)
    m = h[-1]
    w, z, n, nbits, bsb, plb, k, chunk = K.STREAM.unpack_from(buf, BHDR.size)
    o = BHDR.size + K.STREAM.size
    sym = np.frombuffer(buf, np.uint32, k, o)
    codes = np.frombuffer(buf, np.uint32, k, o + 4 * k)
    lens = np.frombuffer(buf, np.uint8, k, o + 8 * k)
    o += 9 * k + 2 * ((n + chunk - 1) // chunk)
    payload = np.frombuffer(buf, np.uint8, plb, o)
    o += plb
    oidx = np.frombuffer(buf, np.uint64, m, o)
    oval = np.frombuffer(buf, np.int64, m, o + 8 * m)
    assert o + 16 * m == len(buf)
    return h, (payload, (sym, codes, lens), n), oidx, oval


@pytest.mark.parametrize("spikes", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(6, 10, 12), (16, 20, 30)])
def test_base_round_trip(gpu, shape, dtype, spikes):
    x = smooth_field(shape, 3, dtype)
    eps = 1e-4
    if spikes:  # residuals beyond 16 bits: the escape and outlier path
        x.ravel()[::97] += dtype(30.0)
    buf, xp = container.base_compress(torch.from_numpy(x).to(gpu), eps, zstd_level=0)
    ref_xp, ref_sym, ref_out, ref_val = reference(x, eps, dtype)
    assert np.array_equal(xp.cpu().numpy(), ref_xp)
    assert np.abs(ref_xp.astype(np.float64) - x).max() <= eps * (1 + 1e-6) + 8 * np.finfo(dtype).eps * np.abs(x).max()
    h, (payload, table, n), oidx, oval = parse_base(buf)
    assert h[0] == b"FFCB" and tuple(h[4:7]) == shape and h[7] == eps
    got_sym = O.huffman_decode(payload, table, n)
    assert np.array_equal(np.asarray(got_sym, np.uint16), ref_sym)
    assert np.array_equal(oidx, ref_out) and np.array_equal(oval, ref_val)
    assert (ref_out.size > 0) == spikes
    info = container.base_info(buf)
    assert info["shape"] == shape and info["outliers"] == ref_out.size
    back = container.base_decompress(buf)
    assert np.array_equal(back, ref_xp)
    # with ZSTD, and into a device buffer
    buf3, _ = container.base_compress(x, eps, zstd_level=3)
    out = torch.empty(shape, dtype=xp.dtype, device=gpu)
    assert np.array_equal(container.base_decompress(buf3, out).cpu().numpy(), ref_xp)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pipeline_base_solver_container(gpu, dtype):
    shape = (16, 20, 30)
    x = smooth_field(shape, 5, dtype)
    eps = 1e-3
    X = np.fft.rfftn(x.astype(np.float64))
    delta = 2e-3 * float(np.abs(X).max())
    base, edits, cert = container.compress_field(torch.from_numpy(x).to(gpu), eps, delta)
    assert cert.ok
    xh = container.decompress_field(base, edits).cpu().numpy().reshape(shape)
    d = xh.astype(np.float64) - x
    assert np.abs(d).max() <= eps * (1 + 1e-6)
    assert O.certify(d, eps * (1 + 1e-6), delta * (1 + 1e-6))["ok"]
