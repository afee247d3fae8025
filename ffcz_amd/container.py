"""Self-describing container of one solve's coded edits, and its decoder.

The reference keeps its Huffman code tables in memory only and never writes
the shape, type, bounds or quantiser range (HuffmanZSTDCoder.cuh:240-246,
projection_example.cu:202-246), so its coded streams cannot be decoded on their
own.  ``pack(solver)`` writes the four streams the reference driver codes
(frequency and spatial flag packs, frequency and spatial quantised edits),
each Huffman-coded bit-exactly as the reference encoder and ZSTD-compressed,
with everything a decoder needs.  ``decode(buf, x_prime)`` rebuilds
x_hat = x' + C2R(deq df)/N + deq(ds) on the GPU.  The byte layout is documented
in ``csrc/container.hip``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, load
from .solver import as_pointer


def pack(solver, zstd_level: int = 3) -> bytes:
    """The solver's extracted edits as container bytes (after
    ``extract_compact_representation``)."""
    lib = load()
    b = C.c_void_p()
    call("ffcz_container_pack", solver._h, int(zstd_level), C.byref(b))
    try:
        n = C.c_size_t()
        call("ffcz_blob_size", b, C.byref(n))
        out = (C.c_uint8 * max(n.value, 1))()
        call("ffcz_blob_copy", b, out, n.value)
    finally:
        lib.ffcz_blob_free(b)
    return bytes(out)[: n.value]


def info(buf: bytes) -> dict:
    """Header fields of a container."""
    a = np.frombuffer(buf, np.uint8)
    dt, q = C.c_int(), C.c_int()
    dims = (C.c_size_t * 3)()
    act = (C.c_size_t * 2)()
    eps, delta = C.c_double(), C.c_double()
    call("ffcz_container_info", a.ctypes.data, a.size, C.byref(dt), dims, C.byref(eps), C.byref(delta),
         C.byref(q), act)
    return {"dtype": np.float32 if dt.value == 0 else np.float64, "shape": tuple(dims), "eps": eps.value,
            "delta": delta.value, "quantizer": q.value, "n_active_freq": act[0], "n_active_spatial": act[1]}


def decode(buf: bytes, x_prime, out=None):
    """x_hat = x' + C2R(deq df)/N + deq(ds).  ``x_prime`` (numpy array or
    torch tensor, host or GPU) holds the base reconstruction in the
    container's dtype; the result goes to ``out`` (same kinds) or a new array
    like ``x_prime``."""
    hdr = info(buf)
    n = int(np.prod(hdr["shape"]))
    try:
        import torch

        is_tensor = isinstance(x_prime, torch.Tensor)
    except ImportError:  # pragma: no cover
        is_tensor = False
    if out is None:
        if is_tensor:
            out = torch.empty_like(x_prime).contiguous()
        else:
            out = np.empty(np.shape(x_prime), dtype=hdr["dtype"])
    xp, keep_in = as_pointer(x_prime, hdr["dtype"])
    xo, keep_out = as_pointer(out, hdr["dtype"])
    if int(np.prod(keep_in.shape)) != n or int(np.prod(keep_out.shape)) != n:
        raise ValueError(f"x_prime and out must hold {n} elements")
    a = np.frombuffer(buf, np.uint8)
    call("ffcz_container_decode", a.ctypes.data, a.size, xp, xo, None)
    return out
