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


def _blob_bytes(b) -> bytes:
    lib = load()
    try:
        n = C.c_size_t()
        call("ffcz_blob_size", b, C.byref(n))
        out = (C.c_uint8 * max(n.value, 1))()
        call("ffcz_blob_copy", b, out, n.value)
    finally:
        lib.ffcz_blob_free(b)
    return bytes(out)[: n.value]


def _like(x, dtype, shape):
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return torch.empty(shape, dtype=getattr(torch, np.dtype(dtype).name), device=x.device)
    except ImportError:  # pragma: no cover
        pass
    return np.empty(shape, dtype=dtype)


def base_compress(x, eps: float, zstd_level: int = 3, x_prime=None):
    """The GPU base compressor (dual-quantisation Lorenzo): returns (stream
    bytes, x') with x' = round(x/(2 eps)) * 2 eps like ``x`` (numpy or torch,
    host or GPU), ready to feed the solver."""
    dt = np.float32 if str(getattr(x, "dtype", "")).endswith("float32") else np.float64
    shape = tuple(x.shape)
    dims = tuple([1] * (3 - len(shape)) + list(shape))
    if x_prime is None:
        x_prime = _like(x, dt, shape)
    xp, keep_x = as_pointer(x, dt)
    op, keep_o = as_pointer(x_prime, dt)
    b = C.c_void_p()
    call("ffcz_base_compress", 0 if dt == np.float32 else 1, xp, dims[0], dims[1], dims[2], float(eps),
         int(zstd_level), op, C.byref(b), None)
    return _blob_bytes(b), x_prime


def base_info(buf: bytes) -> dict:
    a = np.frombuffer(buf, np.uint8)
    dt = C.c_int()
    dims = (C.c_size_t * 3)()
    eps = C.c_double()
    m = C.c_size_t()
    call("ffcz_base_info", a.ctypes.data, a.size, C.byref(dt), dims, C.byref(eps), C.byref(m))
    return {"dtype": np.float32 if dt.value == 0 else np.float64, "shape": tuple(dims), "eps": eps.value,
            "outliers": m.value}


def base_decompress(buf: bytes, out=None):
    """x' from a base stream into ``out`` (numpy or torch, host or GPU), or a
    new numpy array."""
    i = base_info(buf)
    if out is None:
        out = np.empty(i["shape"], dtype=i["dtype"])
    op, keep = as_pointer(out, i["dtype"])
    if int(np.prod(keep.shape)) != int(np.prod(i["shape"])):
        raise ValueError("out has the wrong size")
    a = np.frombuffer(buf, np.uint8)
    call("ffcz_base_decompress", a.ctypes.data, a.size, op, None)
    return out


def compress_field(x, eps: float, delta: float, base_eps: float | None = None, zstd_level: int = 3,
                   max_iter: int = 100, tol: float = 1e-7, margin: float = 0.125, step_quant: bool = True):
    """The whole pipeline on the GPU: the base compressor's x' (bound
    ``base_eps``, default ``eps``) feeds the solver device to device, whose
    certified edits go into the container.  ``x`` is a torch tensor on the GPU
    (or a numpy array, copied there).  Returns (base bytes, container bytes,
    FP64 certification of the quantised edits)."""
    import torch

    from .solver import ProjectionSolver
    from ._lib import QUANT_STEP

    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        x = torch.as_tensor(np.asarray(x)).to("cuda:0")
    shape = tuple([1] * (3 - x.dim()) + list(x.shape))
    dt = np.float32 if x.dtype == torch.float32 else np.float64
    base, xp = base_compress(x, base_eps or eps, zstd_level)
    e0 = xp - x
    with ProjectionSolver(*shape, dtype=dt) as s:
        s.initialize_all_abs(x, eps, delta)
        if margin:
            s.set_margins(margin, margin)
        if step_quant:
            s.set_quantizer(QUANT_STEP, delta / 8, eps / 8)
        s.solve_all_abs(e0, max_iter, tol)
        cert = s.certify(True)
        edits = pack(s, zstd_level)
    return base, edits, cert


def decompress_field(base: bytes, edits: bytes, device="cuda:0"):
    """x_hat from the two streams of ``compress_field``, on the GPU (a torch tensor)."""
    import torch

    i = base_info(base)
    xp = torch.empty(i["shape"], dtype=getattr(torch, np.dtype(i["dtype"]).name), device=device)
    base_decompress(base, xp)
    return decode(edits, xp)
