"""Python mirror of the reference entropy coder
(``SimpleCudaHuffmanZstdCompressor<T>``, HuffmanZSTDCoder.cuh:235-470).

``cuda_huffman_zstd_compress(data, level)`` returns a ``CompressionResult``
with the reference's fields (``compressed_data``, ``code_table`` mapping
symbol -> (code, length), ``original_bitstream_size``, ``num_elements``); the
histogram and bit emission run on the GPU, ZSTD on the host (libzstd).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import call, load
from .solver import as_pointer


@dataclass
class CompressionResult:
    compressed_data: bytes
    code_table: dict
    original_bitstream_size: int
    num_elements: int
    bitstream: bytes = b""
    num_bits: int = 0
    symbols: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    codes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    lengths: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    device_ms: float = 0.0  # the encoder's device work (histogram to emitted bitstream)


def _kind(data) -> str:
    dt = str(getattr(data, "dtype", ""))
    if dt.endswith("uint16"):
        return "u16"
    if dt.endswith("uint8"):
        return "u8"
    raise TypeError("symbols must be uint8 or uint16")


def _collect(h) -> CompressionResult:
    lib = load()
    cb, bb, ne, ns = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
    nbits = C.c_ulonglong()
    try:
        call("ffcz_coded_info", h, C.byref(cb), C.byref(bb), C.byref(ne), C.byref(ns), C.byref(nbits))
        data = (C.c_uint8 * max(cb.value, 1))()
        call("ffcz_coded_data", h, data, cb.value)
        bits = (C.c_uint8 * max(bb.value, 1))()
        call("ffcz_coded_bitstream", h, bits, bb.value)
        sym = np.zeros(ns.value, np.uint32)
        code = np.zeros(ns.value, np.uint32)
        ln = np.zeros(ns.value, np.int32)
        call("ffcz_coded_table", h, sym.ctypes.data, code.ctypes.data, ln.ctypes.data, ns.value)
        dms = C.c_float()
        call("ffcz_coded_device_ms", h, C.byref(dms))
    finally:
        lib.ffcz_coded_free(h)
    table = {int(s): (int(c), int(l)) for s, c, l in zip(sym, code, ln)}
    return CompressionResult(bytes(data)[: cb.value], table, bb.value, ne.value, bytes(bits)[: bb.value],
                             nbits.value, sym, code, ln, float(dms.value))


def cuda_huffman_zstd_compress(data, zstd_level: int = 3) -> CompressionResult:
    """HuffmanZSTDCoder.cuh:454-460 (name kept for drop-in callers)."""
    k = _kind(data)
    p, keep = as_pointer(data)
    n = int(np.prod(keep.shape))
    h = C.c_void_p()
    call(f"ffcz_huffman_zstd_compress_{k}", p, n, int(zstd_level), C.byref(h))
    return _collect(h)


def huffman_encode(data) -> CompressionResult:
    """The same encoder without ZSTD: ``compressed_data`` is the bitstream."""
    k = _kind(data)
    p, keep = as_pointer(data)
    n = int(np.prod(keep.shape))
    h = C.c_void_p()
    call(f"ffcz_huffman_encode_{k}", p, n, C.byref(h))
    return _collect(h)


def cuda_huffman_zstd_decompress(compressed_data: bytes, code_table: dict, num_elements: int,
                                 original_bitstream_size: int, dtype=np.uint16) -> np.ndarray:
    """HuffmanZSTDCoder.cuh:462-470 (host decode)."""
    k = "u16" if np.dtype(dtype) == np.uint16 else "u8"
    items = sorted(code_table.items())
    sym = np.array([s for s, _ in items], np.uint32)
    code = np.array([c for _, (c, _) in items], np.uint32)
    ln = np.array([l for _, (_, l) in items], np.int32)
    comp = np.frombuffer(compressed_data, np.uint8).copy()
    out = np.zeros(num_elements, dtype)
    n = C.c_size_t()
    call(f"ffcz_huffman_zstd_decompress_{k}", comp.ctypes.data, comp.size, sym.ctypes.data, code.ctypes.data,
         ln.ctypes.data, len(sym), int(num_elements), int(original_bitstream_size), out.ctypes.data, C.byref(n))
    return out[: n.value]


cudaHuffmanZstdCompress = cuda_huffman_zstd_compress
cudaHuffmanZstdDecompress = cuda_huffman_zstd_decompress
