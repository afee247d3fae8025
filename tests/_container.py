"""Writer and parser of the container layout (ffcz_amd/csrc/container.hip)
in plain Python over the oracle's Huffman coder: the tests build well-formed
and damaged containers on the CPU with it, and read the library's containers
stream by stream."""
from __future__ import annotations

import struct

import numpy as np

from oracle import ffcz_oracle as O

HDR = struct.Struct("<4sIII3Q2dII6d2Q")  # 128 bytes
STREAM = struct.Struct("<II4QII")
CHUNK = 1024


def build(shape, dtype, eps, delta, quantizer, steps, ranges, fpack, spack, fsym, ssym) -> bytes:
    """A container (payloads stored raw) from flag packs and symbol arrays."""
    nx, ny, nz = shape
    out = [HDR.pack(b"FFCZ", 1, 0 if np.dtype(dtype) == np.float32 else 1, 1, nx, ny, nz, eps, delta, quantizer, 4,
                    steps[0], steps[1], *ranges, int(fsym.size), int(ssym.size))]
    for arr, w in ((fpack, 1), (spack, 1), (fsym, 2), (ssym, 2)):
        arr = np.asarray(arr)
        if arr.size == 0:
            out.append(STREAM.pack(w, 0, 0, 0, 0, 0, 0, CHUNK))
            continue
        bits, (sym, codes, lens) = O.huffman_encode(arr)
        pos = np.searchsorted(sym, arr)
        L = lens[pos].astype(np.int64)
        nbits = int(L.sum())
        nch = (arr.size + CHUNK - 1) // CHUNK
        cb = np.add.reduceat(L, np.arange(0, arr.size, CHUNK)).astype(np.uint16) if nch else np.zeros(0, np.uint16)
        out.append(STREAM.pack(w, 0, arr.size, nbits, bits.size, bits.size, sym.size, CHUNK))
        out += [np.asarray(sym, np.uint32).tobytes(), np.asarray(codes, np.uint32).tobytes(),
                np.asarray(lens, np.uint8).tobytes(), cb.tobytes(), bits.tobytes()]
    return b"".join(out)


def parse(buf: bytes):
    """(header tuple, [stream dicts]) of a container whose payloads are raw."""
    h = HDR.unpack_from(buf, 0)
    o = HDR.size
    streams = []
    for _ in range(4):
        w, z, n, nbits, bsb, plb, k, chunk = STREAM.unpack_from(buf, o)
        o += STREAM.size
        sym = np.frombuffer(buf, np.uint32, k, o)
        o += 4 * k
        codes = np.frombuffer(buf, np.uint32, k, o)
        o += 4 * k
        lens = np.frombuffer(buf, np.uint8, k, o)
        o += k
        nch = (n + chunk - 1) // chunk
        cb = np.frombuffer(buf, np.uint16, nch, o)
        o += 2 * nch
        payload = np.frombuffer(buf, np.uint8, plb, o)
        o += plb
        streams.append({"width": w, "zstd": z, "n": n, "nbits": nbits, "bitstream_bytes": bsb, "sym": sym,
                        "codes": codes, "lens": lens, "chunk_bits": cb, "payload": payload})
    assert o == len(buf)
    return h, streams
