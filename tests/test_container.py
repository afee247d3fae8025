"""The container parser on the CPU: containers written by the oracle's coder
(tests/_container.py) are read by the library; truncated and damaged ones are
rejected by the format checks that run on the host before any device work
(the messages name the check that failed)."""
import struct

import numpy as np
import pytest

from ffcz_amd import container
from ffcz_amd._lib import FFCzError, call
from tests import _container as K

SHAPE = (2, 3, 8)
N = 48
F2 = 2 * (2 * 3 * 5)


def tiny(seed=0, quantizer=2):
    rng = np.random.default_rng(seed)
    ff = rng.random(F2) < 0.3
    sf = rng.random(N) < 0.5
    fsym = rng.integers(1, 9, int(ff.sum())).astype(np.uint16)
    ssym = rng.integers(1, 40, int(sf.sum())).astype(np.uint16)
    return K.build(SHAPE, np.float32, 1e-3, 0.5, quantizer, (0.1, 0.2), (-1.0, 1.0, -2.0, 2.0),
                   np.packbits(ff, bitorder="little"), np.packbits(sf, bitorder="little"), fsym, ssym)


def decode_raw(buf: bytes):
    a = np.frombuffer(buf, np.uint8)
    xp = np.zeros(N, np.float32)
    xo = np.zeros(N, np.float32)
    call("ffcz_container_decode", a.ctypes.data, a.size, xp.ctypes.data, xo.ctypes.data, None)


def test_info_reads_the_header():
    buf = tiny()
    h, streams = K.parse(buf)
    i = container.info(buf)
    assert i["shape"] == SHAPE and i["dtype"] == np.float32
    assert i["eps"] == 1e-3 and i["delta"] == 0.5 and i["quantizer"] == 2
    assert i["n_active_freq"] == streams[2]["n"] and i["n_active_spatial"] == streams[3]["n"]


def test_every_truncation_is_rejected():
    buf = tiny()
    for cut in range(len(buf)):
        with pytest.raises(FFCzError, match="truncated container|not an FFCz container"):
            decode_raw(buf[:cut])


def _patch(buf: bytes, off: int, fmt: str, value) -> bytes:
    b = bytearray(buf)
    struct.pack_into(fmt, b, off, value)
    return bytes(b)


@pytest.mark.parametrize("off, fmt, value, msg", [
    (0, "<I", 0x5A43464E, "not an FFCz container"),
    (4, "<I", 2, "unsupported FFCz container version"),
    (8, "<I", 7, "bad container dtype"),
    (16, "<Q", 0, "bad container shape"),
    (56, "<I", 5, "bad container quantiser"),
    (60, "<I", 3, "bad container stream count"),
    (64, "<d", 0.0, "bad container quantiser steps"),
    (112, "<Q", 10 ** 6, "symbol streams do not match the active counts"),
    (128, "<I", 2, "wrong symbol width"),
    (128 + 4, "<I", 3, "unsupported container stream"),
    (128 + 8, "<Q", 7, "do not match the shape|inconsistent container stream sizes"),
])
def test_damaged_header_fields_are_rejected(off, fmt, value, msg):
    with pytest.raises(FFCzError, match=msg):
        decode_raw(_patch(tiny(), off, fmt, value))


def test_damaged_code_table_and_index_are_rejected():
    buf = tiny()
    _, streams = K.parse(buf)
    # the first stream's code lengths, chunk index, and a bitstream byte
    o = 128 + K.STREAM.size
    k = streams[0]["sym"].size
    len_off = o + 8 * k
    with pytest.raises(FFCzError, match="bad container code length"):
        decode_raw(_patch(buf, len_off, "<B", 0))
    with pytest.raises(FFCzError, match="bad container code length"):
        decode_raw(_patch(buf, len_off, "<B", 33))
    with pytest.raises(FFCzError, match="chunk index does not add up"):
        decode_raw(_patch(buf, len_off + k, "<H", int(streams[0]["chunk_bits"][0]) + 1))


def test_a_code_table_that_is_not_a_prefix_code_is_rejected():
    buf = tiny()
    _, streams = K.parse(buf)
    k = streams[0]["sym"].size
    assert k >= 2
    code_off = 128 + K.STREAM.size + 4 * k
    len_off = code_off + 4 * k
    # the second code made equal to the first (same length and bits)
    b = _patch(buf, code_off + 4, "<I", int(streams[0]["codes"][0]))
    b = _patch(b, len_off + 1, "<B", int(streams[0]["lens"][0]))
    with pytest.raises(FFCzError, match="not a prefix code"):
        decode_raw(b)


def test_base_stream_header_is_checked():
    with pytest.raises(FFCzError, match="not an FFCz base stream"):
        container.base_info(b"FFCZ" + bytes(60))
    with pytest.raises(FFCzError, match="truncated container"):
        container.base_info(b"FFCB" + struct.pack("<I", 1) + bytes(10))
    hdr = struct.pack("<4sIII3QdQ", b"FFCB", 1, 0, 0, 4, 5, 6, 1e-3, 0)
    assert container.base_info(hdr) == {"dtype": np.float32, "shape": (4, 5, 6), "eps": 1e-3, "outliers": 0}
    with pytest.raises(FFCzError, match="bad base stream bound"):
        container.base_info(struct.pack("<4sIII3QdQ", b"FFCB", 1, 0, 0, 4, 5, 6, -1.0, 0))
