"""Device Huffman encoder: bitstream bit-exact against the restated reference
encoder (HuffmanZSTDCoder.cuh:48-357), ZSTD round trip, decode."""
import numpy as np
import pytest
import torch

import ffcz_amd
from ffcz_amd import cuda_huffman_zstd_compress, cuda_huffman_zstd_decompress, huffman_encode
from oracle import ffcz_oracle as O

pytestmark = pytest.mark.gpu

KAT = np.array([1, 2, 3, 1, 2, 1, 4, 5, 1, 2, 3, 3, 3], np.uint16)


def test_known_answer(gpu):
    r = huffman_encode(KAT)
    assert r.bitstream == bytes([0x9E, 0x60, 0x67, 0xF0])
    assert r.num_bits == 28 and r.original_bitstream_size == 4
    assert r.code_table == {1: (2, 2), 2: (1, 2), 3: (3, 2), 4: (0, 3), 5: (1, 3)}


def _cases():
    rng = np.random.default_rng(0)
    yield "single", np.full(1000, 7, np.uint16)
    yield "one", np.array([65535], np.uint16)
    yield "two", np.array([3, 9, 9], np.uint16)
    yield "uniform16", rng.integers(0, 65536, 200_000).astype(np.uint16)
    yield "skewed", np.minimum(rng.geometric(0.3, 300_001), 4000).astype(np.uint16)
    yield "normal", np.clip(rng.normal(32768, 30, 1_000_000), 0, 65535).astype(np.uint16)
    yield "u8", rng.integers(0, 256, 100_003).astype(np.uint8)
    yield "u8skew", np.minimum(rng.geometric(0.5, 50_000), 255).astype(np.uint8)
    yield "fib", np.repeat(np.arange(20), [1, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 233, 377, 610, 987,
                                           1597, 2584, 4181, 6765]).astype(np.uint16)
    # the device codebook holds alphabets of up to 2048 symbols; one more
    # goes to the host codebook: both sides of the boundary, with many ties
    for k in (2047, 2048, 2049):
        counts = 1 + rng.integers(0, 4, k)
        yield f"alphabet{k}", rng.permutation(np.repeat(rng.choice(65536, k, replace=False), counts)).astype(np.uint16)
    yield "ties", np.repeat(np.arange(0, 4000, 3), 5).astype(np.uint16)


@pytest.mark.parametrize("name,data", list(_cases()), ids=[c[0] for c in _cases()])
def test_bitstream_matches_oracle(gpu, name, data):
    ref_bytes, (sym, codes, lens) = O.huffman_encode(data)
    r = huffman_encode(data)
    assert np.array_equal(r.symbols, sym.astype(np.uint32))
    assert np.array_equal(r.lengths, lens.astype(np.int32))
    assert np.array_equal(r.codes, codes.astype(np.uint32))
    assert r.bitstream == ref_bytes.tobytes()


def test_device_input_and_zstd_roundtrip(gpu):
    rng = np.random.default_rng(3)
    data = np.minimum(rng.geometric(0.2, 500_000), 60000).astype(np.uint16)
    r = cuda_huffman_zstd_compress(torch.from_numpy(data.view(np.int16)).to(gpu).view(torch.uint16)
                                   if hasattr(torch, "uint16") else data, 20)
    back = cuda_huffman_zstd_decompress(r.compressed_data, r.code_table, r.num_elements,
                                        r.original_bitstream_size, np.uint16)
    assert np.array_equal(back, data)
    assert r.original_bitstream_size == (r.num_bits + 7) // 8
    assert r.compressed_data[:4] == b"\x28\xb5\x2f\xfd"  # ZSTD frame magic


def test_long_codes_raise(gpu):
    # Fibonacci frequencies give a depth n-1 tree: 35 symbols -> 34-bit codes
    f = [1, 1]
    while len(f) < 35:
        f.append(f[-1] + f[-2])
    data = np.repeat(np.arange(35, dtype=np.uint16), f)
    with pytest.raises(ffcz_amd.FFCzError):
        huffman_encode(data)


def test_empty_raises(gpu):
    with pytest.raises(ffcz_amd.FFCzError, match="Empty data"):
        huffman_encode(np.zeros(0, np.uint16))


def test_calls_of_growing_and_shrinking_size_reuse_the_encoder_buffers(gpu):
    """The encoder keeps its device buffers between calls (growing them when a
    call needs more): results stay exact for calls of any order of sizes."""
    rng = np.random.default_rng(9)
    sizes = [1000, 3_000_000, 10, 700_000, 5, 4_500_000, 64]
    for i, n in enumerate(sizes):
        hi = 300 if i % 2 else 3000
        data = rng.integers(0, hi, n).astype(np.uint16 if i % 3 else np.uint8 if hi <= 256 else np.uint16)
        ref_bytes, (sym, codes, lens) = O.huffman_encode(data)
        r = huffman_encode(data)
        assert np.array_equal(r.symbols, sym.astype(np.uint32)), n
        assert np.array_equal(r.codes, codes.astype(np.uint32)), n
        assert r.bitstream == ref_bytes.tobytes(), n
