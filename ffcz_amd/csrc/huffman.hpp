// Huffman + ZSTD entropy coder, bit-exact with the reference encoder
// (HuffmanZSTDCoder.cuh:48-357) for identical symbols.
#pragma once

#include <vector>

#include "ffcz_common.hpp"

namespace ffcz {

struct HuffmanTable {
    std::vector<uint32_t> sym;   // ascending distinct symbols
    std::vector<uint32_t> code;  // MSB-first code bits
    std::vector<int32_t> len;    // code lengths (1..32)
};

struct HuffmanCoded {
    std::vector<uint8_t> bitstream;  // ceil(bits/8) bytes, MSB first, zero padded
    HuffmanTable table;
    unsigned long long nbits = 0;
    size_t nelem = 0;
    float device_ms = 0.f;  // the encoder's device work (histogram to emitted words), HIP events
};

// Host bitstream buffers are reused between calls (a fresh buffer's pages
// would be faulted in by the copy into it): a coded result's owner hands its
// buffer back here when it is done with it.
void recycle_bitstream(std::vector<uint8_t>&& v);

// Codebook of HuffmanZSTDCoder.cuh:86-178 from (ascending symbols, counts).
// Throws ERR_RANGE if a code would exceed 32 bits.
void huffman_codebook(const std::vector<uint64_t>& freq, std::vector<uint32_t>& code, std::vector<int32_t>& len);

// Histogram, codebook and bit emission; ``data`` is a device pointer.
template <typename S>
HuffmanCoded huffman_encode_device(const S* data, size_t n, hipStream_t s);

template <typename S>
std::vector<S> huffman_decode(const uint8_t* bits, size_t nbytes, const HuffmanTable& t, size_t n);

std::vector<uint8_t> zstd_compress(const uint8_t* src, size_t n, int level);
std::vector<uint8_t> zstd_decompress(const uint8_t* src, size_t n, size_t expected);

}  // namespace ffcz
