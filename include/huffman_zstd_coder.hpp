// huffman_zstd_coder.hpp - C++ drop-in for the reference entropy coder
// (HuffmanZSTDCoder.cuh:235-470), header-only over the C ABI of ffcz.h.
// Histogram and bit emission run on the GPU, the codebook on the host with the
// reference's (frequency, index) merge order, ZSTD on host libzstd.  Codes
// longer than 32 bits raise (the reference stores uint32 codes and silently
// truncates, HuffmanZSTDCoder.cuh:159).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ffcz.h"

namespace ffcz {

template <typename T>
class HuffmanZstdCompressor {
    static_assert(std::is_same<T, uint8_t>::value || std::is_same<T, uint16_t>::value,
                  "symbols are uint8_t or uint16_t");
    static constexpr bool U16 = std::is_same<T, uint16_t>::value;

    static void check(int s) {
        if (s != FFCZ_OK) throw std::runtime_error(ffcz_last_error());
    }

   public:
    struct CompressionResult {
        std::vector<uint8_t> compressed_data;
        std::unordered_map<T, std::pair<uint32_t, int>> code_table;
        size_t original_bitstream_size = 0;
        size_t num_elements = 0;
    };

    CompressionResult compress(const std::vector<T>& data, int zstd_level = 3) {
        if (data.empty()) throw std::runtime_error("Empty data");  // HuffmanZSTDCoder.cuh:250-253
        ffcz_coded* c = nullptr;
        if constexpr (U16) check(ffcz_huffman_zstd_compress_u16(data.data(), data.size(), zstd_level, &c));
        else check(ffcz_huffman_zstd_compress_u8(data.data(), data.size(), zstd_level, &c));
        CompressionResult r;
        size_t cb = 0, bb = 0, ne = 0, ns = 0;
        unsigned long long nbits = 0;
        int st = ffcz_coded_info(c, &cb, &bb, &ne, &ns, &nbits);
        std::vector<uint32_t> sym(ns), code(ns);
        std::vector<int32_t> len(ns);
        if (st == FFCZ_OK) {
            r.compressed_data.resize(cb);
            st = ffcz_coded_data(c, r.compressed_data.data(), cb);
        }
        if (st == FFCZ_OK) st = ffcz_coded_table(c, sym.data(), code.data(), len.data(), ns);
        ffcz_coded_free(c);
        check(st);
        for (size_t i = 0; i < ns; ++i) r.code_table[(T)sym[i]] = {code[i], len[i]};
        r.original_bitstream_size = bb;
        r.num_elements = ne;
        return r;
    }

    std::vector<T> decompress(const std::vector<uint8_t>& compressed_data,
                              const std::unordered_map<T, std::pair<uint32_t, int>>& code_table,
                              size_t num_elements, size_t original_bitstream_size) {
        std::vector<uint32_t> sym, code;
        std::vector<int32_t> len;
        for (const auto& kv : code_table) {
            sym.push_back(kv.first);
            code.push_back(kv.second.first);
            len.push_back(kv.second.second);
        }
        std::vector<T> out(num_elements);
        size_t n = 0;
        if constexpr (U16)
            check(ffcz_huffman_zstd_decompress_u16(compressed_data.data(), compressed_data.size(), sym.data(),
                                                   code.data(), len.data(), sym.size(), num_elements,
                                                   original_bitstream_size, out.data(), &n));
        else
            check(ffcz_huffman_zstd_decompress_u8(compressed_data.data(), compressed_data.size(), sym.data(),
                                                  code.data(), len.data(), sym.size(), num_elements,
                                                  original_bitstream_size, out.data(), &n));
        out.resize(n);
        return out;
    }
};

}  // namespace ffcz

// Names the reference's callers use (HuffmanZSTDCoder.cuh:237, 454-470).
template <typename T>
using SimpleCudaHuffmanZstdCompressor = ffcz::HuffmanZstdCompressor<T>;

template <typename T>
typename ffcz::HuffmanZstdCompressor<T>::CompressionResult cudaHuffmanZstdCompress(const std::vector<T>& data,
                                                                                   int zstd_level = 3) {
    return ffcz::HuffmanZstdCompressor<T>().compress(data, zstd_level);
}

template <typename T>
std::vector<T> cudaHuffmanZstdDecompress(const std::vector<uint8_t>& compressed_data,
                                         const std::unordered_map<T, std::pair<uint32_t, int>>& code_table,
                                         size_t num_elements, size_t original_bitstream_size) {
    return ffcz::HuffmanZstdCompressor<T>().decompress(compressed_data, code_table, num_elements,
                                                       original_bitstream_size);
}
