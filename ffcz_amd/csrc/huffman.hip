// Device Huffman encoder.
//
//  1. histogram: per-workgroup LDS counters (packed 16-bit pairs for the 65536
//     symbol alphabet, chunks of <= 65535 symbols per workgroup), flushed with
//     global atomics;
//  2. codebook on the host from the 64-bit counts (two-queue merge, which pops
//     nodes in the same (freq, index) order as the reference's binary heap);
//  3. lengths -> per-tile bit totals -> scan of tile offsets -> each tile
//     ranks its codes with a workgroup scan, ORs them into an LDS copy of its
//     bit window (MSB-first, stored byte-swapped so memory order is stream
//     order) and writes the window out; the two edge words go through global
//     atomicOr because neighbouring tiles share them.
#include <algorithm>
#include <numeric>

#include "huffman.hpp"

extern "C" {
size_t ZSTD_compressBound(size_t srcSize);
size_t ZSTD_compress(void* dst, size_t dstCapacity, const void* src, size_t srcSize, int compressionLevel);
size_t ZSTD_decompress(void* dst, size_t dstCapacity, const void* src, size_t compressedSize);
unsigned ZSTD_isError(size_t code);
const char* ZSTD_getErrorName(size_t code);
}

namespace ffcz {

constexpr int HUF_TILE = 4096;
constexpr int HIST_CHUNK = 65535;

template <typename S>
__global__ __launch_bounds__(256) void k_hist_small(const S* d, size_t n, unsigned long long* hist) {
    constexpr int NB = 1 << (8 * sizeof(S));
    __shared__ unsigned int h[NB];
    for (int i = threadIdx.x; i < NB; i += 256) h[i] = 0;
    __syncthreads();
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) atomicAdd(&h[d[i]], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < NB; i += 256)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

__global__ __launch_bounds__(1024) void k_hist_u16(const uint16_t* d, size_t n, unsigned long long* hist) {
    __shared__ unsigned int h[32768];  // 65536 packed 16-bit counters
    for (int i = threadIdx.x; i < 32768; i += 1024) h[i] = 0;
    __syncthreads();
    const size_t b0 = (size_t)blockIdx.x * HIST_CHUNK;
    const size_t b1 = b0 + HIST_CHUNK < n ? b0 + HIST_CHUNK : n;
    for (size_t i = b0 + threadIdx.x; i < b1; i += 1024) {
        const unsigned s = d[i];
        atomicAdd(&h[s >> 1], 1u << (16 * (s & 1)));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 32768; i += 1024) {
        const unsigned w = h[i];
        if (w & 0xffffu) atomicAdd(&hist[2 * i], (unsigned long long)(w & 0xffffu));
        if (w >> 16) atomicAdd(&hist[2 * i + 1], (unsigned long long)(w >> 16));
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_len_tiles(const S* d, size_t n, const uint8_t* lens, unsigned int* tsum) {
    const size_t base = (size_t)blockIdx.x * HUF_TILE;
    unsigned int c = 0;
    for (int r = 0; r < HUF_TILE / 256; ++r) {
        const size_t i = base + r * 256 + threadIdx.x;
        if (i < n) c += lens[d[i]];
    }
    __shared__ unsigned int sc[4];
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tsum[blockIdx.x] = sc[0] + sc[1] + sc[2] + sc[3];
}

__global__ __launch_bounds__(1024) void k_scan_u32(const unsigned int* in, unsigned long long* out, size_t ntiles,
                                                   unsigned long long* total) {
    __shared__ unsigned long long wsum[16];
    const int t = threadIdx.x;
    const size_t per = ceil_div(ntiles, (size_t)1024);
    const size_t b0 = t * per, b1 = b0 + per < ntiles ? b0 + per : ntiles;
    unsigned long long s = 0;
    for (size_t i = b0; i < b1; ++i) s += in[i];
    unsigned long long x = s;
    const int lane = t & 63, w = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned long long run = x - s;
    for (int i = 0; i < w; ++i) run += wsum[i];
    for (size_t i = b0; i < b1; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (t == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < 16; ++i) tot += wsum[i];
        *total = tot;
    }
}

__global__ void k_zero_u32(unsigned int* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

__device__ __forceinline__ unsigned int bswap32(unsigned int v) { return __builtin_bswap32(v); }

template <typename S>
__global__ __launch_bounds__(256) void k_emit(const S* d, size_t n, const uint32_t* codes, const uint8_t* lens,
                                              const unsigned long long* toff, const unsigned int* tsum,
                                              unsigned int* out) {
    // tile window: <= 4096*32 bits + 2 edge words
    __shared__ unsigned int win[HUF_TILE + 2];
    __shared__ unsigned int wtot[4];
    const size_t base = (size_t)blockIdx.x * HUF_TILE;
    const unsigned long long B0 = toff[blockIdx.x];
    const unsigned int TB = tsum[blockIdx.x];
    if (TB == 0) return;
    const unsigned long long w0 = B0 >> 5;
    const unsigned long long wlast = (B0 + TB - 1) >> 5;
    const int nw = (int)(wlast - w0 + 1);
    for (int i = threadIdx.x; i < nw; i += 256) win[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long run = B0;
    for (int r = 0; r < HUF_TILE / 256; ++r) {
        const size_t i = base + r * 256 + threadIdx.x;
        unsigned int l = 0, c = 0;
        if (i < n) {
            const unsigned s = d[i];
            l = lens[s];
            c = codes[s];
        }
        // wave-inclusive scan of lengths
        unsigned int x = l;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wtot[w] = x;
        __syncthreads();
        unsigned long long p = run + (x - l);
        for (int k = 0; k < w; ++k) p += wtot[k];
        if (l) {
            const unsigned long long rel = p - (w0 << 5);
            const int o = (int)(rel & 31);
            const int wi = (int)(rel >> 5);
            const unsigned long long v = (unsigned long long)c << (64 - o - l);
            const unsigned int hi = (unsigned int)(v >> 32), lo = (unsigned int)v;
            if (hi) atomicOr(&win[wi], bswap32(hi));
            if (lo) atomicOr(&win[wi + 1], bswap32(lo));
        }
        run += (unsigned long long)wtot[0] + wtot[1] + wtot[2] + wtot[3];
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nw; i += 256) {
        const unsigned int val = win[i];
        if (i == 0 || i == nw - 1) {
            if (val) atomicOr(&out[w0 + i], val);
        } else {
            out[w0 + i] = val;
        }
    }
}

// ---------------------------------------------------------------------------
void huffman_codebook(const std::vector<uint64_t>& freq, std::vector<uint32_t>& code, std::vector<int32_t>& len) {
    const size_t n = freq.size();
    FFCZ_REQUIRE(n > 0, ERR_INVALID, "No symbols found");
    code.assign(n, 0);
    len.assign(n, 0);
    if (n == 1) {  // HuffmanZSTDCoder.cuh:102-111, 168-170
        len[0] = 1;
        return;
    }
    // leaves in (freq, index) order; internal nodes n.. in creation order
    std::vector<uint32_t> leaves(n);
    std::iota(leaves.begin(), leaves.end(), 0u);
    std::stable_sort(leaves.begin(), leaves.end(), [&](uint32_t a, uint32_t b) { return freq[a] < freq[b]; });
    const size_t total = 2 * n - 1;
    std::vector<uint64_t> f(total);
    std::vector<int64_t> left(total, -1), right(total, -1);
    for (size_t i = 0; i < n; ++i) f[i] = freq[i];
    size_t i1 = 0, i2 = n, e2 = n;
    auto pick = [&]() -> size_t {
        if (i1 < n) {
            const size_t a = leaves[i1];
            if (i2 < e2) {
                const size_t b = i2;
                if (f[b] < f[a] || (f[b] == f[a] && b < a)) return i2++;
            }
            ++i1;
            return a;
        }
        return i2++;
    };
    for (size_t k = 0; k + 1 < n; ++k) {
        const size_t a = pick();
        const size_t b = pick();
        f[e2] = f[a] + f[b];
        left[e2] = (int64_t)a;
        right[e2] = (int64_t)b;
        ++e2;
    }
    // DFS from the root: left appends 0, right appends 1 (Huffman:175-176)
    std::vector<std::pair<size_t, std::pair<uint64_t, int>>> st;
    st.push_back({total - 1, {0ull, 0}});
    int maxlen = 0;
    while (!st.empty()) {
        auto [node, cl] = st.back();
        st.pop_back();
        if (node < n) {
            code[node] = (uint32_t)cl.first;
            len[node] = cl.second;
            maxlen = std::max(maxlen, cl.second);
        } else {
            if (cl.second >= 63) throw Error(ERR_RANGE, "Huffman tree too deep");
            st.push_back({(size_t)left[node], {cl.first << 1, cl.second + 1}});
            st.push_back({(size_t)right[node], {(cl.first << 1) | 1ull, cl.second + 1}});
        }
    }
    FFCZ_REQUIRE(maxlen <= 32, ERR_RANGE, "Huffman code longer than 32 bits (undefined in the reference encoder)");
}

template <typename S>
HuffmanCoded huffman_encode_device(const S* d, size_t n, hipStream_t s) {
    FFCZ_REQUIRE(n > 0, ERR_INVALID, "Empty data");
    constexpr size_t NB = (size_t)1 << (8 * sizeof(S));
    HuffmanCoded out;
    out.nelem = n;
    const size_t ntiles = ceil_div(n, (size_t)HUF_TILE);
    unsigned long long* hist = nullptr;
    uint32_t* dcodes = nullptr;
    uint8_t* dlens = nullptr;
    unsigned int* tsum = nullptr;
    unsigned long long* toff = nullptr;
    unsigned long long* dtotal = nullptr;
    unsigned int* words = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(hist);
        (void)hipFree(dcodes);
        (void)hipFree(dlens);
        (void)hipFree(tsum);
        (void)hipFree(toff);
        (void)hipFree(dtotal);
        (void)hipFree(words);
    };
    try {
        FFCZ_HIP(hipMalloc(&hist, NB * sizeof(unsigned long long)));
        FFCZ_HIP(hipMemsetAsync(hist, 0, NB * sizeof(unsigned long long), s));
        if constexpr (sizeof(S) == 1) {
            size_t g = std::min<size_t>(ceil_div(n, (size_t)256 * 16), 4096);
            hipLaunchKernelGGL((k_hist_small<S>), dim3((unsigned)std::max<size_t>(g, 1)), dim3(256), 0, s, d, n, hist);
        } else {
            hipLaunchKernelGGL(k_hist_u16, dim3((unsigned)ceil_div(n, (size_t)HIST_CHUNK)), dim3(1024), 0, s,
                               (const uint16_t*)d, n, hist);
        }
        FFCZ_HIP(hipGetLastError());
        std::vector<unsigned long long> h(NB);
        FFCZ_HIP(hipMemcpyAsync(h.data(), hist, NB * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
        std::vector<uint64_t> freq;
        for (size_t i = 0; i < NB; ++i)
            if (h[i]) {
                out.table.sym.push_back((uint32_t)i);
                freq.push_back(h[i]);
            }
        huffman_codebook(freq, out.table.code, out.table.len);
        std::vector<uint32_t> hc(NB, 0);
        std::vector<uint8_t> hl(NB, 0);
        for (size_t k = 0; k < out.table.sym.size(); ++k) {
            hc[out.table.sym[k]] = out.table.code[k];
            hl[out.table.sym[k]] = (uint8_t)out.table.len[k];
        }
        FFCZ_HIP(hipMalloc(&dcodes, NB * sizeof(uint32_t)));
        FFCZ_HIP(hipMalloc(&dlens, NB));
        FFCZ_HIP(hipMemcpyAsync(dcodes, hc.data(), NB * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        FFCZ_HIP(hipMemcpyAsync(dlens, hl.data(), NB, hipMemcpyHostToDevice, s));
        FFCZ_HIP(hipMalloc(&tsum, ntiles * sizeof(unsigned int)));
        FFCZ_HIP(hipMalloc(&toff, ntiles * sizeof(unsigned long long)));
        FFCZ_HIP(hipMalloc(&dtotal, sizeof(unsigned long long)));
        hipLaunchKernelGGL((k_len_tiles<S>), dim3((unsigned)ntiles), dim3(256), 0, s, d, n, dlens, tsum);
        FFCZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, tsum, toff, ntiles, dtotal);
        FFCZ_HIP(hipGetLastError());
        unsigned long long total = 0;
        FFCZ_HIP(hipMemcpyAsync(&total, dtotal, sizeof(total), hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
        out.nbits = total;
        const size_t nwords = (size_t)(total / 32) + 2;
        FFCZ_HIP(hipMalloc(&words, nwords * sizeof(unsigned int)));
        hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)std::min<size_t>(ceil_div(nwords, (size_t)256), 65536)), dim3(256), 0,
                           s, words, nwords);
        hipLaunchKernelGGL((k_emit<S>), dim3((unsigned)ntiles), dim3(256), 0, s, d, n, dcodes, dlens, toff, tsum, words);
        FFCZ_HIP(hipGetLastError());
        const size_t nbytes = (size_t)((total + 7) / 8);
        out.bitstream.resize(nbytes);
        FFCZ_HIP(hipMemcpyAsync(out.bitstream.data(), words, nbytes, hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    return out;
}

template <typename S>
std::vector<S> huffman_decode(const uint8_t* bits, size_t nbytes, const HuffmanTable& t, size_t n) {
    // binary trie over the codes; the same symbols the reference's bit-serial
    // matcher finds for a prefix code (HuffmanZSTDCoder.cuh:404-447)
    std::vector<int64_t> child(2, -1);  // node*2 + bit
    std::vector<int64_t> leaf(1, -1);
    for (size_t k = 0; k < t.sym.size(); ++k) {
        size_t node = 0;
        for (int b = t.len[k] - 1; b >= 0; --b) {
            const int bit = (t.code[k] >> b) & 1;
            if (child[node * 2 + bit] < 0) {
                child[node * 2 + bit] = (int64_t)leaf.size();
                leaf.push_back(-1);
                child.push_back(-1);
                child.push_back(-1);
            }
            node = (size_t)child[node * 2 + bit];
        }
        leaf[node] = (int64_t)t.sym[k];
    }
    std::vector<S> out;
    out.reserve(n);
    const size_t nb = nbytes * 8;
    size_t p = 0;
    while (out.size() < n && p < nb) {
        size_t node = 0, q = p;
        bool found = false;
        while (q < nb && q - p < 32) {
            const int bit = (bits[q >> 3] >> (7 - (q & 7))) & 1;
            const int64_t c = child[node * 2 + bit];
            ++q;
            if (c < 0) break;
            node = (size_t)c;
            if (leaf[node] >= 0) {
                out.push_back((S)leaf[node]);
                found = true;
                break;
            }
        }
        p = found ? q : p + 1;
    }
    return out;
}

std::vector<uint8_t> zstd_compress(const uint8_t* src, size_t n, int level) {
    std::vector<uint8_t> dst(ZSTD_compressBound(n));
    const size_t r = ZSTD_compress(dst.data(), dst.size(), src, n, level);
    if (ZSTD_isError(r)) throw Error(ERR_ZSTD, std::string("ZSTD compression failed: ") + ZSTD_getErrorName(r));
    dst.resize(r);
    return dst;
}

std::vector<uint8_t> zstd_decompress(const uint8_t* src, size_t n, size_t expected) {
    std::vector<uint8_t> dst(expected);
    const size_t r = ZSTD_decompress(dst.data(), expected, src, n);
    if (ZSTD_isError(r)) throw Error(ERR_ZSTD, std::string("ZSTD decompression failed: ") + ZSTD_getErrorName(r));
    dst.resize(r);
    return dst;
}

template HuffmanCoded huffman_encode_device<uint8_t>(const uint8_t*, size_t, hipStream_t);
template HuffmanCoded huffman_encode_device<uint16_t>(const uint16_t*, size_t, hipStream_t);
template std::vector<uint8_t> huffman_decode<uint8_t>(const uint8_t*, size_t, const HuffmanTable&, size_t);
template std::vector<uint16_t> huffman_decode<uint16_t>(const uint8_t*, size_t, const HuffmanTable&, size_t);

}  // namespace ffcz
