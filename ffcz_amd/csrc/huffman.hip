// Device Huffman encoder.
//
//  1. histogram: per-workgroup LDS counters (packed 16-bit pairs for the 65536
//     symbol alphabet, chunks of <= 65535 symbols per workgroup; one
//     sub-histogram per wave for 8-bit symbols), flushed with global atomics;
//  2. codebook from the 64-bit counts, on the device for alphabets of up to
//     CB_CAP symbols (k_bins, k_codebook), on the host beyond (both a
//     two-queue merge that pops nodes in the same (freq, index) order as the
//     reference's binary heap);
//  3. lengths -> per-tile bit totals -> scan of tile offsets -> each tile
//     ranks its codes with a workgroup scan, ORs them into an LDS copy of its
//     bit window (MSB-first, stored byte-swapped so memory order is stream
//     order) and writes the window out; the two edge words go through global
//     atomicOr because neighbouring tiles share them.
// A call allocates nothing once its per-device buffers are large enough and
// synchronises with the host twice: for the code table and bit count, and for
// the bitstream.
#include <algorithm>
#include <memory>
#include <mutex>
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

// 8-bit symbols: one LDS sub-histogram per wave (less contention between
// waves), 16 symbols per load, and a load whose 16 bytes are one symbol (the
// common case of a skewed stream) added as one count
__device__ __forceinline__ void hist_add16(unsigned int* h, uint4 v) {
    const unsigned int b = v.x & 0xffu;
    const unsigned int rep = b * 0x01010101u;
    if (v.x == rep && v.y == rep && v.z == rep && v.w == rep) {
        atomicAdd(&h[b], 16u);
        return;
    }
    const unsigned int w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) atomicAdd(&h[(w4[k] >> (8 * j)) & 0xffu], 1u);
}
template <typename S>
__global__ __launch_bounds__(256) void k_hist_small(const S* d, size_t n, unsigned long long* hist) {
    static_assert(sizeof(S) == 1, "8-bit symbols");
    __shared__ unsigned int h[4][256];
    for (int i = threadIdx.x; i < 4 * 256; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    unsigned int* hw = h[threadIdx.x >> 6];
    // the 16-byte-aligned body as uint4, the rest byte by byte
    const size_t head = std::min(n, (size_t)((16 - ((uintptr_t)d & 15)) & 15));
    const size_t nv = (n - head) / 16;
    const uint4* dv = reinterpret_cast<const uint4*>(d + head);
    const size_t stride = (size_t)gridDim.x * 256, t0 = blockIdx.x * (size_t)256 + threadIdx.x;
    for (size_t i = t0; i < nv; i += stride) hist_add16(hw, dv[i]);
    for (size_t i = t0; i < head; i += stride) atomicAdd(&hw[d[i]], 1u);
    for (size_t i = head + nv * 16 + t0; i < n; i += stride) atomicAdd(&hw[d[i]], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += 256) {
        const unsigned int c = h[0][i] + h[1][i] + h[2][i] + h[3][i];
        if (c) atomicAdd(&hist[i], (unsigned long long)c);
    }
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

// ---------------------------------------------------------------------------
// Device codebook (alphabets of at most CB_CAP symbols): one workgroup forms
// the same codebook as huffman_codebook.  The symbols present, ascending, are
// leaves 0..n-1; a bitonic sort orders them by (count, leaf index); the
// two-queue merge (one thread) pops nodes in the (count, index) order of the
// reference's binary heap (HuffmanZSTDCoder.cuh:113-150), internal nodes
// numbered from n; a sweep from the root down gives each node its depth and
// code (left 0, right 1, Huffman:161-177).
// ---------------------------------------------------------------------------
constexpr int CB_CAP = 2048;
constexpr int CB_THREADS = 1024;
constexpr int BIN_BLOCK = 1024;  // histogram bins per compaction workgroup

struct CbHeader {
    unsigned long long nbits;  // total coded bits (after the scan)
    unsigned int n;            // symbols present
    unsigned int on_device;    // 1: the codebook below was formed on the device
    unsigned int err;          // 1: a code longer than 32 bits
    unsigned int pad;
};

// The symbols present, per block of BIN_BLOCK bins: each workgroup writes its
// present bins (ascending) and their counts to its own segment and its count
// of them; it also clears the dense length table over its bins.
__global__ __launch_bounds__(BIN_BLOCK) void k_bins(const unsigned long long* hist, size_t nb, uint16_t* segsym,
                                                    unsigned long long* segcnt, unsigned int* blkn, uint8_t* dlens) {
    __shared__ unsigned int wsum[BIN_BLOCK / 64];
    const size_t b = blockIdx.x * (size_t)BIN_BLOCK + threadIdx.x;
    const unsigned long long c = b < nb ? hist[b] : 0ull;
    if (b < nb) dlens[b] = 0;
    const unsigned long long m = __ballot(c != 0ull);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = (unsigned int)__popcll(m);
    __syncthreads();
    unsigned int pos = (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
    for (int i = 0; i < w; ++i) pos += wsum[i];
    if (c != 0ull) {
        segsym[blockIdx.x * (size_t)BIN_BLOCK + pos] = (uint16_t)b;
        segcnt[blockIdx.x * (size_t)BIN_BLOCK + pos] = c;
    }
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int i = 0; i < BIN_BLOCK / 64; ++i) t += wsum[i];
        blkn[blockIdx.x] = t;
    }
}

// One workgroup: the symbols present, ascending, become leaves 0..n-1; a
// bitonic sort orders them by (count, leaf index); the two-queue merge (one
// thread) pops nodes in the (count, index) order of the reference's binary
// heap (HuffmanZSTDCoder.cuh:113-150), internal nodes numbered from n; a
// sweep from the root down gives every node its depth and code (left 0,
// right 1, Huffman:161-177).  Alphabets beyond CB_CAP are left to the host.
__global__ __launch_bounds__(CB_THREADS) void k_codebook(const uint16_t* segsym, const unsigned long long* segcnt,
                                                         const unsigned int* blkn, int nblk, uint32_t* dcodes,
                                                         uint8_t* dlens, CbHeader* hdr, uint32_t* tsym,
                                                         uint32_t* tcode, uint8_t* tlen) {
    __shared__ unsigned long long key[CB_CAP];      // (count << 16 | leaf), sorted; then the internal codes
    __shared__ unsigned long long fint[CB_CAP];     // internal node counts, by creation
    __shared__ uint32_t lcode[CB_CAP];              // leaf codes
    __shared__ uint16_t kid[2][CB_CAP];             // children of internal node n + k
    __shared__ uint16_t sym[CB_CAP];                // leaf -> symbol
    __shared__ uint8_t depth[2 * CB_CAP];
    __shared__ unsigned int boff[64], total_n, bad;
    const int t = threadIdx.x;
    if (t < 64) {
        // block offsets: a wave scan over at most 64 blocks
        const unsigned int c = t < nblk ? blkn[t] : 0u;
        unsigned int x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int y = __shfl_up(x, o, 64);
            if (t >= o) x += y;
        }
        boff[t] = x - c;
        if (t == 63) total_n = x;
    }
    __syncthreads();
    const unsigned int n = total_n;
    if (t == 0) {
        hdr->n = n;
        hdr->on_device = (n <= (unsigned)CB_CAP && n > 0) ? 1u : 0u;
        hdr->err = 0u;
    }
    if (n > (unsigned)CB_CAP || n == 0) return;
    unsigned int P = 1;
    while (P < n) P <<= 1;
    for (int b = 0; b < nblk; ++b) {
        const unsigned int cnt = blkn[b], o = boff[b];
        for (unsigned int i = t; i < cnt; i += CB_THREADS) {
            const unsigned int leaf = o + i;
            sym[leaf] = segsym[b * (size_t)BIN_BLOCK + i];
            key[leaf] = (segcnt[b * (size_t)BIN_BLOCK + i] << 16) | leaf;
        }
    }
    for (unsigned int i = n + t; i < P; i += CB_THREADS) key[i] = ~0ull;
    __syncthreads();
    for (unsigned int k = 2; k <= P; k <<= 1) {
        for (unsigned int j = k >> 1; j > 0; j >>= 1) {
            for (unsigned int i = t; i < P; i += CB_THREADS) {
                const unsigned int l = i ^ j;
                if (l > i) {
                    const unsigned long long a = key[i], b = key[l];
                    if ((a > b) == ((i & k) == 0)) {
                        key[i] = b;
                        key[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (t == 0) {
        bad = 0u;
        if (n == 1) {  // HuffmanZSTDCoder.cuh:102-111, 168-170: code 0, length 1
            depth[0] = 1;
            lcode[0] = 0u;
        } else {
            // two-queue merge: leaves by (count, leaf), internal nodes by
            // creation (their counts never decrease); a tie goes to the leaf,
            // whose index is the smaller.  A leaf's count travels in its key.
            unsigned int i1 = 0, i2 = 0, e2 = 0;  // internal queue [i2, e2) as k = node - n
            unsigned long long kl = key[0];
            auto pick = [&](unsigned long long& cnt) -> unsigned int {
                if (i1 < n) {
                    const unsigned long long lc = kl >> 16;
                    if (i2 < e2 && fint[i2] < lc) {
                        cnt = fint[i2];
                        return n + i2++;
                    }
                    const unsigned int a = (unsigned int)(kl & 0xffffull);
                    cnt = lc;
                    ++i1;
                    kl = i1 < n ? key[i1] : 0ull;
                    return a;
                }
                cnt = fint[i2];
                return n + i2++;
            };
            for (unsigned int k = 0; k + 1 < n; ++k) {
                unsigned long long ca, cb;
                const unsigned int a = pick(ca);
                const unsigned int b = pick(cb);
                fint[e2] = ca + cb;
                kid[0][e2] = (uint16_t)a;
                kid[1][e2] = (uint16_t)b;
                ++e2;
            }
            // depths and codes from the root down: children precede their
            // parent, so a sweep by decreasing index meets parents first; the
            // sort keys are spent, so key[] holds the internal nodes' codes
            const unsigned int root = 2 * n - 2;
            depth[root] = 0;
            key[root - n] = 0ull;
            for (unsigned int q = root; q >= n; --q) {
                const unsigned long long v = key[q - n];
                const uint8_t d = depth[q];
                if (d >= 32) bad = 1u;
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const unsigned int c = kid[side][q - n];
                    const unsigned long long cv = (v << 1) | (unsigned long long)side;
                    depth[c] = (uint8_t)(d + 1);
                    if (c < n) lcode[c] = (uint32_t)cv;
                    else key[c - n] = cv;
                }
            }
        }
        hdr->err = bad;
    }
    __syncthreads();
    if (bad) return;
    for (unsigned int i = t; i < n; i += CB_THREADS) {
        const uint32_t cv = lcode[i];
        const uint8_t l = depth[i];
        dcodes[sym[i]] = cv;
        dlens[sym[i]] = l;
        tsym[i] = sym[i];
        tcode[i] = cv;
        tlen[i] = l;
    }
}

// the edge words of each tile (shared with its neighbours, filled by
// atomicOr in k_emit) start at zero; interior words are stored whole
__global__ void k_zero_edges(const unsigned long long* toff, const unsigned int* tsum, size_t ntiles,
                             unsigned int* out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ntiles; i += (size_t)gridDim.x * blockDim.x) {
        const unsigned int tb = tsum[i];
        if (tb == 0) continue;
        const unsigned long long b0 = toff[i];
        out[b0 >> 5] = 0u;
        out[(b0 + tb - 1) >> 5] = 0u;
    }
}

namespace {
// Encoder buffers, kept per device and grown when a call needs more, so a
// call allocates nothing once they are large enough.  Calls on one device
// take turns (each finishes its stream before returning).
struct HuffWorkspace {
    std::mutex mu;
    unsigned char* d = nullptr;
    size_t dcap = 0;
    unsigned char* h = nullptr;  // pinned: header and compact table
    size_t hcap = 0;
    unsigned char* dev(size_t bytes) {
        if (bytes > dcap) {
            const size_t cap = std::max(bytes, dcap + dcap / 2);
            if (d) FFCZ_HIP(hipFree(d));
            d = nullptr;
            dcap = 0;
            FFCZ_HIP(hipMalloc(&d, cap));
            dcap = cap;
        }
        return d;
    }
    hipEvent_t ev[4] = {};
    hipEvent_t* events() {
        if (!ev[0])
            for (auto& e : ev) FFCZ_HIP(hipEventCreate(&e));
        return ev;
    }
    unsigned char* host(size_t bytes) {
        if (bytes > hcap) {
            const size_t cap = std::max(bytes, hcap + hcap / 2);
            if (h) FFCZ_HIP(hipHostFree(h));
            h = nullptr;
            hcap = 0;
            FFCZ_HIP(hipHostMalloc(&h, cap, hipHostMallocDefault));
            hcap = cap;
        }
        return h;
    }
};
HuffWorkspace& workspace() {
    static std::mutex mu;
    static std::vector<std::unique_ptr<HuffWorkspace>> ws;
    int dev = 0;
    FFCZ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if ((size_t)dev >= ws.size()) ws.resize((size_t)dev + 1);
    if (!ws[dev]) ws[dev].reset(new HuffWorkspace());
    return *ws[dev];
}
size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// a few host buffers, kept with their pages mapped
std::mutex g_pool_mu;
std::vector<std::vector<uint8_t>> g_pool;
constexpr size_t POOL_MAX = 4, POOL_MAX_BYTES = (size_t)1 << 28;
std::vector<uint8_t> take_bitstream(size_t bytes) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (size_t i = 0; i < g_pool.size(); ++i)
        if (g_pool[i].capacity() >= bytes) {
            std::vector<uint8_t> v = std::move(g_pool[i]);
            g_pool.erase(g_pool.begin() + (long)i);
            v.clear();
            return v;
        }
    return std::vector<uint8_t>();
}
}  // namespace

void recycle_bitstream(std::vector<uint8_t>&& v) {
    if (v.capacity() == 0 || v.capacity() > POOL_MAX_BYTES) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (g_pool.size() >= POOL_MAX) {
        // keep the larger buffers
        auto it = std::min_element(g_pool.begin(), g_pool.end(),
                                   [](const std::vector<uint8_t>& a, const std::vector<uint8_t>& b) {
                                       return a.capacity() < b.capacity();
                                   });
        if (it->capacity() >= v.capacity()) return;
        g_pool.erase(it);
    }
    g_pool.push_back(std::move(v));
}

template <typename S>
HuffmanCoded huffman_encode_device(const S* d, size_t n, hipStream_t s) {
    FFCZ_REQUIRE(n > 0, ERR_INVALID, "Empty data");
    constexpr size_t NB = (size_t)1 << (8 * sizeof(S));
    HuffmanCoded out;
    out.nelem = n;
    const size_t ntiles = ceil_div(n, (size_t)HUF_TILE);
    HuffWorkspace& ws = workspace();
    std::lock_guard<std::mutex> lk(ws.mu);
    // device layout: histogram, dense code and length tables, per-block
    // present bins, tile bit counts and offsets; the header and compact table
    // (copied back in one transfer, the same layout on the host); then the
    // coded words
    constexpr size_t NBLK = (NB + BIN_BLOCK - 1) / BIN_BLOCK;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += al256(bytes);
        return o;
    };
    const size_t o_hist = take(NB * 8), o_codes = take(NB * 4), o_lens = take(NB), o_segsym = take(NB * 2),
                 o_segcnt = take(NB * 8), o_blkn = take(NBLK * 4), o_tsum = take(ntiles * 4),
                 o_toff = take(ntiles * 8);
    const size_t r_tsym = al256(sizeof(CbHeader)), r_tcode = r_tsym + CB_CAP * 4, r_tlen = r_tcode + CB_CAP * 4,
                 r_bytes = r_tlen + CB_CAP;
    const size_t o_res = take(r_bytes), o_words = off;
    unsigned char* base = ws.dev(o_words + 64);
    auto* hist = reinterpret_cast<unsigned long long*>(base + o_hist);
    auto* dcodes = reinterpret_cast<uint32_t*>(base + o_codes);
    auto* dlens = reinterpret_cast<uint8_t*>(base + o_lens);
    auto* segsym = reinterpret_cast<uint16_t*>(base + o_segsym);
    auto* segcnt = reinterpret_cast<unsigned long long*>(base + o_segcnt);
    auto* blkn = reinterpret_cast<unsigned int*>(base + o_blkn);
    auto* tsum = reinterpret_cast<unsigned int*>(base + o_tsum);
    auto* toff = reinterpret_cast<unsigned long long*>(base + o_toff);
    auto* dhdr = reinterpret_cast<CbHeader*>(base + o_res);
    unsigned char* hb = ws.host(r_bytes);
    auto* hhdr = reinterpret_cast<CbHeader*>(hb);

    hipEvent_t* ev = ws.events();
    FFCZ_HIP(hipEventRecord(ev[0], s));
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)ceil_div(NB * 2, (size_t)256)), dim3(256), 0, s,
                       reinterpret_cast<unsigned int*>(hist), NB * 2);
    if constexpr (sizeof(S) == 1) {
        size_t g = std::min<size_t>(ceil_div(n, (size_t)256 * 64), 2048);
        hipLaunchKernelGGL((k_hist_small<S>), dim3((unsigned)std::max<size_t>(g, 1)), dim3(256), 0, s, d, n, hist);
    } else {
        hipLaunchKernelGGL(k_hist_u16, dim3((unsigned)ceil_div(n, (size_t)HIST_CHUNK)), dim3(1024), 0, s,
                           (const uint16_t*)d, n, hist);
    }
    FFCZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bins, dim3((unsigned)NBLK), dim3(BIN_BLOCK), 0, s, hist, NB, segsym, segcnt, blkn, dlens);
    hipLaunchKernelGGL(k_codebook, dim3(1), dim3(CB_THREADS), 0, s, segsym, segcnt, blkn, (int)NBLK, dcodes, dlens,
                       dhdr, reinterpret_cast<uint32_t*>(base + o_res + r_tsym),
                       reinterpret_cast<uint32_t*>(base + o_res + r_tcode),
                       reinterpret_cast<uint8_t*>(base + o_res + r_tlen));
    FFCZ_HIP(hipGetLastError());
    auto lengths_and_scan = [&] {
        hipLaunchKernelGGL((k_len_tiles<S>), dim3((unsigned)ntiles), dim3(256), 0, s, d, n, dlens, tsum);
        FFCZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, s, tsum, toff, ntiles, &dhdr->nbits);
        FFCZ_HIP(hipGetLastError());
    };
    // a device codebook: the lengths and the scan follow at once; the header
    // and compact table come back with one synchronisation
    lengths_and_scan();
    FFCZ_HIP(hipEventRecord(ev[1], s));
    FFCZ_HIP(hipMemcpyAsync(hb, base + o_res, r_bytes, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    const CbHeader hd = *hhdr;
    FFCZ_REQUIRE(hd.n > 0, ERR_INVALID, "No symbols found");
    if (hd.on_device) {
        FFCZ_REQUIRE(!hd.err, ERR_RANGE, "Huffman code longer than 32 bits (undefined in the reference encoder)");
        const auto* hs = reinterpret_cast<const uint32_t*>(hb + r_tsym);
        const auto* hc = reinterpret_cast<const uint32_t*>(hb + r_tcode);
        const auto* hl = hb + r_tlen;
        out.table.sym.assign(hs, hs + hd.n);
        out.table.code.assign(hc, hc + hd.n);
        out.table.len.assign(hl, hl + hd.n);
        out.nbits = hd.nbits;
    } else {
        // a larger alphabet: the codebook on the host from the histogram (the
        // lengths launched above read the zeroed length table and are redone)
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
        FFCZ_HIP(hipMemcpyAsync(dcodes, hc.data(), NB * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        FFCZ_HIP(hipMemcpyAsync(dlens, hl.data(), NB, hipMemcpyHostToDevice, s));
        lengths_and_scan();
        FFCZ_HIP(hipMemcpyAsync(hb, dhdr, sizeof(CbHeader), hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
        out.nbits = hhdr->nbits;
    }
    const unsigned long long total = out.nbits;
    const size_t nwords = (size_t)(total / 32) + 2;
    // the coded words after the tables (growing the buffer keeps nothing the
    // emission needs: tables and offsets are written again below if it moved)
    unsigned char* base2 = ws.dev(o_words + nwords * 4);
    if (base2 != base) {
        // a new allocation: the tables were lost with the old one
        base = base2;
        hist = reinterpret_cast<unsigned long long*>(base + o_hist);
        dcodes = reinterpret_cast<uint32_t*>(base + o_codes);
        dlens = reinterpret_cast<uint8_t*>(base + o_lens);
        tsum = reinterpret_cast<unsigned int*>(base + o_tsum);
        toff = reinterpret_cast<unsigned long long*>(base + o_toff);
        dhdr = reinterpret_cast<CbHeader*>(base + o_res);
        std::vector<uint32_t> hc(NB, 0);
        std::vector<uint8_t> hl(NB, 0);
        for (size_t k = 0; k < out.table.sym.size(); ++k) {
            hc[out.table.sym[k]] = out.table.code[k];
            hl[out.table.sym[k]] = (uint8_t)out.table.len[k];
        }
        FFCZ_HIP(hipMemcpyAsync(dcodes, hc.data(), NB * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        FFCZ_HIP(hipMemcpyAsync(dlens, hl.data(), NB, hipMemcpyHostToDevice, s));
        lengths_and_scan();
    }
    auto* words = reinterpret_cast<unsigned int*>(base + o_words);
    FFCZ_HIP(hipEventRecord(ev[2], s));
    hipLaunchKernelGGL(k_zero_edges, dim3((unsigned)std::min<size_t>(ceil_div(ntiles, (size_t)256), 4096)), dim3(256),
                       0, s, toff, tsum, ntiles, words);
    hipLaunchKernelGGL((k_emit<S>), dim3((unsigned)ntiles), dim3(256), 0, s, d, n, dcodes, dlens, toff, tsum, words);
    FFCZ_HIP(hipGetLastError());
    FFCZ_HIP(hipEventRecord(ev[3], s));
    const size_t nbytes = (size_t)((total + 7) / 8);
    // through the pinned buffer: a copy into pageable memory runs at a
    // fraction of the link's rate
    unsigned char* hbits = ws.host(std::max(nbytes, r_bytes));
    FFCZ_HIP(hipMemcpyAsync(hbits, words, nbytes, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    out.bitstream = take_bitstream(nbytes);
    out.bitstream.assign(hbits, hbits + nbytes);
    // the device work: up to the scan, and the emission (the host codebook
    // of a large alphabet runs in between and is not counted)
    float t1 = 0.f, t2 = 0.f;
    FFCZ_HIP(hipEventElapsedTime(&t1, ev[0], ev[1]));
    FFCZ_HIP(hipEventElapsedTime(&t2, ev[2], ev[3]));
    out.device_ms = t1 + t2;
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
