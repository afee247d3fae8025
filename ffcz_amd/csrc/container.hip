// Container of one solve's coded edits, and the GPU reconstruction.
//
// Layout (little-endian; offsets in bytes):
//   header (128):
//     0 "FFCZ"   4 u32 version (1)   8 u32 dtype (0 float, 1 double)   12 u32 mode
//     16 u64 nx  24 u64 ny  32 u64 nz
//     40 f64 eps  48 f64 delta (the abs / ptw bound; NaN for per-bin bounds)
//     56 u32 quantiser (1 reference 16-bit, 2 step)  60 u32 streams (4)
//     64 f64 frequency step  72 f64 spatial step  (step quantiser)
//     80 f64 frequency min, max  96 f64 spatial min, max  (the reference quantiser's ranges)
//     112 u64 active frequency edits  120 u64 active spatial edits
//   four streams, in the order the reference driver codes them
//   (projection_example.cu:202-246): frequency flags (u8 bytes of the
//   LSB-first flag pack, ceil(2F/8)), spatial flags (ceil(N/8)), frequency
//   symbols (u16), spatial symbols (u16).  Each stream:
//     u32 symbol width (1, 2)  u32 zstd (payload is ZSTD of the bitstream)
//     u64 symbols  u64 coded bits  u64 bitstream bytes  u64 payload bytes
//     u32 table size k  u32 chunk (symbols per index entry, 1024)
//     u32 symbol[k]  u32 code[k]  u8 length[k]   (the encoder's own codes)
//     u16 coded bits of each chunk of symbols[ceil(symbols/chunk)]
//     payload
// The bitstream is the reference encoder's (HuffmanZSTDCoder.cuh:307-357);
// the chunk index is side information that lets a decoder start anywhere.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "container.hpp"
#include "huffman.hpp"
#include "quant.hpp"

namespace ffcz {

namespace {

constexpr char MAGIC[4] = {'F', 'F', 'C', 'Z'};
constexpr uint32_t VERSION = 1;
constexpr size_t HDR_BYTES = 128;
constexpr uint32_t CHUNK = 1024;  // 1024 codes of <= 32 bits: a chunk's bit count fits 16 bits

struct Writer {
    std::vector<uint8_t>& b;
    template <typename V>
    void put(V v) {
        const size_t o = b.size();
        b.resize(o + sizeof(V));
        std::memcpy(b.data() + o, &v, sizeof(V));
    }
    void bytes(const void* p, size_t n) {
        const size_t o = b.size();
        b.resize(o + n);
        if (n) std::memcpy(b.data() + o, p, n);
    }
};

struct Reader {
    const uint8_t* p;
    size_t n;
    size_t o = 0;
    const uint8_t* take(size_t k) {
        FFCZ_REQUIRE(o <= n && n - o >= k, ERR_INVALID, "truncated container");
        const uint8_t* q = p + o;
        o += k;
        return q;
    }
    template <typename V>
    V get() {
        V v;
        std::memcpy(&v, take(sizeof(V)), sizeof(V));
        return v;
    }
};

// device memory owned for the length of one call
struct DBuf {
    void* p = nullptr;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename V>
    V* alloc(size_t n) {
        FFCZ_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(V)));
        return static_cast<V*>(p);
    }
};

// Waits for the stream when it goes out of scope, so that kernels still
// reading call-owned buffers (an exception's way out included) finish before
// the buffers declared ahead of it are freed.
struct StreamWait {
    hipStream_t s;
    ~StreamWait() { (void)hipStreamSynchronize(s); }
};

bool on_device(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // a host pointer is not an error here
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// coded bits of each CHUNK-symbol chunk (one workgroup per chunk)
template <typename S>
__global__ __launch_bounds__(256) void k_chunk_bits(const S* d, size_t n, const uint8_t* lens, uint16_t* out) {
    __shared__ unsigned int red[256];
    const size_t c0 = (size_t)blockIdx.x * CHUNK;
    unsigned int b = 0;
    for (unsigned int i = threadIdx.x; i < CHUNK; i += 256) {
        const size_t j = c0 + i;
        if (j < n) b += lens[d[j]];
    }
    red[threadIdx.x] = b;
    __syncthreads();
    for (unsigned int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = (uint16_t)red[0];
}

template <typename S>
void pack_stream(Writer& w, const S* d, size_t n, int level, hipStream_t s) {
    w.put<uint32_t>((uint32_t)sizeof(S));
    if (n == 0) {
        w.put<uint32_t>(0);
        for (int i = 0; i < 4; ++i) w.put<uint64_t>(0);
        w.put<uint32_t>(0);
        w.put<uint32_t>(CHUNK);
        return;
    }
    const HuffmanCoded h = huffman_encode_device<S>(d, n, s);
    // the chunk index from the code lengths of the symbols in place
    constexpr size_t NB = (size_t)1 << (8 * sizeof(S));
    std::vector<uint8_t> hl(NB, 0);
    for (size_t k = 0; k < h.table.sym.size(); ++k) hl[h.table.sym[k]] = (uint8_t)h.table.len[k];
    const size_t nchunks = ceil_div(n, (size_t)CHUNK);
    std::vector<uint16_t> cbits(nchunks);
    {
        DBuf dl, dc;
        const StreamWait wait{s};
        uint8_t* lens = dl.alloc<uint8_t>(NB);
        uint16_t* cb = dc.alloc<uint16_t>(nchunks);
        FFCZ_HIP(hipMemcpyAsync(lens, hl.data(), NB, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_chunk_bits<S>), dim3((unsigned)nchunks), dim3(256), 0, s, d, n, lens, cb);
        FFCZ_HIP(hipGetLastError());
        FFCZ_HIP(hipMemcpyAsync(cbits.data(), cb, nchunks * sizeof(uint16_t), hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
    }
    const bool z = level > 0;
    const std::vector<uint8_t> payload = z ? zstd_compress(h.bitstream.data(), h.bitstream.size(), level) : h.bitstream;
    w.put<uint32_t>(z ? 1u : 0u);
    w.put<uint64_t>(n);
    w.put<uint64_t>(h.nbits);
    w.put<uint64_t>(h.bitstream.size());
    w.put<uint64_t>(payload.size());
    const size_t k = h.table.sym.size();
    w.put<uint32_t>((uint32_t)k);
    w.put<uint32_t>(CHUNK);
    w.bytes(h.table.sym.data(), k * sizeof(uint32_t));
    w.bytes(h.table.code.data(), k * sizeof(uint32_t));
    for (size_t i = 0; i < k; ++i) w.put<uint8_t>((uint8_t)h.table.len[i]);
    w.bytes(cbits.data(), nchunks * sizeof(uint16_t));
    w.bytes(payload.data(), payload.size());
}

struct Stream {
    uint32_t width = 0, zstd = 0, chunk = 0;
    uint64_t nsym = 0, nbits = 0, bsbytes = 0, plbytes = 0;
    HuffmanTable t;
    std::vector<uint16_t> cbits;
    const uint8_t* payload = nullptr;
};

Stream read_stream(Reader& r, uint32_t width) {
    Stream st;
    st.width = r.get<uint32_t>();
    FFCZ_REQUIRE(st.width == width, ERR_INVALID, "container stream has the wrong symbol width");
    st.zstd = r.get<uint32_t>();
    st.nsym = r.get<uint64_t>();
    st.nbits = r.get<uint64_t>();
    st.bsbytes = r.get<uint64_t>();
    st.plbytes = r.get<uint64_t>();
    const uint32_t k = r.get<uint32_t>();
    st.chunk = r.get<uint32_t>();
    FFCZ_REQUIRE(st.zstd <= 1 && st.chunk == CHUNK, ERR_INVALID, "unsupported container stream");
    FFCZ_REQUIRE(st.nsym <= ((uint64_t)1 << 44), ERR_INVALID, "container stream too large");
    FFCZ_REQUIRE(st.bsbytes == (st.nbits + 7) / 8 && st.nbits <= 32 * st.nsym && st.nbits >= st.nsym, ERR_INVALID,
                 "inconsistent container stream sizes");
    FFCZ_REQUIRE((st.nsym == 0) == (k == 0) && k <= ((uint64_t)1 << (8 * width)), ERR_INVALID,
                 "bad container code table size");
    const uint8_t* sym = r.take((size_t)k * 4);
    const uint8_t* code = r.take((size_t)k * 4);
    const uint8_t* len = r.take(k);
    st.t.sym.resize(k);
    st.t.code.resize(k);
    st.t.len.resize(k);
    if (k) {
        std::memcpy(st.t.sym.data(), sym, (size_t)k * 4);
        std::memcpy(st.t.code.data(), code, (size_t)k * 4);
    }
    for (uint32_t i = 0; i < k; ++i) {
        st.t.len[i] = len[i];
        FFCZ_REQUIRE(len[i] >= 1 && len[i] <= 32, ERR_INVALID, "bad container code length");
        FFCZ_REQUIRE(len[i] == 32 || (st.t.code[i] >> len[i]) == 0, ERR_INVALID, "bad container code");
        FFCZ_REQUIRE(st.t.sym[i] < ((uint64_t)1 << (8 * width)), ERR_INVALID, "bad container symbol");
    }
    const size_t nchunks = ceil_div((size_t)st.nsym, (size_t)CHUNK);
    const uint8_t* cb = r.take(nchunks * sizeof(uint16_t));
    st.cbits.resize(nchunks);
    if (nchunks) std::memcpy(st.cbits.data(), cb, nchunks * sizeof(uint16_t));
    uint64_t tot = 0;
    for (uint16_t c : st.cbits) tot += c;
    FFCZ_REQUIRE(tot == st.nbits, ERR_INVALID, "container chunk index does not add up to the stream");
    st.payload = r.take(st.plbytes);
    FFCZ_REQUIRE(st.zstd || st.plbytes == st.bsbytes, ERR_INVALID, "inconsistent container payload size");
    return st;
}

// A stream made ready for the device decoder on the host: its bitstream
// (ZSTD undone), decode tables and chunk bit offsets.  Building the tables
// checks that the code table is a prefix code.
constexpr int LUT_BITS = 12;
constexpr uint32_t LUT_BAD = 0xFFFFFFFFu;

struct Prepared {
    const Stream* st = nullptr;
    std::vector<uint8_t> raw;  // decompressed bitstream (ZSTD payloads)
    const uint8_t* bits = nullptr;
    // lut[next 12 bits]: (length << 16) | symbol for codes of <= 12 bits, or
    // bit 31 | the trie node reached after those 12 bits for longer codes
    std::vector<uint32_t> lut;
    std::vector<int32_t> child;  // trie: node*2 + bit -> node (-1: none)
    std::vector<int32_t> leaf;   // node -> symbol (-1: inner node)
    std::vector<unsigned long long> off;  // bit offset of each chunk, and the total
};

Prepared prepare(const Stream& st) {
    Prepared pr;
    pr.st = &st;
    if (st.nsym == 0) return pr;
    pr.bits = st.payload;
    if (st.zstd) {
        pr.raw = zstd_decompress(st.payload, st.plbytes, st.bsbytes);
        FFCZ_REQUIRE(pr.raw.size() == st.bsbytes, ERR_INVALID, "container payload does not decompress to its bitstream");
        pr.bits = pr.raw.data();
    }
    pr.child.assign(2, -1);
    pr.leaf.assign(1, -1);
    for (size_t k = 0; k < st.t.sym.size(); ++k) {
        size_t node = 0;
        for (int b = st.t.len[k] - 1; b >= 0; --b) {
            FFCZ_REQUIRE(pr.leaf[node] < 0, ERR_INVALID, "container code table is not a prefix code");
            const int bit = (st.t.code[k] >> b) & 1;
            if (pr.child[node * 2 + bit] < 0) {
                pr.child[node * 2 + bit] = (int32_t)pr.leaf.size();
                pr.leaf.push_back(-1);
                pr.child.push_back(-1);
                pr.child.push_back(-1);
            }
            node = (size_t)pr.child[node * 2 + bit];
        }
        FFCZ_REQUIRE(pr.leaf[node] < 0 && pr.child[node * 2] < 0 && pr.child[node * 2 + 1] < 0, ERR_INVALID,
                     "container code table is not a prefix code");
        pr.leaf[node] = (int32_t)st.t.sym[k];
    }
    pr.lut.assign((size_t)1 << LUT_BITS, LUT_BAD);
    for (size_t k = 0; k < st.t.sym.size(); ++k) {
        const int len = st.t.len[k];
        const uint32_t code = st.t.code[k];
        if (len <= LUT_BITS) {
            const uint32_t base = code << (LUT_BITS - len), span = 1u << (LUT_BITS - len);
            for (uint32_t i = 0; i < span; ++i) pr.lut[base + i] = ((uint32_t)len << 16) | st.t.sym[k];
        } else {
            const uint32_t pre = code >> (len - LUT_BITS);
            size_t node = 0;
            for (int b = LUT_BITS - 1; b >= 0; --b) node = (size_t)pr.child[node * 2 + ((pre >> b) & 1)];
            pr.lut[pre] = 0x80000000u | (uint32_t)node;
        }
    }
    pr.off.assign(st.cbits.size() + 1, 0);
    for (size_t c = 0; c < st.cbits.size(); ++c) pr.off[c + 1] = pr.off[c] + st.cbits[c];
    return pr;
}

// One thread per chunk of CHUNK symbols, from the chunk's bit offset: a
// 64-bit MSB-first window over the big-endian bitstream words, codes of up to
// 12 bits in one LDS lookup, longer ones down the trie.  A thread that meets
// an invalid code, or whose codes do not span exactly its chunk's bits, sets
// *err.
template <typename S>
__global__ __launch_bounds__(256) void k_huff_decode(const uint32_t* words, size_t nwords,
                                                     const unsigned long long* off, size_t nchunks, size_t nsym,
                                                     const uint32_t* glut, const int32_t* child, const int32_t* leaf,
                                                     S* out, unsigned int* err) {
    __shared__ uint32_t lut[1 << LUT_BITS];
    for (unsigned int i = threadIdx.x; i < (1u << LUT_BITS); i += 256) lut[i] = glut[i];
    __syncthreads();
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= nchunks) return;
    const unsigned long long p0 = off[c], p1 = off[c + 1];
    const size_t s0 = c * CHUNK;
    const int cnt = (int)(nsym - s0 < (size_t)CHUNK ? nsym - s0 : (size_t)CHUNK);
    // reads past the padded end repeat the last word (a damaged stream then
    // fails the bit-count check below instead of reading out of bounds)
    auto word = [&](size_t i) -> uint64_t { return __builtin_bswap32(words[i < nwords ? i : nwords - 1]); };
    size_t wi = (size_t)(p0 >> 5);
    uint64_t acc = (word(wi) << 32) | word(wi + 1);
    wi += 2;
    const int sk = (int)(p0 & 31);
    acc <<= sk;
    int nb = 64 - sk;
    unsigned long long used = 0;
    bool bad = false;
    for (int i = 0; i < cnt; ++i) {
        if (nb < 32) {  // at least 32 valid bits: a whole code
            acc |= word(wi++) << (32 - nb);
            nb += 32;
        }
        const uint32_t e = lut[acc >> (64 - LUT_BITS)];
        uint32_t sym;
        int len;
        if (!(e & 0x80000000u)) {
            len = (int)((e >> 16) & 31u);
            sym = e & 0xFFFFu;
        } else {
            if (e == LUT_BAD) {
                bad = true;
                break;
            }
            int node = (int)(e & 0x7FFFFFFFu), lf = -1;
            len = LUT_BITS;
            while (len < 32) {
                const int bit = (int)((acc >> (63 - len)) & 1u);
                ++len;
                node = child[node * 2 + bit];
                if (node < 0) break;
                lf = leaf[node];
                if (lf >= 0) break;
            }
            if (lf < 0) {
                bad = true;
                break;
            }
            sym = (uint32_t)lf;
        }
        acc <<= len;
        nb -= len;
        used += (unsigned long long)len;
        out[s0 + (size_t)i] = (S)sym;
    }
    if (bad || used != p1 - p0) atomicOr(err, 1u);
}

// the symbols of a prepared stream into dout (device, nsym elements)
template <typename S>
void gpu_decode(const Prepared& pr, S* dout, hipStream_t s) {
    const Stream& st = *pr.st;
    if (st.nsym == 0) return;
    const size_t nwords = ceil_div((size_t)st.bsbytes, (size_t)4) + 2;
    const size_t nch = pr.off.size() - 1;
    DBuf bw, bo, bl, bc, bf, be;
    const StreamWait wait{s};
    uint32_t* w = bw.alloc<uint32_t>(nwords);
    FFCZ_HIP(hipMemsetAsync(w, 0, nwords * 4, s));
    FFCZ_HIP(hipMemcpyAsync(w, pr.bits, st.bsbytes, hipMemcpyHostToDevice, s));
    unsigned long long* off = bo.alloc<unsigned long long>(nch + 1);
    FFCZ_HIP(hipMemcpyAsync(off, pr.off.data(), (nch + 1) * 8, hipMemcpyHostToDevice, s));
    uint32_t* lut = bl.alloc<uint32_t>(pr.lut.size());
    FFCZ_HIP(hipMemcpyAsync(lut, pr.lut.data(), pr.lut.size() * 4, hipMemcpyHostToDevice, s));
    int32_t* child = bc.alloc<int32_t>(pr.child.size());
    FFCZ_HIP(hipMemcpyAsync(child, pr.child.data(), pr.child.size() * 4, hipMemcpyHostToDevice, s));
    int32_t* leaf = bf.alloc<int32_t>(pr.leaf.size());
    FFCZ_HIP(hipMemcpyAsync(leaf, pr.leaf.data(), pr.leaf.size() * 4, hipMemcpyHostToDevice, s));
    unsigned int* err = be.alloc<unsigned int>(1);
    FFCZ_HIP(hipMemsetAsync(err, 0, 4, s));
    hipLaunchKernelGGL((k_huff_decode<S>), dim3((unsigned)ceil_div(nch, (size_t)256)), dim3(256), 0, s,
                       (const uint32_t*)w, nwords, (const unsigned long long*)off, nch, (size_t)st.nsym,
                       (const uint32_t*)lut, (const int32_t*)child, (const int32_t*)leaf, dout, err);
    FFCZ_HIP(hipGetLastError());
    unsigned int herr = 0;
    FFCZ_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    FFCZ_REQUIRE(herr == 0, ERR_INVALID, "corrupt Huffman stream in container");
}

// ---- reconstruction kernels -------------------------------------------------
template <typename T>
struct SymDesc {
    int kind;
    double step;
    T mn, mx;
};

// Value of a coded edit: the reference quantiser's dequantisation
// (projection_algorithm.cu:629-643, in T, no contraction) or q*step for the
// zigzag step symbols.
template <typename T, typename S>
__device__ __forceinline__ T sym_value(S z, const SymDesc<T>& q) {
#pragma clang fp contract(off)
    if (q.kind == Q_REF16) {
        const T qv = (T)z;
        return qv / (T)65535 * (q.mx - q.mn) + q.mn;
    }
    const long long k = (long long)(z >> 1) ^ -(long long)(z & 1);
    return (T)((double)k * q.step);
}

// word i of a flag pack with the bits at or beyond n cleared
__device__ __forceinline__ uint32_t flag_word(const uint32_t* w, size_t i, size_t n) {
    uint32_t v = w[i];
    const size_t b0 = i * 32;
    if (b0 + 32 > n) v &= n > b0 ? ((1u << (uint32_t)(n - b0)) - 1u) : 0u;
    return v;
}

__global__ __launch_bounds__(256) void k_block_popc(const uint32_t* w, size_t nwords, size_t n, uint32_t* bsum) {
    __shared__ uint32_t red[256];
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    red[threadIdx.x] = i < nwords ? (uint32_t)__popc(flag_word(w, i, n)) : 0u;
    __syncthreads();
    for (unsigned int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// exclusive prefix of the block counts (and their total in boff[nb]), one workgroup
__global__ __launch_bounds__(1024) void k_scan_blocks(const uint32_t* bsum, size_t nb, unsigned long long* boff) {
    __shared__ unsigned long long s[1024];
    const size_t per = (nb + 1023) / 1024;
    const size_t a = (size_t)threadIdx.x * per, e = a + per < nb ? a + per : nb;
    unsigned long long t = 0;
    for (size_t i = a; i < e; ++i) t += bsum[i];
    s[threadIdx.x] = t;
    __syncthreads();
    for (unsigned int off = 1; off < 1024; off <<= 1) {
        const unsigned long long v = threadIdx.x >= off ? s[threadIdx.x - off] : 0ull;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = threadIdx.x ? s[threadIdx.x - 1] : 0ull;
    for (size_t i = a; i < e; ++i) {
        boff[i] = run;
        run += bsum[i];
    }
    if (threadIdx.x == 1023) boff[nb] = s[1023];  // the total
}

// dense edits: out[i] = value of the next symbol where flag i is set, else 0
// (one thread per 32-flag word; symbols are taken in index order)
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_expand(const uint32_t* w, size_t nwords, size_t n,
                                                const unsigned long long* boff, const S* sym, size_t m,
                                                SymDesc<T> q, T* out) {
    __shared__ uint32_t sc[256];
    const size_t wi = (size_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t v = wi < nwords ? flag_word(w, wi, n) : 0u;
    const uint32_t c = (uint32_t)__popc(v);
    sc[threadIdx.x] = c;
    __syncthreads();
    for (unsigned int off = 1; off < 256; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? sc[threadIdx.x - off] : 0u;
        __syncthreads();
        sc[threadIdx.x] += t;
        __syncthreads();
    }
    if (wi >= nwords) return;
    unsigned long long pos = boff[blockIdx.x] + sc[threadIdx.x] - c;
    const size_t i0 = wi * 32;
    for (int j = 0; j < 32; ++j) {
        const size_t i = i0 + (size_t)j;
        if (i >= n) break;
        T val = T(0);
        if ((v >> j) & 1u) {
            if (pos < m) val = sym_value<T, S>(sym[pos], q);
            ++pos;
        }
        out[i] = val;
    }
}

// split [Re | Im] frequency edits into the padded half spectrum
template <typename T>
__global__ void k_split_to_spec(const T* dfd, size_t F, size_t rows, size_t nh, size_t mp, cplx<T>* spec) {
    const size_t total = rows * mp;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
        const size_t r = q / mp, k = q - r * mp;
        cplx<T> v = mk(T(0), T(0));
        if (k < nh) {
            const size_t i = r * nh + k;
            v = mk(dfd[i], dfd[F + i]);
        }
        spec[q] = v;
    }
}

// x_hat = (x' + C2R(deq df)/N) + deq(ds)
template <typename T>
__global__ void k_sum_edits(const T* xp, const T* rf, const T* ds, T* xh, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        xh[i] = rf ? (xp[i] + rf[i]) + ds[i] : xp[i] + ds[i];
}

unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>(std::max<size_t>(ceil_div(n, (size_t)256), 1), 65536); }

// dense edits from a device flag pack (zero-padded to whole 32-bit words)
// and its m device symbols; the flags must mark exactly m elements
template <typename T, typename S>
void expand(const uint32_t* w, size_t n, const S* dsym, size_t m, const SymDesc<T>& q, T* out, hipStream_t s) {
    const size_t nwords = ceil_div(n, (size_t)32), nblk = ceil_div(nwords, (size_t)256);
    DBuf db, dboff;
    const StreamWait wait{s};
    uint32_t* bsum = db.alloc<uint32_t>(nblk);
    unsigned long long* boff = dboff.alloc<unsigned long long>(nblk + 1);
    hipLaunchKernelGGL(k_block_popc, dim3((unsigned)nblk), dim3(256), 0, s, w, nwords, n, bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, (const uint32_t*)bsum, nblk, boff);
    FFCZ_HIP(hipGetLastError());
    unsigned long long total = 0;
    FFCZ_HIP(hipMemcpyAsync(&total, boff + nblk, 8, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    FFCZ_REQUIRE(total == m, ERR_INVALID, "container flags do not match the active counts");
    hipLaunchKernelGGL((k_expand<T, S>), dim3((unsigned)nblk), dim3(256), 0, s, w, nwords, n,
                       (const unsigned long long*)boff, dsym, m, q, out);
    FFCZ_HIP(hipGetLastError());
    FFCZ_HIP(hipStreamSynchronize(s));
}

// a flag stream and its symbol stream, decoded on the device and expanded
template <typename T>
void decode_edits(const Prepared& flags, const Prepared& syms, size_t n, const SymDesc<T>& q, T* out, hipStream_t s) {
    DBuf bp, bs;
    const StreamWait wait{s};
    uint32_t* w = bp.alloc<uint32_t>(ceil_div(n, (size_t)32));
    FFCZ_HIP(hipMemsetAsync(w, 0, ceil_div(n, (size_t)32) * 4, s));
    gpu_decode<uint8_t>(flags, reinterpret_cast<uint8_t*>(w), s);
    uint16_t* sym = bs.alloc<uint16_t>(syms.st->nsym);
    gpu_decode<uint16_t>(syms, sym, s);
    expand<T, uint16_t>(w, n, sym, (size_t)syms.st->nsym, q, out, s);
}

template <typename T>
void decode_typed(const ContainerInfo& in, const Prepared* pr, const SymDesc<T>& qf, const SymDesc<T>& qs,
                  const void* x_prime, void* x_hat, hipStream_t s) {
    Fft3d<T> f;
    f.create(in.nx, in.ny, in.nz);
    struct FftGuard {
        Fft3d<T>& f;
        ~FftGuard() { f.destroy(); }
    } fg{f};
    const size_t N = f.N, F = f.F;
    DBuf bxp, bxh, bds, bdf, bsp, brf;
    const StreamWait wait{s};
    const T* xp = static_cast<const T*>(x_prime);
    if (!on_device(x_prime)) {
        T* t = bxp.alloc<T>(N);
        FFCZ_HIP(hipMemcpyAsync(t, x_prime, N * sizeof(T), hipMemcpyHostToDevice, s));
        xp = t;
    }
    const bool dev_out = on_device(x_hat);
    T* xh = dev_out ? static_cast<T*>(x_hat) : bxh.alloc<T>(N);
    T* dsd = bds.alloc<T>(N);
    decode_edits<T>(pr[1], pr[3], N, qs, dsd, s);
    T* rf = nullptr;
    if (pr[2].st->nsym > 0) {
        // C2R(deq df)/N: the split edits into the padded half spectrum, then
        // the library's inverse transform scaled by 1/N (impl:242)
        T* dfd = bdf.alloc<T>(2 * F);
        decode_edits<T>(pr[0], pr[2], 2 * F, qf, dfd, s);
        cplx<T>* spec = bsp.alloc<cplx<T>>(f.spec_elems());
        hipLaunchKernelGGL((k_split_to_spec<T>), dim3(grid_for(f.spec_elems())), dim3(256), 0, s, (const T*)dfd, F,
                           f.rows, f.Nh, f.Mp, spec);
        FFCZ_HIP(hipGetLastError());
        rf = brf.alloc<T>(N);
        fft_c2r<T>(f, spec, rf, T(1) / static_cast<T>(N), s);
    }
    hipLaunchKernelGGL((k_sum_edits<T>), dim3(grid_for(N)), dim3(256), 0, s, xp, (const T*)rf, (const T*)dsd, xh, N);
    FFCZ_HIP(hipGetLastError());
    if (!dev_out) FFCZ_HIP(hipMemcpyAsync(x_hat, xh, N * sizeof(T), hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
}

struct Header {
    ContainerInfo info;
    double fstep = 0, sstep = 0, fmin = 0, fmax = 0, smin = 0, smax = 0;
};

Header read_header(Reader& r) {
    Header h;
    const uint8_t* m = r.take(4);
    FFCZ_REQUIRE(std::memcmp(m, MAGIC, 4) == 0, ERR_INVALID, "not an FFCz container");
    FFCZ_REQUIRE(r.get<uint32_t>() == VERSION, ERR_INVALID, "unsupported FFCz container version");
    h.info.dtype = (int)r.get<uint32_t>();
    h.info.mode = (int)r.get<uint32_t>();
    h.info.nx = r.get<uint64_t>();
    h.info.ny = r.get<uint64_t>();
    h.info.nz = r.get<uint64_t>();
    h.info.eps = r.get<double>();
    h.info.delta = r.get<double>();
    h.info.qkind = (int)r.get<uint32_t>();
    const uint32_t ns = r.get<uint32_t>();
    h.fstep = r.get<double>();
    h.sstep = r.get<double>();
    h.fmin = r.get<double>();
    h.fmax = r.get<double>();
    h.smin = r.get<double>();
    h.smax = r.get<double>();
    h.info.n_freq = r.get<uint64_t>();
    h.info.n_spatial = r.get<uint64_t>();
    const auto& in = h.info;
    FFCZ_REQUIRE(in.dtype == 0 || in.dtype == 1, ERR_INVALID, "bad container dtype");
    FFCZ_REQUIRE(in.qkind == Q_REF16 || in.qkind == Q_STEP, ERR_INVALID, "bad container quantiser");
    FFCZ_REQUIRE(ns == 4, ERR_INVALID, "bad container stream count");
    FFCZ_REQUIRE(in.nx >= 1 && in.ny >= 1 && in.nz >= 1 && in.nx <= (1u << 20) && in.ny <= (1u << 20) &&
                     in.nz <= (1u << 20) && in.nx * in.ny <= ((size_t)1 << 40) / in.nz,
                 ERR_INVALID, "bad container shape");
    FFCZ_REQUIRE(in.qkind != Q_STEP || (h.fstep > 0 && h.sstep > 0), ERR_INVALID, "bad container quantiser steps");
    return h;
}

// ---- base compressor: dual-quantisation Lorenzo ---------------------------
// q = round(x/(2 eps)) as int32 (x' = q * 2 eps, |x' - x| <= eps up to the
// rounding of the product); the symbols are the zigzag of the 3D Lorenzo
// residual of q (the mixed difference along x, y and z, zero outside the
// field), exact in integers, so decoding is three prefix sums.  Residuals
// whose zigzag does not fit 16 bits take the escape symbol and an outlier
// entry (index, residual).
constexpr char BMAGIC[4] = {'F', 'F', 'C', 'B'};
constexpr size_t BHDR_BYTES = 56;
constexpr uint16_t ESC = 65535;

unsigned grid_rows(size_t n) { return (unsigned)std::min<size_t>(std::max<size_t>(n, 1), 65535); }

template <typename T>
__global__ void k_prequant(const T* x, size_t n, double inv2e, double two_eps, int32_t* q, T* xp, unsigned int* bad) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = rint((double)x[i] * inv2e);
        int32_t qi = 0;
        if (fabs(v) < 1073741824.0) qi = (int32_t)v;
        else atomicOr(bad, 1u);
        q[i] = qi;
        xp[i] = (T)((double)qi * two_eps);
    }
}

// one workgroup per row (ix, iy) of nz elements
__global__ __launch_bounds__(256) void k_lorenzo(const int32_t* q, size_t nx, size_t ny, size_t nz, uint16_t* sym,
                                                 unsigned long long* ocount, unsigned long long* oidx,
                                                 long long* oval, size_t ocap) {
    for (size_t row = blockIdx.x; row < nx * ny; row += gridDim.x) {
        const size_t ix = row / ny, iy = row - ix * ny;
        const bool x0 = ix > 0, y0 = iy > 0;
        const int32_t* c = q + row * nz;
        const int32_t* cx = x0 ? c - ny * nz : nullptr;
        const int32_t* cy = y0 ? c - nz : nullptr;
        const int32_t* cxy = (x0 && y0) ? c - ny * nz - nz : nullptr;
        for (size_t k = threadIdx.x; k < nz; k += 256) {
            auto at = [&](const int32_t* p, size_t kk) -> long long { return p ? (long long)p[kk] : 0ll; };
            long long r = (long long)c[k] - at(cx, k) - at(cy, k) + at(cxy, k);
            if (k > 0) r -= (long long)c[k - 1] - at(cx, k - 1) - at(cy, k - 1) + at(cxy, k - 1);
            const unsigned long long z = zigzag64(r);
            const size_t i = row * nz + k;
            if (z < ESC) {
                sym[i] = (uint16_t)z;
            } else {
                sym[i] = ESC;
                const unsigned long long slot = atomicAdd(ocount, 1ull);
                if (slot < ocap) {
                    oidx[slot] = i;
                    oval[slot] = r;
                }
            }
        }
    }
}

// residuals from the symbols; counts the escapes
__global__ void k_unsym(const uint16_t* sym, size_t n, long long* r, unsigned long long* nesc) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const unsigned int z = sym[i];
        if (z == ESC) ++c;
        r[i] = (long long)(z >> 1) ^ -(long long)(z & 1u);
    }
    if (c) atomicAdd(nesc, c);
}

// the outliers into place; an outlier whose symbol is not the escape is an error
__global__ void k_outliers(const uint16_t* sym, const unsigned long long* oidx, const long long* oval, size_t m,
                           long long* r, unsigned int* err) {
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < m; s += (size_t)gridDim.x * blockDim.x) {
        const size_t i = oidx[s];
        if (sym[i] != ESC) atomicOr(err, 1u);
        r[i] = oval[s];
    }
}

// inclusive prefix sums along z (a workgroup per row), y and x (a thread per column)
__global__ __launch_bounds__(256) void k_scan_z(long long* v, size_t rows, size_t nz) {
    __shared__ long long sh[256];
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        long long carry = 0;
        for (size_t c0 = 0; c0 < nz; c0 += 256) {
            const size_t k = c0 + threadIdx.x;
            sh[threadIdx.x] = k < nz ? v[row * nz + k] : 0ll;
            __syncthreads();
            for (unsigned int off = 1; off < 256; off <<= 1) {
                const long long t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0ll;
                __syncthreads();
                sh[threadIdx.x] += t;
                __syncthreads();
            }
            if (k < nz) v[row * nz + k] = sh[threadIdx.x] + carry;
            carry += sh[255];
            __syncthreads();
        }
    }
}
__global__ void k_scan_strided(long long* v, size_t outer, size_t len, size_t inner) {
    // columns (o, i) of length len, element (o, l, i) at (o*len + l)*inner + i
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < outer * inner;
         c += (size_t)gridDim.x * blockDim.x) {
        const size_t o = c / inner, i = c - o * inner;
        long long acc = 0;
        for (size_t l = 0; l < len; ++l) {
            const size_t at = (o * len + l) * inner + i;
            acc += v[at];
            v[at] = acc;
        }
    }
}

template <typename T>
__global__ void k_dequant_base(const long long* q, size_t n, double two_eps, T* xp) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        xp[i] = (T)((double)(int32_t)q[i] * two_eps);
}

}  // namespace

template <typename T>
std::vector<uint8_t> container_pack(Solver<T>& s, int zstd_level) {
    FFCZ_REQUIRE(s.extracted(), ERR_STATE, "extract_compact_representation() has not run");
    const size_t nf = s.n_active_freq(), ns = s.n_active_spatial();  // waits for the extraction
    const QuantDesc qf = s.freq_quant(), qs = s.spatial_quant();
    std::vector<uint8_t> buf;
    buf.reserve(HDR_BYTES);
    Writer w{buf};
    w.bytes(MAGIC, 4);
    w.put<uint32_t>(VERSION);
    w.put<uint32_t>(sizeof(T) == 4 ? 0u : 1u);
    w.put<uint32_t>((uint32_t)s.solve_mode());
    w.put<uint64_t>(s.nx());
    w.put<uint64_t>(s.ny());
    w.put<uint64_t>(s.nz());
    w.put<double>((double)s.eps());
    w.put<double>(s.solve_mode() == MODE_DIRECT ? std::nan("") : (double)s.delta());
    w.put<uint32_t>((uint32_t)qf.kind);
    w.put<uint32_t>(4);
    w.put<double>(qf.kind == Q_STEP ? qf.step : 0.0);
    w.put<double>(qs.kind == Q_STEP ? qs.step : 0.0);
    w.put<double>(s.freq_min());
    w.put<double>(s.freq_max());
    w.put<double>(s.spatial_min());
    w.put<double>(s.spatial_max());
    w.put<uint64_t>(nf);
    w.put<uint64_t>(ns);
    const hipStream_t st = s.stream();
    pack_stream<uint8_t>(w, s.d_freq_pack(), ceil_div(2 * s.F(), (size_t)8), zstd_level, st);
    pack_stream<uint8_t>(w, s.d_spatial_pack(), ceil_div(s.N(), (size_t)8), zstd_level, st);
    pack_stream<uint16_t>(w, s.d_freq_quant(), nf, zstd_level, st);
    pack_stream<uint16_t>(w, s.d_spatial_quant(), ns, zstd_level, st);
    return buf;
}

ContainerInfo container_info(const uint8_t* buf, size_t len) {
    FFCZ_REQUIRE(buf != nullptr, ERR_INVALID, "null container");
    Reader r{buf, len};
    return read_header(r).info;
}

void container_decode(const uint8_t* buf, size_t len, const void* x_prime, void* x_hat, hipStream_t s) {
    FFCZ_REQUIRE(buf != nullptr && x_prime != nullptr && x_hat != nullptr, ERR_INVALID, "null argument");
    Reader r{buf, len};
    const Header h = read_header(r);
    const ContainerInfo& in = h.info;
    const size_t N = in.nx * in.ny * in.nz, F = in.nx * in.ny * (in.nz / 2 + 1);
    const Stream sf = read_stream(r, 1), ss = read_stream(r, 1), qf = read_stream(r, 2), qs = read_stream(r, 2);
    FFCZ_REQUIRE(sf.nsym == ceil_div(2 * F, (size_t)8) && ss.nsym == ceil_div(N, (size_t)8), ERR_INVALID,
                 "container flag streams do not match the shape");
    FFCZ_REQUIRE(qf.nsym == in.n_freq && qs.nsym == in.n_spatial, ERR_INVALID,
                 "container symbol streams do not match the active counts");
    // every stream checked and made ready on the host before any device work
    const Prepared pr[4] = {prepare(sf), prepare(ss), prepare(qf), prepare(qs)};
    if (in.dtype == 0) {
        const SymDesc<float> df{in.qkind, h.fstep, (float)h.fmin, (float)h.fmax};
        const SymDesc<float> ds{in.qkind, h.sstep, (float)h.smin, (float)h.smax};
        decode_typed<float>(in, pr, df, ds, x_prime, x_hat, s);
    } else {
        const SymDesc<double> df{in.qkind, h.fstep, h.fmin, h.fmax};
        const SymDesc<double> ds{in.qkind, h.sstep, h.smin, h.smax};
        decode_typed<double>(in, pr, df, ds, x_prime, x_hat, s);
    }
}

template <typename T>
std::vector<uint8_t> base_compress(const T* x, size_t nx, size_t ny, size_t nz, double eps, int zstd_level,
                                   T* x_prime, hipStream_t s) {
    FFCZ_REQUIRE(x != nullptr && x_prime != nullptr, ERR_INVALID, "null argument");
    FFCZ_REQUIRE(eps > 0 && std::isfinite(eps), ERR_INVALID, "the base bound must be positive");
    FFCZ_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nx * ny <= ((size_t)1 << 40) / nz, ERR_INVALID, "bad shape");
    const size_t n = nx * ny * nz, ocap = n / 16 + 4096;
    DBuf bx, bxp, bq, bs, bc, bi, bv, bb;
    const StreamWait wait{s};
    const T* dx = x;
    if (!on_device(x)) {
        T* t = bx.alloc<T>(n);
        FFCZ_HIP(hipMemcpyAsync(t, x, n * sizeof(T), hipMemcpyHostToDevice, s));
        dx = t;
    }
    const bool dev_out = on_device(x_prime);
    T* dxp = dev_out ? x_prime : bxp.alloc<T>(n);
    int32_t* q = bq.alloc<int32_t>(n);
    uint16_t* sym = bs.alloc<uint16_t>(n);
    unsigned long long* cnt = bc.alloc<unsigned long long>(1);
    unsigned long long* oidx = bi.alloc<unsigned long long>(ocap);
    long long* oval = bv.alloc<long long>(ocap);
    unsigned int* bad = bb.alloc<unsigned int>(1);
    FFCZ_HIP(hipMemsetAsync(cnt, 0, 8, s));
    FFCZ_HIP(hipMemsetAsync(bad, 0, 4, s));
    hipLaunchKernelGGL((k_prequant<T>), dim3(grid_for(n)), dim3(256), 0, s, dx, n, 1.0 / (2.0 * eps), 2.0 * eps, q,
                       dxp, bad);
    hipLaunchKernelGGL(k_lorenzo, dim3(grid_rows(nx * ny)), dim3(256), 0, s, (const int32_t*)q, nx, ny, nz, sym, cnt,
                       oidx, oval, ocap);
    FFCZ_HIP(hipGetLastError());
    unsigned long long m = 0;
    unsigned int hbad = 0;
    FFCZ_HIP(hipMemcpyAsync(&m, cnt, 8, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    FFCZ_REQUIRE(hbad == 0, ERR_RANGE, "the base bound is too small for the field's range (|x|/(2 eps) >= 2^30)");
    FFCZ_REQUIRE(m <= ocap, ERR_RANGE, "too many base residuals outside 16 bits (the bound is too small)");
    std::vector<unsigned long long> hi(m);
    std::vector<long long> hv(m);
    if (m) {
        FFCZ_HIP(hipMemcpyAsync(hi.data(), oidx, m * 8, hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipMemcpyAsync(hv.data(), oval, m * 8, hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
    }
    // index order (the atomics filled the list in any order)
    std::vector<size_t> ord(m);
    for (size_t i = 0; i < m; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return hi[a] < hi[b]; });
    std::vector<uint8_t> buf;
    Writer w{buf};
    w.bytes(BMAGIC, 4);
    w.put<uint32_t>(VERSION);
    w.put<uint32_t>(sizeof(T) == 4 ? 0u : 1u);
    w.put<uint32_t>(0);
    w.put<uint64_t>(nx);
    w.put<uint64_t>(ny);
    w.put<uint64_t>(nz);
    w.put<double>(eps);
    w.put<uint64_t>(m);
    pack_stream<uint16_t>(w, sym, n, zstd_level, s);
    for (size_t i : ord) w.put<uint64_t>(hi[i]);
    for (size_t i : ord) w.put<int64_t>(hv[i]);
    if (!dev_out) {
        FFCZ_HIP(hipMemcpyAsync(x_prime, dxp, n * sizeof(T), hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
    }
    return buf;
}

ContainerInfo base_info(const uint8_t* buf, size_t len) {
    FFCZ_REQUIRE(buf != nullptr, ERR_INVALID, "null base stream");
    Reader r{buf, len};
    FFCZ_REQUIRE(std::memcmp(r.take(4), BMAGIC, 4) == 0, ERR_INVALID, "not an FFCz base stream");
    FFCZ_REQUIRE(r.get<uint32_t>() == VERSION, ERR_INVALID, "unsupported FFCz base stream version");
    ContainerInfo in;
    in.dtype = (int)r.get<uint32_t>();
    (void)r.get<uint32_t>();
    in.nx = r.get<uint64_t>();
    in.ny = r.get<uint64_t>();
    in.nz = r.get<uint64_t>();
    in.eps = r.get<double>();
    in.n_spatial = r.get<uint64_t>();  // outliers
    FFCZ_REQUIRE(in.dtype == 0 || in.dtype == 1, ERR_INVALID, "bad base stream dtype");
    FFCZ_REQUIRE(in.nx >= 1 && in.ny >= 1 && in.nz >= 1 && in.nx <= (1u << 20) && in.ny <= (1u << 20) &&
                     in.nz <= (1u << 20) && in.nx * in.ny <= ((size_t)1 << 40) / in.nz,
                 ERR_INVALID, "bad base stream shape");
    FFCZ_REQUIRE(in.eps > 0 && std::isfinite(in.eps), ERR_INVALID, "bad base stream bound");
    return in;
}

void base_decompress(const uint8_t* buf, size_t len, void* x_prime, hipStream_t s) {
    FFCZ_REQUIRE(x_prime != nullptr, ERR_INVALID, "null argument");
    const ContainerInfo in = base_info(buf, len);
    Reader r{buf, len};
    (void)r.take(BHDR_BYTES);
    const size_t n = in.nx * in.ny * in.nz, m = in.n_spatial;
    const Stream st = read_stream(r, 2);
    FFCZ_REQUIRE(st.nsym == n, ERR_INVALID, "base stream does not match the shape");
    FFCZ_REQUIRE(m <= n, ERR_INVALID, "bad base outlier count");
    const uint8_t* pi = r.take(m * 8);
    const uint8_t* pv = r.take(m * 8);
    std::vector<unsigned long long> hi(m);
    std::vector<long long> hv(m);
    if (m) {
        std::memcpy(hi.data(), pi, m * 8);
        std::memcpy(hv.data(), pv, m * 8);
    }
    for (size_t i = 0; i < m; ++i)
        FFCZ_REQUIRE(hi[i] < n && (i == 0 || hi[i] > hi[i - 1]), ERR_INVALID, "bad base outlier index");
    const Prepared pr = prepare(st);
    DBuf bs, br, bi, bv, bc, be, bxp;
    const StreamWait wait{s};
    uint16_t* sym = bs.alloc<uint16_t>(n);
    gpu_decode<uint16_t>(pr, sym, s);
    long long* v = br.alloc<long long>(n);
    unsigned long long* nesc = bc.alloc<unsigned long long>(1);
    unsigned int* err = be.alloc<unsigned int>(1);
    FFCZ_HIP(hipMemsetAsync(nesc, 0, 8, s));
    FFCZ_HIP(hipMemsetAsync(err, 0, 4, s));
    hipLaunchKernelGGL(k_unsym, dim3(grid_for(n)), dim3(256), 0, s, (const uint16_t*)sym, n, v, nesc);
    if (m) {
        unsigned long long* di = bi.alloc<unsigned long long>(m);
        long long* dv = bv.alloc<long long>(m);
        FFCZ_HIP(hipMemcpyAsync(di, hi.data(), m * 8, hipMemcpyHostToDevice, s));
        FFCZ_HIP(hipMemcpyAsync(dv, hv.data(), m * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_outliers, dim3(grid_for(m)), dim3(256), 0, s, (const uint16_t*)sym,
                           (const unsigned long long*)di, (const long long*)dv, m, v, err);
    }
    FFCZ_HIP(hipGetLastError());
    unsigned long long hesc = 0;
    unsigned int herr = 0;
    FFCZ_HIP(hipMemcpyAsync(&hesc, nesc, 8, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    FFCZ_REQUIRE(herr == 0 && hesc == m, ERR_INVALID, "base outliers do not match the escaped symbols");
    // the inverse of the mixed difference: prefix sums along z, y and x
    hipLaunchKernelGGL(k_scan_z, dim3(grid_rows(in.nx * in.ny)), dim3(256), 0, s, v, in.nx * in.ny, in.nz);
    hipLaunchKernelGGL(k_scan_strided, dim3(grid_for(in.nx * in.nz)), dim3(256), 0, s, v, in.nx, in.ny, in.nz);
    hipLaunchKernelGGL(k_scan_strided, dim3(grid_for(in.ny * in.nz)), dim3(256), 0, s, v, (size_t)1, in.nx,
                       in.ny * in.nz);
    FFCZ_HIP(hipGetLastError());
    const bool dev_out = on_device(x_prime);
    const double two_eps = 2.0 * in.eps;
    if (in.dtype == 0) {
        float* xp = dev_out ? static_cast<float*>(x_prime) : bxp.alloc<float>(n);
        hipLaunchKernelGGL((k_dequant_base<float>), dim3(grid_for(n)), dim3(256), 0, s, (const long long*)v, n, two_eps,
                           xp);
        if (!dev_out) FFCZ_HIP(hipMemcpyAsync(x_prime, xp, n * 4, hipMemcpyDeviceToHost, s));
    } else {
        double* xp = dev_out ? static_cast<double*>(x_prime) : bxp.alloc<double>(n);
        hipLaunchKernelGGL((k_dequant_base<double>), dim3(grid_for(n)), dim3(256), 0, s, (const long long*)v, n,
                           two_eps, xp);
        if (!dev_out) FFCZ_HIP(hipMemcpyAsync(x_prime, xp, n * 8, hipMemcpyDeviceToHost, s));
    }
    FFCZ_HIP(hipGetLastError());
    FFCZ_HIP(hipStreamSynchronize(s));
}

template std::vector<uint8_t> base_compress<float>(const float*, size_t, size_t, size_t, double, int, float*,
                                                   hipStream_t);
template std::vector<uint8_t> base_compress<double>(const double*, size_t, size_t, size_t, double, int, double*,
                                                    hipStream_t);
template std::vector<uint8_t> container_pack<float>(Solver<float>&, int);
template std::vector<uint8_t> container_pack<double>(Solver<double>&, int);

}  // namespace ffcz
