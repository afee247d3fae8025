// Edit extraction kernels: one read of the edit array.
//
// Tiles of 256 threads x 8 consecutive elements x 4 rounds, held in
// registers.  k_tile_compact forms the flags, stores them LSB-first
// (four threads' bytes combined into one word), ranks them with wave scans,
// finds the tile's output offset by decoupled look-back over the tiles before
// it, and writes compact values (and step-quantised symbols) in index order.
// k_extract_finalize reduces count, min and max; with the reference
// quantiser, k_quant_compact then quantises the compact values (the global
// min/max are known only after the pass).
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <string>

#include "extract.hpp"
#include "quant.hpp"

namespace ffcz {

constexpr int X_EPT = 8;                          // elements per thread per round
constexpr int X_ROUNDS = 4;                      // rounds of the smallest (FP64) tile
constexpr int X_TILE = 256 * X_EPT * X_ROUNDS;   // 8192: scratch is sized for it
static_assert(X_TILE == EXTRACT_TILE, "tile size");
// rounds per tile (8192 elements; twice that for FP32 was measured slower)
template <typename T>
constexpr int x_rounds() { return X_ROUNDS; }
template <typename T>
constexpr int x_tile() { return 256 * X_EPT * x_rounds<T>(); }

size_t extract_scratch_bytes(size_t n) {
    // [0, 64): results of the synchronous entry; [128]: tile ticket;
    // [256, ...): tile status words, tile minima, tile maxima
    const size_t ntiles = ceil_div(n ? n : 1, (size_t)X_TILE);
    return 256 + ntiles * (sizeof(unsigned long long) + 2 * sizeof(double));
}

// The edit's flag where comparisons decide it: 1 active, 0 inactive, 2 for a
// step-quantised value within 1e-5 (relative) of half a step, which the
// quantiser itself decides (edit_flag_exact).
template <typename T>
__device__ __forceinline__ unsigned int edit_flag3(T v, const QuantDesc& q, double lo, double hi) {
    if (q.kind == Q_STEP) {
        const double a = fabs((double)v);
        return a < lo ? 0u : (a > hi ? 1u : 2u);
    }
    return fabs(v) > (T)q.threshold ? 1u : 0u;  // projection_algorithm.cu:377
}
__device__ __noinline__ bool edit_flag_exact(double v, double step) { return quant_step_f(v, step) != 0.0; }

// 8 consecutive elements starting at i (a multiple of 8); lanes past n get 0
template <typename T>
__device__ __forceinline__ void load8(const T* v, size_t i, size_t n, T (&x)[X_EPT]) {
    if (i + X_EPT <= n) {
        if constexpr (sizeof(T) == 4) {
            const float4* p = reinterpret_cast<const float4*>(v + i);
            const float4 a = p[0], b = p[1];
            x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
            x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
        } else {
            const double2* p = reinterpret_cast<const double2*>(v + i);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double2 a = p[k];
                x[2 * k] = a.x;
                x[2 * k + 1] = a.y;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < X_EPT; ++k) x[k] = (i + k < n) ? v[i + k] : T(0);
    }
}

// Tile status words of the decoupled look-back (Merrill & Garland, "Single-
// pass Parallel Prefix Scan with Decoupled Look-back", 2016): the top two bits
// say whether the low 62 hold the tile's own flag count (A) or the inclusive
// prefix up to and including it (P); 0 = not yet published.
constexpr unsigned long long ST_A = 1ull << 62;
constexpr unsigned long long ST_P = 2ull << 62;
constexpr unsigned long long ST_VAL = ST_A - 1;
constexpr unsigned int LOOKBACK_SPIN_LIMIT = 1u << 22;

__global__ void k_extract_init(unsigned long long* status, size_t ntiles, unsigned int* ticket, ExtractGlobals* g) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ntiles; i += (size_t)gridDim.x * blockDim.x)
        status[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *ticket = 0u;
        *g = ExtractGlobals{};
    }
}

// One pass over the edits: every workgroup takes the next tile in dispatch
// order (a ticket, so every tile it waits for belongs to a workgroup that is
// already running), forms its flags, packs them, counts them, publishes the
// count, looks back over its predecessors' published counts/prefixes for its
// output offset, then writes the compact values (and, for the step quantiser,
// the zigzag symbols) in index order.  The look-back's wait is bounded: past
// LOOKBACK_SPIN_LIMIT polls it records an error instead of waiting on.
// `tile` numbers the tile within its edit array; status, pmin and pmax point
// at that array's tile 0.
template <typename T>
__device__ __forceinline__ void tile_compact(const unsigned int tile, const T* v, size_t n, const QuantDesc& q,
                                             unsigned long long* status, double* pmin, double* pmax,
                                             ExtractGlobals* g, unsigned int* pack, T* compact, uint16_t* quant) {
    __shared__ unsigned long long s_pre;
    constexpr int RND = x_rounds<T>();
    __shared__ unsigned int wtot[RND][4];
    __shared__ T smn[4], smx[4];
    // the tile's active values in index order, half a tile at a time (a
    // full tile of values would cost occupancy)
    constexpr unsigned int SVAL = X_TILE / 2;
    __shared__ T sval[SVAL];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const size_t base = (size_t)tile * x_tile<T>();
    T x[RND][X_EPT];
    unsigned int bits[RND], excl[RND];
    T mn = (T)INFINITY, mx = (T)-INFINITY;
    // the step quantiser's flag by comparisons, away from half a step (the
    // margin is far wider than the rounding of the thresholds and quotients)
    const double flo = q.step * 0.49999, fhi = q.step * 0.50001;
    // every round's loads first: the flag words stored below could alias
    // the edits for all the compiler knows, which would hold each round's
    // loads back until the previous round's stores
#pragma unroll
    for (int r = 0; r < RND; ++r) {
        const size_t i = base + ((size_t)r * 256 + tid) * X_EPT;
        if (i < n) load8(v, i, n, x[r]);
    }
#pragma unroll
    for (int r = 0; r < RND; ++r) {
        const size_t i = base + ((size_t)r * 256 + tid) * X_EPT;
        bits[r] = 0;
        if (i < n) {
            unsigned int unsure = 0;
#pragma unroll
            for (int k = 0; k < X_EPT; ++k) {
                if (i + k < n) {
                    const unsigned int f = edit_flag3(x[r][k], q, flo, fhi);
                    bits[r] |= (f & 1u) << k;
                    unsure |= (f >> 1) << k;
                    mn = x[r][k] < mn ? x[r][k] : mn;
                    mx = x[r][k] > mx ? x[r][k] : mx;
                }
            }
            // the rare values near half a step: the quantiser decides (one
            // copy of the division, the value read again from memory)
            while (unsure) {
                const int k = __builtin_ctz(unsure);
                unsure &= unsure - 1u;
                if (edit_flag_exact((double)v[i + k], q.step)) bits[r] |= 1u << k;
            }
        }
        // four consecutive threads' flag bytes form one little-endian word
        const unsigned int b1 = __shfl_down(bits[r], 1, 64), b2 = __shfl_down(bits[r], 2, 64),
                           b3 = __shfl_down(bits[r], 3, 64);
        if ((lane & 3) == 0 && i < n) pack[i >> 5] = bits[r] | (b1 << 8) | (b2 << 16) | (b3 << 24);
        const unsigned int c = (unsigned int)__popc(bits[r]);
        unsigned int xs = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int y = __shfl_up(xs, o, 64);
            if (lane >= o) xs += y;
        }
        excl[r] = xs - c;
        if (lane == 63) wtot[r][w] = xs;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 0) {
        smn[w] = mn;
        smx[w] = mx;
    }
    __syncthreads();
    if (w == 0) {
        // tile aggregate, min and max
        unsigned long long agg = 0;
        T a = smn[0], b = smx[0];
        for (int r = 0; r < RND; ++r)
            for (int k = 0; k < 4; ++k) agg += wtot[r][k];
        for (int k = 1; k < 4; ++k) {
            a = smn[k] < a ? smn[k] : a;
            b = smx[k] > b ? smx[k] : b;
        }
        unsigned long long pre = 0;
        if (lane == 0) {
            pmin[tile] = (double)a;
            pmax[tile] = (double)b;
            __hip_atomic_store(&status[tile], (tile == 0 ? ST_P : ST_A) | agg, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        if (tile > 0) {
            // the wave reads 64 predecessors at once, nearest first; it sums
            // them up to and including the nearest inclusive prefix (P), and
            // polls again while any tile before that one is unpublished
            long j = (long)tile - 1 - lane;
            unsigned int spins = 0;
            for (;;) {
                const unsigned long long sw =
                    j >= 0 ? __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ST_P;
                const unsigned long long pm = __ballot((sw & ST_P) != 0ull);
                const unsigned long long xm = __ballot((sw & ~ST_VAL) == 0ull);
                const int fp = pm ? __builtin_ctzll(pm) : 64;
                const unsigned long long upto = fp >= 63 ? ~0ull : ((1ull << (fp + 1)) - 1ull);
                if (xm & upto) {
                    if (++spins > LOOKBACK_SPIN_LIMIT) {
                        if (lane == 0) atomicOr(&g->overflow, 2u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);  // fewer polls while predecessors are still running
                    continue;
                }
                pre += wave_sum((lane <= fp && j >= 0) ? (sw & ST_VAL) : 0ull);
                if (fp < 64) break;
                j -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&status[tile], ST_P | (pre + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_pre = pre;
    }
    __syncthreads();  // s_pre is read by every wave below
    // active values to their tile-local rank, staged in LDS half a tile at a
    // time (one chunk, or two for a tile more than half active), then the
    // contiguous output range [s_pre, s_pre + count) is written with
    // coalesced stores
    unsigned int total = 0;
#pragma unroll
    for (int r = 0; r < RND; ++r) total += wtot[r][0] + wtot[r][1] + wtot[r][2] + wtot[r][3];
    const unsigned long long out0 = s_pre;
    bool ovf = false;
    unsigned int rank0[RND];  // this thread's first rank in each round
    {
        unsigned int run = 0;
#pragma unroll
        for (int r = 0; r < RND; ++r) {
            unsigned int pos = run + excl[r];
            for (int k = 0; k < w; ++k) pos += wtot[r][k];
            rank0[r] = pos;
            run += wtot[r][0] + wtot[r][1] + wtot[r][2] + wtot[r][3];
        }
    }
    for (unsigned int lo = 0; lo < total; lo += SVAL) {
        if (lo > 0) __syncthreads();  // the previous chunk has been copied out
#pragma unroll
        for (int r = 0; r < RND; ++r) {
            unsigned int pos = rank0[r];
            if (bits[r]) {
#pragma unroll
                for (int k = 0; k < X_EPT; ++k) {
                    if (bits[r] & (1u << k)) {
                        if (pos - lo < SVAL) sval[pos - lo] = x[r][k];  // pos >= lo (unsigned wrap otherwise)
                        ++pos;
                    }
                }
            }
        }
        __syncthreads();
        const unsigned int cnt = min(total - lo, SVAL);
        for (unsigned int i = tid; i < cnt; i += 256) {
            const unsigned long long at = out0 + lo + i;
            const T val = sval[i];
            compact[at] = val;
            if (q.kind == Q_STEP) {
                const unsigned long long z = zigzag64((long long)quant_step_fi((double)val, q.step, q.inv_step));
                if (z > 65535ull) ovf = true;
                quant[at] = (uint16_t)z;
            }
        }
    }
    if (ovf) atomicOr(&g->overflow, 1u);
}

// tiles per CU: six for FP32 edits (at most 80 VGPRs), four for FP64; the
// batched kernel's segment lookup needs a few more registers (five)
template <typename T>
constexpr int x_waves(bool batched = false) { return sizeof(T) == 4 ? (batched ? 5 : 6) : 4; }

// Each workgroup takes one ticket (an atomic on one counter) for one tile.
// Several tiles per ticket would chain each workgroup's first tile to the
// previous workgroup's last one.
template <typename T>
__global__ __launch_bounds__(256, x_waves<T>()) void k_tile_compact(const T* v, size_t n, QuantDesc q, unsigned long long* status,
                                                      unsigned int* ticket, double* pmin, double* pmax,
                                                      ExtractGlobals* g, unsigned int* pack, T* compact,
                                                      uint16_t* quant) {
    __shared__ unsigned int s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    tile_compact<T>(s_tile, v, n, q, status, pmin, pmax, g, pack, compact, quant);
}

// Many edit arrays (segments) in one launch: the tickets number the tiles of
// segment 0, then those of segment 1, and so on, so a tile's predecessors in
// its own segment hold smaller tickets and belong to workgroups already
// running, as in the single-array pass.
template <typename T>
__global__ __launch_bounds__(256, x_waves<T>(true)) void k_tile_compact_b(const ExtractSeg<T>* segs, int nseg,
                                                        unsigned long long* status, unsigned int* ticket,
                                                        double* pmin, double* pmax) {
    __shared__ unsigned int s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const unsigned int t = s_tile;
    // the last segment whose first tile is <= t (segments without tiles share
    // their first tile with the next one and are passed over)
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (segs[m].tile0 <= t) lo = m;
        else hi = m - 1;
    }
    const ExtractSeg<T>& sg = segs[lo];
    const unsigned int b = sg.tile0;
    tile_compact<T>(t - b, sg.v, sg.n, sg.q, status + b, pmin + b, pmax + b, sg.g, sg.pack, sg.compact, sg.quant);
}

// Count (the last tile's inclusive prefix), min and max over the tiles of one
// array, by a workgroup of 1024 threads; a second copy of the results goes to
// g2 when it is set.
__device__ __forceinline__ void finalize_tiles(const unsigned long long* status, const double* pmin,
                                               const double* pmax, size_t ntiles, ExtractGlobals* g,
                                               ExtractGlobals* g2) {
    __shared__ double wmn[16], wmx[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    double mn = INFINITY, mx = -INFINITY;
    for (size_t i = t; i < ntiles; i += 1024) {
        mn = pmin[i] < mn ? pmin[i] : mn;
        mx = pmax[i] > mx ? pmax[i] : mx;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 0) {
        wmn[w] = mn;
        wmx[w] = mx;
    }
    __syncthreads();
    if (t == 0) {
        double a = INFINITY, b = -INFINITY;
        for (int i = 0; i < 16; ++i) {
            a = wmn[i] < a ? wmn[i] : a;
            b = wmx[i] > b ? wmx[i] : b;
        }
        ExtractGlobals r = *g;
        if (ntiles == 0) {
            r.count = 0;
            r.vmin = r.vmax = 0.0;
        } else {
            const unsigned long long last = status[ntiles - 1];
            if ((last & ST_P) == 0ull) r.overflow |= 2u;
            r.count = last & ST_VAL;
            r.vmin = a;
            r.vmax = b;
        }
        *g = r;
        if (g2) *g2 = r;
    }
}

__global__ __launch_bounds__(1024) void k_extract_finalize(const unsigned long long* status, const double* pmin,
                                                           const double* pmax, size_t ntiles, ExtractGlobals* g) {
    finalize_tiles(status, pmin, pmax, ntiles, g, nullptr);
}

// One workgroup per segment; every segment's results also go to out[segment].
template <typename T>
__global__ __launch_bounds__(1024) void k_extract_finalize_b(const ExtractSeg<T>* segs, int nseg,
                                                             const unsigned long long* status, const double* pmin,
                                                             const double* pmax, size_t ntot, ExtractGlobals* out) {
    const int k = blockIdx.x;
    const size_t b = segs[k].tile0, e = k + 1 < nseg ? segs[k + 1].tile0 : ntot;
    finalize_tiles(status + b, pmin + b, pmax + b, e - b, segs[k].g, out + k);
}

__global__ void k_extract_init_b(ExtractGlobals* const* gs, int nseg, unsigned long long* status, size_t ntot,
                                 unsigned int* ticket) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ntot; i += (size_t)gridDim.x * blockDim.x)
        status[i] = 0ull;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) *ticket = 0u;
        for (int k = threadIdx.x; k < nseg; k += blockDim.x) *gs[k] = ExtractGlobals{};
    }
}

// The reference quantiser needs the global min/max: quantise the compact
// values once the pass above has them (projection_algorithm.cu:447-455).
template <typename T>
__global__ __launch_bounds__(256) void k_quant_compact(const T* compact, uint16_t* quant, const ExtractGlobals* g) {
    const size_t cnt = g->count;
    const T mn = (T)g->vmin, mx = (T)g->vmax;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        quant[i] = (uint16_t)quant_ref16(compact[i], mn, mx);
}

// The same for the segments with the reference quantiser (grid.y = segment).
template <typename T>
__global__ __launch_bounds__(256) void k_quant_compact_b(const ExtractSeg<T>* segs) {
    const ExtractSeg<T>& sg = segs[blockIdx.y];
    if (sg.q.kind == Q_STEP) return;
    const ExtractGlobals* g = sg.g;
    const size_t cnt = g->count;
    const T mn = (T)g->vmin, mx = (T)g->vmax;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        sg.quant[i] = (uint16_t)quant_ref16(sg.compact[i], mn, mx);
}

template <typename T>
__global__ void k_dequant_dense(const T* v, size_t n, QuantDesc q, T* out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (T)dequant(v[i], q);
}

template <typename T>
static void extract_launch(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                           void* scratch, ExtractGlobals* g, hipStream_t s) {
    FFCZ_REQUIRE(((uintptr_t)v % 16) == 0, ERR_INVALID, "edit array must be 16-byte aligned");
    FFCZ_REQUIRE(((uintptr_t)pack % 4) == 0, ERR_INVALID, "flag buffer must be 4-byte aligned");
    const size_t ntiles = ceil_div(n, (size_t)x_tile<T>());
    FFCZ_REQUIRE(ntiles < (1u << 31), ERR_INVALID, "array too large for extraction");
    unsigned char* p = (unsigned char*)scratch;
    unsigned int* ticket = (unsigned int*)(p + 128);
    unsigned long long* status = (unsigned long long*)(p + 256);
    double* pmin = (double*)(status + ntiles);
    double* pmax = pmin + ntiles;
    const unsigned ig = (unsigned)std::min<size_t>(ceil_div(ntiles, (size_t)256), 1024);
    hipLaunchKernelGGL(k_extract_init, dim3(ig), dim3(256), 0, s, status, ntiles, ticket, g);
    FFCZ_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_tile_compact<T>), dim3((unsigned)ntiles), dim3(256), 0, s, v, n, q, status, ticket, pmin,
                       pmax, g, (unsigned int*)pack, compact, quant);
    FFCZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_extract_finalize, dim3(1), dim3(1024), 0, s, (const unsigned long long*)status,
                       (const double*)pmin, (const double*)pmax, ntiles, g);
    FFCZ_HIP(hipGetLastError());
    if (q.kind != Q_STEP) {
        const unsigned qg = (unsigned)std::min<size_t>(ceil_div(n, (size_t)256), 4096);
        hipLaunchKernelGGL((k_quant_compact<T>), dim3(qg), dim3(256), 0, s, (const T*)compact, quant,
                           (const ExtractGlobals*)g);
        FFCZ_HIP(hipGetLastError());
    }
}

template <typename T>
ExtractResult extract_edits(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                            void* scratch, hipStream_t s) {
    ExtractResult res;
    if (n == 0) return res;
    ExtractGlobals* g = (ExtractGlobals*)scratch;
    extract_launch(v, n, q, pack, compact, quant, scratch, g, s);
    ExtractGlobals h{};
    FFCZ_HIP(hipMemcpyAsync(&h, g, sizeof(h), hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    FFCZ_REQUIRE((h.overflow & 2u) == 0, ERR_HIP, "extraction look-back did not complete");
    FFCZ_REQUIRE(h.overflow == 0, ERR_RANGE, "quantised edit does not fit a 16-bit symbol (raise the quantiser step)");
    res.count = (size_t)h.count;
    res.vmin = h.vmin;
    res.vmax = h.vmax;
    return res;
}

// Scratch of the allocation-free extraction: results and ticket (256 bytes),
// then the status word and tile minimum and maximum of every tile of an array
// of EXTRACT_NOALLOC_MAX elements (extract_scratch_bytes).  In host code this
// name is not a device address: noalloc_scratch_ptr() gives the current
// device's copy.
constexpr size_t NOALLOC_SCRATCH_BYTES =
    256 + (EXTRACT_NOALLOC_MAX / X_TILE) * (sizeof(unsigned long long) + 2 * sizeof(double));
static __device__ __align__(256) unsigned char noalloc_scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::once_flag g_noalloc_once[NOALLOC_MAX_DEVICES];
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES];
hipError_t g_noalloc_err[NOALLOC_MAX_DEVICES];

// The current device's address of noalloc_scratch, looked up once per device
// (with that device current, as the calling thread has it).
void* noalloc_scratch_ptr() {
    int dev = 0;
    FFCZ_HIP(hipGetDevice(&dev));
    FFCZ_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, ERR_INVALID, "device index beyond the scratch table");
    std::call_once(g_noalloc_once[dev], [dev] {
        void* p = nullptr;
        g_noalloc_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
        g_noalloc_ptr[dev] = g_noalloc_err[dev] == hipSuccess ? p : nullptr;
    });
    FFCZ_REQUIRE(g_noalloc_err[dev] == hipSuccess && g_noalloc_ptr[dev], ERR_HIP,
                 std::string("hipGetSymbolAddress of the extraction scratch failed: ") +
                     hipGetErrorName(g_noalloc_err[dev]));
    return g_noalloc_ptr[dev];
}
}  // namespace

template <typename T>
ExtractResult extract_edits_noalloc(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                                    hipStream_t s) {
    FFCZ_REQUIRE(n <= EXTRACT_NOALLOC_MAX, ERR_CAPACITY,
                 "array larger than the allocation-free extraction holds (2^30 elements)");
    static_assert(NOALLOC_SCRATCH_BYTES >= 256 + (EXTRACT_NOALLOC_MAX / X_TILE) * 24, "scratch size");
    if (n == 0) return ExtractResult{};
    // extract_launch clears the ticket, status words and results first
    return extract_edits(v, n, q, pack, compact, quant, noalloc_scratch_ptr(), s);
}

__global__ void k_clear_globals(ExtractGlobals* g) {
    if (threadIdx.x == 0) *g = ExtractGlobals{};
}

template <typename T>
void extract_edits_async(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                         void* scratch, ExtractGlobals* dglob, ExtractGlobals* hglob, hipStream_t s) {
    if (n == 0) {
        hipLaunchKernelGGL(k_clear_globals, dim3(1), dim3(64), 0, s, dglob);
        FFCZ_HIP(hipGetLastError());
    } else {
        extract_launch(v, n, q, pack, compact, quant, scratch, dglob, s);
    }
    if (hglob) FFCZ_HIP(hipMemcpyAsync(hglob, dglob, sizeof(ExtractGlobals), hipMemcpyDeviceToHost, s));
}

__global__ void k_extract_zeros(unsigned long long* pack, size_t words, ExtractGlobals* g) {
    const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i0 == 0) *g = ExtractGlobals{};
    for (size_t i = i0; i < words; i += (size_t)gridDim.x * blockDim.x) pack[i] = 0ull;
}

void extract_zeros_async(uint8_t* pack, size_t n, ExtractGlobals* dglob, hipStream_t s) {
    const size_t words = ceil_div(n, (size_t)64);
    const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>(ceil_div(words, (size_t)256), 4096));
    hipLaunchKernelGGL(k_extract_zeros, dim3(g), dim3(256), 0, s, (unsigned long long*)pack, words, dglob);
    FFCZ_HIP(hipGetLastError());
}

void ExtractBatchWork::release() {
    if (dev) (void)hipFree(dev);
    if (host) (void)hipHostFree(host);
    if (done) (void)hipEventDestroy(done);
    dev = host = nullptr;
    dev_cap = host_cap = 0;
    done = nullptr;
}

template <typename T>
const ExtractGlobals* extract_edits_batch(std::vector<ExtractSeg<T>>& segs, ExtractBatchWork& w, hipStream_t s) {
    const int nseg = (int)segs.size();
    FFCZ_REQUIRE(nseg > 0, ERR_INVALID, "no edit arrays to extract");
    size_t ntot = 0, maxn = 0;
    bool ref = false;
    for (ExtractSeg<T>& sg : segs) {
        FFCZ_REQUIRE(((uintptr_t)sg.v % 16) == 0, ERR_INVALID, "edit array must be 16-byte aligned");
        FFCZ_REQUIRE(((uintptr_t)sg.pack % 4) == 0, ERR_INVALID, "flag buffer must be 4-byte aligned");
        sg.tile0 = (unsigned int)ntot;
        ntot += ceil_div(sg.n, (size_t)x_tile<T>());
        FFCZ_REQUIRE(ntot < (1u << 31), ERR_INVALID, "batch too large for extraction");
        maxn = std::max(maxn, sg.n);
        ref = ref || (sg.q.kind != Q_STEP && sg.n > 0);
    }
    auto r256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // device: segment table | result pointers | ticket | status, pmin, pmax | results
    // pinned host: segment table | result pointers | results
    const size_t tab_b = r256(nseg * sizeof(ExtractSeg<T>)), gs_b = r256(nseg * sizeof(ExtractGlobals*));
    const size_t ticket_off = tab_b + gs_b, status_off = ticket_off + 256;
    const size_t out_off = r256(status_off + ntot * (sizeof(unsigned long long) + 2 * sizeof(double)));
    const size_t dev_b = out_off + nseg * sizeof(ExtractGlobals), host_b = tab_b + gs_b + nseg * sizeof(ExtractGlobals);
    // the staging area and results are reused: the previous batch must be done with them
    if (w.done) FFCZ_HIP(hipEventSynchronize(w.done));
    else FFCZ_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
    if (dev_b > w.dev_cap) {
        if (w.dev) FFCZ_HIP(hipFree(w.dev));
        w.dev = nullptr;
        w.dev_cap = 0;
        FFCZ_HIP(hipMalloc((void**)&w.dev, dev_b));
        w.dev_cap = dev_b;
    }
    if (host_b > w.host_cap) {
        if (w.host) FFCZ_HIP(hipHostFree(w.host));
        w.host = nullptr;
        w.host_cap = 0;
        FFCZ_HIP(hipHostMalloc((void**)&w.host, host_b, hipHostMallocDefault));
        w.host_cap = host_b;
    }
    std::memcpy(w.host, segs.data(), nseg * sizeof(ExtractSeg<T>));
    ExtractGlobals** hg = (ExtractGlobals**)(w.host + tab_b);
    for (int k = 0; k < nseg; ++k) hg[k] = segs[k].g;
    FFCZ_HIP(hipMemcpyAsync(w.dev, w.host, tab_b + gs_b, hipMemcpyHostToDevice, s));
    const ExtractSeg<T>* dsegs = (const ExtractSeg<T>*)w.dev;
    ExtractGlobals* const* dgs = (ExtractGlobals* const*)(w.dev + tab_b);
    unsigned int* ticket = (unsigned int*)(w.dev + ticket_off);
    unsigned long long* status = (unsigned long long*)(w.dev + status_off);
    double* pmin = (double*)(status + ntot);
    double* pmax = pmin + ntot;
    ExtractGlobals* dout = (ExtractGlobals*)(w.dev + out_off);
    ExtractGlobals* hout = (ExtractGlobals*)(w.host + tab_b + gs_b);
    const unsigned ig = (unsigned)std::min<size_t>(std::max<size_t>(ceil_div(ntot, (size_t)256), 1), 1024);
    hipLaunchKernelGGL(k_extract_init_b, dim3(ig), dim3(256), 0, s, dgs, nseg, status, ntot, ticket);
    FFCZ_HIP(hipGetLastError());
    if (ntot > 0) {
        hipLaunchKernelGGL((k_tile_compact_b<T>), dim3((unsigned)ntot), dim3(256), 0, s, dsegs, nseg, status, ticket,
                           pmin, pmax);
        FFCZ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL((k_extract_finalize_b<T>), dim3((unsigned)nseg), dim3(1024), 0, s, dsegs, nseg,
                       (const unsigned long long*)status, (const double*)pmin, (const double*)pmax, ntot, dout);
    FFCZ_HIP(hipGetLastError());
    if (ref) {
        const unsigned qx = (unsigned)std::min<size_t>(ceil_div(maxn, (size_t)256), 256);
        FFCZ_REQUIRE(nseg < 65536, ERR_INVALID, "too many edit arrays in one batch");
        hipLaunchKernelGGL((k_quant_compact_b<T>), dim3(qx, (unsigned)nseg), dim3(256), 0, s, dsegs);
        FFCZ_HIP(hipGetLastError());
    }
    FFCZ_HIP(hipMemcpyAsync(hout, dout, nseg * sizeof(ExtractGlobals), hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipEventRecord(w.done, s));
    return hout;
}

template <typename T>
void dequantize_dense(const T* v, size_t n, QuantDesc q, T* out, hipStream_t s) {
    size_t gsz = ceil_div(n ? n : 1, (size_t)256);
    if (gsz > (1u << 20)) gsz = 1u << 20;
    hipLaunchKernelGGL((k_dequant_dense<T>), dim3((unsigned)gsz), dim3(256), 0, s, v, n, q, out);
    FFCZ_HIP(hipGetLastError());
}

template ExtractResult extract_edits<float>(const float*, size_t, QuantDesc, uint8_t*, float*, uint16_t*, void*, hipStream_t);
template ExtractResult extract_edits<double>(const double*, size_t, QuantDesc, uint8_t*, double*, uint16_t*, void*, hipStream_t);
template ExtractResult extract_edits_noalloc<float>(const float*, size_t, QuantDesc, uint8_t*, float*, uint16_t*,
                                                   hipStream_t);
template ExtractResult extract_edits_noalloc<double>(const double*, size_t, QuantDesc, uint8_t*, double*, uint16_t*,
                                                    hipStream_t);
template void extract_edits_async<float>(const float*, size_t, QuantDesc, uint8_t*, float*, uint16_t*, void*,
                                         ExtractGlobals*, ExtractGlobals*, hipStream_t);
template void extract_edits_async<double>(const double*, size_t, QuantDesc, uint8_t*, double*, uint16_t*, void*,
                                          ExtractGlobals*, ExtractGlobals*, hipStream_t);
template const ExtractGlobals* extract_edits_batch<float>(std::vector<ExtractSeg<float>>&, ExtractBatchWork&,
                                                         hipStream_t);
template const ExtractGlobals* extract_edits_batch<double>(std::vector<ExtractSeg<double>>&, ExtractBatchWork&,
                                                          hipStream_t);
template void dequantize_dense<float>(const float*, size_t, QuantDesc, float*, hipStream_t);
template void dequantize_dense<double>(const double*, size_t, QuantDesc, double*, hipStream_t);

}  // namespace ffcz
