// Edit extraction kernels.
//
// Tiles of 8192 elements: 256 threads x 8 consecutive elements x 4 rounds.
//  1. k_tile_stats   flag count, min and max per tile (vector loads);
//  2. k_tile_scan    one workgroup scans the tile counts in coalesced chunks
//                    and reduces min/max;
//  3. k_tile_scatter forms the flags again, ranks them with a wave scan of
//                    the per-thread counts, writes compact values and
//                    quantised symbols in index order, and stores the flags
//                    LSB-first (four threads' bytes combined into one word).
#include <cfloat>

#include "extract.hpp"
#include "quant.hpp"

namespace ffcz {

constexpr int X_EPT = 8;                          // elements per thread per round
constexpr int X_ROUNDS = 4;
constexpr int X_TILE = 256 * X_EPT * X_ROUNDS;   // 8192
static_assert(X_TILE == EXTRACT_TILE, "tile size");

size_t extract_scratch_bytes(size_t n) {
    const size_t ntiles = ceil_div(n ? n : 1, (size_t)X_TILE);
    return 256 + ntiles * (sizeof(unsigned int) + sizeof(unsigned long long) + 2 * sizeof(double));
}

template <typename T>
__device__ __forceinline__ bool edit_flag(T v, const QuantDesc& q, long long& r) {
    if (q.kind == Q_STEP) {
        r = quant_step((double)v, q.step);
        return r != 0;
    }
    r = 0;
    return fabs(v) > (T)q.threshold;  // projection_algorithm.cu:377
}

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

template <typename T>
__global__ __launch_bounds__(256) void k_tile_stats(const T* v, size_t n, QuantDesc q, unsigned int* counts,
                                                    double* pmin, double* pmax) {
    const size_t base = (size_t)blockIdx.x * X_TILE;
    unsigned int c = 0;
    T mn = (T)INFINITY, mx = (T)-INFINITY;
#pragma unroll
    for (int r = 0; r < X_ROUNDS; ++r) {
        const size_t i = base + ((size_t)r * 256 + threadIdx.x) * X_EPT;
        if (i < n) {
            T x[X_EPT];
            load8(v, i, n, x);
#pragma unroll
            for (int k = 0; k < X_EPT; ++k) {
                if (i + k < n) {
                    long long qq;
                    c += edit_flag(x[k], q, qq) ? 1u : 0u;
                    mn = x[k] < mn ? x[k] : mn;
                    mx = x[k] > mx ? x[k] : mx;
                }
            }
        }
    }
    __shared__ unsigned int sc[4];
    __shared__ T smn[4], smx[4];
    c = wave_sum(c);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sc[w] = c;
        smn[w] = mn;
        smx[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int tc = 0;
        T a = smn[0], b = smx[0];
        for (int i = 0; i < 4; ++i) {
            tc += sc[i];
            a = smn[i] < a ? smn[i] : a;
            b = smx[i] > b ? smx[i] : b;
        }
        counts[blockIdx.x] = tc;
        pmin[blockIdx.x] = (double)a;
        pmax[blockIdx.x] = (double)b;
    }
}

// Exclusive scan of the tile counts in chunks of 1024 (one element per
// thread per chunk, coalesced), plus the global min / max.
__global__ __launch_bounds__(1024) void k_tile_scan(const unsigned int* counts, unsigned long long* offsets,
                                                    const double* pmin, const double* pmax, size_t ntiles,
                                                    ExtractGlobals* g) {
    __shared__ unsigned long long wsum[16];
    __shared__ double wmn[16], wmx[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned long long run = 0;
    double mn = INFINITY, mx = -INFINITY;
    for (size_t c0 = 0; c0 < ntiles; c0 += 1024) {
        const size_t i = c0 + t;
        const unsigned long long v = i < ntiles ? counts[i] : 0ull;
        if (i < ntiles) {
            mn = pmin[i] < mn ? pmin[i] : mn;
            mx = pmax[i] > mx ? pmax[i] : mx;
        }
        unsigned long long x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        unsigned long long woff = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            woff += (k < w) ? wsum[k] : 0ull;
            tot += wsum[k];
        }
        if (i < ntiles) offsets[i] = run + woff + x - v;
        run += tot;
        __syncthreads();
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
        g->count = run;
        g->vmin = ntiles ? a : 0.0;
        g->vmax = ntiles ? b : 0.0;
        g->overflow = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tile_scatter(const T* v, size_t n, QuantDesc q,
                                                      const unsigned long long* offsets, ExtractGlobals* g,
                                                      unsigned int* pack, T* compact, uint16_t* quant) {
    __shared__ unsigned int wtot[4];
    const size_t base = (size_t)blockIdx.x * X_TILE;
    unsigned long long run = offsets[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const T mn = (T)g->vmin, mx = (T)g->vmax;
    bool ovf = false;
#pragma unroll 1
    for (int r = 0; r < X_ROUNDS; ++r) {
        const size_t i = base + ((size_t)r * 256 + threadIdx.x) * X_EPT;
        T x[X_EPT];
        long long qq[X_EPT];
        unsigned int bits = 0;
        if (i < n) {
            load8(v, i, n, x);
#pragma unroll
            for (int k = 0; k < X_EPT; ++k)
                if (i + k < n && edit_flag(x[k], q, qq[k])) bits |= 1u << k;
        }
        // four consecutive threads' flag bytes form one little-endian word
        const unsigned int b1 = __shfl_down(bits, 1, 64), b2 = __shfl_down(bits, 2, 64),
                           b3 = __shfl_down(bits, 3, 64);
        if ((lane & 3) == 0 && i < n) pack[i >> 5] = bits | (b1 << 8) | (b2 << 16) | (b3 << 24);
        const unsigned int c = (unsigned int)__popc(bits);
        unsigned int xs = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int y = __shfl_up(xs, o, 64);
            if (lane >= o) xs += y;
        }
        if (lane == 63) wtot[w] = xs;
        __syncthreads();
        unsigned long long pos = run + (xs - c);
        for (int k = 0; k < w; ++k) pos += wtot[k];
        const unsigned int rtot = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        if (bits) {
#pragma unroll
            for (int k = 0; k < X_EPT; ++k) {
                if (bits & (1u << k)) {
                    compact[pos] = x[k];
                    if (q.kind == Q_STEP) {
                        const unsigned long long z = zigzag64(qq[k]);
                        if (z > 65535ull) ovf = true;
                        quant[pos] = (uint16_t)z;
                    } else {
                        quant[pos] = (uint16_t)quant_ref16(x[k], mn, mx);
                    }
                    ++pos;
                }
            }
        }
        run += rtot;
        __syncthreads();
    }
    if (ovf) atomicOr(&g->overflow, 1u);
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
    const size_t ntiles = ceil_div(n, (size_t)X_TILE);
    FFCZ_REQUIRE(ntiles < (1u << 31), ERR_INVALID, "array too large for extraction");
    unsigned char* p = (unsigned char*)scratch;
    unsigned long long* offsets = (unsigned long long*)(p + 256);
    double* pmin = (double*)(offsets + ntiles);
    double* pmax = pmin + ntiles;
    unsigned int* counts = (unsigned int*)(pmax + ntiles);
    hipLaunchKernelGGL((k_tile_stats<T>), dim3((unsigned)ntiles), dim3(256), 0, s, v, n, q, counts, pmin, pmax);
    FFCZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, counts, offsets, pmin, pmax, ntiles, g);
    FFCZ_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_tile_scatter<T>), dim3((unsigned)ntiles), dim3(256), 0, s, v, n, q, offsets, g,
                       (unsigned int*)pack, compact, quant);
    FFCZ_HIP(hipGetLastError());
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
    FFCZ_REQUIRE(h.overflow == 0, ERR_RANGE, "quantised edit does not fit a 16-bit symbol (raise the quantiser step)");
    res.count = (size_t)h.count;
    res.vmin = h.vmin;
    res.vmax = h.vmax;
    return res;
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
    FFCZ_HIP(hipMemcpyAsync(hglob, dglob, sizeof(ExtractGlobals), hipMemcpyDeviceToHost, s));
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
template void extract_edits_async<float>(const float*, size_t, QuantDesc, uint8_t*, float*, uint16_t*, void*,
                                         ExtractGlobals*, ExtractGlobals*, hipStream_t);
template void extract_edits_async<double>(const double*, size_t, QuantDesc, uint8_t*, double*, uint16_t*, void*,
                                          ExtractGlobals*, ExtractGlobals*, hipStream_t);
template void dequantize_dense<float>(const float*, size_t, QuantDesc, float*, hipStream_t);
template void dequantize_dense<double>(const double*, size_t, QuantDesc, double*, hipStream_t);

}  // namespace ffcz
