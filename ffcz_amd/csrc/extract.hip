// Edit extraction kernels: a tile-statistics pass (flag count, min, max), a
// single-workgroup scan of the tile counts, and a scatter pass that forms the
// flags again, ranks them with wave ballots, writes compact values and
// quantised symbols in index order and stores each wave's 64-bit ballot as
// eight LSB-first flag bytes.
#include <cfloat>

#include "extract.hpp"
#include "quant.hpp"

namespace ffcz {

struct ExtractGlobals {
    unsigned long long count;
    double vmin, vmax;
    unsigned int overflow;
    unsigned int pad;
};

size_t extract_scratch_bytes(size_t n) {
    const size_t ntiles = ceil_div(n ? n : 1, (size_t)EXTRACT_TILE);
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

template <typename T>
__global__ __launch_bounds__(256) void k_tile_stats(const T* v, size_t n, QuantDesc q, unsigned int* counts,
                                                    double* pmin, double* pmax) {
    const size_t base = (size_t)blockIdx.x * EXTRACT_TILE;
    unsigned int c = 0;
    T mn = (T)INFINITY, mx = (T)-INFINITY;
    for (int r = 0; r < EXTRACT_TILE / 256; ++r) {
        const size_t i = base + r * 256 + threadIdx.x;
        if (i < n) {
            const T x = v[i];
            long long qq;
            c += edit_flag(x, q, qq) ? 1u : 0u;
            mn = x < mn ? x : mn;
            mx = x > mx ? x : mx;
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

__global__ __launch_bounds__(1024) void k_tile_scan(const unsigned int* counts, unsigned long long* offsets,
                                                    const double* pmin, const double* pmax, size_t ntiles,
                                                    ExtractGlobals* g) {
    __shared__ unsigned long long wsum[16];
    __shared__ double wmn[16], wmx[16];
    const int t = threadIdx.x;
    const size_t per = ceil_div(ntiles, (size_t)1024);
    const size_t b0 = t * per, b1 = b0 + per < ntiles ? b0 + per : ntiles;
    unsigned long long s = 0;
    double mn = INFINITY, mx = -INFINITY;
    for (size_t i = b0; i < b1; ++i) {
        s += counts[i];
        mn = pmin[i] < mn ? pmin[i] : mn;
        mx = pmax[i] > mx ? pmax[i] : mx;
    }
    // inclusive scan across the workgroup of the per-thread sums
    unsigned long long x = s;
    const int lane = t & 63, w = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 63) wsum[w] = x;
    if (lane == 0) {
        wmn[w] = mn;
        wmx[w] = mx;
    }
    __syncthreads();
    unsigned long long woff = 0;
    for (int i = 0; i < w; ++i) woff += wsum[i];
    unsigned long long run = woff + x - s;  // exclusive prefix of this thread's chunk
    for (size_t i = b0; i < b1; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    if (t == 0) {
        unsigned long long tot = 0;
        double a = INFINITY, b = -INFINITY;
        for (int i = 0; i < 16; ++i) {
            tot += wsum[i];
            a = wmn[i] < a ? wmn[i] : a;
            b = wmx[i] > b ? wmx[i] : b;
        }
        g->count = tot;
        g->vmin = ntiles ? a : 0.0;
        g->vmax = ntiles ? b : 0.0;
        g->overflow = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tile_scatter(const T* v, size_t n, QuantDesc q,
                                                      const unsigned long long* offsets, ExtractGlobals* g,
                                                      unsigned long long* pack, T* compact, uint16_t* quant) {
    __shared__ unsigned int wtot[4];
    const size_t base = (size_t)blockIdx.x * EXTRACT_TILE;
    unsigned long long run = offsets[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const T mn = (T)g->vmin, mx = (T)g->vmax;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    bool ovf = false;
    for (int r = 0; r < EXTRACT_TILE / 256; ++r) {
        const size_t i = base + r * 256 + threadIdx.x;
        T x = 0;
        long long qq = 0;
        bool f = false;
        if (i < n) {
            x = v[i];
            f = edit_flag(x, q, qq);
        }
        const unsigned long long m = __ballot(f);
        const unsigned int wc = (unsigned int)__popcll(m);
        if (lane == 0) wtot[w] = wc;
        const size_t wbase = base + r * 256 + w * 64;
        if (lane == 0 && wbase < n) pack[wbase >> 6] = m;
        __syncthreads();
        unsigned long long off = run;
        for (int k = 0; k < w; ++k) off += wtot[k];
        const unsigned int rtot = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        if (f) {
            const unsigned long long pos = off + (unsigned long long)__popcll(m & below);
            compact[pos] = x;
            if (q.kind == Q_STEP) {
                const unsigned long long z = zigzag64(qq);
                if (z > 65535ull) ovf = true;
                quant[pos] = (uint16_t)z;
            } else {
                quant[pos] = (uint16_t)quant_ref16(x, mn, mx);
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
ExtractResult extract_edits(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                            void* scratch, hipStream_t s) {
    ExtractResult res;
    if (n == 0) return res;
    const size_t ntiles = ceil_div(n, (size_t)EXTRACT_TILE);
    FFCZ_REQUIRE(ntiles < (1u << 31), ERR_INVALID, "array too large for extraction");
    unsigned char* p = (unsigned char*)scratch;
    ExtractGlobals* g = (ExtractGlobals*)p;
    unsigned long long* offsets = (unsigned long long*)(p + 256);
    double* pmin = (double*)(offsets + ntiles);
    double* pmax = pmin + ntiles;
    unsigned int* counts = (unsigned int*)(pmax + ntiles);
    hipLaunchKernelGGL((k_tile_stats<T>), dim3((unsigned)ntiles), dim3(256), 0, s, v, n, q, counts, pmin, pmax);
    FFCZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, counts, offsets, pmin, pmax, ntiles, g);
    FFCZ_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_tile_scatter<T>), dim3((unsigned)ntiles), dim3(256), 0, s, v, n, q, offsets, g,
                       (unsigned long long*)pack, compact, quant);
    FFCZ_HIP(hipGetLastError());
    ExtractGlobals h{};
    FFCZ_HIP(hipMemcpyAsync(&h, g, sizeof(h), hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    FFCZ_REQUIRE(h.overflow == 0, ERR_RANGE, "quantised edit does not fit a 16-bit symbol (raise the quantiser step)");
    res.count = (size_t)h.count;
    res.vmin = h.vmin;
    res.vmax = h.vmax;
    return res;
}

template <typename T>
void dequantize_dense(const T* v, size_t n, QuantDesc q, T* out, hipStream_t s) {
    size_t g = ceil_div(n ? n : 1, (size_t)256);
    if (g > (1u << 20)) g = 1u << 20;
    hipLaunchKernelGGL((k_dequant_dense<T>), dim3((unsigned)g), dim3(256), 0, s, v, n, q, out);
    FFCZ_HIP(hipGetLastError());
}

template ExtractResult extract_edits<float>(const float*, size_t, QuantDesc, uint8_t*, float*, uint16_t*, void*, hipStream_t);
template ExtractResult extract_edits<double>(const double*, size_t, QuantDesc, uint8_t*, double*, uint16_t*, void*, hipStream_t);
template void dequantize_dense<float>(const float*, size_t, QuantDesc, float*, hipStream_t);
template void dequantize_dense<double>(const double*, size_t, QuantDesc, double*, hipStream_t);

}  // namespace ffcz
