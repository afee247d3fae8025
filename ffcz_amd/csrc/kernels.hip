// Stand-alone element-wise kernels of the reference (per-kernel entries of the
// C ABI).  The solver runs the same arithmetic fused into the FFT passes; these
// exist so each clamp / reduction / quantiser can be checked on its own.
#include "kernels.hpp"
#include "quant.hpp"

namespace ffcz {

template <typename T>
__global__ void k_project_spatial(T* e, T eps, T* ds, size_t n) {
    // projection_algorithm.cu:335-361
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const T v = e[i];
        if (v > eps) {
            ds[i] += eps - v;
            e[i] = eps;
        } else if (v < -eps) {
            ds[i] += -eps - v;
            e[i] = -eps;
        }
    }
}

template <typename T>
__global__ void k_project_freq_abs(cplx<T>* E, T d, T* df, size_t F) {
    // projection_algorithm.cu:147-188 (one thread per bin, Re then Im)
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < F; i += (size_t)gridDim.x * blockDim.x) {
        cplx<T> v = E[i];
        if (v.x > d) {
            df[i] += d - v.x;
            v.x = d;
        } else if (v.x < -d) {
            df[i] += -d - v.x;
            v.x = -d;
        }
        if (v.y > d) {
            df[i + F] += d - v.y;
            v.y = d;
        } else if (v.y < -d) {
            df[i + F] += -d - v.y;
            v.y = -d;
        }
        E[i] = v;
    }
}

template <typename T>
__global__ void k_convergence(const T* e, T eps, size_t n, unsigned long long* out) {
    // projection_algorithm.cu:586-612 reduced to one max
    T m = T(0);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const T a = fabs(e[i]);
        if (a > eps) m = fmax(m, a - eps);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0 && m > T(0)) atomicMax(out, ordbits(m));
}

template <typename T>
__global__ void k_quant_dequant(const T* v, T* out, T mn, T mx, size_t n, T thr) {
    // projection_algorithm.cu:615-644
#pragma clang fp contract(off)
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const T x = v[i];
        if (fabs(x) > thr) {
            const T q = quant_ref16(x, mn, mx);
            out[i] = q / (T)65535 * (mx - mn) + mn;
        } else {
            out[i] = T(0);
        }
    }
}

static unsigned grid1(size_t n) {
    size_t g = ceil_div(n ? n : 1, (size_t)256);
    return (unsigned)(g > 65536 ? 65536 : g);
}

template <typename T>
void launch_project_spatial(T* e, T eps, T* ds, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_project_spatial<T>), dim3(grid1(n)), dim3(256), 0, s, e, eps, ds, n);
    FFCZ_HIP(hipGetLastError());
}
template <typename T>
void launch_project_freq_abs(cplx<T>* E, T d, T* df, size_t F, hipStream_t s) {
    hipLaunchKernelGGL((k_project_freq_abs<T>), dim3(grid1(F)), dim3(256), 0, s, E, d, df, F);
    FFCZ_HIP(hipGetLastError());
}
template <typename T>
double convergence_metric(const T* e, T eps, size_t n, hipStream_t s) {
    unsigned long long* d = nullptr;
    FFCZ_HIP(hipMalloc(&d, sizeof(unsigned long long)));
    unsigned long long h = 0;
    try {
        FFCZ_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL((k_convergence<T>), dim3(grid1(n)), dim3(256), 0, s, e, eps, n, d);
        FFCZ_HIP(hipGetLastError());
        FFCZ_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipFree(d);
        throw;
    }
    (void)hipFree(d);
    if (sizeof(T) == 4) return (double)__builtin_bit_cast(float, (uint32_t)h);
    return __builtin_bit_cast(double, h);
}
template <typename T>
void launch_quant_dequant(const T* v, T* out, T mn, T mx, size_t n, T thr, hipStream_t s) {
    hipLaunchKernelGGL((k_quant_dequant<T>), dim3(grid1(n)), dim3(256), 0, s, v, out, mn, mx, n, thr);
    FFCZ_HIP(hipGetLastError());
}

#define FFCZ_INST(T)                                                                       \
    template void launch_project_spatial<T>(T*, T, T*, size_t, hipStream_t);              \
    template void launch_project_freq_abs<T>(cplx<T>*, T, T*, size_t, hipStream_t);       \
    template double convergence_metric<T>(const T*, T, size_t, hipStream_t);              \
    template void launch_quant_dequant<T>(const T*, T*, T, T, size_t, T, hipStream_t);
FFCZ_INST(float)
FFCZ_INST(double)

}  // namespace ffcz
