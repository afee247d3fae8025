// Common definitions for the FFCz MI355X library: error handling, complex type,
// fast integer division, device helpers.  gfx950 only (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace ffcz {

// FFT pass variants of the FP32 iteration: vector-ALU FP32, or radix-16
// stages as DFT-matrix GEMMs on MFMA in split BF16x3 / power-of-two-scaled
// split FP16x3 with FP32 accumulation.
enum FftPrec : int { PREC_VALU = 0, PREC_BF16X3 = 1, PREC_FP16X3 = 2 };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes of the C ABI (include/ffcz.h mirrors them).
enum Status : int {
    OK = 0,
    ERR_INVALID = 1,      // bad argument / unsupported shape
    ERR_STATE = 2,        // e.g. solve before initialize
    ERR_HIP = 3,          // HIP runtime failure
    ERR_RANGE = 4,        // value does not fit (Huffman code > 32 bits, quantised symbol > 16 bits)
    ERR_ALLOC = 5,        // device allocation failed
    ERR_ZSTD = 6,         // libzstd failure
    ERR_NOT_CONVERGED = 7, // informational: loop hit max_iterations
    ERR_CAPACITY = 8      // input larger than a fixed-size buffer holds
};

#define FFCZ_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t _e = (call);                                                                 \
        if (_e != hipSuccess)                                                                   \
            throw ::ffcz::Error(::ffcz::ERR_HIP, std::string("HIP error ") + hipGetErrorName(_e) + \
                                                     " (" + hipGetErrorString(_e) + ") at " +   \
                                                     __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

#define FFCZ_REQUIRE(cond, code, msg)                      \
    do {                                                   \
        if (!(cond)) throw ::ffcz::Error((code), (msg));   \
    } while (0)

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cplx<R> mk(R a, R b) {
    cplx<R> c;
    c.x = a;
    c.y = b;
    return c;
}
template <typename R>
__device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> w) {
    return mk(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
template <typename R>
__device__ __forceinline__ cplx<R> conjg(cplx<R> a) { return mk(a.x, -a.y); }
template <typename R>
__device__ __forceinline__ cplx<R> scal(cplx<R> a, R s) { return mk(a.x * s, a.y * s); }

// Unsigned division by a run-time constant d (1 <= d < 2^31) for dividends
// n < 2^31: q = (umulhi(n, m) + n) >> l  with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1  (Granlund & Montgomery, PLDI 1994, §4).
struct FastDiv {
    uint32_t d, m, l;
    FastDiv() : d(1), m(1), l(0) {}
    explicit FastDiv(uint32_t dv) : d(dv) {
        l = 0;
        while ((1ull << l) < dv) ++l;
        uint64_t two32 = 1ull << 32;
        m = (uint32_t)((two32 * ((1ull << l) - dv)) / dv + 1);
    }
};
__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t t = __umulhi(n, f.m);
#else
    uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
    return (t + n) >> f.l;
}

// Order-preserving map of a non-negative floating value to u64 (for atomicMax).
__device__ __forceinline__ unsigned long long ordbits(float v) { return (unsigned long long)__float_as_uint(v); }
__device__ __forceinline__ unsigned long long ordbits(double v) { return (unsigned long long)__double_as_longlong(v); }
template <typename R>
__device__ __forceinline__ R from_ordbits(unsigned long long b);
template <>
__device__ __forceinline__ float from_ordbits<float>(unsigned long long b) { return __uint_as_float((uint32_t)b); }
template <>
__device__ __forceinline__ double from_ordbits<double>(unsigned long long b) { return __longlong_as_double((long long)b); }

// Wave64 / workgroup reductions.
template <typename V>
__device__ __forceinline__ V wave_max(V v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        V w = __shfl_xor(v, o, 64);
        v = v > w ? v : w;
    }
    return v;
}
template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__host__ __device__ inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace ffcz
