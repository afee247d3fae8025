// Edit quantisers shared by extraction and certification.
#pragma once

#include "fft3d.hpp"

namespace ffcz {

// Reference 16-bit quantiser, projection_algorithm.cu:447-455, in type T:
// n = (v-min)/(max-min); q = n > 0.5 ? ceil(n*65535) : floor(n*65535).
template <typename T>
__device__ __forceinline__ T quant_ref16(T v, T mn, T mx) {
#pragma clang fp contract(off)
    const T n = (v - mn) / (mx - mn);
    const T s = n * (T)65535;
    return n > (T)0.5 ? ceil(s) : floor(s);
}

// Step quantiser: llround(v/step) in FP64, ties away from zero, as an
// integer-valued double (|n| <= 4e15 < 2^53, so every value is exact).
__device__ __forceinline__ double quant_step_f(double v, double step) {
    double r = v / step;
    const double lim = 4.0e15;
    r = r > lim ? lim : (r < -lim ? -lim : r);
    return r < 0.0 ? -floor(-r + 0.5) : floor(r + 0.5);
}
__device__ __forceinline__ long long quant_step(double v, double step) { return (long long)quant_step_f(v, step); }

// quant_step_f(v, step) with the quotient taken through inv = 1/step (a
// multiply instead of an FP64 division) where that cannot change the result.
// r = v*inv is within 1.5*2^-52 (relative) of the rounded quotient v/step, so
// below 2^24 in magnitude the two differ by less than 6e-9; where r lies
// farther than 1e-6 from a half-integer, both round to the same integer, which
// rint(r) gives exactly (no tie to break there).  Near half-integers, for
// large or non-finite values, and without a reciprocal (inv <= 0) the
// division decides.  A zero result may carry either sign.
__device__ __forceinline__ bool quant_step_fast(double r, double n, double inv) {
    return inv > 0.0 && fabs(r) < 16777216.0 && fabs(r - n) < 0.499999;
}
__device__ __forceinline__ double quant_step_fi(double v, double step, double inv) {
    const double r = v * inv, n = rint(r);
    if (quant_step_fast(r, n, inv)) return n;
    return quant_step_f(v, step);
}

// quant_step_f(v, step) != 0 without the division where the answer is
// clear: |v/step| < 1/2 rounds to 0, above it to at least 1.  The relative
// margin 1e-5 is far wider than the rounding of the thresholds and of the
// quotient, so only |v| within it of step/2 takes the quantiser itself.
__device__ __forceinline__ bool quant_step_nonzero(double v, double step) {
    const double a = fabs(v);
    if (a < step * 0.49999) return false;
    if (a > step * 0.50001) return true;
    return quant_step_f(v, step) != 0.0;
}

__device__ __forceinline__ unsigned long long zigzag64(long long q) {
    return ((unsigned long long)q << 1) ^ (unsigned long long)(q >> 63);
}

// Value a decoder reconstructs for edit v (0 for inactive edits).
template <typename T>
__device__ __forceinline__ double dequant(T v, const QuantDesc& q) {
#pragma clang fp contract(off)
    if (q.kind == Q_NONE) return (double)v;
    if (q.kind == Q_REF16) {
        // projection_algorithm.cu:629-643 in type T
        if (!(fabs(v) > (T)q.threshold)) return 0.0;
        const T mn = (T)(q.dminmax ? q.dminmax[0] : q.vmin), mx = (T)(q.dminmax ? q.dminmax[1] : q.vmax);
        const T qv = quant_ref16(v, mn, mx);
        const T r = qv / (T)65535 * (mx - mn) + mn;
        return (double)r;
    }
    // edits below half a step dequantise to zero (with the quantiser's sign);
    // the reciprocal decides wherever quant_step_fi's would, the division
    // elsewhere, and the symbol stays an integer-valued double
    const double dv = (double)v, r = dv * q.inv_step, n = rint(r);
    if (quant_step_fast(r, n, q.inv_step)) return n != 0.0 ? n * q.step : (v < (T)0 ? -0.0 : 0.0);
    if (!quant_step_nonzero(dv, q.step)) return v < (T)0 ? -0.0 : 0.0;
    return quant_step_f(dv, q.step) * q.step;
}

}  // namespace ffcz
