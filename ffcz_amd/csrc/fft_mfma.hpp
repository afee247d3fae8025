// Radix-16 Stockham stages as DFT-matrix GEMMs on the gfx950 matrix cores.
//
// A radix-16 stage of a tile of NPEN pencils holds NPEN*L/16 butterflies.
// After the twiddle multiply (VALU, FP32) the butterflies are the rows of a
// real matrix A [butterflies x 32], row = (Re x_0, Im x_0, ..., Re x_15,
// Im x_15), and the stage is the real GEMM  Y = A * B  with the 32x32 matrix
// B[2r+c][2m+c'] that maps (Re, Im) of input r to (Re, Im) of output m of
// the forward 16-point DFT.  One wave computes 16 butterflies per pair of
// v_mfma_f32_16x16x32 instructions (outputs m = 0..7 and 8..15).  The inverse
// uses DFT^+(x) = conj(DFT^-(conj x)), so one B serves both directions.
//
// Operand lane layout of v_mfma_f32_16x16x32_{bf16,f16} (wave64, checked on
// the device): lane l holds A[l%16][8*(l/16) + 0..7], B[8*(l/16) + 0..7][l%16]
// and D[4*(l/16) + 0..3][l%16].  With the interleaved K order above, lane l
// owns inputs r = 4*(l/16) .. +3 of butterfly l%16, so each complex input is
// read from LDS exactly once.
//
// Precision variants (FP32 accumulation throughout):
//   PREC_BF16X3  x = hi + lo, B = Bhi + Blo in bfloat16; Y = hi*Bhi + lo*Bhi + hi*Blo
//   PREC_FP16X3  the same split in float16, each butterfly's 16 inputs first
//                scaled by a power of two so that max |x| lies in [2^14, 2^15)
//                (the dynamic range of a pencil segment then fits float16);
//                the scale is undone exactly on the FP32 outputs.
// The stages before the radix-16 ones (the length's other factors) stay on the
// vector ALU (fft_pencil.hpp); so do the twiddles and the first-stage loads
// and last-stage stores with their fused element-wise work.
#pragma once

#include "fft_pencil.hpp"

namespace ffcz {

namespace mx {

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct Op {
    using E = float;
    using V = int;
};
template <>
struct Op<PREC_BF16X3> {
    using E = __bf16;
    typedef __bf16 V __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ f32x4 mma(V a, V b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Op<PREC_FP16X3> {
    using E = _Float16;
    typedef _Float16 V __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ f32x4 mma(V a, V b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// radix-16 factors of L done on the matrix cores, and the VALU remainder
constexpr int n16(int L) {
    int a = 0;
    while (L > 1 && L % 16 == 0) {
        L /= 16;
        ++a;
    }
    return a;
}
constexpr int pow16(int a) {
    int p = 1;
    for (int i = 0; i < a; ++i) p *= 16;
    return p;
}
constexpr int rest_of(int L) { return L / pow16(n16(L)); }

// cos(2 pi t / 16) from its four distinct magnitudes
__device__ __forceinline__ float cos16(int t) {
    int u = t & 15;
    if (u > 8) u = 16 - u;
    float s = 1.0f;
    if (u > 4) {
        u = 8 - u;
        s = -1.0f;
    }
    const float v = u == 0 ? 1.0f
                    : u == 1 ? 0.92387953251128675613f
                    : u == 2 ? 0.70710678118654752440f
                    : u == 3 ? 0.38268343236508977173f
                             : 0.0f;
    return s * v;
}
__device__ __forceinline__ float sin16(int t) { return cos16(t - 4); }

template <int PREC>
struct Bmat {
    typename Op<PREC>::V hi[2], lo[2];
};

template <int PREC>
__device__ __forceinline__ void split(float x, typename Op<PREC>::E& h, typename Op<PREC>::E& l) {
    using E = typename Op<PREC>::E;
    h = (E)x;
    l = (E)(x - (float)h);
}

// The workgroup's copy of every lane's DFT-matrix slice (hi/lo for both output
// halves, 4 KB): written once per transform by the first wave, read by each
// stage just before its MFMAs.  Keeping it in LDS rather than registers holds
// the kernels to two 1024-thread workgroups per CU.
template <int PREC>
__device__ __forceinline__ typename Op<PREC>::V* b_table() {
    __shared__ typename Op<PREC>::V tab[4][64];
    return &tab[0][0];
}

// This lane's slice of the forward 16-point DFT matrix (both output halves).
template <int PREC>
__device__ __forceinline__ Bmat<PREC> dft16_operand(int lane) {
    Bmat<PREC> b{};
    if constexpr (PREC != PREC_VALU) {
        const int col = lane & 15, g = (lane >> 4) & 3;
        const int cp = col & 1;
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            const int m = (col >> 1) + 8 * jb;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int r = 4 * g + (e >> 1), c = e & 1;
                const int t = r * m;
                const float wre = cos16(t), wim = -sin16(t);
                const float v = c == 0 ? (cp == 0 ? wre : wim) : (cp == 0 ? -wim : wre);
                typename Op<PREC>::E h, l;
                split<PREC>(v, h, l);
                b.hi[jb][e] = h;
                b.lo[jb][e] = l;
            }
        }
    }
    return b;
}

// One in-place radix-16 Stockham stage (sub-transform length NS before the
// stage) over the tile buf[n*SW + p], p < NPEN, n < L, run by NT threads.
// The caller barriers before (inputs visible) and after (outputs visible).
template <int PREC, int L, int NS, int SW, int NPEN, int NT>
__device__ __forceinline__ void r16_stage(cplx<float>* buf, const cplx<float>* __restrict__ tw, int tid, bool inv) {
    using O = Op<PREC>;
    using V = typename O::V;
    constexpr int LR = L / 16;
    constexpr int TWS = L / (NS * 16);
    constexpr int NBF = NPEN * LR;
    static_assert(NBF % 16 == 0, "butterflies must fill whole MFMA batches");
    static_assert(NT % 64 == 0, "MFMA stages need whole waves");
    constexpr int NQ = NBF / 16;
    constexpr int NW = NT / 64;
    constexpr int QPW = (NQ + NW - 1) / NW;
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    V ahi[QPW], alo[QPW];
    float unsc[QPW];
#pragma unroll
    for (int qq = 0; qq < QPW; ++qq) {
        const int q = wave + qq * NW;
        unsc[qq] = 1.0f;
        if ((NQ % NW) != 0 && q >= NQ) continue;  // uniform per wave
        const int b = q * 16 + i;
        const int p = b % NPEN, j = b / NPEN;
        const int k = j % NS;
        float v[8];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = 4 * g + u;
            cplx<float> x = buf[(j + r * LR) * SW + p];
            if (inv) x.y = -x.y;
            if constexpr (NS > 1) {
                if (r > 0) x = cmul(x, tw[k * r * TWS]);
            }
            v[2 * u] = x.x;
            v[2 * u + 1] = x.y;
        }
        if constexpr (PREC == PREC_FP16X3) {
            float am = 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) am = fmaxf(am, fabsf(v[e]));
            am = fmaxf(am, __shfl_xor(am, 16));
            am = fmaxf(am, __shfl_xor(am, 32));
            // am < 2^ex with ex the frexp exponent; scale 2^(15-ex), clamped to normal powers of two
            const int ex = (int)((__float_as_uint(am) >> 23) & 255u) - 126;
            int se = 15 - ex;
            se = se > 126 ? 126 : (se < -126 ? -126 : se);
            const float sc = __uint_as_float((uint32_t)(se + 127) << 23);
            unsc[qq] = __uint_as_float((uint32_t)(127 - se) << 23);
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] *= sc;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            typename O::E h, l;
            split<PREC>(v[e], h, l);
            ahi[qq][e] = h;
            alo[qq][e] = l;
        }
    }
    __syncthreads();
    const V* bt = b_table<PREC>();
    Bmat<PREC> bm;
    bm.hi[0] = bt[0 * 64 + lane];
    bm.hi[1] = bt[1 * 64 + lane];
    bm.lo[0] = bt[2 * 64 + lane];
    bm.lo[1] = bt[3 * 64 + lane];
    float* fb = reinterpret_cast<float*>(buf);
    const int col = lane & 15, cp = col & 1, m0 = col >> 1;
#pragma unroll
    for (int qq = 0; qq < QPW; ++qq) {
        const int q = wave + qq * NW;
        if ((NQ % NW) != 0 && q >= NQ) continue;
        f32x4 d[2];
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = O::mma(alo[qq], bm.hi[jb], c);
            c = O::mma(ahi[qq], bm.lo[jb], c);
            d[jb] = O::mma(ahi[qq], bm.hi[jb], c);
        }
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int ib = 4 * g + rr;
            float us = 1.0f;
            if constexpr (PREC == PREC_FP16X3) us = __shfl(unsc[qq], ib);
            const int b = q * 16 + ib;
            const int p = b % NPEN, j = b / NPEN;
            const int k = j % NS;
            const int base = (j - k) * 16 + k;
#pragma unroll
            for (int jb = 0; jb < 2; ++jb) {
                const int m = m0 + 8 * jb;
                float val = d[jb][rr] * us;
                if (inv && cp) val = -val;
                fb[2 * ((base + m * NS) * SW + p) + cp] = val;
            }
        }
    }
}

template <typename R, int L, int TP, int REST, int S, bool FIRST_IN_BUF, class GLD, class LLD, class LST>
__device__ __forceinline__ void valu_stages(const GLD& gld, const LLD& lld, const LST& lst,
                                            const cplx<R>* __restrict__ tw, int t, int sign) {
    constexpr int NSTR = pen::nstages(REST);
    if constexpr (S < NSTR) {
        constexpr int RAD = pen::radix(REST, S), NS = pen::nsub(REST, S);
        if constexpr (S == 0) pen::stage_rn<R, L, TP, RAD, NS, FIRST_IN_BUF>(gld, lst, tw, t, sign);
        else pen::stage_rn<R, L, TP, RAD, NS, true>(lld, lst, tw, t, sign);
        __syncthreads();
        valu_stages<R, L, TP, REST, S + 1, FIRST_IN_BUF>(gld, lld, lst, tw, t, sign);
    }
}

template <int PREC, int L, int REST, int S, int SW, int NPEN, int NT>
__device__ __forceinline__ void mfma_stages(cplx<float>* buf, const cplx<float>* __restrict__ tw, int tid, bool inv) {
    if constexpr (S < n16(L)) {
        r16_stage<PREC, L, REST * pow16(S), SW, NPEN, NT>(buf, tw, tid, inv);
        __syncthreads();
        mfma_stages<PREC, L, REST, S + 1, SW, NPEN, NT>(buf, tw, tid, inv);
    }
}

// Full transform of the tile's pencils: the VALU stages of the non-16 factors
// (or a copy into LDS), the radix-16 stages on MFMA, then a copy out through
// gst.  IN_BUF: the input already sits in buf (gld reads it).  OUT_BUF: the
// caller wants the result in buf itself (gst is not called).  Same calling
// convention as pen::transform: t is the thread's lane within its pencil,
// tid its workgroup thread index; no trailing barrier after the copy out.
template <int PREC, int L, int TP, int SW, int NPEN, int NT, bool IN_BUF, bool OUT_BUF, class GLD, class GST,
          class LLD, class LST>
__device__ __forceinline__ void transform(const GLD& gld, const GST& gst, const LLD& lld, const LST& lst,
                                          cplx<float>* buf, const cplx<float>* __restrict__ tw, int tid, int t,
                                          int sign) {
    static_assert(PREC != PREC_VALU && n16(L) > 0, "MFMA transform needs a radix-16 factor");
    constexpr int REST = rest_of(L);
    if (tid < 64) {
        // read only after the barrier that ends the first MFMA stage's loads
        const Bmat<PREC> bm = dft16_operand<PREC>(tid);
        typename Op<PREC>::V* bt = b_table<PREC>();
        bt[0 * 64 + tid] = bm.hi[0];
        bt[1 * 64 + tid] = bm.hi[1];
        bt[2 * 64 + tid] = bm.lo[0];
        bt[3 * 64 + tid] = bm.lo[1];
    }
    if constexpr (REST > 1) {
        valu_stages<float, L, TP, REST, 0, IN_BUF>(gld, lld, lst, tw, t, sign);
    } else if constexpr (!IN_BUF) {
        for (int n = t; n < L; n += TP) lst(n, gld(n));
        __syncthreads();
    }
    mfma_stages<PREC, L, REST, 0, SW, NPEN, NT>(buf, tw, tid, sign > 0);
    if constexpr (!OUT_BUF) {
        for (int n = t; n < L; n += TP) gst(n, lld(n));
    }
}

}  // namespace mx
}  // namespace ffcz
