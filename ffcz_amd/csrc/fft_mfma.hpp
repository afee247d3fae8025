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
// read from LDS exactly once.  The same registers are the lane's slices of
// B^T (as the first operand) and A^T (as the second), so the instruction
// computes Y^T with lane l receiving two whole outputs of butterfly l%16,
// stored to LDS as complex values (16 lanes write 128 contiguous bytes).
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

// One radix-16 Stockham stage (sub-transform length NS before the stage) over
// the tile's NPEN pencils of length L, run by NT threads.  The caller
// barriers before (inputs visible) and after (outputs visible).
// ld(n, p) reads element n of pencil p (from LDS or straight from HBM),
// st(n, p, v) stores output n of pencil p.  The 16 butterflies of an MFMA
// batch are 16 adjacent pencils at one position (ROWS false: pencils are the
// tile's columns, adjacent in memory) or 16 adjacent positions of one pencil
// (ROWS true: pencils are rows, contiguous in memory).  The stage barriers
// between its loads and its stores, so ld and st may address the same buffer.
// TWO: ld is a pair {pre, fin}: pre(n, p) issues the element's loads and
// returns them raw, fin(n, p, raw) does its element-wise work; every pre of
// the thread runs before the first fin (the work's stores would otherwise hold
// later loads back).
template <int PREC, int L, int NS, int NPEN, int NT, bool ROWS, bool TWO = false, class LD, class ST>
__device__ __forceinline__ void r16_stage(const LD& ld, const ST& st, const cplx<float>* __restrict__ tw, int tid,
                                          bool inv) {
    using O = Op<PREC>;
    using V = typename O::V;
    constexpr int LR = L / 16;
    constexpr int TWS = L / (NS * 16);
    constexpr int NBF = NPEN * LR;
    static_assert(NBF % 16 == 0, "butterflies must fill whole MFMA batches");
    static_assert(!ROWS || LR % 16 == 0, "row batches need whole groups of 16 butterflies per row");
    static_assert(NT % 64 == 0, "MFMA stages need whole waves");
    constexpr int NQ = NBF / 16;
    constexpr int NW = NT / 64;
    constexpr int QPW = (NQ + NW - 1) / NW;
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    V ahi[QPW], alo[QPW];
    float unsc[QPW];
    auto raw_of = [&]() {
        if constexpr (TWO) {
            using RawT = decltype(ld.pre(0, 0));
            struct Arr {
                RawT w[QPW][4];
            } a{};
#pragma unroll
            for (int qq = 0; qq < QPW; ++qq) {
                const int q = wave + qq * NW;
                if ((NQ % NW) != 0 && q >= NQ) continue;
                const int b = q * 16 + i;
                const int p = ROWS ? b / LR : b % NPEN, j = ROWS ? b % LR : b / NPEN;
#pragma unroll
                for (int u = 0; u < 4; ++u) a.w[qq][u] = ld.pre(j + (4 * g + u) * LR, p);
            }
            return a;
        } else {
            return 0;
        }
    };
    const auto raw = raw_of();
#pragma unroll
    for (int qq = 0; qq < QPW; ++qq) {
        const int q = wave + qq * NW;
        unsc[qq] = 1.0f;
        if ((NQ % NW) != 0 && q >= NQ) continue;  // uniform per wave
        const int b = q * 16 + i;
        const int p = ROWS ? b / LR : b % NPEN, j = ROWS ? b % LR : b / NPEN;
        const int k = j % NS;
        float v[8];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = 4 * g + u;
            cplx<float> x;
            if constexpr (TWO) x = ld.fin(j + r * LR, p, raw.w[qq][u]);
            else x = ld(j + r * LR, p);
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
#pragma unroll
    for (int qq = 0; qq < QPW; ++qq) {
        const int q = wave + qq * NW;
        if ((NQ % NW) != 0 && q >= NQ) continue;
        // Y^T = B^T A^T: the lane's DFT slice is its slice of the first
        // operand B^T and its butterfly slice that of A^T, so the product's
        // lane layout gives lane l the outputs 2(l/16), 2(l/16)+1 (+8 jb) of
        // its own butterfly l%16 as whole complex values (D^T rows 4(l/16)..+3
        // = Re, Im, Re, Im)
        f32x4 d[2];
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = O::mma(bm.hi[jb], alo[qq], c);
            c = O::mma(bm.lo[jb], ahi[qq], c);
            d[jb] = O::mma(bm.hi[jb], ahi[qq], c);
        }
        const float us = PREC == PREC_FP16X3 ? unsc[qq] : 1.0f;  // this lane's butterfly's scale
        const int b = q * 16 + i;
        const int p = ROWS ? b / LR : b % NPEN, j = ROWS ? b % LR : b / NPEN;
        const int k = j % NS;
        const int base = (j - k) * 16 + k;
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int m = 2 * g + h + 8 * jb;
                cplx<float> v = mk(d[jb][2 * h] * us, d[jb][2 * h + 1] * us);
                if (inv) v.y = -v.y;
                st(base + m * NS, p, v);
            }
        }
    }
}

// The stages after the radix-16 ones: Stockham stages of the remaining
// factors REST on the vector ALU, with sub-transform lengths NSM (the product
// of the radix-16 stages before them) times their own.  The last one stores
// through gst (GST_LDS: gst itself writes buf, so the stage barriers between
// its reads and its writes); OUT_BUF: the last one writes buf.
template <int L, int TP, int REST, int NSM, int S, bool OUT_BUF, bool GST_LDS, class GST, class LLD, class LST>
__device__ __forceinline__ void valu_tail(const GST& gst, const LLD& lld, const LST& lst,
                                          const cplx<float>* __restrict__ tw, int t, int sign) {
    constexpr int NSTR = pen::nstages(REST);
    if constexpr (S < NSTR) {
        constexpr int RAD = pen::radix(REST, S), NS = NSM * pen::nsub(REST, S);
        if constexpr (S == NSTR - 1) {
            if constexpr (OUT_BUF) pen::stage_rn<float, L, TP, RAD, NS, true>(lld, lst, tw, t, sign);
            else pen::stage_rn<float, L, TP, RAD, NS, GST_LDS>(lld, gst, tw, t, sign);
        } else {
            pen::stage_rn<float, L, TP, RAD, NS, true>(lld, lst, tw, t, sign);
            __syncthreads();
            valu_tail<L, TP, REST, NSM, S + 1, OUT_BUF, GST_LDS>(gst, lld, lst, tw, t, sign);
        }
    }
}

template <int PREC, int L, int S, int SW, int NPEN, int NT, bool IN_BUF, bool OUT_BUF, class GLDP, class GSTP>
__device__ __forceinline__ void mfma_stages(cplx<float>* buf, const cplx<float>* __restrict__ tw, int tid, bool inv,
                                            const GLDP& gldp, const GSTP& gstp) {
    constexpr int NM = n16(L);
    if constexpr (S < NM) {
        constexpr bool LAST = S == NM - 1 && rest_of(L) == 1;
        auto bld = [&](int n, int p) { return buf[n * SW + p]; };
        auto bst = [&](int n, int p, cplx<float> v) { buf[n * SW + p] = v; };
        constexpr bool FROM_G = S == 0 && !IN_BUF, TO_G = LAST && !OUT_BUF;
        if constexpr (FROM_G && TO_G) r16_stage<PREC, L, pow16(S), NPEN, NT, false>(gldp, gstp, tw, tid, inv);
        else if constexpr (FROM_G) r16_stage<PREC, L, pow16(S), NPEN, NT, false>(gldp, bst, tw, tid, inv);
        else if constexpr (TO_G) r16_stage<PREC, L, pow16(S), NPEN, NT, false>(bld, gstp, tw, tid, inv);
        else r16_stage<PREC, L, pow16(S), NPEN, NT, false>(bld, bst, tw, tid, inv);
        if constexpr (!LAST) __syncthreads();
        mfma_stages<PREC, L, S + 1, SW, NPEN, NT, IN_BUF, OUT_BUF>(buf, tw, tid, inv, gldp, gstp);
    }
}

// Full transform of the tile's NPEN pencils buf[n*SW + p] by NT threads: the
// radix-16 stages first on the matrix cores (the first one loads straight
// from HBM through gldp(n, p) unless IN_BUF; a radix-16 last stage stores
// through gstp(n, p, v) unless OUT_BUF), then the vector-ALU stages of the
// remaining factors by thread t of the TP threads of pencil pth, the last of
// them storing through gstp(n, pth, v) (or into buf when OUT_BUF).  GST_LDS:
// gstp writes buf itself.  No trailing barrier, like pen::transform.
// TABLE: write the DFT-matrix table first (false: an earlier transform of
// the same kernel left it in LDS).
template <int PREC, int L, int TP, int SW, int NPEN, int NT, bool IN_BUF, bool OUT_BUF, bool GST_LDS,
          bool TABLE = true, class GLDP, class GSTP>
__device__ __forceinline__ void transform(const GLDP& gldp, const GSTP& gstp, cplx<float>* buf,
                                          const cplx<float>* __restrict__ tw, int tid, int pth, int t, int sign) {
    static_assert(PREC != PREC_VALU && n16(L) > 0, "MFMA transform needs a radix-16 factor");
    constexpr int REST = rest_of(L);
    if (TABLE && tid < 64) {
        // read by the first stage after its loads' barrier
        const Bmat<PREC> bm = dft16_operand<PREC>(tid);
        typename Op<PREC>::V* bt = b_table<PREC>();
        bt[0 * 64 + tid] = bm.hi[0];
        bt[1 * 64 + tid] = bm.hi[1];
        bt[2 * 64 + tid] = bm.lo[0];
        bt[3 * 64 + tid] = bm.lo[1];
    }
    mfma_stages<PREC, L, 0, SW, NPEN, NT, IN_BUF, OUT_BUF>(buf, tw, tid, sign > 0, gldp, gstp);
    if constexpr (REST > 1) {
        auto lld = [&](int n) { return buf[n * SW + pth]; };
        auto lst = [&](int n, cplx<float> v) { buf[n * SW + pth] = v; };
        auto gst = [&](int n, cplx<float> v) { gstp(n, pth, v); };
        valu_tail<L, TP, REST, pow16(n16(L)), 0, OUT_BUF, GST_LDS>(gst, lld, lst, tw, t, sign);
    }
}

// The transform of length 256 = 16 x 16 over the rows of a tile: both stages
// on the matrix cores with row batches (16 adjacent positions of a row per
// MFMA batch), the exchange between them in LDS in the row-major layout
// rows_idx (row stride RS, room for X[0..LZ]).  Within each group of 16
// positions the slot is XOR-swizzled by the group index, so the first
// stage's stores (16j + m over lanes j), the second stage's loads and stores
// (j + 16r over lanes j) and the R2C/C2R passes over adjacent positions all
// fall on distinct LDS banks, without pad slots.  ld0 feeds the first stage
// (straight from HBM, or from buf in that layout); st1 takes the second
// stage's outputs (to HBM, or into buf).  No trailing barrier.
template <int LZ>
__host__ __device__ constexpr int rows_stride() { return LZ + 1; }
__device__ __forceinline__ int rows_idx(int n, int p, int rs) { return p * rs + (n ^ ((n >> 4) & 15)); }

// TWO0: ld0 is a {pre, fin} pair (r16_stage's TWO).
template <int PREC, int LZ, int NPEN, int NT, bool TWO0 = false, class LD0, class ST1>
__device__ __forceinline__ void rows_transform(const LD0& ld0, const ST1& st1, cplx<float>* buf,
                                               const cplx<float>* __restrict__ tw, int tid, int sign) {
    static_assert(PREC != PREC_VALU && n16(LZ) == 2 && rest_of(LZ) == 1, "rows_transform is the 16 x 16 transform");
    constexpr int RS = rows_stride<LZ>();
    if (tid < 64) {
        const Bmat<PREC> bm = dft16_operand<PREC>(tid);
        typename Op<PREC>::V* bt = b_table<PREC>();
        bt[0 * 64 + tid] = bm.hi[0];
        bt[1 * 64 + tid] = bm.hi[1];
        bt[2 * 64 + tid] = bm.lo[0];
        bt[3 * 64 + tid] = bm.lo[1];
    }
    auto bld = [&](int n, int p) { return buf[rows_idx(n, p, RS)]; };
    auto bst = [&](int n, int p, cplx<float> v) { buf[rows_idx(n, p, RS)] = v; };
    r16_stage<PREC, LZ, 1, NPEN, NT, true, TWO0>(ld0, bst, tw, tid, sign > 0);
    __syncthreads();
    r16_stage<PREC, LZ, 16, NPEN, NT, true>(bld, st1, tw, tid, sign > 0);
}

}  // namespace mx
}  // namespace ffcz
