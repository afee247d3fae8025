// Length-specialised z / y / x pass kernels built on the pencil engine.
// Included by the per-precision translation units (fft_fast_*.hip), which
// instantiate the dispatchers for their (R, T) pair and pass modes.
#pragma once

#include <type_traits>

#include "fft_fast.hpp"
#include "fft_mfma.hpp"

namespace ffcz {

// Supported lengths: y/x axes, and half lengths nz/2 of the z axis.  They
// cover the 2/3/5-smooth axes of the benchmark configurations
// {60,72,96,120,150,180,240,250,256,300,512,600} and powers of two.
#define FFCZ_FAST_S_LENGTHS(X) \
    X(1) X(8) X(16) X(32) X(60) X(64) X(72) X(96) X(120) X(128) X(150) X(180) X(240) X(250) X(256) X(300) X(512) X(600)
#define FFCZ_FAST_Z_LENGTHS(X) \
    X(4) X(8) X(16) X(30) X(32) X(36) X(48) X(60) X(64) X(75) X(90) X(120) X(125) X(128) X(150) X(256) X(300)

// Tile variants (chosen per pass by the autotuner, fft3d.hpp TileVars):
//   y/x passes: V=0 columns of 128 bytes per tile, V=1 64 bytes (half the LDS),
//               V=2 128-byte tiles with the length's radices in reversed order
//               (vector-ALU iteration passes of lengths whose order is not a
//               palindrome; a different rounding, not bit-identical to V=0)
//   z passes:   V=0 up to 256 threads per workgroup, V=1 up to 128
template <typename R>
constexpr int s_width(int V = 0) { return (sizeof(R) == 4 ? 16 : 8) >> (V & 1); }
constexpr bool s_rev(int V) { return (V & 2) != 0; }
constexpr int z_tp(int LZ) { return pen::threads_per_pencil(LZ); }
// threads per pencil of the y/x kernels: FP64 pencils keep ~8 elements per
// thread at every length (more threads in flight for the long FP64 axes)
template <typename R>
constexpr int s_tp(int L) { return sizeof(R) == 8 ? (L + 7) / 8 : pen::threads_per_pencil(L); }
constexpr int z_rows(int LZ, int V = 0) {
    int p = 1;
    while (p * 2 * z_tp(LZ) <= (V ? 256 : 512)) p *= 2;
    return p;
}
// The z tile variant of the batched launches: the narrower tiles (variant 1)
// for lengths that are not powers of two, whose kernels need more than 64
// registers, so that more workgroups share a CU (batched 48-field step
// 11.3 -> 10.7 ms, most of it in the FP64 certification's z passes).
constexpr int zb_var(int LZ) { return (LZ & (LZ - 1)) != 0 ? 1 : 0; }

// The vector-ALU z tiles are row-major with a pad slot after every 8
// elements: a radix-8 stage's stores (positions 8t + r of thread t) and
// loads (t + L/8 r) then fall on distinct LDS banks across a row's threads.
// The stride holds X[0..LZ] (the C2R input's Nyquist bin included).
constexpr int zrows8_stride(int LZ) { return LZ + LZ / 8 + 1; }
__device__ __forceinline__ int zrows8_idx(int n, int p, int rs) { return p * rs + n + (n >> 3); }

// Element o of a tile that starts at base.  The byte offset is formed in
// 32 bits (a tile spans far less than 4 GiB), so the access is a uniform
// base plus a 32-bit per-lane offset instead of 64-bit address arithmetic
// per element.
template <typename E>
__device__ __forceinline__ E& tile_at(E* base, uint32_t o) {
    return *reinterpret_cast<E*>(reinterpret_cast<char*>(base) + o * (uint32_t)sizeof(E));
}
template <typename E>
__device__ __forceinline__ const E& tile_at(const E* base, uint32_t o) {
    return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(base) + o * (uint32_t)sizeof(E));
}

// W_nz^k for k <= M = nz/2 from the stored half k <= M/2:
// W^(M-k) = -conj(W^k) (the host table is mirrored the same way)
template <typename R, int M>
__device__ __forceinline__ cplx<R> half_twiddle(const cplx<R>* tw2, int k) {
    if (k <= M / 2) return tw2[k];
    const cplx<R> w = tw2[M - k];
    return mk(-w.x, w.y);
}

// the z passes of half length 256 in the MFMA variants: both radix-16 stages
// on row batches over a row-major tile (mx::rows_transform)
template <int PREC, int LZ>
constexpr bool ZROWS = PREC != PREC_VALU && mx::n16(LZ) == 2 && mx::rest_of(LZ) == 1;

// ---------------------------------------------------------------------------
// forward z (R2C, nz even): loads fuse the deferred spatial clamp (ITER) or
// the FP64 reconstruction and spatial check (CERT); the half-length complex
// FFT ends in LDS and the R2C split X[k] = A[k] + W_nz^k B[k] writes the
// padded half spectrum.
// ---------------------------------------------------------------------------
template <typename R, typename T, int LZ, int MODE, int PREC, int V>
__device__ __forceinline__ void zfwd_body(const ZArgs<R, T>& a, uint32_t blk) {
    constexpr int TP = z_tp(LZ), P = z_rows(LZ, V), NT = TP * P;
    constexpr int NZ = 2 * LZ, M = LZ, NH = M + 1, SN = P + 1;
    // tile layouts: ZL_ROWS8 (vector ALU), mx rows (the 16 x 16 MFMA
    // transform), columns n*SN + p (other MFMA lengths)
    constexpr bool ROWS = ZROWS<PREC, LZ>, COLS = PREC != PREC_VALU && !ROWS;
    constexpr int RS = ROWS ? mx::rows_stride<LZ>() : zrows8_stride(LZ);
    __shared__ cplx<R> tw[LZ];
    __shared__ cplx<R> tw2[M / 2 + 1];
    __shared__ cplx<R> buf[COLS ? LZ * SN : P * RS];
    auto bidx = [&](int n, int pp) {
        return COLS ? n * SN + pp : ROWS ? mx::rows_idx(n, pp, RS) : zrows8_idx(n, pp, RS);
    };
    auto tw2v = [&](int k) { return half_twiddle<R, M>(tw2, k); };
    if (MODE == Z_ITER && a.st->done) return;
    const int tid = threadIdx.x;
    for (int i = tid; i < LZ; i += NT) tw[i] = a.tz[i];
    for (int i = tid; i <= M / 2; i += NT) tw2[i] = a.tz2[i];
    __syncthreads();
    const int t = tid % TP, p = tid / TP;
    const size_t row0 = (size_t)blk * P;
    const bool rv = row0 + p < a.rows;
    // the tile's first element: the per-tile bases are uniform, the offsets
    // of the tile's elements from them 32-bit (tile_at)
    const size_t t0 = row0 * (size_t)NZ;
    const bool clamp = (MODE == Z_ITER) && a.st->iter > 0;
    const bool fresh = clamp && a.ds_fresh && a.st->iter == 1;
    // the first iteration reads the initial error itself (e is written by the inverse pass)
    const R* esrc = (MODE == Z_ITER && !clamp && a.e_init) ? a.e_init : a.e;
    R* ecopy = (MODE == Z_ITER && !clamp && a.e_init) ? a.e_copy : nullptr;
    unsigned long long nviol = 0;
    double vmax = 0.0;
    // The loads of element pair n of tile row pp (ok: the row is in range)
    // and the pass's fused element-wise work on them, kept apart so a thread
    // can issue all of its first-stage loads before any of that work (its
    // stores to ds / e0 could alias later loads for all the compiler knows,
    // and its branches would keep each load waiting for the previous element).
    struct Raw {
        cplx<T> e, s;      // iterate (ITER: e, in s's place unused) / e0 and ds (CERT)
        cplx<double> f;    // C2R(df)/N (CERT with frequency edits)
    };
    auto rawr = [&](int n, int pp, bool ok) -> Raw {
        Raw w{};
        if (!ok) return w;
        const uint32_t o2 = (uint32_t)(pp * LZ + n);  // element pairs as vectors
        if constexpr (MODE == Z_ITER) {
            const cplx<R> v = tile_at(reinterpret_cast<const cplx<R>*>(esrc + t0), o2);
            w.e = mk((T)v.x, (T)v.y);
        } else if constexpr (MODE == Z_PLAIN) {
            const cplx<R> v = tile_at(reinterpret_cast<const cplx<R>*>(a.in + t0), o2);
            w.f = mk((double)v.x, (double)v.y);
        } else {
            w.e = tile_at(reinterpret_cast<const cplx<T>*>(a.e0 + t0), o2);
            w.s = tile_at(reinterpret_cast<const cplx<T>*>(a.cds + t0), o2);
            if (a.dfi) w.f = tile_at(reinterpret_cast<const cplx<double>*>(a.dfi + t0), o2);
        }
        return w;
    };
    auto applyr = [&](int n, int pp, bool ok, const Raw& w) -> cplx<R> {
        if (!ok) return mk(R(0), R(0));
        const uint32_t o2 = (uint32_t)(pp * LZ + n);  // pair index within the tile
        if constexpr (MODE == Z_ITER) {
            cplx<R> v = mk((R)w.e.x, (R)w.e.y);
            if (ecopy) tile_at(reinterpret_cast<cplx<R>*>(ecopy + t0), o2) = v;
            if (fresh) {
                cplx<R> d;
                d.x = clamp_edit(v.x, a.eps);
                d.y = clamp_edit(v.y, a.eps);
                tile_at(reinterpret_cast<cplx<R>*>(a.ds + t0), o2) = d;
            } else if (clamp) {
                R* dsp = &tile_at(reinterpret_cast<cplx<R>*>(a.ds + t0), o2).x;
                clamp_acc(v.x, a.eps, dsp);
                clamp_acc(v.y, a.eps, dsp + 1);
            }
            return v;
        } else if constexpr (MODE == Z_PLAIN) {
            return mk((R)w.f.x, (R)w.f.y);
        } else {
            // x_hat - x = e0 + C2R(df)/N + ds, element pairs as vectors
            double d0 = (double)w.e.x, d1 = (double)w.e.y;
            if (a.dfi) {
                d0 += w.f.x;
                d1 += w.f.y;
            }
            d0 += dequant(w.s.x, a.qs);
            d1 += dequant(w.s.y, a.qs);
            const double x0 = fabs(d0) - a.ceps, x1 = fabs(d1) - a.ceps;
            nviol += (x0 > 0.0 ? 1u : 0u) + (x1 > 0.0 ? 1u : 0u);
            vmax = fmax(vmax, fmax(x0, x1));
            return mk((R)d0, (R)d1);
        }
    };
    auto gld = [&](int n) { return applyr(n, p, rv, rawr(n, p, rv)); };
    auto lld = [&](int n) { return buf[bidx(n, p)]; };
    auto lst = [&](int n, cplx<R> v) { buf[bidx(n, p)] = v; };
    if constexpr (PREC == PREC_VALU && MODE == Z_PLAIN) {
        pen::transform<R, LZ, TP, false, true>(gld, lst, lld, lst, tw, t, -1);
    } else if constexpr (PREC == PREC_VALU) {
        // every first-stage load of the thread first, then the fused work as
        // the stage consumes them
        constexpr int NB0 = pen::first_nb<LZ, TP>(), R0 = pen::radix(LZ, 0);
        Raw raw[NB0][R0];
#pragma unroll
        for (int b = 0; b < NB0; ++b) {
            const bool has = pen::first_has<LZ, TP>(t, b);
#pragma unroll
            for (int r = 0; r < R0; ++r) raw[b][r] = rawr(pen::first_elem<LZ>(t, b, r, TP), p, rv && has);
        }
        auto gld_br = [&](int n, int b, int r) { return applyr(n, p, rv, raw[b][r]); };
        pen::transform<R, LZ, TP, false, true, true>(gld_br, lst, lld, lst, tw, t, -1);
    } else if constexpr (ROWS) {
        // the first stage loads 16 adjacent pairs of a row per MFMA batch,
        // all of the thread's loads before the fused work
        auto pre = [&](int n, int pp) { return rawr(n, pp, row0 + pp < a.rows); };
        auto fin = [&](int n, int pp, const Raw& w) { return applyr(n, pp, row0 + pp < a.rows, w); };
        struct {
            decltype(pre)& pre;
            decltype(fin)& fin;
        } ld0{pre, fin};
        auto bst = [&](int n, int pp, cplx<float> v) { buf[mx::rows_idx(n, pp, RS)] = v; };
        mx::rows_transform<PREC, LZ, P, NT, true>(ld0, bst, buf, tw, tid, -1);
    } else {
        // rows are contiguous along n: a coalesced copy into the tile first
        for (int n = t; n < LZ; n += TP) lst(n, gld(n));
        __syncthreads();
        auto none_ld = [](int, int) { return mk(0.0f, 0.0f); };
        auto none_st = [](int, int, cplx<float>) {};
        mx::transform<PREC, LZ, TP, SN, P, NT, true, true, false>(none_ld, none_st, buf, tw, tid, p, t, -1);
    }
    if constexpr (MODE == Z_CERT) {
        block_count_atomic(nviol, &a.stats->spatial_viol);
        block_max_atomic<double, NT>(vmax, &a.stats->spatial_max_bits);
    }
    __syncthreads();
    cplx<R>* spec_t = a.spec + row0 * a.Mp;
    const uint32_t mp = (uint32_t)a.Mp;
    for (int q = tid; q < P * NH; q += NT) {
        const int pp = q / NH;
        const int k = q - pp * NH;
        if (row0 + pp >= a.rows) continue;
        const int ik = (k == M) ? 0 : k;
        const int im = (k == 0) ? 0 : M - k;
        const cplx<R> zk = buf[bidx(ik, pp)];
        const cplx<R> zm = buf[bidx(im, pp)];
        const cplx<R> A = mk(R(0.5) * (zk.x + zm.x), R(0.5) * (zk.y - zm.y));
        const cplx<R> B = mk(R(0.5) * (zk.y + zm.y), R(0.5) * (zm.x - zk.x));
        tile_at(spec_t, (uint32_t)pp * mp + (uint32_t)k) = A + cmul(tw2v(k), B);
    }
}

// ---------------------------------------------------------------------------
// inverse z (C2R, nz even): the tile's rows X[0..M] are staged in LDS with
// coalesced loads (each element read from HBM once), turned in place into
// Z[n] = (X[n] + conj X[M-n]) + i W^-n (X[n] - conj X[M-n]) (imaginary parts
// of the DC and Nyquist bins dropped), transformed, and the last stage writes
// the real row scaled by 1/N; ITER keeps the workgroup's spatial violation.
// ---------------------------------------------------------------------------
template <typename R, typename T, int LZ, int MODE, int PREC, int V>
__device__ __forceinline__ void zinv_body(const ZArgs<R, T>& a, uint32_t blk) {
    constexpr int TP = z_tp(LZ), P = z_rows(LZ, V), NT = TP * P;
    constexpr int NZ = 2 * LZ, M = LZ, NH = M + 1, SN = P + 1, NPAIR = M / 2 + 1;
    constexpr bool ROWS = ZROWS<PREC, LZ>, COLS = PREC != PREC_VALU && !ROWS;
    // the row strides hold X[0..M] with the pad slots
    constexpr int RS = ROWS ? mx::rows_stride<LZ>() : zrows8_stride(LZ);
    __shared__ cplx<R> tw[LZ];
    __shared__ cplx<R> tw2[M / 2 + 1];
    __shared__ cplx<R> buf[COLS ? NH * SN : P * RS];
    auto bidx = [&](int n, int pp) {
        return COLS ? n * SN + pp : ROWS ? mx::rows_idx(n, pp, RS) : zrows8_idx(n, pp, RS);
    };
    auto tw2v = [&](int k) { return half_twiddle<R, M>(tw2, k); };
    if (MODE == Z_ITER && a.st->done) return;
    const int tid = threadIdx.x;
    for (int i = tid; i < LZ; i += NT) tw[i] = a.tz[i];
    for (int i = tid; i <= M / 2; i += NT) tw2[i] = a.tz2[i];
    const size_t row0 = (size_t)blk * P;
    // the tile's half-spectrum rows X[0..M] into LDS, each element loaded
    // once: every load of the thread is issued before the first LDS store
    // (a load-store loop would keep one load in flight per thread)
    {
        constexpr int NQ = (P * NH + NT - 1) / NT;
        const cplx<R>* spec_t = a.spec + row0 * a.Mp;
        const uint32_t mp = (uint32_t)a.Mp;
        cplx<R> xv[NQ];
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            const int q = tid + u * NT;
            const int pp = q / NH;
            const int k = q - pp * NH;
            xv[u] = (q < P * NH && row0 + pp < a.rows) ? tile_at(spec_t, (uint32_t)pp * mp + (uint32_t)k)
                                                       : mk(R(0), R(0));
        }
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            const int q = tid + u * NT;
            const int pp = q / NH;
            if (q < P * NH) buf[bidx(q - pp * NH, pp)] = xv[u];
        }
    }
    __syncthreads();
    // Z[n] = (X[n] + conj X[M-n]) + i W^-n (X[n] - conj X[M-n]) in place, by
    // pairs (n, M-n); the imaginary parts of the DC and Nyquist bins are dropped
    auto zpre = [&](cplx<R> xk, cplx<R> xm, int n) {
        const cplx<R> s = xk + conjg(xm);
        const cplx<R> d = cmul(xk - conjg(xm), conjg(tw2v(n)));
        return mk(s.x - d.y, s.y + d.x);
    };
    for (int q = tid; q < P * NPAIR; q += NT) {
        const int pp = q / NPAIR;
        const int n = q - pp * NPAIR;
        const int m = M - n;
        cplx<R> xn = buf[bidx(n, pp)];
        cplx<R> xm = buf[bidx(m, pp)];
        if (n == 0) {
            xn.y = R(0);
            xm.y = R(0);
        }
        buf[bidx(n, pp)] = zpre(xn, xm, n);
        if (m != n && m < M) buf[bidx(m, pp)] = zpre(xm, xn, m);
    }
    __syncthreads();
    const int t = tid % TP, p = tid / TP;
    const bool rv = row0 + p < a.rows;
    // the tile's first element: uniform bases, 32-bit offsets (tile_at)
    const size_t t0 = row0 * (size_t)NZ;
    R vm = R(0);
    // output pair n of tile row pp (ok: the row is in range); ITER keeps the violation
    auto gstr = [&](int n, int pp, bool ok, cplx<R> z) {
        if (!ok) return;
        const uint32_t o2 = (uint32_t)(pp * LZ + n);
        const R v0 = z.x * a.scale, v1 = z.y * a.scale;
        if constexpr (MODE == Z_ITER) {
            tile_at(reinterpret_cast<cplx<R>*>(a.e + t0), o2) = mk(v0, v1);
            // max(|v| - eps, 0) without branches: vm >= 0 absorbs the
            // non-positive excesses and fmax ignores NaN, as the test did
            vm = fmax(vm, fmax(fabs(v0) - a.eps, fabs(v1) - a.eps));
        } else {
            tile_at(reinterpret_cast<cplx<R>*>(a.out + t0), o2) = mk(v0, v1);
        }
    };
    auto gst = [&](int n, cplx<R> z) { gstr(n, p, rv, z); };
    auto lld = [&](int n) { return buf[bidx(n, p)]; };
    auto lst = [&](int n, cplx<R> v) { buf[bidx(n, p)] = v; };
    if constexpr (PREC == PREC_VALU) {
        pen::transform<R, LZ, TP, true, false>(lld, gst, lld, lst, tw, t, +1);
    } else if constexpr (ROWS) {
        // the second stage stores 16 adjacent pairs of a row per MFMA batch
        auto bld = [&](int n, int pp) { return buf[mx::rows_idx(n, pp, RS)]; };
        auto gstp = [&](int n, int pp, cplx<float> z) { gstr(n, pp, row0 + pp < a.rows, z); };
        mx::rows_transform<PREC, LZ, P, NT>(bld, gstp, buf, tw, tid, +1);
    } else {
        // a vector-ALU last stage stores the thread's own row (coalesced); a
        // radix-16 last stage leaves the tile in LDS for a coalesced copy out
        constexpr bool COPY_OUT = mx::rest_of(LZ) == 1;
        auto none_ld = [](int, int) { return mk(0.0f, 0.0f); };
        auto gstp = [&](int n, int, cplx<float> z) { gst(n, z); };
        mx::transform<PREC, LZ, TP, SN, P, NT, true, COPY_OUT, false>(none_ld, gstp, buf, tw, tid, p, t, +1);
        if constexpr (COPY_OUT) {
            __syncthreads();
            for (int n = t; n < LZ; n += TP) gst(n, lld(n));
        }
    }
    if constexpr (MODE == Z_ITER) block_max_store<R, NT>(vm, &a.vpart[a.row_off / P + blk]);
}

// ---------------------------------------------------------------------------
// The iteration's z pass: inverse z of iteration k, its spatial violation,
// the spatial clamp of iteration k and the forward z of iteration k + 1, in
// place on the tile's half-spectrum rows (the iterate e never goes to HBM).
//
// The loop stops before the spatial projection of a converged iteration
// (projection_algorithm_impl.cuh:845-851), which only the finalisation after
// this pass can tell, so the clamp is speculative.  A converged pass has
// every excess |e| - eps below tol (the finalisation's test), so only clamps
// with an excess below tol can belong to one.  Those are deferred: the
// spectrum gets the clamped value, but instead of adding the edit to ds the
// pass logs the element (tile-relative offset, unclamped value) in its
// workgroup's segment.  The next pass over the same tile, which runs only
// if this one did not converge, first adds the logged edits to ds (each
// element's edits still accumulate in iteration order); a converged loop
// drops the log, and one stopped by max_iter applies it (k_zz_apply).
// Clamps with a larger excess go into ds at once.
// ---------------------------------------------------------------------------
// First slot of this lane's n entries in the workgroup's log: a prefix sum
// of the counts over the wave and one LDS atomic per wave (by its last lane;
// the lanes of a workgroup's last wave may not all exist).
__device__ __forceinline__ uint32_t wave_log_base(uint32_t n, uint32_t* counter) {
    const int lane = threadIdx.x & 63;
    uint32_t x = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    const int last = 63 - __clzll((long long)__ballot(1));
    uint32_t base = 0u;
    if (lane == last) base = atomicAdd(counter, x);
    base = __shfl(base, last, 64);
    return base + x - n;
}

template <typename R, int LZ, int V>
__device__ __forceinline__ void zz_body(const ZArgs<R, R>& a, uint32_t blk) {
    constexpr int TP = z_tp(LZ), P = z_rows(LZ, V), NT = TP * P;
    constexpr int NZ = 2 * LZ, M = LZ, NH = M + 1;
    constexpr int RS = zrows8_stride(LZ);
    constexpr int NST = pen::nstages(LZ), SL = NST - 1;
    // The inverse transform runs the radices in the usual order and the
    // forward one in reversed order, so the forward's first stage reads
    // exactly the positions (pairs j + r L/R) the inverse's last stage leaves
    // in a thread's registers: the spatial clamp works there, with no LDS
    // round trip between the transforms.
    using DL = pen::StageDim<LZ, TP, SL, false>;
    using DF = pen::StageDim<LZ, TP, 0, true>;
    static_assert(DL::RAD == DF::RAD && DL::NB == DF::NB && DL::NS == DF::LR, "the transforms meet in registers");
    __shared__ cplx<R> tw[LZ];
    __shared__ cplx<R> tw2[M / 2 + 1];
    __shared__ cplx<R> buf[P * RS];
    __shared__ uint32_t nlog;
    auto bidx = [&](int n, int pp) { return zrows8_idx(n, pp, RS); };
    auto tw2v = [&](int k) { return half_twiddle<R, M>(tw2, k); };
    if (a.st->done) return;
    const int tid = threadIdx.x;
    for (int i = tid; i < LZ; i += NT) tw[i] = a.tz[i];
    for (int i = tid; i <= M / 2; i += NT) tw2[i] = a.tz2[i];
    if (tid == 0) nlog = 0u;
    const int t = tid % TP, p = tid / TP;
    const size_t row0 = (size_t)blk * P;
    const bool rv = row0 + p < a.rows;
    cplx<R>* spec_t = a.spec + row0 * a.Mp;
    const uint32_t mp = (uint32_t)a.Mp;
    const R eps = a.eps;
    const size_t t0 = row0 * (size_t)NZ;
    R* dse = a.ds + t0;
    cplx<R>* dsp = reinterpret_cast<cplx<R>*>(dse);
    const size_t gb = a.row_off / P + blk;  // the workgroup's log segment and violation slot
    const size_t seg = gb * (size_t)(P * NZ);
    const uint32_t nvalid = (uint32_t)(row0 + P <= a.rows ? P * LZ : (a.rows - row0) * LZ);  // pairs
    if (a.st->iter > 0) {
        // the previous pass's deferred clamps of this tile (it did not
        // converge); counts and offsets are held to the tile, whatever wrote
        // the log
        const uint32_t npend = min(a.log_cnt[gb], 2u * nvalid);
        for (uint32_t i = tid; i < npend; i += NT) {
            const uint32_t o = a.log_off[seg + i];
            if (o >= 2u * nvalid) continue;
            R v = a.log_v[seg + i];
            const R ed = clamp_edit(v, eps);
            dse[o] = dse[o] + ed;
        }
    }
    __syncthreads();
    // Inverse C2R along z.  The first stage's inputs Z[n] = (X[n] + conj
    // X[M-n]) + i W^-n (X[n] - conj X[M-n]) (imaginary parts of the DC and
    // Nyquist bins dropped) come straight from the row: every X[n] and
    // X[M-n] of the thread is loaded before any is used.
    constexpr int NB0 = pen::first_nb<LZ, TP>(), R0 = pen::radix(LZ, 0);
    cplx<R> xa[NB0][R0], xb[NB0][R0];
    // A row past the field's end reads the field's last row instead (no
    // per-lane zeroing or masked loads): its values go through the transforms
    // into its own tile slots, and it stores, logs and counts nothing.
    const uint32_t lro = (uint32_t)(rv ? p : (int)(a.rows - 1 - row0)) * mp;
#pragma unroll
    for (int b = 0; b < NB0; ++b) {
        const bool has = pen::first_has<LZ, TP>(t, b);
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int n = pen::first_elem<LZ>(t, b, r, TP);
            xa[b][r] = has ? tile_at(spec_t, lro + (uint32_t)n) : mk(R(0), R(0));
            xb[b][r] = has ? tile_at(spec_t, lro + (uint32_t)(M - n)) : mk(R(0), R(0));
        }
    }
    auto zload = [&](int n, int b, int r) {
        cplx<R> xn = xa[b][r], xm = xb[b][r];
        if (n == 0) {
            xn.y = R(0);
            xm.y = R(0);
        }
        const cplx<R> sm = xn + conjg(xm);
        const cplx<R> d = cmul(xn - conjg(xm), conjg(tw2v(n)));
        return mk(sm.x - d.y, sm.y + d.x);
    };
    auto lld = [&](int n) { return buf[bidx(n, p)]; };
    auto lst = [&](int n, cplx<R> v) { buf[bidx(n, p)] = v; };
    cplx<R> v[DL::NB][DL::RAD];
    if constexpr (NST == 1) {
        pen::stage_compute<R, LZ, TP, 0, false, true>(zload, v, tw, t, +1);
    } else {
        // stages 0 .. SL-1 through the tile, then the last one into registers
        pen::transform<R, LZ, TP, false, true, true, 0, false, SL>(zload, lst, lld, lst, tw, t, +1);
        pen::stage_compute<R, LZ, TP, SL, false>(lld, v, tw, t, +1);
    }
    // The spatial clamp on the thread's pairs n = j + r L/R of its row (pair
    // n is elements 2n, 2n+1: consecutive threads, consecutive pairs, so ds
    // moves coalesced).  The first clamp since the accumulators were cleared
    // stores ds whole; later ones add their edits where they fire, as atomic
    // adds that return nothing (each element has one writer per pass, so the
    // sum is the plain old + edit, and no register waits for the old value).
    const bool fresh = a.ds_fresh && a.st->iter == 0;
    R tolr;
    if constexpr (sizeof(R) == 4) tolr = a.st->tol_f;
    else tolr = (R)a.st->tol;
    R vm = R(0);
    // deferred clamps (excess below tol) of this thread, one bit per component
    // (bit 2(b RAD + r) + c), so the log below writes exactly what it counts
    using DMask = std::conditional_t<(2 * DL::NB * DL::RAD <= 32), uint32_t, uint64_t>;
    static_assert(2 * DL::NB * DL::RAD <= 64, "deferred-clamp mask too narrow");
    DMask dmask = 0;
    // Branch-free tests (bitwise, not short-circuit, so no per-element exec
    // mask branches): a component fires where |u| > eps (u > eps or
    // u < -eps) at a valid position.  Butterflies beyond the stage's count
    // (j >= LR) hold no values; a row past the field's end is dropped from vm
    // once, after the loop.
#pragma unroll
    for (int b = 0; b < DL::NB; ++b) {
        const int j = t + b * TP;
        const bool ok = rv && ((DL::LR % TP == 0) || j < DL::LR);
#pragma unroll
        for (int r = 0; r < DL::RAD; ++r) {
            v[b][r].x *= a.scale;
            v[b][r].y *= a.scale;
            const R x0 = fabs(v[b][r].x) - eps, x1 = fabs(v[b][r].y) - eps;
            if constexpr (DL::LR % TP == 0) vm = fmax(vm, fmax(x0, x1));
            else vm = (j < DL::LR) ? fmax(vm, fmax(x0, x1)) : vm;
            // x0 > 0 is |u| > eps (the excess of a value not above eps is
            // never positive: denormals are kept); the clamp below tests
            // |u| > eps itself, so these masks end here
            const bool l0 = ok & (x0 > R(0)) & (x0 < tolr);
            const bool l1 = ok & (x1 > R(0)) & (x1 < tolr);
            const int bit = 2 * (b * DL::RAD + r);
            dmask |= (l0 ? DMask(1) << bit : DMask(0)) | (l1 ? DMask(2) << bit : DMask(0));
        }
    }
    // deferred clamps: counted per thread, one wave-uniform test, and only
    // then slots (one reservation per wave) and log entries of the unclamped
    // values
    if (__ballot(dmask != 0)) {
        const uint32_t nl = (uint32_t)__popcll((unsigned long long)dmask);
        uint32_t slot = wave_log_base(nl, &nlog);
#pragma unroll
        for (int b = 0; b < DL::NB; ++b) {
            const int j = t + b * TP;
#pragma unroll
            for (int r = 0; r < DL::RAD; ++r) {
                const uint32_t q = (uint32_t)(p * LZ + j + r * DL::NS);
                const int bit = 2 * (b * DL::RAD + r);
                if ((dmask >> bit) & 1u) {
                    a.log_off[seg + slot] = 2u * q;
                    a.log_v[seg + slot] = v[b][r].x;
                    ++slot;
                }
                if ((dmask >> (bit + 1)) & 1u) {
                    a.log_off[seg + slot] = 2u * q + 1u;
                    a.log_v[seg + slot] = v[b][r].y;
                    ++slot;
                }
            }
        }
    }
    // the clamp itself (v clamped in place for the forward transform), with
    // the edits of the clamps not deferred: the store of a fresh pass and the
    // atomic adds of a later one in loops of their own
    // the bound held in a vector register once (copysign's bit select cannot
    // take it and its mask both from scalar registers)
    R epsv = eps;
    if constexpr (sizeof(R) == 4) asm volatile("v_mov_b32 %0, %1" : "=v"(epsv) : "s"(eps));
    auto clamp_pair = [&](int b, int r, bool ok, R& e0, R& e1, bool& f0, bool& f1) {
        const int bit = 2 * (b * DL::RAD + r);
        f0 = ok & !((dmask >> bit) & 1u);
        f1 = ok & !((dmask >> (bit + 1)) & 1u);
        e0 = clamp_edit_if(v[b][r].x, epsv, f0);
        e1 = clamp_edit_if(v[b][r].y, epsv, f1);
    };
    if (fresh) {
#pragma unroll
        for (int b = 0; b < DL::NB; ++b) {
            const int j = t + b * TP;
            const bool ok = rv && ((DL::LR % TP == 0) || j < DL::LR);
#pragma unroll
            for (int r = 0; r < DL::RAD; ++r) {
                R e0, e1;
                bool f0, f1;
                clamp_pair(b, r, ok, e0, e1, f0, f1);
                if (ok) tile_at(dsp, (uint32_t)(p * LZ + j + r * DL::NS)) = mk(e0, e1);
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < DL::NB; ++b) {
            const int j = t + b * TP;
            const bool ok = rv && ((DL::LR % TP == 0) || j < DL::LR);
#pragma unroll
            for (int r = 0; r < DL::RAD; ++r) {
                R e0, e1;
                bool f0, f1;
                clamp_pair(b, r, ok, e0, e1, f0, f1);
                const uint32_t q = (uint32_t)(p * LZ + j + r * DL::NS);
                if (f0) atomicAdd(&tile_at(dse, 2u * q), e0);
                if (f1) atomicAdd(&tile_at(dse, 2u * q + 1u), e1);
            }
        }
    }
    block_max_store<R, NT>(rv ? vm : R(0), &a.vpart[gb]);
    // every thread is past its reads of the tile (block_max_store's barrier)
    if (tid == 0) a.log_cnt[gb] = nlog;
    // Forward transform, radices reversed: its first stage butterflies the
    // clamped pairs in registers (no twiddles: one sub-transform), and the
    // rest runs through the tile.  Its twiddles and tile addresses are reached
    // through an offset the compiler cannot see is zero, so they are formed
    // again rather than kept live from the inverse transform (which would
    // double the kernel's registers).
    // The thread's and row's indices go through the same zero (the thread
    // index kept in [0, TP) where TP is a power of two), so no address of the
    // inverse half stays live into this one.
    int zoff;
    asm volatile("s_mov_b32 %0, 0" : "=s"(zoff));
    cplx<R>* bf = buf + zoff;
    const int tf = (TP & (TP - 1)) == 0 ? ((t + zoff) & (TP - 1)) : t + zoff;
    const int pf = p + zoff;
    auto lldf = [&](int n) { return bf[bidx(n, pf)]; };
    auto lstf = [&](int n, cplx<R> u) { bf[bidx(n, pf)] = u; };
#pragma unroll
    for (int b = 0; b < DF::NB; ++b) {
        const int j = tf + b * TP;
        if ((DF::LR % TP == 0) || j < DF::LR) dftR<R, DF::RAD>(v[b], -1);
    }
    pen::stage_store<R, LZ, TP, 0, true>(lstf, v, tf);
    __syncthreads();
    if constexpr (NST > 1) {
        pen::transform<R, LZ, TP, true, true, false, 1, true>(lldf, lstf, lldf, lstf, tw + zoff, tf, -1);
        __syncthreads();
    }
    // R2C split X[k] = A[k] + W_nz^k B[k] into the thread's own row, bins k
    // and M-k together: with A, B from Z[k] and Z[M-k], X[M-k] is
    // conj(A - W^k B) (W^(M-k) = -conj W^k), so one product serves both; k = 0
    // gives the DC and Nyquist bins.  Rounds of TP bins unrolled; for an even
    // M with TP dividing M/2 the last round is the lone bin M/2 (one lane),
    // and every store is the row's uniform base plus a 32-bit offset.
    if (rv) {
        const uint32_t ro = (uint32_t)pf * mp;
        auto split = [&](int k) {
            const cplx<R> zk = bf[bidx(k, pf)];
            const cplx<R> zm = bf[bidx(k == 0 ? 0 : M - k, pf)];
            const cplx<R> A = mk(R(0.5) * (zk.x + zm.x), R(0.5) * (zk.y - zm.y));
            const cplx<R> B = mk(R(0.5) * (zk.y + zm.y), R(0.5) * (zm.x - zk.x));
            const cplx<R> wb = cmul(tw2[k], B);
            tile_at(spec_t, ro + (uint32_t)k) = A + wb;
            if (k != M - k) tile_at(spec_t, ro + (uint32_t)(M - k)) = conjg(A - wb);
        };
        constexpr int KB = M / 2 + 1;  // bins k = 0 .. M/2
#pragma unroll
        for (int i = 0; i < KB / TP; ++i) split(tf + i * TP);
        if constexpr (KB % TP != 0) {
            const int k = tf + (KB / TP) * TP;
            if (k < KB) split(k);
        }
    }
}

// ---------------------------------------------------------------------------
// y / x passes over 128-byte column tiles.
// ---------------------------------------------------------------------------
template <typename R, typename T, int L, int MODE, int PREC, int V>
__device__ __forceinline__ void strided_body(const SArgs<R, T>& a, uint32_t blk) {
    constexpr int TP = s_tp<R>(L), W = s_width<R>(V), NT = TP * W;
    constexpr bool RV = s_rev(V);
    static_assert(!RV || PREC == PREC_VALU, "radix order variants are vector-ALU transforms");
    __shared__ cplx<R> tw[L];
    __shared__ cplx<R> buf[L * W];
    if (a.st && a.st->done) return;
    const int tid = threadIdx.x;
    for (int i = tid; i < L; i += NT) tw[i] = a.tw[i];
    __syncthreads();
    const int p = tid % W, t = tid / W;
    const uint32_t ob_ = fdiv(blk, a.dNct);
    const int ct = (int)(blk - ob_ * (uint32_t)a.nct);
    const size_t o = ob_;
    const int col = ct * W + p;
    const bool valid = col < a.ncols;
    cplx<R>* base = a.data + o * a.ob + col;
    auto gld = [&](int n) -> cplx<R> {
        if (!valid) return mk(R(0), R(0));
        if constexpr (MODE == S_LOAD_DF) {
            const size_t idx = (size_t)n * a.cs_n + o * a.cs_o + col;
            const cplx<T> d = a.dfsrc[idx];
            return mk((R)dequant(d.x, a.qf), (R)dequant(d.y, a.qf));
        } else {
            return base[(size_t)n * a.es];
        }
    };
    auto gst = [&](int n, cplx<R> v) {
        if (valid) base[(size_t)n * a.es] = v;
    };
    auto lld = [&](int n) { return buf[n * W + p]; };
    auto lst = [&](int n, cplx<R> v) { buf[n * W + p] = v; };
    // element n of column pp of the tile (the radix-16 stages serve
    // butterflies of any column of the tile, 16 adjacent columns at a time)
    auto gldp = [&](int n, int pp) -> cplx<R> {
        const int c = ct * W + pp;
        if (c >= a.ncols) return mk(R(0), R(0));
        return a.data[o * a.ob + c + (size_t)n * a.es];
    };
    auto gstp = [&](int n, int pp, cplx<R> v) {
        const int c = ct * W + pp;
        if (c < a.ncols) a.data[o * a.ob + c + (size_t)n * a.es] = v;
    };
    if constexpr (MODE == S_PLAIN || MODE == S_LOAD_DF) {
        const int sg = MODE == S_LOAD_DF ? +1 : a.sign;
        if constexpr (PREC == PREC_VALU) pen::transform<R, L, TP, false, false, false, 0, RV>(gld, gst, lld, lst, tw, t, sg);
        else mx::transform<PREC, L, TP, W, W, NT, false, false, false>(gldp, gstp, buf, tw, tid, p, t, sg);
    } else if constexpr (MODE == S_CLAMP) {
        bool fired = false;
        auto cst = [&](int n, cplx<R> v) {
            if (valid) fired |= freq_clamp_bin(v, (size_t)n * a.cs_n + o * a.cs_o + col, a);
            buf[n * W + p] = v;
        };
        auto cstp = [&](int n, int pp, cplx<R> v) {
            const int c = ct * W + pp;
            if (c < a.ncols) fired |= freq_clamp_bin(v, (size_t)n * a.cs_n + o * a.cs_o + c, a);
            buf[n * W + pp] = v;
        };
        if constexpr (PREC == PREC_VALU) pen::transform<R, L, TP, false, true, false, 0, RV>(gld, cst, lld, lst, tw, t, -1);
        else mx::transform<PREC, L, TP, W, W, NT, false, false, true>(gldp, cstp, buf, tw, tid, p, t, -1);
        if (fired) a.st->freq_clamped = 1u;
        __syncthreads();
        if constexpr (PREC == PREC_VALU) pen::transform<R, L, TP, true, false, false, 0, RV>(lld, gst, lld, lst, tw, t, +1);
        else mx::transform<PREC, L, TP, W, W, NT, true, false, false, false>(gldp, gstp, buf, tw, tid, p, t, +1);
    } else if constexpr (MODE == S_CLAMP_PASS && PREC != PREC_VALU) {
        // The MFMA fused x pass (every bound kind): the forward transform ends
        // in the tile and the clamp runs as a pass of its own over it, one bin
        // at a time (freq_clamp_bin).  Clamping in the transform's last vector
        // stage kept the clamp's state live there and pushed the kernel, held
        // to 64 VGPRs, into spills.
        auto bstp = [&](int n, int pp, cplx<R> v) { buf[n * W + pp] = v; };
        mx::transform<PREC, L, TP, W, W, NT, false, false, true>(gldp, bstp, buf, tw, tid, p, t, -1);
        __syncthreads();
        constexpr int PER = (L * W + NT - 1) / NT;
        const size_t obase = o * a.cs_o + (size_t)ct * W;
        bool fired = false;
        for (int k = 0; k < PER; ++k) {
            const int q = tid + k * NT;
            const int pp = q % W;
            if (q >= L * W || ct * W + pp >= a.ncols) continue;
            cplx<R> v = buf[q];
            fired |= freq_clamp_bin(v, (size_t)(q / W) * a.cs_n + obase + pp, a);
            buf[q] = v;
        }
        if (fired) a.st->freq_clamped = 1u;
        __syncthreads();
        mx::transform<PREC, L, TP, W, W, NT, true, false, false, false>(gldp, gstp, buf, tw, tid, p, t, +1);
    } else if constexpr (MODE == S_CLAMP_PASS) {
        // Per-bin and ptw bounds: the forward transform ends in the tile, then
        // the clamp runs as a pass of its own over it.  Each thread issues all
        // of its bins' bound loads (delta_k, or F(x) for ptw) first; the edits
        // of the bins that clamp go to the accumulator as atomic adds that
        // return nothing (one writer per bin and pass: the sum is the plain
        // old + edit), so no load of the accumulator holds the pass up.  The
        // arithmetic is freq_clamp_bin's: the edit formed, then accumulated;
        // ptw as ptw_clamp (the DC bin's real part zeroed).
        static_assert(PREC == PREC_VALU, "the clamp pass follows the vector-ALU transform");
        constexpr int PER = (L * W + NT - 1) / NT;
        const bool ptw = a.fmode == FM_PTW;
        const size_t obase = o * a.cs_o + (size_t)ct * W;
        auto bin_ok = [&](int k) {
            const int q = tid + k * NT;
            return k < PER && q < L * W && ct * W + q % W < a.ncols;
        };
        auto bin_idx = [&](int k) {
            const int q = tid + k * NT;
            return (size_t)(q / W) * a.cs_n + obase + q % W;
        };
        // the bins' bounds are loaded before the forward transform, so their
        // latency overlaps it (the tiles of the long FP64 axes leave room for
        // one workgroup per CU, too few to hide it otherwise)
        cplx<R> bnd[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const size_t idx = bin_idx(k);
            bnd[k] = !bin_ok(k) ? mk(R(0), R(0))
                                : (ptw ? a.orig[idx] : mk(a.delta_dense[idx] * a.dscale, R(0)));
        }
        pen::transform<R, L, TP, false, true, false, 0, RV>(gld, lst, lld, lst, tw, t, -1);
        __syncthreads();
        // the ptw constants, as ptw_clamp forms them (projection_algorithm.cu:199-258)
        const R sp = sqrt(R(1) + a.delta_ptw), sm = sqrt(R(1) - a.delta_ptw);
        const R dptw = fmin(R(1) - sm, sp - R(1)) / (R)sqrtf(2.0f);
        const R dcorr = fmin(R(1) - sm, sp - R(1)) / sqrt(R(2));
        bool fired = false;
        R* dfr = &a.df[0].x;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (!bin_ok(k)) continue;
            const int q = tid + k * NT;
            const size_t idx = bin_idx(k);
            const cplx<R> v = buf[q];
            const cplx<R> b = bnd[k];
            cplx<R> c = v;
            R d = b.x;
            bool active = true;
            if (ptw) {
                // as ptw_clamp (reference or corrected rule per bin)
                const int rule = ptw_rule<R>(idx, a.ptw_mask, a.ptw_corr);
                if (rule == PTW_NONE) continue;
                if (rule == PTW_DC) {
                    const R lo = b.x * (R(1) / sp - R(1)), hi = b.x * (R(1) / sm - R(1));
                    if (v.x > hi || v.x < lo) {
                        atomicAdd(dfr + 2 * idx, -v.x);
                        c.x = R(0);
                        fired = true;
                    }
                    buf[q] = c;
                    continue;
                }
                const R omag = b.x * b.x + b.y * b.y;
                const R cr = b.x + v.x, ci = b.y + v.y;
                const R ratio = (cr * cr + ci * ci) / omag;
                if (a.ptw_corr) {
                    active = !(ratio <= R(1) + a.delta_ptw && ratio >= R(1) - a.delta_ptw);
                    d = dcorr * sqrt(omag);
                } else {
                    active = !(ratio <= sp && ratio >= sm);
                    d = dptw;
                }
            }
            const bool xh = v.x > d, xl = v.x < -d, yh = v.y > d, yl = v.y < -d;
            if (active && (xh || xl)) {
                c.x = xh ? d : -d;
                atomicAdd(dfr + 2 * idx, c.x - v.x);
                fired = true;
            }
            if (active && (yh || yl)) {
                c.y = yh ? d : -d;
                atomicAdd(dfr + 2 * idx + 1, c.y - v.y);
                fired = true;
            }
            buf[q] = c;
        }
        if (fired) a.st->freq_clamped = 1u;
        __syncthreads();
        pen::transform<R, L, TP, true, false, false, 0, RV>(lld, gst, lld, lst, tw, t, +1);
    } else {
        static_assert(PREC == PREC_VALU, "certification runs on the vector ALU");  // S_CERT_CHECK
        unsigned long long nv = 0;
        double vmax = 0.0;
        auto chk = [&](int n, cplx<R> v) {
            if (!valid) return;
            const size_t idx = (size_t)n * a.cs_n + o * a.cs_o + col;
            const double d = a.cdelta_dense ? (double)a.cdelta_dense[idx] : a.cdelta;
            const double ex = fmax(fabs((double)v.x), fabs((double)v.y)) - d;
            if (ex > 0.0) {
                ++nv;
                vmax = ex > vmax ? ex : vmax;
            }
        };
        pen::transform<R, L, TP, false, false, false, 0, RV>(gld, chk, lld, lst, tw, t, -1);
        block_count_atomic(nv, &a.stats->freq_viol);
        block_max_atomic<double, NT>(vmax, &a.stats->freq_max_bits);
    }
}

// ---------------------------------------------------------------------------
// kernels: one field (arguments in the kernarg segment), or a batch of fields
// of the same axis length in one launch (workgroup -> field by the tile
// prefix table; each field's own state gates its tiles)
// ---------------------------------------------------------------------------
template <typename A>
__device__ __forceinline__ uint32_t batch_field(const BatchTab<A>& t, uint32_t blk) {
    int lo = 0, hi = t.n - 1;  // largest f with start[f] <= blk
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.start[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return (uint32_t)lo;
}

template <typename R, typename T, int LZ, int MODE, int PREC = PREC_VALU, int V = 0>
__global__ __launch_bounds__(z_tp(LZ) * z_rows(LZ, V)) void k2_zfwd(ZArgs<R, T> a) {
    zfwd_body<R, T, LZ, MODE, PREC, V>(a, blockIdx.x);
}
template <typename R, typename T, int LZ, int MODE, int PREC = PREC_VALU, int V = 0>
__global__ __launch_bounds__(z_tp(LZ) * z_rows(LZ, V), PREC != PREC_VALU ? 8 : 1) void k2_zinv(ZArgs<R, T> a) {
    zinv_body<R, T, LZ, MODE, PREC, V>(a, blockIdx.x);
}
// FP32 power-of-two lengths at eight waves per SIMD (at most 64 VGPRs, no
// spills), as the separate z kernels run; other lengths and FP64 spill under
// that bound and take the registers they need
template <typename R, int LZ>
constexpr int zz_min_waves() { return (sizeof(R) == 4 && (LZ & (LZ - 1)) == 0) ? 8 : 1; }
template <typename R, int LZ, int V = 0>
__global__ __launch_bounds__(z_tp(LZ) * z_rows(LZ, V), (zz_min_waves<R, LZ>())) void k2_zz(ZArgs<R, R> a) {
    zz_body<R, LZ, V>(a, blockIdx.x);
}
// The MFMA variants ask for 8 waves per SIMD (at most 64 VGPRs): two
// 1024-thread tiles per CU, like the vector-ALU kernels.  The fused clamp
// kernel spills a few registers under that bound and still runs faster than
// at one tile per CU (0.41 against 0.53 ms at 512^3).
// Workgroup i of a launch runs on XCD i mod 8 (each XCD has its own L2).
// Neighbouring column tiles share cache lines: of a tile narrower than a
// 128-byte line, and of the per-bin arrays in compact order (bounds, edit
// accumulators), whose rows are not line-aligned.  So each XCD takes a
// contiguous run of tiles, in order: workgroup i gets tile (i mod 8) n/8 +
// i/8, and a tile's neighbours run on the same XCD at about the same time and
// find the shared lines in its L2.  A bijection on [0, n) (the last n mod 8
// workgroups keep their tiles).  Tiles a whole 128-byte line wide that read
// no per-bin array share no line with a neighbour and run in dispatch order,
// which is faster for them (a load-tile / store-tile kernel over the 512^3
// f32 half spectrum's y columns: 0.234 against 0.251 ms).
__device__ __forceinline__ uint32_t xcd_chunk_tile(uint32_t i, uint32_t n) {
    const uint32_t per = n >> 3;
    if (i >= (per << 3)) return i;
    return (i & 7u) * per + (i >> 3);
}
template <typename R, typename T, int MODE, int V>
__device__ __forceinline__ bool s_chunked(const SArgs<R, T>& a) {
    if ((V & 1) != 0) return true;
    switch (MODE) {
        case S_CLAMP_PASS:
        case S_LOAD_DF:
            return true;
        case S_CLAMP:
            return a.fmode != FM_ABS;
        case S_CERT_CHECK:
            return a.cdelta_dense != nullptr;
        default:
            return false;
    }
}
template <typename R, typename T, int L, int MODE, int PREC = PREC_VALU, int V = 0>
__global__ __launch_bounds__(s_tp<R>(L) * s_width<R>(V),
                             (PREC != PREC_VALU) ? (V == 0 ? 8 : 6) : 1) void k2_strided(SArgs<R, T> a) {
    strided_body<R, T, L, MODE, PREC, V>(a, s_chunked<R, T, MODE, V>(a) ? xcd_chunk_tile(blockIdx.x, gridDim.x)
                                                                         : blockIdx.x);
}
// Two forward z passes over the same rows in one workgroup (the paired
// certification): the second re-reads the rows the first just read, from cache.
template <typename R, typename T, int LZ, int MODE>
__global__ __launch_bounds__(z_tp(LZ) * z_rows(LZ)) void k2_zfwd_pair(ZArgs<R, T> a, ZArgs<R, T> b) {
    zfwd_body<R, T, LZ, MODE, PREC_VALU, 0>(a, blockIdx.x);
    __syncthreads();
    zfwd_body<R, T, LZ, MODE, PREC_VALU, 0>(b, blockIdx.x);
}
template <typename R, typename T, int LZ, int MODE>
__global__ __launch_bounds__(z_tp(LZ) * z_rows(LZ, zb_var(LZ))) void k2_zfwd_b(BatchTab<ZArgs<R, T>> t) {
    const uint32_t f = batch_field(t, blockIdx.x);
    const ZArgs<R, T> a = t.args[f];
    zfwd_body<R, T, LZ, MODE, PREC_VALU, zb_var(LZ)>(a, blockIdx.x - t.start[f]);
}
template <typename R, typename T, int LZ, int MODE>
__global__ __launch_bounds__(z_tp(LZ) * z_rows(LZ, zb_var(LZ))) void k2_zinv_b(BatchTab<ZArgs<R, T>> t) {
    const uint32_t f = batch_field(t, blockIdx.x);
    const ZArgs<R, T> a = t.args[f];
    zinv_body<R, T, LZ, MODE, PREC_VALU, zb_var(LZ)>(a, blockIdx.x - t.start[f]);
}
template <typename R, int LZ>
__global__ __launch_bounds__(z_tp(LZ) * z_rows(LZ, zb_var(LZ)), (zz_min_waves<R, LZ>())) void k2_zz_b(BatchTab<ZArgs<R, R>> t) {
    const uint32_t f = batch_field(t, blockIdx.x);
    const ZArgs<R, R> a = t.args[f];
    zz_body<R, LZ, zb_var(LZ)>(a, blockIdx.x - t.start[f]);
}
template <typename R, typename T, int L, int MODE>
__global__ __launch_bounds__(s_tp<R>(L) * s_width<R>()) void k2_strided_b(BatchTab<SArgs<R, T>> t) {
    const uint32_t f = batch_field(t, blockIdx.x);
    const SArgs<R, T> a = t.args[f];
    strided_body<R, T, L, MODE, PREC_VALU, 0>(a, blockIdx.x - t.start[f]);
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <typename R, typename T, int MODE, int LZ, int PREC = PREC_VALU, int V = 0>
inline void launch_zfwd_len(const ZArgs<R, T>& a, hipStream_t s) {
    const unsigned grid = (unsigned)ceil_div(a.rows, (size_t)z_rows(LZ, V));
    hipLaunchKernelGGL((k2_zfwd<R, T, LZ, MODE, PREC, V>), dim3(grid), dim3(z_tp(LZ) * z_rows(LZ, V)), 0, s, a);
}
template <typename R, typename T, int MODE, int LZ, int PREC = PREC_VALU, int V = 0>
inline void launch_zinv_len(const ZArgs<R, T>& a, hipStream_t s) {
    const unsigned grid = (unsigned)ceil_div(a.rows, (size_t)z_rows(LZ, V));
    hipLaunchKernelGGL((k2_zinv<R, T, LZ, MODE, PREC, V>), dim3(grid), dim3(z_tp(LZ) * z_rows(LZ, V)), 0, s, a);
}
template <typename R, typename T, int MODE, int L, int PREC = PREC_VALU, int V = 0>
inline void launch_strided_len(SArgs<R, T> a, size_t outer, hipStream_t s) {
    constexpr int W = s_width<R>(V);
    a.W = W;
    a.nct = (int)ceil_div((size_t)a.ncols, (size_t)W);
    a.dNct = FastDiv((uint32_t)a.nct);
    const unsigned grid = (unsigned)(outer * (size_t)a.nct);
    hipLaunchKernelGGL((k2_strided<R, T, L, MODE, PREC, V>), dim3(grid), dim3(s_tp<R>(L) * W), 0, s, a);
}

template <typename R, typename T, int MODE>
bool zfwd_by_len(const ZArgs<R, T>& a, hipStream_t s) {
    if (!a.half) return false;
    switch (a.Lz) {
#define FFCZ_ZF(Lv)                                                              \
    case Lv:                                                                     \
        if (a.var) launch_zfwd_len<R, T, MODE, Lv, PREC_VALU, 1>(a, s);          \
        else launch_zfwd_len<R, T, MODE, Lv, PREC_VALU, 0>(a, s);                \
        return true;
        FFCZ_FAST_Z_LENGTHS(FFCZ_ZF)
#undef FFCZ_ZF
        default:
            return false;
    }
}
template <typename R, int LZ, int V = 0>
inline void launch_zz_len(const ZArgs<R, R>& a, hipStream_t s) {
    const unsigned grid = (unsigned)ceil_div(a.rows, (size_t)z_rows(LZ, V));
    hipLaunchKernelGGL((k2_zz<R, LZ, V>), dim3(grid), dim3(z_tp(LZ) * z_rows(LZ, V)), 0, s, a);
}
template <typename R>
bool zz_by_len(const ZArgs<R, R>& a, hipStream_t s) {
    if (!a.half) return false;
    switch (a.Lz) {
#define FFCZ_ZZ(Lv)                                        \
    case Lv:                                               \
        if (a.var) launch_zz_len<R, Lv, 1>(a, s);          \
        else launch_zz_len<R, Lv, 0>(a, s);                \
        return true;
        FFCZ_FAST_Z_LENGTHS(FFCZ_ZZ)
#undef FFCZ_ZZ
        default:
            return false;
    }
}
template <typename R, typename T, int MODE>
bool zinv_by_len(const ZArgs<R, T>& a, hipStream_t s) {
    if (!a.half) return false;
    switch (a.Lz) {
#define FFCZ_ZI(Lv)                                                              \
    case Lv:                                                                     \
        if (a.var) launch_zinv_len<R, T, MODE, Lv, PREC_VALU, 1>(a, s);          \
        else launch_zinv_len<R, T, MODE, Lv, PREC_VALU, 0>(a, s);                \
        return true;
        FFCZ_FAST_Z_LENGTHS(FFCZ_ZI)
#undef FFCZ_ZI
        default:
            return false;
    }
}
// the pass modes of the iteration take the radix order variant (V=2)
template <int MODE>
constexpr bool s_rev_variant() {
    return MODE == S_PLAIN || MODE == S_CLAMP || MODE == S_CLAMP_PASS;
}
template <typename R, typename T, int MODE>
bool strided_by_len(const SArgs<R, T>& a, size_t outer, hipStream_t s) {
    switch (a.L) {
#define FFCZ_SL(Lv)                                                                 \
    case Lv:                                                                        \
        if constexpr (s_rev_variant<MODE>() && pen::rev_differs(Lv)) {               \
            if (a.var == 2) {                                                       \
                launch_strided_len<R, T, MODE, Lv, PREC_VALU, 2>(a, outer, s);      \
                return true;                                                        \
            }                                                                       \
        }                                                                           \
        if (a.var & 1) launch_strided_len<R, T, MODE, Lv, PREC_VALU, 1>(a, outer, s); \
        else launch_strided_len<R, T, MODE, Lv, PREC_VALU, 0>(a, outer, s);         \
        return true;
        FFCZ_FAST_S_LENGTHS(FFCZ_SL)
#undef FFCZ_SL
        default:
            return false;
    }
}

// MASK bits select the pass modes a translation unit instantiates.
template <typename R, typename T, unsigned MASK>
bool zfwd_dispatch(int mode, const ZArgs<R, T>& a, hipStream_t s) {
    // the specialised kernels move element pairs as one vector
    if (mode == Z_PLAIN && ((uintptr_t)a.in % (2 * sizeof(R))) != 0) return false;
    if (mode == Z_PLAIN) {
        if constexpr ((MASK & (1u << Z_PLAIN)) != 0) return zfwd_by_len<R, T, Z_PLAIN>(a, s);
    } else if (mode == Z_ITER) {
        if constexpr ((MASK & (1u << Z_ITER)) != 0) return zfwd_by_len<R, T, Z_ITER>(a, s);
    } else if (mode == Z_CERT) {
        if constexpr ((MASK & (1u << Z_CERT)) != 0) return zfwd_by_len<R, T, Z_CERT>(a, s);
    }
    return false;
}
template <typename R, typename T, unsigned MASK>
bool zinv_dispatch(int mode, const ZArgs<R, T>& a, hipStream_t s) {
    if (mode == Z_PLAIN && ((uintptr_t)a.out % (2 * sizeof(R))) != 0) return false;
    if (mode == Z_PLAIN) {
        if constexpr ((MASK & (1u << Z_PLAIN)) != 0) return zinv_by_len<R, T, Z_PLAIN>(a, s);
    } else if (mode == Z_ITER) {
        if constexpr ((MASK & (1u << Z_ITER)) != 0) return zinv_by_len<R, T, Z_ITER>(a, s);
    } else if (mode == Z_FUSED) {
        if constexpr ((MASK & (1u << Z_FUSED)) != 0 && std::is_same<R, T>::value) return zz_by_len<R>(a, s);
    }
    return false;
}
template <typename R, typename T, unsigned MASK>
bool strided_dispatch(int mode, const SArgs<R, T>& a, size_t outer, hipStream_t s) {
    if (mode == S_PLAIN) {
        if constexpr ((MASK & (1u << S_PLAIN)) != 0) return strided_by_len<R, T, S_PLAIN>(a, outer, s);
    } else if (mode == S_CLAMP) {
        if constexpr ((MASK & (1u << S_CLAMP)) != 0) return strided_by_len<R, T, S_CLAMP>(a, outer, s);
    } else if (mode == S_CERT_CHECK) {
        if constexpr ((MASK & (1u << S_CERT_CHECK)) != 0) return strided_by_len<R, T, S_CERT_CHECK>(a, outer, s);
    } else if (mode == S_LOAD_DF) {
        if constexpr ((MASK & (1u << S_LOAD_DF)) != 0) return strided_by_len<R, T, S_LOAD_DF>(a, outer, s);
    } else if (mode == S_CLAMP_PASS) {
        if constexpr ((MASK & (1u << S_CLAMP_PASS)) != 0) return strided_by_len<R, T, S_CLAMP_PASS>(a, outer, s);
    }
    return false;
}

// ---------------------------------------------------------------------------
// MFMA split-precision variants (R = T = float): lengths with a radix-16
// factor whose tiles fill whole waves.  The z kernels keep the VALU kernels'
// rows per workgroup, so the per-workgroup violation slots line up.
// ---------------------------------------------------------------------------
#define FFCZ_MFMA_S_LENGTHS(X) X(32) X(64) X(96) X(128) X(256) X(512)
#define FFCZ_MFMA_Z_LENGTHS(X) X(16) X(32) X(48) X(64) X(128) X(256)

// The MFMA passes run the full-width tiles only: the tile variants apply to
// the vector-ALU passes (the fused z pass included, which is a vector-ALU
// kernel in every FFT variant).

template <int PREC>
bool mfma_zfwd_dispatch(int mode, const ZArgs<float, float>& a, hipStream_t s) {
    if (!a.half || (mode != Z_PLAIN && mode != Z_ITER)) return false;
    if (mode == Z_PLAIN && ((uintptr_t)a.in % 8) != 0) return false;
    switch (a.Lz) {
#define FFCZ_MZF(Lv)                                                                                      \
    case Lv:                                                                                              \
        if (mode == Z_ITER) launch_zfwd_len<float, float, Z_ITER, Lv, PREC>(a, s);                        \
        else launch_zfwd_len<float, float, Z_PLAIN, Lv, PREC>(a, s);                                      \
        return true;
        FFCZ_MFMA_Z_LENGTHS(FFCZ_MZF)
#undef FFCZ_MZF
        default:
            return false;
    }
}
template <int PREC>
bool mfma_zinv_dispatch(int mode, const ZArgs<float, float>& a, hipStream_t s) {
    if (!a.half || (mode != Z_PLAIN && mode != Z_ITER)) return false;
    if (mode == Z_PLAIN && ((uintptr_t)a.out % 8) != 0) return false;
    switch (a.Lz) {
#define FFCZ_MZI(Lv)                                                                                      \
    case Lv:                                                                                              \
        if (mode == Z_ITER) launch_zinv_len<float, float, Z_ITER, Lv, PREC>(a, s);                        \
        else launch_zinv_len<float, float, Z_PLAIN, Lv, PREC>(a, s);                                      \
        return true;
        FFCZ_MFMA_Z_LENGTHS(FFCZ_MZI)
#undef FFCZ_MZI
        default:
            return false;
    }
}
template <int PREC>
bool mfma_strided_dispatch(int mode, const SArgs<float, float>& a, size_t outer, hipStream_t s) {
    // the MFMA fused x pass clamps in a pass of its own over the tile, for
    // every bound kind
    if (mode != S_PLAIN && mode != S_CLAMP && mode != S_CLAMP_PASS) return false;
    const bool clamp = mode != S_PLAIN;
    switch (a.L) {
#define FFCZ_MSL(Lv)                                                                                 \
    case Lv:                                                                                         \
        if (clamp) launch_strided_len<float, float, S_CLAMP_PASS, Lv, PREC>(a, outer, s);            \
        else launch_strided_len<float, float, S_PLAIN, Lv, PREC>(a, outer, s);                       \
        return true;
        FFCZ_MFMA_S_LENGTHS(FFCZ_MSL)
#undef FFCZ_MSL
        default:
            return false;
    }
}

template <typename R, typename T, int MODE>
bool zfwd_pair_dispatch(const ZArgs<R, T>& a, const ZArgs<R, T>& b, hipStream_t s) {
    if (!a.half) return false;
    switch (a.Lz) {
#define FFCZ_PZF(Lv)                                                                                  \
    case Lv:                                                                                          \
        hipLaunchKernelGGL((k2_zfwd_pair<R, T, Lv, MODE>), dim3((unsigned)ceil_div(a.rows, (size_t)z_rows(Lv))), \
                           dim3(z_tp(Lv) * z_rows(Lv)), 0, s, a, b);                                  \
        return true;
        FFCZ_FAST_Z_LENGTHS(FFCZ_PZF)
#undef FFCZ_PZF
        default:
            return false;
    }
}

// ---------------------------------------------------------------------------
// batched launches (vector-ALU kernels): tiles of every field in the table
// ---------------------------------------------------------------------------
template <typename R, typename T, int MODE>
bool zfwd_batch_dispatch(int Lz, const BatchTab<ZArgs<R, T>>& t, unsigned grid, hipStream_t s) {
    switch (Lz) {
#define FFCZ_BZF(Lv)                                                                                        \
    case Lv:                                                                                                \
        hipLaunchKernelGGL((k2_zfwd_b<R, T, Lv, MODE>), dim3(grid), dim3(z_tp(Lv) * z_rows(Lv, zb_var(Lv))), 0, s, t);  \
        return true;
        FFCZ_FAST_Z_LENGTHS(FFCZ_BZF)
#undef FFCZ_BZF
        default:
            return false;
    }
}
template <typename R, typename T, int MODE>
bool zinv_batch_dispatch(int Lz, const BatchTab<ZArgs<R, T>>& t, unsigned grid, hipStream_t s) {
    switch (Lz) {
#define FFCZ_BZI(Lv)                                                                                        \
    case Lv:                                                                                                \
        hipLaunchKernelGGL((k2_zinv_b<R, T, Lv, MODE>), dim3(grid), dim3(z_tp(Lv) * z_rows(Lv, zb_var(Lv))), 0, s, t);  \
        return true;
        FFCZ_FAST_Z_LENGTHS(FFCZ_BZI)
#undef FFCZ_BZI
        default:
            return false;
    }
}
template <typename R>
bool zz_batch_dispatch(int Lz, const BatchTab<ZArgs<R, R>>& t, unsigned grid, hipStream_t s) {
    switch (Lz) {
#define FFCZ_BZZ(Lv)                                                                                   \
    case Lv:                                                                                           \
        hipLaunchKernelGGL((k2_zz_b<R, Lv>), dim3(grid), dim3(z_tp(Lv) * z_rows(Lv, zb_var(Lv))), 0, s, t);        \
        return true;
        FFCZ_FAST_Z_LENGTHS(FFCZ_BZZ)
#undef FFCZ_BZZ
        default:
            return false;
    }
}
template <typename R, typename T, int MODE>
bool strided_batch_dispatch(int L, const BatchTab<SArgs<R, T>>& t, unsigned grid, hipStream_t s) {
    switch (L) {
#define FFCZ_BSL(Lv)                                                                                      \
    case Lv:                                                                                              \
        hipLaunchKernelGGL((k2_strided_b<R, T, Lv, MODE>), dim3(grid),                                   \
                           dim3(s_tp<R>(Lv) * s_width<R>()), 0, s, t);                                    \
        return true;
        FFCZ_FAST_S_LENGTHS(FFCZ_BSL)
#undef FFCZ_BSL
        default:
            return false;
    }
}

}  // namespace ffcz
