// Kernel argument blocks and device helpers shared by the generic and the
// length-specialised FFT pass kernels.
#pragma once

#include "fft3d.hpp"
#include "quant.hpp"

namespace ffcz {

template <typename R>
__device__ __forceinline__ void clamp_acc(R& v, R d, R* acc) {
    // projection_algorithm.cu:161-172 / :351-360: the edit is formed first
    // and then accumulated (a single writer per address).  The clamped value
    // is selected without branches (d - v above the box, -d - v below it);
    // only the accumulator's read-modify-write is conditional.
    const bool hi = v > d, lo = v < -d;
    const R c = hi ? d : (lo ? -d : v);
    if (hi || lo) *acc += c - v;
    v = c;
}

// The same clamp when the accumulator is known to be zero: returns the edit
// (0 where nothing is clamped) for a plain store, bit-identical to adding it to 0.
template <typename R>
__device__ __forceinline__ R clamp_edit(R& v, R d) {
    const bool hi = v > d, lo = v < -d;
    const R c = hi ? d : (lo ? -d : v);
    const R e = (hi || lo) ? c - v : R(0);
    v = c;
    return e;
}

// clamp_edit without branches, for an edit that is kept only where `keep`
// holds: v is clamped in any case; the edit (c - v) is returned, and `keep`
// left true, only where v was clamped and keep held (0 and false otherwise).
// |v| > d is the same test as v > d || v < -d, and copysign(d, v) the same
// bound as hi ? d : -d, so v and the kept edit are those of clamp_edit.
template <typename R>
__device__ __forceinline__ R clamp_edit_if(R& v, R d, bool& keep) {
    const bool h = fabs(v) > d;
    const R c = h ? copysign(d, v) : v;
    keep = keep & h;
    const R e = keep ? c - v : R(0);
    v = c;
    return e;
}

template <typename R, int NT = FFT_THREADS>
__device__ __forceinline__ void block_max_atomic(R v, unsigned long long* target) {
    __shared__ R red[(NT + 63) / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        R m = red[0];
        for (int i = 1; i < (NT + 63) / 64; ++i) m = red[i] > m ? red[i] : m;
        if (m > R(0)) atomicMax(target, ordbits(m));
    }
}

// Workgroup max written to its own slot (no atomics on a shared address).
template <typename R, int NT>
__device__ __forceinline__ void block_max_store(R v, R* slot) {
    __shared__ R red[(NT + 63) / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        R m = red[0];
        for (int i = 1; i < (NT + 63) / 64; ++i) m = red[i] > m ? red[i] : m;
        *slot = m;
    }
}

// Run by one thread of the last workgroup of an iteration's inverse z pass:
// publish the max spatial violation, count the iteration and decide
// termination (projection_algorithm_impl.cuh:841-848).
template <typename R>
__device__ __forceinline__ void finalize_iteration(DevState* st) {
    __threadfence();
    const unsigned long long b = atomicAdd(&st->viol_bits, 0ull);
    const R v = from_ordbits<R>(b);
    st->iter += 1;
    st->last_viol = (double)v;
    bool conv;
    if constexpr (sizeof(R) == 4) conv = v < st->tol_f;
    else conv = v < (R)st->tol;
    if (conv) {
        st->converged = 1;
        st->done = 1;
    } else if (st->iter >= st->max_iter) {
        st->done = 1;
    }
    atomicExch(&st->viol_bits, 0ull);
    atomicExch(&st->blocks_done, 0u);
    __threadfence();
}

__device__ __forceinline__ void block_count_atomic(unsigned long long c, unsigned long long* target) {
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(target, c);
}

enum ZMode : int { Z_PLAIN = 0, Z_ITER = 1, Z_CERT = 2, Z_FUSED = 3 };

template <typename R, typename T>
struct ZArgs {
    int nz, Lz, M, half, P, sn;
    int prec;             // FftPrec (float passes only)
    int var;              // tile variant of the length-specialised kernel
    FastDiv dP, dM, dNh, dNp, dNz, dFill;  // P, M, nz/2+1, M/2+1, nz, nz-(nz/2+1)
    size_t rows, Mp;
    FftPlan pl;
    const cplx<R>* tz;
    const cplx<R>* tz2;
    const R* in;          // Z_PLAIN forward input
    R* out;               // Z_PLAIN inverse output
    cplx<R>* spec;
    R scale;              // inverse scale
    // Z_ITER
    R* e;
    const R* e_init;  // read instead of e by the first iteration (iter == 0), if set
    R* e_copy;        // the first iteration also stores what it read from e_init here, if set
    int ds_fresh;     // ds holds no edits yet: the first clamp (iter == 1) stores ds whole
    R* ds;
    R eps;
    DevState* st;
    R* vpart;  // per-workgroup max violation of the inverse z pass
    size_t row_off;  // first row of this launch within the field (slab launches)
    // Z_FUSED log of deferred clamps: per workgroup a segment of P*nz entries
    // (tile-relative element offset, the unclamped value) and a count
    uint32_t* log_off;
    R* log_v;
    uint32_t* log_cnt;
    // Z_CERT
    const T* e0;
    const T* cds;
    const double* dfi;
    QuantDesc qs;
    double ceps;
    CertStats* stats;
};

// Arguments of one launch over several fields: field f owns workgroups
// [start[f], start[f+1]).
template <typename A>
struct BatchTab {
    const A* args;
    const uint32_t* start;
    int n;
};

// S_CLAMP_PASS: S_CLAMP with per-bin or ptw bounds, the clamp done as a pass
// of its own over the transformed tile (vector-ALU kernels; the generic
// kernels run it as S_CLAMP)
enum SMode : int { S_PLAIN = 0, S_CLAMP = 1, S_CERT_CHECK = 2, S_LOAD_DF = 3, S_CLAMP_PASS = 4 };

template <typename R, typename T>
struct SArgs {
    cplx<R>* data;
    FftPlan pl;
    const cplx<R>* tw;
    int L, W, wshift, ncols, nct;
    int prec;             // FftPrec (float passes only)
    int var;              // tile variant of the length-specialised kernel
    FastDiv dW, dNct;
    size_t ob, es;
    int sign;
    DevState* st;
    // clamp / check / load: bin index = n*cs_n + o*cs_o + col (reference layout)
    size_t F, cs_n, cs_o;
    int fmode;
    R delta;
    const R* delta_dense;
    R dscale;
    cplx<R>* df;          // frequency edits, one (Re, Im) pair per bin (the solver's working layout)
    const cplx<R>* orig;
    R delta_ptw;
    const uint8_t* ptw_mask;  // FM_PTW: per-bin rule (PTW_*; null: every bin, DC by the DC rule)
    int ptw_corr;             // FM_PTW: the corrected power-spectrum bound instead of the reference's
    // certification
    double cdelta;
    const T* cdelta_dense;
    CertStats* stats;
    const cplx<T>* dfsrc;  // frequency edits, one (Re, Im) pair per bin
    QuantDesc qf;
};

template <typename R>
__device__ __forceinline__ int ptw_rule(size_t idx, const uint8_t* mask, int corr) {
    if (mask) return mask[idx];
    return (idx == 0 && !corr) ? PTW_DC : PTW_RATIO;
}

// The pointwise bound on one bin.  Reference semantics (projection_algorithm.
// cu:199-258): the DC rule zeroes the DC error outside [X0(1/sqrt(1+d)-1),
// X0(1/sqrt(1-d)-1)]; the ratio rule tests |X+E|^2/|X|^2 against sqrt(1+-d)
// and clamps Re, Im to the absolute d' = min(1-sqrt(1-d), sqrt(1+d)-1) /
// sqrtf(2).  Corrected semantics (corr; D11, not the reference's): the power
// ratio is tested against [1-d, 1+d] and Re, Im are clamped to d'|X| (sqrt(2)
// in R), for every constrained bin, DC included.
template <typename R>
__device__ __forceinline__ void ptw_clamp(cplx<R>& v, size_t idx, const cplx<R>* orig, R dptw, cplx<R>* df,
                                          const uint8_t* mask, int corr) {
    const int rule = ptw_rule<R>(idx, mask, corr);
    if (rule == PTW_NONE) return;
    if (rule == PTW_DC) {
        const R lo = orig[0].x * (R(1) / sqrt(R(1) + dptw) - R(1));
        const R hi = orig[0].x * (R(1) / sqrt(R(1) - dptw) - R(1));
        if (v.x > hi || v.x < lo) {
            df[0].x += -v.x;
            v.x = R(0);
        }
        return;
    }
    const cplx<R> o = orig[idx];
    const R omag = o.x * o.x + o.y * o.y;
    const R cr = o.x + v.x, ci = o.y + v.y;
    const R cmag = cr * cr + ci * ci;
    const R ratio = cmag / omag;
    R d;
    if (corr) {
        if (ratio <= R(1) + dptw && ratio >= R(1) - dptw) return;
        d = fmin(R(1) - sqrt(R(1) - dptw), sqrt(R(1) + dptw) - R(1)) / sqrt(R(2)) * sqrt(omag);
    } else {
        if (ratio <= sqrt(R(1) + dptw) && ratio >= sqrt(R(1) - dptw)) return;
        d = fmin(R(1) - sqrt(R(1) - dptw), sqrt(R(1) + dptw) - R(1)) / (R)sqrtf(2.0f);
    }
    clamp_acc(v.x, d, &df[idx].x);
    clamp_acc(v.y, d, &df[idx].y);
}

// The frequency projection of one half-spectrum bin for the pass's mode
// (projection_algorithm.cu:147-188 abs, 93-145 per-bin, 190-259 ptw).
// Returns true when a clamp changed the bin.
template <typename R, typename T>
__device__ __forceinline__ bool freq_clamp_bin(cplx<R>& v, size_t idx, const SArgs<R, T>& a) {
    const cplx<R> v0 = v;
    if (a.fmode == FM_ABS) {
        clamp_acc(v.x, a.delta, &a.df[idx].x);
        clamp_acc(v.y, a.delta, &a.df[idx].y);
    } else if (a.fmode == FM_DENSE) {
        const R d = a.delta_dense[idx] * a.dscale;
        clamp_acc(v.x, d, &a.df[idx].x);
        clamp_acc(v.y, d, &a.df[idx].y);
    } else if (a.fmode == FM_PTW) {
        ptw_clamp(v, idx, a.orig, a.delta_ptw, a.df, a.ptw_mask, a.ptw_corr);
    }
    return v.x != v0.x || v.y != v0.y;
}

}  // namespace ffcz
