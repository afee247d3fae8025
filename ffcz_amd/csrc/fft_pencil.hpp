// Compile-time-specialised pencil FFT engine.
//
// For a length L known at compile time the radix sequence (8s, then 4 or 2,
// then 5s, then 3s — the same sequence as the run-time planner) and every
// stage constant are constexpr, so all index arithmetic folds to shifts and
// multiplies and every loop unrolls.  A pencil is transformed by TP threads;
// in stage s each thread owns ceil((L/R_s)/TP) butterflies of radix R_s in
// registers (Stockham DIT, twiddle before butterfly, as in fft_core.hpp).
// The first stage reads its inputs through a caller-supplied loader (straight
// from HBM, with the iteration's fused element-wise work), the last stage
// writes through a caller-supplied storer; only the exchanges between stages
// go through LDS.
#pragma once

#include "fft_core.hpp"

namespace ffcz {
namespace pen {

constexpr int fac_next(int L) {
    return L % 8 == 0 ? 8 : L % 4 == 0 ? 4 : L % 2 == 0 ? 2 : L % 5 == 0 ? 5 : L % 3 == 0 ? 3 : 0;
}
constexpr int nstages(int L) {
    int n = 0;
    while (L > 1) {
        const int r = fac_next(L);
        if (!r) return -1;
        L /= r;
        ++n;
    }
    return n;
}
constexpr int radix(int L, int s) {
    int r = 1;
    for (int i = 0; i <= s; ++i) {
        r = fac_next(L);
        L /= r;
    }
    return r;
}
constexpr int nsub(int L, int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) {
        const int r = fac_next(L);
        L /= r;
        ns *= r;
    }
    return ns;
}
// The radices in forward or reversed order.  A transform with the reversed
// order starts with the radix the normal order ends with, so its first stage
// reads exactly the positions (j + r L/R) the normal order's last stage
// writes: one transform can feed the other from registers.
constexpr int radix_o(int L, int s, bool rev) { return rev ? radix(L, nstages(L) - 1 - s) : radix(L, s); }
constexpr int nsub_o(int L, int s, bool rev) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= radix_o(L, i, rev);
    return ns;
}
// whether the reversed radix order is a different plan (not a palindrome)
constexpr bool rev_differs(int L) {
    const int n = nstages(L);
    for (int s = 0; s < n; ++s)
        if (radix(L, s) != radix(L, n - 1 - s)) return true;
    return false;
}
// threads per pencil: ~8 elements per thread up to L = 512, ~16 beyond
constexpr int threads_per_pencil(int L) { return L <= 1 ? 1 : (L <= 512 ? (L + 7) / 8 : (L + 15) / 16); }

// One Stockham stage of radix RAD with sub-transform length NS before it.
// BR: the loader also takes the (butterfly, input) position (b, r) of the
// element, both compile-time constants after unrolling, so a caller can serve
// the stage from values it loaded beforehand.
template <typename R, int L, int TP, int RAD, int NS, bool SYNC_MID, bool BR = false, class LD, class ST>
__device__ __forceinline__ void stage_rn(const LD& ld, const ST& st, const cplx<R>* __restrict__ tw, int t,
                                         int sign) {
    constexpr int LR = L / RAD;
    constexpr int NB = (LR + TP - 1) / TP;
    constexpr int TWS = L / (NS * RAD);
    cplx<R> v[NB][RAD];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + b * TP;
        if ((LR % TP == 0) || j < LR) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) {
                if constexpr (BR) v[b][r] = ld(j + r * LR, b, r);
                else v[b][r] = ld(j + r * LR);
            }
            if constexpr (NS > 1) {
                const int k = j % NS;
#pragma unroll
                for (int r = 1; r < RAD; ++r) {
                    cplx<R> w = tw[k * r * TWS];
                    w.y = sign > 0 ? -w.y : w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dftR<R, RAD>(v[b], sign);
        }
    }
    if constexpr (SYNC_MID) __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + b * TP;
        if ((LR % TP == 0) || j < LR) {
            const int k = j % NS;
            const int base = (j - k) * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) st(base + r * NS, v[b][r]);
        }
    }
}

template <typename R, int L, int TP, int S, bool SYNC_MID, bool BR = false, bool REV = false, class LD, class ST>
__device__ __forceinline__ void run_stage(const LD& ld, const ST& st, const cplx<R>* __restrict__ tw, int t,
                                          int sign) {
    stage_rn<R, L, TP, radix_o(L, S, REV), nsub_o(L, S, REV), SYNC_MID, BR>(ld, st, tw, t, sign);
}

// Stage S (radix order REV) split in two: its loads, twiddles and
// butterflies into registers, and its stores from them.
template <int L, int TP, int S, bool REV>
struct StageDim {
    static constexpr int RAD = radix_o(L, S, REV), NS = nsub_o(L, S, REV), LR = L / RAD;
    static constexpr int NB = (LR + TP - 1) / TP, TWS = L / (NS * RAD);
};
template <typename R, int L, int TP, int S, bool REV, bool BR = false, class LD>
__device__ __forceinline__ void stage_compute(const LD& ld,
                                              cplx<R> (&v)[StageDim<L, TP, S, REV>::NB][StageDim<L, TP, S, REV>::RAD],
                                              const cplx<R>* __restrict__ tw, int t, int sign) {
    using D = StageDim<L, TP, S, REV>;
#pragma unroll
    for (int b = 0; b < D::NB; ++b) {
        const int j = t + b * TP;
        if ((D::LR % TP == 0) || j < D::LR) {
#pragma unroll
            for (int r = 0; r < D::RAD; ++r) {
                if constexpr (BR) v[b][r] = ld(j + r * D::LR, b, r);
                else v[b][r] = ld(j + r * D::LR);
            }
            if constexpr (D::NS > 1) {
                const int k = j % D::NS;
#pragma unroll
                for (int r = 1; r < D::RAD; ++r) {
                    cplx<R> w = tw[k * r * D::TWS];
                    w.y = sign > 0 ? -w.y : w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dftR<R, D::RAD>(v[b], sign);
        }
    }
}
template <typename R, int L, int TP, int S, bool REV, class ST>
__device__ __forceinline__ void stage_store(const ST& st,
                                            const cplx<R> (&v)[StageDim<L, TP, S, REV>::NB][StageDim<L, TP, S, REV>::RAD],
                                            int t) {
    using D = StageDim<L, TP, S, REV>;
#pragma unroll
    for (int b = 0; b < D::NB; ++b) {
        const int j = t + b * TP;
        if ((D::LR % TP == 0) || j < D::LR) {
            const int k = j % D::NS;
            const int base = (j - k) * D::RAD + k;
#pragma unroll
            for (int r = 0; r < D::RAD; ++r) st(base + r * D::NS, v[b][r]);
        }
    }
}

// The first stage's loads of thread t: position (b, r) is element
// t + b*TP + r*L/R0 (R0 the first radix) when b < first_nb<L, TP>() and that
// element's butterfly exists (first_has<L, TP>(t, b)).
template <int L, int TP>
constexpr int first_nb() { return (L / radix(L, 0) + TP - 1) / TP; }
template <int L, int TP>
__device__ __forceinline__ bool first_has(int t, int b) {
    constexpr int LR = L / radix(L, 0);
    return (LR % TP == 0) || t + b * TP < LR;
}
template <int L>
__device__ __forceinline__ int first_elem(int t, int b, int r, int TP) {
    return t + b * TP + r * (L / radix(L, 0));
}

// Full transform of one pencil.  gld/gst: first-stage loader and last-stage
// storer; lld/lst: LDS exchange between stages.  IN_LDS / OUT_LDS say that
// gld / gst themselves address the exchange buffer (so the stage must barrier
// between its reads and its writes).  Returns after the last stage's stores
// without a trailing barrier.
// BR0: the first stage's loader takes (n, b, r) (stage_rn's BR).
// REV: the radices in reversed order.  SE >= 0 stops before stage SE (its
// last stage run then stores through lst and ends with a barrier, like a
// middle stage), for a caller that runs the rest itself.
template <typename R, int L, int TP, bool IN_LDS, bool OUT_LDS, bool BR0 = false, int S = 0, bool REV = false,
          int SE = -1, class GLD, class GST, class LLD, class LST>
__device__ __forceinline__ void transform(const GLD& gld, const GST& gst, const LLD& lld, const LST& lst,
                                          const cplx<R>* __restrict__ tw, int t, int sign) {
    constexpr int NST = SE < 0 ? nstages(L) : SE;
    constexpr bool STOPS = SE >= 0 && SE < nstages(L);
    static_assert(NST >= 0, "length must be 2/3/5-smooth");
    if constexpr (NST == 0) {
        if (t == 0) {
            cplx<R> v;
            if constexpr (BR0) v = gld(0, 0, 0);
            else v = gld(0);
            if constexpr (IN_LDS && OUT_LDS) __syncthreads();
            gst(0, v);
        } else if constexpr (IN_LDS && OUT_LDS) {
            __syncthreads();
        }
    } else if constexpr (S < NST) {
        constexpr bool FIRST = (S == 0), LAST = (S == NST - 1) && !STOPS;
        if constexpr (FIRST && LAST) {
            run_stage<R, L, TP, S, IN_LDS && OUT_LDS, BR0, REV>(gld, gst, tw, t, sign);
        } else if constexpr (FIRST) {
            run_stage<R, L, TP, S, IN_LDS, BR0, REV>(gld, lst, tw, t, sign);
            __syncthreads();
        } else if constexpr (LAST) {
            run_stage<R, L, TP, S, OUT_LDS, false, REV>(lld, gst, tw, t, sign);
        } else {
            run_stage<R, L, TP, S, true, false, REV>(lld, lst, tw, t, sign);
            __syncthreads();
        }
        transform<R, L, TP, IN_LDS, OUT_LDS, BR0, S + 1, REV, SE>(gld, gst, lld, lst, tw, t, sign);
    }
}

}  // namespace pen
}  // namespace ffcz
