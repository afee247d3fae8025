// LDS-resident mixed-radix (2/3/4/5/8, and primes up to 31) Stockham FFT over
// a tile of pencils.
//
// A tile holds P pencils of length L in LDS; element (p, n) sits at
// buf[p*sp + n*sn].  Every stage is the self-sorting Stockham step (decimation
// in time, twiddle before butterfly): for butterfly j of radix R with the
// current sub-transform length Ns,
//     k = j mod Ns,  v_r = in[j + r L/R] * W_{Ns R}^{s k r}   (r = 0..R-1)
//     v = DFT_R(v);   out[(j-k) R + k + r Ns] = v_r
// (J. Stockham 1966; the GPU form follows E. Bainville, "OpenCL Fast Fourier
// Transform", 2010).  Inputs are read into registers, the workgroup barriers,
// and outputs are written back in place, so one LDS buffer suffices.
// Twiddles come from a per-length LDS table W_L^m, m < L, built on the host in
// FP64.  sign = -1: forward (cuFFT/numpy convention), +1: inverse (unscaled).
#pragma once

#include "ffcz_common.hpp"

namespace ffcz {

constexpr int FFT_MAX_STAGES = 16;
constexpr int FFT_THREADS = 256;     // threads per workgroup for every FFT kernel
constexpr int FFT_EPT = 16;          // max tile elements per thread (tile <= 4096 elements)
constexpr int FFT_MAX_TILE = FFT_THREADS * FFT_EPT;
constexpr int FFT_MAX_PRIME = 31;    // largest prime radix (direct DFT stage)

struct FftPlan {
    int L;                 // transform length
    int nst;               // number of stages
    int radix[FFT_MAX_STAGES];
    int Ns[FFT_MAX_STAGES];       // sub-transform length before the stage
    int twstride[FFT_MAX_STAGES]; // L / (Ns * R)
    FastDiv ns_div[FFT_MAX_STAGES];
    FastDiv lr_div[FFT_MAX_STAGES];  // L / R
};

// ---------------------------------------------------------------------------
// Radix-R DFT kernels on registers.  SIGN = -1 forward, +1 inverse.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ cplx<R> mul_si(cplx<R> a, int sign) {
    // a * (sign * i)
    return sign < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

template <typename R>
__device__ __forceinline__ void dft2(cplx<R>* v) {
    cplx<R> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <typename R>
__device__ __forceinline__ void dft3(cplx<R>* v, int sign) {
    const R c = R(-0.5);
    const R s = R(0.86602540378443864676372317075294);  // sqrt(3)/2
    cplx<R> t = v[1] + v[2];
    cplx<R> u = v[1] - v[2];
    cplx<R> a = mk(v[0].x + c * t.x, v[0].y + c * t.y);
    cplx<R> b = mul_si(scal(u, s), sign);
    v[0] = v[0] + t;
    v[1] = a + b;
    v[2] = a - b;
}

template <typename R>
__device__ __forceinline__ void dft4(cplx<R>* v, int sign) {
    cplx<R> t0 = v[0] + v[2];
    cplx<R> t1 = v[0] - v[2];
    cplx<R> t2 = v[1] + v[3];
    cplx<R> t3 = mul_si(v[1] - v[3], sign);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
}

template <typename R>
__device__ __forceinline__ void dft5(cplx<R>* v, int sign) {
    const R c1 = R(0.30901699437494742410229341718282);    // cos(2pi/5)
    const R c2 = R(-0.80901699437494742410229341718282);   // cos(4pi/5)
    const R s1 = R(0.95105651629515357211643933337938);    // sin(2pi/5)
    const R s2 = R(0.58778525229247312916870595463907);    // sin(4pi/5)
    cplx<R> t1 = v[1] + v[4], t2 = v[2] + v[3];
    cplx<R> u1 = v[1] - v[4], u2 = v[2] - v[3];
    cplx<R> a0 = v[0];
    cplx<R> a1 = mk(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
    cplx<R> a2 = mk(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
    cplx<R> b1 = mul_si(mk(s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y), sign);
    cplx<R> b2 = mul_si(mk(s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y), sign);
    v[0] = a0 + t1 + t2;
    v[1] = a1 + b1;
    v[4] = a1 - b1;
    v[2] = a2 + b2;
    v[3] = a2 - b2;
}

template <typename R>
__device__ __forceinline__ void dft8(cplx<R>* v, int sign) {
    const R h = R(0.70710678118654752440084436210485);  // 1/sqrt(2)
    cplx<R> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<R> o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e, sign);
    dft4(o, sign);
    // o[k] *= W8^{k}: W8 = (1 + sign*i)/sqrt(2)
    cplx<R> o1 = o[1], o3 = o[3];
    o[1] = sign < 0 ? mk(h * (o1.x + o1.y), h * (o1.y - o1.x)) : mk(h * (o1.x - o1.y), h * (o1.y + o1.x));
    o[2] = mul_si(o[2], sign);
    // W8^3 = (-1 + sign*i)/sqrt(2)
    o[3] = sign < 0 ? mk(h * (-o3.x + o3.y), h * (-o3.y - o3.x)) : mk(h * (-o3.x - o3.y), h * (-o3.y + o3.x));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = e[k] + o[k];
        v[k + 4] = e[k] - o[k];
    }
}

template <typename R, int RAD>
__device__ __forceinline__ void dftR(cplx<R>* v, int sign) {
    if constexpr (RAD == 2) dft2(v);
    else if constexpr (RAD == 3) dft3(v, sign);
    else if constexpr (RAD == 4) dft4(v, sign);
    else if constexpr (RAD == 5) dft5(v, sign);
    else if constexpr (RAD == 8) dft8(v, sign);
}

// One Stockham stage of radix RAD over the whole tile.
//   pfast: thread->butterfly mapping with the pencil index fastest (use when
//   sn == 1 is false, i.e. element-major layouts); otherwise j fastest.
template <typename R, int RAD>
__device__ __forceinline__ void fft_stage(cplx<R>* buf, const cplx<R>* tw, int L, int P, int sp, int sn,
                                          int Ns, int twstride, const FastDiv& nsdiv,
                                          const FastDiv& pdiv, const FastDiv& lrdiv, bool pfast, int sign) {
    constexpr int MB = (FFT_EPT + RAD - 1) / RAD;
    const int T = FFT_THREADS;
    const int t = threadIdx.x;
    const int LR = L / RAD;
    const int nb = LR * P;
    cplx<R> v[MB][RAD];
    int obase[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int bi = t + m * T;
        if (bi < nb) {
            int p, j;
            if (pfast) {
                j = (int)fdiv((uint32_t)bi, pdiv);
                p = bi - j * P;
            } else {
                p = (int)fdiv((uint32_t)bi, lrdiv);
                j = bi - p * LR;
            }
            const int q = (int)fdiv((uint32_t)j, nsdiv);
            const int k = j - q * Ns;
            const int ib = p * sp + j * sn;
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[m][r] = buf[ib + r * LR * sn];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r) {
                    cplx<R> w = tw[k * r * twstride];
                    if (sign > 0) w.y = -w.y;
                    v[m][r] = cmul(v[m][r], w);
                }
            }
            dftR<R, RAD>(v[m], sign);
            obase[m] = p * sp + (q * Ns * RAD + k) * sn;
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int bi = t + m * T;
        if (bi < nb) {
#pragma unroll
            for (int r = 0; r < RAD; ++r) buf[obase[m] + r * Ns * sn] = v[m][r];
        }
    }
    __syncthreads();
}

// One Stockham stage of a prime radix 7 <= rad <= FFT_MAX_PRIME (a factor the
// fixed radices leave): the butterfly's DFT is evaluated directly,
//     out_m = sum_r v_r W_rad^(-s r m),   W_rad^q = W_L^(q L/rad) from the table,
// rad^2 complex products per butterfly.  One butterfly per thread: the plan
// keeps (L/rad) * P <= FFT_THREADS, so every input is in a register before the
// barrier and each output is formed and stored after it.
template <typename R>
__device__ __forceinline__ void fft_stage_prime(cplx<R>* buf, const cplx<R>* tw, int L, int P, int sp, int sn, int rad,
                                                 int Ns, int twstride, const FastDiv& nsdiv, const FastDiv& pdiv,
                                                 const FastDiv& lrdiv, bool pfast, int sign) {
    const int t = threadIdx.x;
    const int LR = L / rad;
    const bool on = t < LR * P;
    cplx<R> v[FFT_MAX_PRIME];
    int obase = 0;
    if (on) {
        int p, j;
        if (pfast) {
            j = (int)fdiv((uint32_t)t, pdiv);
            p = t - j * P;
        } else {
            p = (int)fdiv((uint32_t)t, lrdiv);
            j = t - p * LR;
        }
        const int q = (int)fdiv((uint32_t)j, nsdiv);
        const int k = j - q * Ns;
        const int ib = p * sp + j * sn;
#pragma unroll
        for (int r = 0; r < FFT_MAX_PRIME; ++r) {
            if (r < rad) {
                cplx<R> x = buf[ib + r * LR * sn];
                if (r > 0 && Ns > 1) {
                    cplx<R> w = tw[k * r * twstride];
                    if (sign > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[r] = x;
            }
        }
        obase = p * sp + (q * Ns * rad + k) * sn;
    }
    __syncthreads();
    if (on) {
        for (int m = 0; m < rad; ++m) {
            cplx<R> acc = v[0];
            int e = 0;  // r * m mod rad, stepped
#pragma unroll
            for (int r = 1; r < FFT_MAX_PRIME; ++r) {
                if (r < rad) {
                    e += m;
                    if (e >= rad) e -= rad;
                    cplx<R> w = tw[e * LR];
                    if (sign > 0) w.y = -w.y;
                    acc = acc + cmul(v[r], w);
                }
            }
            buf[obase + m * Ns * sn] = acc;
        }
    }
    __syncthreads();
}

// Full transform of all P pencils of the tile (caller has synchronised after
// filling buf and tw).  Leaves the result in natural order in buf; ends with a
// barrier.
template <typename R>
__device__ void tile_fft(cplx<R>* buf, const cplx<R>* tw, const FftPlan& pl, int P, const FastDiv& pdiv,
                         int sp, int sn, bool pfast, int sign) {
    for (int s = 0; s < pl.nst; ++s) {
        const int rad = pl.radix[s];
        const FastDiv& lrdiv = pl.lr_div[s];
        switch (rad) {
            case 8: fft_stage<R, 8>(buf, tw, pl.L, P, sp, sn, pl.Ns[s], pl.twstride[s], pl.ns_div[s], pdiv, lrdiv, pfast, sign); break;
            case 4: fft_stage<R, 4>(buf, tw, pl.L, P, sp, sn, pl.Ns[s], pl.twstride[s], pl.ns_div[s], pdiv, lrdiv, pfast, sign); break;
            case 2: fft_stage<R, 2>(buf, tw, pl.L, P, sp, sn, pl.Ns[s], pl.twstride[s], pl.ns_div[s], pdiv, lrdiv, pfast, sign); break;
            case 5: fft_stage<R, 5>(buf, tw, pl.L, P, sp, sn, pl.Ns[s], pl.twstride[s], pl.ns_div[s], pdiv, lrdiv, pfast, sign); break;
            case 3: fft_stage<R, 3>(buf, tw, pl.L, P, sp, sn, pl.Ns[s], pl.twstride[s], pl.ns_div[s], pdiv, lrdiv, pfast, sign); break;
            default:
                fft_stage_prime<R>(buf, tw, pl.L, P, sp, sn, rad, pl.Ns[s], pl.twstride[s], pl.ns_div[s], pdiv, lrdiv,
                                   pfast, sign);
                break;
        }
    }
}

}  // namespace ffcz
