// 3D R2C/C2R passes, the fused projection iteration and the FP64
// certification transform.  See fft3d.hpp for the layouts.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fft3d.hpp"
#include "fft_kernels_common.hpp"
#include "fft_fast.hpp"

namespace ffcz {

// ===========================================================================
// Host planning
// ===========================================================================
static FftPlan make_plan(int L) {
    FFCZ_REQUIRE(L >= 1, ERR_INVALID, "FFT length must be >= 1");
    FftPlan p{};
    p.L = L;
    p.nst = 0;
    int rem = L, Ns = 1;
    auto push = [&](int r) {
        FFCZ_REQUIRE(p.nst < FFT_MAX_STAGES, ERR_INVALID, "FFT length has too many factors");
        p.radix[p.nst] = r;
        p.Ns[p.nst] = Ns;
        p.twstride[p.nst] = L / (Ns * r);
        p.ns_div[p.nst] = FastDiv((uint32_t)Ns);
        p.lr_div[p.nst] = FastDiv((uint32_t)(L / r));
        ++p.nst;
        Ns *= r;
    };
    while (rem % 8 == 0) { push(8); rem /= 8; }
    while (rem % 4 == 0) { push(4); rem /= 4; }
    while (rem % 2 == 0) { push(2); rem /= 2; }
    while (rem % 5 == 0) { push(5); rem /= 5; }
    while (rem % 3 == 0) { push(3); rem /= 3; }
    // primes up to FFT_MAX_PRIME (composites are gone by now): direct DFT stages
    for (int q = 7; q <= FFT_MAX_PRIME; q += 2)
        while (rem % q == 0) { push(q); rem /= q; }
    FFCZ_REQUIRE(rem == 1, ERR_INVALID,
                 "FFT axis length " + std::to_string(L) + " has a prime factor above " +
                     std::to_string(FFT_MAX_PRIME));
    return p;
}

// Butterflies per pencil of the plan's widest prime stage (0 without one): a
// prime stage runs one butterfly per thread, so a tile of P pencils needs
// span * P <= FFT_THREADS.
static int prime_span(const FftPlan& p) {
    int span = 0;
    for (int s = 0; s < p.nst; ++s)
        if (p.radix[s] > 5 && p.radix[s] != 8) span = std::max(span, p.L / p.radix[s]);
    return span;
}

template <typename R>
static cplx<R>* upload_table(int L, int count, bool half_mirror = false) {
    std::vector<cplx<R>> h((size_t)count);
    const double two_pi = 6.283185307179586476925286766559;
    for (int m = 0; m < count; ++m) {
        double a = -two_pi * (double)m / (double)L;
        h[m].x = (R)std::cos(a);
        h[m].y = (R)std::sin(a);
    }
    if (half_mirror) {
        // W_L^(M-k) = -conj(W_L^k) with M = L/2: the upper half is stored as
        // the exact mirror of the lower, so kernels that keep only k <= M/2
        // and derive the rest see the same values as the table
        const int M = L / 2;
        for (int m = M / 2 + 1; m < count; ++m) {
            h[m].x = -h[M - m].x;
            h[m].y = h[M - m].y;
        }
    }
    cplx<R>* d = nullptr;
    FFCZ_HIP(hipMalloc(&d, sizeof(cplx<R>) * (size_t)count));
    FFCZ_HIP(hipMemcpy(d, h.data(), sizeof(cplx<R>) * (size_t)count, hipMemcpyHostToDevice));
    return d;
}

constexpr size_t LDS_BUDGET = 48 * 1024;

template <typename R>
void Fft3d<R>::create(size_t nx_, size_t ny_, size_t nz_) {
    FFCZ_REQUIRE(nx_ >= 1 && ny_ >= 1 && nz_ >= 1, ERR_INVALID, "field dimensions must be >= 1");
    FFCZ_REQUIRE(nx_ <= 65536 && ny_ <= 65536 && nz_ <= 131072, ERR_INVALID, "field dimension too large");
    nx = nx_; ny = ny_; nz = nz_;
    N = nx * ny * nz;
    rows = nx * ny;
    Nh = nz / 2 + 1;
    F = rows * Nh;
    FFCZ_REQUIRE(2 * F < (1ull << 40) && N < (1ull << 40), ERR_INVALID, "field too large");
    half = (nz % 2 == 0);
    Lz = half ? (int)(nz / 2) : (int)nz;
    Mp = (Nh + 15) / 16 * 16;
    pz = make_plan(Lz);
    py = make_plan((int)ny);
    px = make_plan((int)nx);
    const size_t cs = sizeof(cplx<R>);
    // z tile: P pencils, element-major with pencil stride P+1 (odd => no LDS bank aliasing)
    const size_t zrows = half ? (size_t)Lz + 1 : nz;
    const size_t ztab = ((size_t)Lz + nz / 2 + 1) * cs;
    Pz = 0;
    const size_t span_z = (size_t)prime_span(pz);
    for (int P = 256; P >= 1; P >>= 1) {
        if ((size_t)Lz * P <= (size_t)FFT_MAX_TILE && zrows * (P + 1) * cs + ztab <= LDS_BUDGET &&
            span_z * P <= (size_t)FFT_THREADS) { Pz = P; break; }
    }
    FFCZ_REQUIRE(Pz >= 1, ERR_INVALID, "z axis too long for one LDS tile");
    auto pickW = [&](size_t L, const FftPlan& pl) {
        const size_t span = (size_t)prime_span(pl);
        for (int W = 16; W >= 1; W >>= 1)
            if (L * W <= (size_t)FFT_MAX_TILE && (L * W + L) * cs <= LDS_BUDGET && span * W <= (size_t)FFT_THREADS)
                return W;
        return 0;
    };
    Wy = pickW(ny, py);
    Wx = pickW(nx, px);
    FFCZ_REQUIRE(Wy >= 1 && Wx >= 1, ERR_INVALID, "x/y axis too long for one LDS tile");
    d_tz = upload_table<R>(Lz, Lz);
    d_tz2 = upload_table<R>((int)nz, (int)(nz / 2 + 1), nz % 2 == 0);
    d_ty = upload_table<R>((int)ny, (int)ny);
    d_tx = upload_table<R>((int)nx, (int)nx);
}

template <typename R>
void Fft3d<R>::destroy() {
    if (d_tz) (void)hipFree(d_tz);
    if (d_tz2) (void)hipFree(d_tz2);
    if (d_ty) (void)hipFree(d_ty);
    if (d_tx) (void)hipFree(d_tx);
    d_tz = d_tz2 = d_ty = d_tx = nullptr;
}

// ===========================================================================
// Device helpers
// ===========================================================================
// ===========================================================================
// z pass kernels (contiguous rows)
// ===========================================================================
template <typename R, typename T, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_zfwd(ZArgs<R, T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (MODE == Z_ITER && a.st->done) return;
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem);
    cplx<R>* tw2 = tw + a.Lz;
    cplx<R>* buf = tw2 + (a.nz / 2 + 1);
    const int t = threadIdx.x;
    for (int i = t; i < a.Lz; i += FFT_THREADS) tw[i] = a.tz[i];
    for (int i = t; i <= a.nz / 2; i += FFT_THREADS) tw2[i] = a.tz2[i];
    const size_t row0 = (size_t)blockIdx.x * a.P;
    const int P = a.P, sn = a.sn;
    const bool clamp = (MODE == Z_ITER) && a.st->iter > 0;
    const bool fresh = clamp && a.ds_fresh && a.st->iter == 1;
    // the first iteration reads the initial error itself (e is written by the inverse pass)
    const R* esrc = (MODE == Z_ITER && !clamp && a.e_init) ? a.e_init : a.e;
    R* ecopy = (MODE == Z_ITER && !clamp && a.e_init) ? a.e_copy : nullptr;
    unsigned long long nviol = 0;
    double vmax = 0.0;
    auto load = [&](size_t gi) -> R {
        if constexpr (MODE == Z_PLAIN) {
            return a.in[gi];
        } else if constexpr (MODE == Z_ITER) {
            R v = esrc[gi];
            if (ecopy) ecopy[gi] = v;
            if (fresh) a.ds[gi] = clamp_edit(v, a.eps);
            else if (clamp) clamp_acc(v, a.eps, &a.ds[gi]);
            return v;
        } else {
            double d = (double)a.e0[gi];
            if (a.dfi) d += a.dfi[gi];
            d += dequant(a.cds[gi], a.qs);
            const double ex = fabs(d) - a.ceps;
            if (ex > 0.0) {
                ++nviol;
                vmax = ex > vmax ? ex : vmax;
            }
            return (R)d;
        }
    };
    if (a.half) {
        const int M = a.M;
        const FastDiv& md = a.dM;
        for (int q = t; q < P * M; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, md);
            const int n = q - p * M;
            const size_t row = row0 + p;
            R v0 = 0, v1 = 0;
            if (row < a.rows) {
                const size_t gi = row * (size_t)a.nz + 2 * (size_t)n;
                v0 = load(gi);
                v1 = load(gi + 1);
            }
            buf[n * sn + p] = mk(v0, v1);
        }
    } else {
        const FastDiv& zd = a.dNz;
        for (int q = t; q < P * a.nz; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, zd);
            const int n = q - p * a.nz;
            const size_t row = row0 + p;
            R v0 = 0;
            if (row < a.rows) v0 = load(row * (size_t)a.nz + n);
            buf[n * sn + p] = mk(v0, R(0));
        }
    }
    if constexpr (MODE == Z_CERT) {
        block_count_atomic(nviol, &a.stats->spatial_viol);
        block_max_atomic(vmax, &a.stats->spatial_max_bits);
    }
    __syncthreads();
    tile_fft(buf, tw, a.pl, P, a.dP, 1, sn, true, -1);
    if (a.half) {
        const int M = a.M;
        const int nh = M + 1;
        const FastDiv& hd = a.dNh;
        for (int q = t; q < P * nh; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, hd);
            const int k = q - p * nh;
            const size_t row = row0 + p;
            if (row >= a.rows) continue;
            const int ik = (k == M) ? 0 : k;
            const int im = (k == 0) ? 0 : M - k;
            const cplx<R> zk = buf[ik * sn + p];
            const cplx<R> zm = buf[im * sn + p];
            // X[k] = A[k] + W_nz^k B[k], A = (Zk + conj Zm)/2, B = (Zk - conj Zm)/(2i)
            const cplx<R> A = mk(R(0.5) * (zk.x + zm.x), R(0.5) * (zk.y - zm.y));
            const cplx<R> B = mk(R(0.5) * (zk.y + zm.y), R(0.5) * (zm.x - zk.x));
            a.spec[row * a.Mp + k] = A + cmul(tw2[k], B);
        }
    } else {
        const int nh = a.nz / 2 + 1;
        const FastDiv& hd = a.dNh;
        for (int q = t; q < P * nh; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, hd);
            const int k = q - p * nh;
            const size_t row = row0 + p;
            if (row >= a.rows) continue;
            a.spec[row * a.Mp + k] = buf[k * sn + p];
        }
    }
}

template <typename R, typename T, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_zinv(ZArgs<R, T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (MODE == Z_ITER && a.st->done) return;
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem);
    cplx<R>* tw2 = tw + a.Lz;
    cplx<R>* buf = tw2 + (a.nz / 2 + 1);
    const int t = threadIdx.x;
    for (int i = t; i < a.Lz; i += FFT_THREADS) tw[i] = a.tz[i];
    for (int i = t; i <= a.nz / 2; i += FFT_THREADS) tw2[i] = a.tz2[i];
    const size_t row0 = (size_t)blockIdx.x * a.P;
    const int P = a.P, sn = a.sn;
    const int nh = a.nz / 2 + 1;
    {
        const FastDiv& hd = a.dNh;
        for (int q = t; q < P * nh; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, hd);
            const int k = q - p * nh;
            const size_t row = row0 + p;
            cplx<R> v = row < a.rows ? a.spec[row * a.Mp + k] : mk(R(0), R(0));
            // a C2R reads only the real part of a row's DC and Nyquist bins
            // (the Hermitian part of the half spectrum, as pocketfft/FFTW do)
            if (k == 0 || (a.half && k == a.M)) v.y = R(0);
            buf[k * sn + p] = v;
        }
    }
    __syncthreads();
    if (a.half) {
        // Z[k] = (X[k] + conj X[M-k]) + i W_nz^{-k} (X[k] - conj X[M-k]); the
        // inverse length-M transform of Z gives nz*(x[2n] + i x[2n+1]).
        const int M = a.M;
        const int np = M / 2 + 1;
        const FastDiv& pd = a.dNp;
        for (int q = t; q < P * np; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, pd);
            const int k = q - p * np;
            const cplx<R> xk = buf[k * sn + p];
            const cplx<R> xm = buf[(M - k) * sn + p];
            {
                const cplx<R> s = xk + conjg(xm);
                const cplx<R> d = cmul(xk - conjg(xm), conjg(tw2[k]));
                buf[k * sn + p] = mk(s.x - d.y, s.y + d.x);
            }
            if (k != 0 && M - k != k) {
                const cplx<R> s = xm + conjg(xk);
                const cplx<R> d = cmul(xm - conjg(xk), conjg(tw2[M - k]));
                buf[(M - k) * sn + p] = mk(s.x - d.y, s.y + d.x);
            }
        }
    } else {
        const int nz = a.nz;
        const int nfill = nz - nh;
        if (nfill > 0) {
            const FastDiv& fd = a.dFill;
            for (int q = t; q < P * nfill; q += FFT_THREADS) {
                const int p = (int)fdiv((uint32_t)q, fd);
                const int k = nh + (q - p * nfill);
                buf[k * sn + p] = conjg(buf[(nz - k) * sn + p]);
            }
        }
    }
    __syncthreads();
    tile_fft(buf, tw, a.pl, P, a.dP, 1, sn, true, +1);
    R vmax = R(0);
    if (a.half) {
        const int M = a.M;
        const FastDiv& md = a.dM;
        for (int q = t; q < P * M; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, md);
            const int n = q - p * M;
            const size_t row = row0 + p;
            if (row >= a.rows) continue;
            const cplx<R> z = buf[n * sn + p];
            const size_t gi = row * (size_t)a.nz + 2 * (size_t)n;
            if constexpr (MODE == Z_ITER) {
                const R v0 = z.x * a.scale, v1 = z.y * a.scale;
                a.e[gi] = v0;
                a.e[gi + 1] = v1;
                const R a0 = fabs(v0), a1 = fabs(v1);
                if (a0 > a.eps) vmax = fmax(vmax, a0 - a.eps);
                if (a1 > a.eps) vmax = fmax(vmax, a1 - a.eps);
            } else {
                a.out[gi] = z.x * a.scale;
                a.out[gi + 1] = z.y * a.scale;
            }
        }
    } else {
        const FastDiv& zd = a.dNz;
        for (int q = t; q < P * a.nz; q += FFT_THREADS) {
            const int p = (int)fdiv((uint32_t)q, zd);
            const int n = q - p * a.nz;
            const size_t row = row0 + p;
            if (row >= a.rows) continue;
            const R v0 = buf[n * sn + p].x * a.scale;
            const size_t gi = row * (size_t)a.nz + n;
            if constexpr (MODE == Z_ITER) {
                a.e[gi] = v0;
                const R a0 = fabs(v0);
                if (a0 > a.eps) vmax = fmax(vmax, a0 - a.eps);
            } else {
                a.out[gi] = v0;
            }
        }
    }
    if constexpr (MODE == Z_ITER) {
        // convergence metric (projection_algorithm.cu:586-612): one partial
        // per workgroup, reduced by k_finalize_iteration
        block_max_store<R, FFT_THREADS>(vmax, &a.vpart[a.row_off / a.P + blockIdx.x]);
    }
}

// ===========================================================================
// strided (y / x) pass kernels
// ===========================================================================
template <typename R, typename T, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_strided(SArgs<R, T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (a.st && a.st->done) return;
    cplx<R>* tw = reinterpret_cast<cplx<R>*>(smem);
    cplx<R>* buf = tw + a.L;
    const int t = threadIdx.x;
    for (int i = t; i < a.L; i += FFT_THREADS) tw[i] = a.tw[i];
    const uint32_t ob_ = fdiv(blockIdx.x, a.dNct);
    const int ct = (int)(blockIdx.x - ob_ * (uint32_t)a.nct);
    const size_t o = ob_;
    const int c0 = ct * a.W;
    const int W = a.W, L = a.L, ws = a.wshift;
    const size_t base = o * a.ob + c0;
    for (int q = t; q < L * W; q += FFT_THREADS) {
        const int n = q >> ws, c = q & (W - 1);
        cplx<R> v = mk(R(0), R(0));
        if (c0 + c < a.ncols) {
            if constexpr (MODE == S_LOAD_DF) {
                const size_t idx = (size_t)n * a.cs_n + o * a.cs_o + c0 + c;
                const cplx<T> d = a.dfsrc[idx];
                v = mk((R)dequant(d.x, a.qf), (R)dequant(d.y, a.qf));
            } else {
                v = a.data[base + (size_t)n * a.es + c];
            }
        }
        buf[q] = v;
    }
    __syncthreads();
    const int sign = (MODE == S_CLAMP || MODE == S_CERT_CHECK) ? -1 : a.sign;
    tile_fft(buf, tw, a.pl, W, a.dW, 1, W, true, sign);
    if constexpr (MODE == S_CLAMP) {
        bool fired = false;
        for (int q = t; q < L * W; q += FFT_THREADS) {
            const int n = q >> ws, c = q & (W - 1);
            if (c0 + c >= a.ncols) continue;
            const size_t idx = (size_t)n * a.cs_n + o * a.cs_o + c0 + c;
            cplx<R> v = buf[q];
            fired |= freq_clamp_bin(v, idx, a);
            buf[q] = v;
        }
        if (fired) a.st->freq_clamped = 1u;
        __syncthreads();
        tile_fft(buf, tw, a.pl, W, a.dW, 1, W, true, +1);
    }
    if constexpr (MODE == S_CERT_CHECK) {
        unsigned long long nv = 0;
        double vmax = 0.0;
        for (int q = t; q < L * W; q += FFT_THREADS) {
            const int n = q >> ws, c = q & (W - 1);
            if (c0 + c >= a.ncols) continue;
            const size_t idx = (size_t)n * a.cs_n + o * a.cs_o + c0 + c;
            const cplx<R> v = buf[q];
            const double d = a.cdelta_dense ? (double)a.cdelta_dense[idx] : a.cdelta;
            const double m = fmax(fabs((double)v.x), fabs((double)v.y));
            const double ex = m - d;
            if (ex > 0.0) {
                ++nv;
                vmax = ex > vmax ? ex : vmax;
            }
        }
        block_count_atomic(nv, &a.stats->freq_viol);
        block_max_atomic(vmax, &a.stats->freq_max_bits);
        return;
    }
    for (int q = t; q < L * W; q += FFT_THREADS) {
        const int n = q >> ws, c = q & (W - 1);
        if (c0 + c < a.ncols) a.data[base + (size_t)n * a.es + c] = buf[q];
    }
}

// ===========================================================================
// small kernels
// ===========================================================================
// Reduces the inverse z pass's per-workgroup violations and finalises the
// iteration (projection_algorithm_impl.cuh:741-770 and 841-848, on the device).
template <typename R>
__global__ __launch_bounds__(1024) void k_finalize_iteration(const R* part, size_t n, DevState* st) {
    if (st->done) return;
    R m = R(0);
    for (size_t i = threadIdx.x; i < n; i += 1024) m = part[i] > m ? part[i] : m;
    __shared__ R red[16];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 16; ++i) m = red[i] > m ? red[i] : m;
        st->viol_bits = ordbits(red[0] > m ? red[0] : m);
        finalize_iteration<R>(st);
    }
}

__device__ __forceinline__ void reset_one(DevState* st, unsigned int max_iter, double tol) {
    st->done = max_iter == 0 ? 1u : 0u;
    st->converged = 0;
    st->freq_clamped = 0;
    st->iter = 0;
    st->max_iter = max_iter;
    st->blocks_done = 0;
    st->viol_bits = 0;
    st->tol = tol;
    st->tol_f = (float)tol;
    st->last_viol = 0.0;
}

__global__ void k_reset_state(DevState* st, unsigned int max_iter, double tol) {
    if (threadIdx.x == 0 && blockIdx.x == 0) reset_one(st, max_iter, tol);
}

__global__ void k_reset_states(DevState* const* sts, size_t n, unsigned int max_iter, double tol) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        reset_one(sts[i], max_iter, tol);
}

template <typename R>
__global__ void k_epilogue_clamp(R* e, R* ds, R eps, size_t n, const DevState* st, int ds_fresh) {
    if (!(st->done && !st->converged && st->iter > 0)) return;
    // first clamp since the accumulators were reset: store ds whole
    const bool fresh = ds_fresh && st->iter == 1;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        R v = e[i];
        if (fresh) {
            ds[i] = clamp_edit(v, eps);
            e[i] = v;
        } else if (v > eps || v < -eps) {
            clamp_acc(v, eps, &ds[i]);
            e[i] = v;
        }
    }
}

template <typename R>
__global__ void k_spec_to_compact(const cplx<R>* spec, cplx<R>* comp, size_t rows, size_t Mp, size_t Nh) {
    const size_t total = rows * Nh;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / Nh, k = i - r * Nh;
        comp[i] = spec[r * Mp + k];
    }
}
template <typename R>
__global__ void k_compact_to_spec(const cplx<R>* comp, cplx<R>* spec, size_t rows, size_t Mp, size_t Nh) {
    const size_t total = rows * Nh;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / Nh, k = i - r * Nh;
        spec[r * Mp + k] = comp[i];
    }
}

__global__ void k_reset_stats(CertStats* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->spatial_viol = 0;
        s->freq_viol = 0;
        s->spatial_max_bits = 0;
        s->freq_max_bits = 0;
    }
}
__global__ void k_reset_stats2(CertStats* a, CertStats* b) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *a = CertStats{};
        *b = CertStats{};
    }
}

// ===========================================================================
// Host launchers
// ===========================================================================
static inline unsigned grid_for(size_t n, unsigned block = 256) {
    size_t g = (n + block - 1) / block;
    if (g > (1u << 20)) g = 1u << 20;
    if (g == 0) g = 1;
    return (unsigned)g;
}

template <typename R, typename T>
static ZArgs<R, T> zargs(const Fft3d<R>& f) {
    ZArgs<R, T> a{};
    a.nz = (int)f.nz;
    a.Lz = f.Lz;
    a.M = (int)(f.nz / 2);
    a.half = f.half ? 1 : 0;
    a.P = f.Pz;
    a.sn = f.Pz + 1;
    a.rows = f.rows;
    a.Mp = f.Mp;
    a.pl = f.pz;
    a.tz = f.d_tz;
    a.tz2 = f.d_tz2;
    a.prec = f.prec;
    const uint32_t M = (uint32_t)(f.nz / 2), nh = M + 1;
    a.dP = FastDiv((uint32_t)f.Pz);
    a.dM = FastDiv(M ? M : 1);
    a.dNh = FastDiv(nh);
    a.dNp = FastDiv(M / 2 + 1);
    a.dNz = FastDiv((uint32_t)f.nz);
    a.dFill = FastDiv((uint32_t)f.nz > nh ? (uint32_t)f.nz - nh : 1u);
    return a;
}
template <typename R>
static size_t zlds(const Fft3d<R>& f) {
    const size_t zrows = f.half ? (size_t)f.Lz + 1 : f.nz;
    return ((size_t)f.Lz + f.nz / 2 + 1 + zrows * (f.Pz + 1)) * sizeof(cplx<R>);
}
template <typename R>
static unsigned zgrid(const Fft3d<R>& f) { return (unsigned)ceil_div(f.rows, (size_t)f.Pz); }

template <typename R, typename T, int MODE>
static void launch_zfwd(const Fft3d<R>& f, const ZArgs<R, T>& z0, hipStream_t s) {
    ZArgs<R, T> z = z0;
    z.var = f.prec == 0 ? f.zf_var : 0;
    if (!(fast_fft_enabled() && fast_zfwd<R, T>(MODE, z, s)))
        hipLaunchKernelGGL((k_zfwd<R, T, MODE>), dim3(zgrid(f)), dim3(FFT_THREADS), zlds(f), s, z);
    FFCZ_HIP(hipGetLastError());
}
template <typename R, typename T, int MODE>
static void launch_zinv(const Fft3d<R>& f, const ZArgs<R, T>& z0, hipStream_t s) {
    ZArgs<R, T> z = z0;
    z.var = f.zinv_var();
    if (!(fast_fft_enabled() && fast_zinv<R, T>(MODE, z, s)))
        hipLaunchKernelGGL((k_zinv<R, T, MODE>), dim3(zgrid(f)), dim3(FFT_THREADS), zlds(f), s, z);
    FFCZ_HIP(hipGetLastError());
}

static int ilog2(int w) {
    int s = 0;
    while ((1 << s) < w) ++s;
    return s;
}

template <typename R, typename T>
static SArgs<R, T> sargs(const Fft3d<R>& f, cplx<R>* data, bool xaxis) {
    SArgs<R, T> a{};
    a.data = data;
    if (xaxis) {
        a.pl = f.px;
        a.tw = f.d_tx;
        a.L = (int)f.nx;
        a.W = f.Wx;
        a.ob = f.Mp;
        a.es = f.ny * f.Mp;
    } else {
        a.pl = f.py;
        a.tw = f.d_ty;
        a.L = (int)f.ny;
        a.W = f.Wy;
        a.ob = f.ny * f.Mp;
        a.es = f.Mp;
    }
    a.wshift = ilog2(a.W);
    a.ncols = (int)f.Nh;
    a.nct = (int)ceil_div(f.Nh, (size_t)a.W);
    a.dW = FastDiv((uint32_t)a.W);
    a.dNct = FastDiv((uint32_t)a.nct);
    a.F = f.F;
    a.cs_n = f.ny * f.Nh;
    a.cs_o = f.Nh;
    a.prec = f.prec;
    return a;
}
template <typename R, typename T>
static void launch_strided(const Fft3d<R>& f, const SArgs<R, T>& a0, int mode, bool xaxis, hipStream_t s,
                           size_t outer_override = 0) {
    const size_t outer = outer_override ? outer_override : (xaxis ? f.ny : f.nx);
    SArgs<R, T> a = a0;
    a.var = f.prec == 0 ? (xaxis ? f.x_var : f.y_var) : 0;
    if (fast_fft_enabled() && fast_strided<R, T>(mode, a, outer, s)) {
        FFCZ_HIP(hipGetLastError());
        return;
    }
    const unsigned grid = (unsigned)(outer * (size_t)a.nct);
    const size_t lds = ((size_t)a.L + (size_t)a.L * a.W) * sizeof(cplx<R>);
    switch (mode) {
        case S_PLAIN: hipLaunchKernelGGL((k_strided<R, T, S_PLAIN>), dim3(grid), dim3(FFT_THREADS), lds, s, a); break;
        case S_CLAMP:
        case S_CLAMP_PASS: hipLaunchKernelGGL((k_strided<R, T, S_CLAMP>), dim3(grid), dim3(FFT_THREADS), lds, s, a); break;
        case S_CERT_CHECK: hipLaunchKernelGGL((k_strided<R, T, S_CERT_CHECK>), dim3(grid), dim3(FFT_THREADS), lds, s, a); break;
        case S_LOAD_DF: hipLaunchKernelGGL((k_strided<R, T, S_LOAD_DF>), dim3(grid), dim3(FFT_THREADS), lds, s, a); break;
    }
    FFCZ_HIP(hipGetLastError());
}

template <typename R>
void fft_r2c(const Fft3d<R>& f, const R* in, cplx<R>* spec, hipStream_t s) {
    auto z = zargs<R, R>(f);
    z.in = in;
    z.spec = spec;
    launch_zfwd<R, R, Z_PLAIN>(f, z, s);
    auto y = sargs<R, R>(f, spec, false);
    y.sign = -1;
    launch_strided(f, y, S_PLAIN, false, s);
    auto x = sargs<R, R>(f, spec, true);
    x.sign = -1;
    launch_strided(f, x, S_PLAIN, true, s);
}

template <typename R>
void fft_c2r(const Fft3d<R>& f, cplx<R>* spec, R* out, R scale, hipStream_t s) {
    auto x = sargs<R, R>(f, spec, true);
    x.sign = +1;
    launch_strided(f, x, S_PLAIN, true, s);
    auto y = sargs<R, R>(f, spec, false);
    y.sign = +1;
    launch_strided(f, y, S_PLAIN, false, s);
    auto z = zargs<R, R>(f);
    z.spec = spec;
    z.out = out;
    z.scale = scale;
    launch_zinv<R, R, Z_PLAIN>(f, z, s);
}

template <typename R>
void spec_to_compact(const Fft3d<R>& f, const cplx<R>* spec, cplx<R>* compact, hipStream_t s) {
    hipLaunchKernelGGL((k_spec_to_compact<R>), dim3(grid_for(f.F)), dim3(256), 0, s, spec, compact, f.rows, f.Mp, f.Nh);
    FFCZ_HIP(hipGetLastError());
}
template <typename R>
void compact_to_spec(const Fft3d<R>& f, const cplx<R>* compact, cplx<R>* spec, hipStream_t s) {
    hipLaunchKernelGGL((k_compact_to_spec<R>), dim3(grid_for(f.F)), dim3(256), 0, s, compact, spec, f.rows, f.Mp, f.Nh);
    FFCZ_HIP(hipGetLastError());
}

template <typename R>
void reset_state(DevState* st, int max_iter, double tol, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(64), 0, s, st, (unsigned)(max_iter < 0 ? 0 : max_iter), tol);
    FFCZ_HIP(hipGetLastError());
}

// x-planes per slab: the forward z and y passes (and the inverse y and z
// passes) run slab by slab so the half spectrum a z pass just wrote is still
// in the 256 MB memory-side cache when the y pass reads it.  The slab row
// count is a multiple of 512 so every z-pass workgroup maps to one partial.
template <typename R>
static size_t slab_planes(const Fft3d<R>& f) {
    const long mb = tuning().slab_mb.load();
    if (mb <= 0) return f.nx;
    const size_t plane = f.ny * f.Mp * sizeof(cplx<R>);
    size_t xs = std::max<size_t>(1, ((size_t)mb << 20) / plane);
    while (xs < f.nx && (xs * f.ny) % 512 != 0) ++xs;
    return xs >= f.nx ? f.nx : xs;
}

// the fused x pass's mode: per-bin and ptw bounds clamp in a pass of their
// own over the vector-ALU tile (S_CLAMP_PASS); abs bounds, and the MFMA
// variants, clamp in the last stage
template <typename R>
static int x_clamp_mode(const Fft3d<R>& f, int fmode) {
    if (fmode == FM_NONE) return S_PLAIN;
    return (fmode == FM_DENSE || fmode == FM_PTW) && f.prec == PREC_VALU ? S_CLAMP_PASS : S_CLAMP;
}

template <typename R>
void iteration(const Fft3d<R>& f, const IterArgs<R>& ia, hipStream_t s, EventLog* log) {
    auto mark = [&](int kind) {
        if (log) log->mark(kind, s);
    };
    const size_t xs = slab_planes(f);
    mark(-1);
    auto zslab = [&](size_t x0, size_t x1) {
        auto z = zargs<R, R>(f);
        const size_t r0 = x0 * f.ny;
        z.e = ia.e + r0 * f.nz;
        z.e_init = ia.e_init ? ia.e_init + r0 * f.nz : nullptr;
        z.e_copy = ia.e_copy ? ia.e_copy + r0 * f.nz : nullptr;
        z.ds_fresh = ia.ds_fresh;
        z.ds = ia.ds + r0 * f.nz;
        z.eps = ia.eps;
        z.st = ia.st;
        z.spec = ia.spec + r0 * f.Mp;
        z.scale = ia.norm;
        z.vpart = ia.vpart;
        z.row_off = r0;
        z.rows = (x1 - x0) * f.ny;
        return z;
    };
    auto yslab = [&](size_t x0, size_t x1, int sign) {
        auto y = sargs<R, R>(f, ia.spec + x0 * f.ny * f.Mp, false);
        y.st = ia.st;
        y.sign = sign;
        launch_strided(f, y, S_PLAIN, false, s, x1 - x0);
    };
    for (size_t x0 = 0; x0 < f.nx; x0 += xs) {
        const size_t x1 = std::min(f.nx, x0 + xs);
        launch_zfwd<R, R, Z_ITER>(f, zslab(x0, x1), s);
        mark(0);
        yslab(x0, x1, -1);
        mark(1);
    }
    auto x = sargs<R, R>(f, ia.spec, true);
    x.st = ia.st;
    x.fmode = ia.fmode;
    x.delta = ia.delta;
    x.delta_dense = ia.delta_dense;
    x.dscale = ia.dscale;
    x.df = ia.df;
    x.orig = ia.orig;
    x.delta_ptw = ia.delta_ptw;
    x.ptw_mask = ia.ptw_mask;
    x.ptw_corr = ia.ptw_corr;
    launch_strided(f, x, x_clamp_mode(f, ia.fmode), true, s);
    mark(2);
    for (size_t x0 = 0; x0 < f.nx; x0 += xs) {
        const size_t x1 = std::min(f.nx, x0 + xs);
        yslab(x0, x1, +1);
        mark(3);
        launch_zinv<R, R, Z_ITER>(f, zslab(x0, x1), s);
        mark(4);
    }
    const int zr = fast_z_rows(f.Lz, f.half, f.zinv_var());
    const size_t nparts = ceil_div(f.rows, (size_t)(zr ? zr : f.Pz));
    hipLaunchKernelGGL((k_finalize_iteration<R>), dim3(1), dim3(1024), 0, s, (const R*)ia.vpart, nparts, ia.st);
    FFCZ_HIP(hipGetLastError());
    mark(5);
}

template <typename R>
bool fused_iteration_available(const Fft3d<R>& f) {
    return fast_fft_enabled() && tuning().slab_mb.load() <= 0 && fast_z_rows(f.Lz, f.half, f.zi_var) > 0;
}
template <typename R>
size_t zz_blocks(const Fft3d<R>& f) {
    const int zr = fast_z_rows(f.Lz, f.half, f.zi_var);
    return zr ? ceil_div(f.rows, (size_t)zr) : 0;
}
template <typename R>
size_t zz_log_entries(const Fft3d<R>& f) {
    // a segment of P rows' elements per workgroup, for either tile variant
    size_t n = 0;
    for (int v = 0; v < 2; ++v) {
        const int zr = fast_z_rows(f.Lz, f.half, v);
        if (zr) n = std::max(n, ceil_div(f.rows, (size_t)zr) * (size_t)zr * f.nz);
    }
    return n;
}
template <typename R>
size_t zz_blocks_max(const Fft3d<R>& f) {
    size_t n = 0;
    for (int v = 0; v < 2; ++v) {
        const int zr = fast_z_rows(f.Lz, f.half, v);
        if (zr) n = std::max(n, ceil_div(f.rows, (size_t)zr));
    }
    return n;
}

template <typename R>
void iteration_fused(const Fft3d<R>& f, const IterArgs<R>& ia, bool first, hipStream_t s, EventLog* log) {
    auto mark = [&](int kind) {
        if (log) log->mark(kind, s);
    };
    mark(-1);
    auto z = zargs<R, R>(f);
    z.e = ia.e;
    z.e_init = ia.e_init;
    z.e_copy = ia.e_copy;
    z.ds_fresh = ia.ds_fresh;
    z.ds = ia.ds;
    z.eps = ia.eps;
    z.st = ia.st;
    z.spec = ia.spec;
    z.scale = ia.norm;
    z.vpart = ia.vpart;
    z.row_off = 0;
    z.log_off = ia.log_off;
    z.log_v = ia.log_v;
    z.log_cnt = ia.log_cnt;
    if (first) {
        // the first iteration's forward z transform of the initial error
        launch_zfwd<R, R, Z_ITER>(f, z, s);
        mark(0);
    }
    auto y = sargs<R, R>(f, ia.spec, false);
    y.st = ia.st;
    y.sign = -1;
    launch_strided(f, y, S_PLAIN, false, s);
    mark(1);
    auto x = sargs<R, R>(f, ia.spec, true);
    x.st = ia.st;
    x.fmode = ia.fmode;
    x.delta = ia.delta;
    x.delta_dense = ia.delta_dense;
    x.dscale = ia.dscale;
    x.df = ia.df;
    x.orig = ia.orig;
    x.delta_ptw = ia.delta_ptw;
    x.ptw_mask = ia.ptw_mask;
    x.ptw_corr = ia.ptw_corr;
    launch_strided(f, x, x_clamp_mode(f, ia.fmode), true, s);
    mark(2);
    y.sign = +1;
    launch_strided(f, y, S_PLAIN, false, s);
    mark(3);
    z.var = f.zi_var;
    const bool launched = fast_zinv<R, R>(Z_FUSED, z, s);
    FFCZ_REQUIRE(launched, ERR_INVALID, "no fused z kernel for this z length");
    FFCZ_HIP(hipGetLastError());
    mark(4);
    hipLaunchKernelGGL((k_finalize_iteration<R>), dim3(1), dim3(1024), 0, s, (const R*)ia.vpart, zz_blocks(f), ia.st);
    FFCZ_HIP(hipGetLastError());
    mark(5);
}

// Adds the last z fused pass's deferred clamps to ds after a loop that
// stopped at max_iter without converging (a converged loop drops them).
// Counts and offsets are held to each workgroup's segment and to the field.
template <typename R>
__global__ void k_zz_apply(R* ds, R eps, const uint32_t* off, const R* v, const uint32_t* cnt, size_t nblk,
                           size_t seg, size_t n_elem, const DevState* st) {
    if (st->converged || st->iter == 0) return;
    for (size_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint32_t n = (uint32_t)min((size_t)cnt[b], seg);
        const size_t l0 = b * seg;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const size_t j = l0 + off[l0 + i];
            if (off[l0 + i] >= seg || j >= n_elem) continue;
            R x = v[l0 + i];
            const R ed = clamp_edit(x, eps);
            ds[j] = ds[j] + ed;
        }
    }
}
template <typename R>
__global__ void k_zz_patch_e(R* e, const uint32_t* off, const R* v, const uint32_t* cnt, size_t nblk, size_t seg,
                             size_t n_elem, const DevState* st) {
    if (!st->converged) return;
    for (size_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint32_t n = (uint32_t)min((size_t)cnt[b], seg);
        const size_t l0 = b * seg;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const size_t j = l0 + off[l0 + i];
            if (off[l0 + i] < seg && j < n_elem) e[j] = v[l0 + i];
        }
    }
}
template <typename R>
static size_t zz_seg(const Fft3d<R>& f, int var) {
    return (size_t)fast_z_rows(f.Lz, f.half, var) * f.nz;
}
template <typename R>
static size_t zz_nblk(const Fft3d<R>& f, int var) {
    return ceil_div(f.rows, (size_t)fast_z_rows(f.Lz, f.half, var));
}
template <typename R>
void zz_apply(const Fft3d<R>& f, const IterArgs<R>& a, int var, hipStream_t s) {
    const size_t nb = zz_nblk(f, var);
    const unsigned g = (unsigned)std::min<size_t>(nb, 4096);
    hipLaunchKernelGGL((k_zz_apply<R>), dim3(g), dim3(256), 0, s, a.ds, a.eps, (const uint32_t*)a.log_off,
                       (const R*)a.log_v, (const uint32_t*)a.log_cnt, nb, zz_seg(f, var), f.N,
                       (const DevState*)a.st);
    FFCZ_HIP(hipGetLastError());
}
template <typename R>
void zz_recover_e(const Fft3d<R>& f, const IterArgs<R>& a, int var, hipStream_t s) {
    // the rows hold the z transform of the clamped iterate: C2R along z
    // returns it times nz
    auto z = zargs<R, R>(f);
    z.spec = a.spec;
    z.out = a.e;
    z.scale = R(1) / static_cast<R>(f.nz);
    launch_zinv<R, R, Z_PLAIN>(f, z, s);
    const size_t nb = zz_nblk(f, var);
    const unsigned g = (unsigned)std::min<size_t>(nb, 4096);
    hipLaunchKernelGGL((k_zz_patch_e<R>), dim3(g), dim3(256), 0, s, a.e, (const uint32_t*)a.log_off,
                       (const R*)a.log_v, (const uint32_t*)a.log_cnt, nb, zz_seg(f, var), f.N,
                       (const DevState*)a.st);
    FFCZ_HIP(hipGetLastError());
}

template <typename R>
void iteration_kernel(const Fft3d<R>& f, const IterArgs<R>& ia, hipStream_t s, int kind) {
    auto z = zargs<R, R>(f);
    z.e = ia.e;
    z.e_init = ia.e_init;
    z.e_copy = ia.e_copy;
    z.ds_fresh = ia.ds_fresh;
    z.ds = ia.ds;
    z.eps = ia.eps;
    z.st = ia.st;
    z.spec = ia.spec;
    z.scale = ia.norm;
    z.vpart = ia.vpart;
    if (kind == 0) {
        launch_zfwd<R, R, Z_ITER>(f, z, s);
    } else if (kind == 1 || kind == 3) {
        auto y = sargs<R, R>(f, ia.spec, false);
        y.st = ia.st;
        y.sign = kind == 1 ? -1 : +1;
        launch_strided(f, y, S_PLAIN, false, s);
    } else if (kind == 2) {
        auto x = sargs<R, R>(f, ia.spec, true);
        x.st = ia.st;
        x.fmode = ia.fmode;
        x.delta = ia.delta;
        x.delta_dense = ia.delta_dense;
        x.dscale = ia.dscale;
        x.df = ia.df;
        x.orig = ia.orig;
        x.delta_ptw = ia.delta_ptw;
        x.ptw_mask = ia.ptw_mask;
        x.ptw_corr = ia.ptw_corr;
        launch_strided(f, x, x_clamp_mode(f, ia.fmode), true, s);
    } else {
        launch_zinv<R, R, Z_ITER>(f, z, s);
    }
}

template <typename R>
void epilogue_clamp(const Fft3d<R>& f, const IterArgs<R>& a, hipStream_t s) {
    // a small grid: the kernel usually returns at once (converged loop)
    const unsigned g = (unsigned)std::min<size_t>(ceil_div(f.N, (size_t)256), 2048);
    hipLaunchKernelGGL((k_epilogue_clamp<R>), dim3(g), dim3(256), 0, s, a.e, a.ds, a.eps, f.N,
                       (const DevState*)a.st, a.ds_fresh);
    FFCZ_HIP(hipGetLastError());
}

template <typename T>
void certify(const Fft3d<double>& f, const CertArgs<T>& c, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_stats, dim3(1), dim3(64), 0, s, c.stats);
    FFCZ_HIP(hipGetLastError());
    const double* dfi = nullptr;
    if (c.have_df) {
        // C2R(deq df)/N in FP64: x pass loads the split accumulator.
        auto x = sargs<double, T>(f, c.work_spec, true);
        x.sign = +1;
        x.dfsrc = c.df;
        x.qf = c.qf;
        launch_strided(f, x, S_LOAD_DF, true, s);
        auto y = sargs<double, T>(f, c.work_spec, false);
        y.sign = +1;
        launch_strided(f, y, S_PLAIN, false, s);
        auto z = zargs<double, T>(f);
        z.spec = c.work_spec;
        z.out = c.work_real;
        z.scale = 1.0 / (double)f.N;
        launch_zinv<double, T, Z_PLAIN>(f, z, s);
        dfi = c.work_real;
    }
    auto z = zargs<double, T>(f);
    z.spec = c.work_spec;
    z.e0 = c.e0;
    z.cds = c.ds;
    z.dfi = dfi;
    z.qs = c.qs;
    z.ceps = c.eps;
    z.stats = c.stats;
    launch_zfwd<double, T, Z_CERT>(f, z, s);
    auto y = sargs<double, T>(f, c.work_spec, false);
    y.sign = -1;
    launch_strided(f, y, S_PLAIN, false, s);
    auto x = sargs<double, T>(f, c.work_spec, true);
    x.cdelta = c.delta;
    x.cdelta_dense = c.delta_dense;
    x.stats = c.stats;
    launch_strided(f, x, S_CERT_CHECK, true, s);
}

// explicit instantiations
// ===========================================================================
// batched iteration
// ===========================================================================
template <typename R>
struct FinEntry {
    const R* vpart;
    size_t n;
    DevState* st;
};

// one workgroup per field: the reduction and bookkeeping of k_finalize_iteration
template <typename R>
__global__ __launch_bounds__(1024) void k_finalize_batch(const FinEntry<R>* t) {
    const FinEntry<R> e = t[blockIdx.x];
    if (e.st->done) return;
    R m = R(0);
    for (size_t i = threadIdx.x; i < e.n; i += 1024) m = e.vpart[i] > m ? e.vpart[i] : m;
    __shared__ R red[16];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 16; ++i) m = red[i] > m ? red[i] : m;
        e.st->viol_bits = ordbits(red[0] > m ? red[0] : m);
        finalize_iteration<R>(e.st);
    }
}

__global__ void k_gather_states(DevState* const* sts, size_t n, DevState* out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = *sts[i];
}

template <typename R>
static bool batchable(const Fft3d<R>& f) {
    if constexpr (sizeof(R) != 4) {
        return false;
    } else {
        return fast_fft_enabled() && f.prec == PREC_VALU && tuning().slab_mb.load() <= 0 && f.half &&
               fast_z_rows(f.Lz, f.half) > 0 && fast_s_supported((int)f.ny) && fast_s_supported((int)f.nx);
    }
}

template <typename R>
void batch_plan_free(BatchPlan<R>& bp) {
    if (bp.dmem) (void)hipFree(bp.dmem);
    bp = BatchPlan<R>{};
}

template <typename R>
void batch_plan_build(BatchPlan<R>& bp, const std::vector<const Fft3d<R>*>& fields,
                      const std::vector<IterArgs<R>>& args) {
    FFCZ_REQUIRE(fields.size() == args.size() && !fields.empty(), ERR_INVALID, "empty batch");
    unsigned char* keep = bp.dmem;
    const size_t keep_cap = bp.dcap;
    bp = BatchPlan<R>{};
    bp.dmem = keep;
    bp.dcap = keep_cap;
    bp.nfields = fields.size();
    // host staging of every device table; offsets are patched into pointers
    // once the device block exists
    std::vector<unsigned char> host;
    auto put = [&](const void* p, size_t bytes) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        std::memcpy(host.data() + off, p, bytes);
        return off;
    };
    struct Pending {
        typename BatchPlan<R>::Group g;
        size_t args_off, start_off;
    };
    std::vector<Pending> pend;
    std::vector<size_t> batched;
    for (size_t i = 0; i < fields.size(); ++i) {
        if (batchable(*fields[i])) batched.push_back(i);
        else bp.fallback.emplace_back(fields[i], args[i]);
    }
    if constexpr (sizeof(R) == 4) {
        // group the batched fields by the axis length of each pass
        auto group_by = [&](auto key) {
            std::vector<std::pair<int, std::vector<size_t>>> gs;
            for (size_t i : batched) {
                const int k = key(*fields[i]);
                auto it = std::find_if(gs.begin(), gs.end(), [&](auto& g) { return g.first == k; });
                if (it == gs.end()) gs.push_back({k, {i}});
                else it->second.push_back(i);
            }
            return gs;
        };
        auto zgroups = group_by([](const Fft3d<R>& f) { return f.Lz; });
        auto ygroups = group_by([](const Fft3d<R>& f) { return (int)f.ny; });
        auto xgroups = group_by([](const Fft3d<R>& f) { return (int)f.nx; });
        auto add_z = [&](int pass, const std::pair<int, std::vector<size_t>>& g) {
            std::vector<ZArgs<R, R>> tab;
            std::vector<uint32_t> start;
            uint64_t tiles = 0;
            for (size_t i : g.second) {
                const Fft3d<R>& f = *fields[i];
                const IterArgs<R>& ia = args[i];
                auto z = zargs<R, R>(f);
                z.e = ia.e;
                z.e_init = ia.e_init;
                z.e_copy = ia.e_copy;
                z.ds_fresh = ia.ds_fresh;
                z.ds = ia.ds;
                z.eps = ia.eps;
                z.st = ia.st;
                z.spec = ia.spec;
                z.scale = ia.norm;
                z.vpart = ia.vpart;
                z.row_off = 0;
                z.log_off = ia.log_off;
                z.log_v = ia.log_v;
                z.log_cnt = ia.log_cnt;
                tab.push_back(z);
                start.push_back((uint32_t)tiles);
                tiles += ceil_div(f.rows, (size_t)fast_z_rows(f.Lz, f.half, fast_zb_var(f.Lz)));
            }
            FFCZ_REQUIRE(tiles < (1ull << 31), ERR_INVALID, "batch too large for one launch");
            Pending p{};
            p.g.pass = pass;
            p.g.L = g.first;
            p.g.grid = (unsigned)tiles;
            p.g.n = (int)tab.size();
            p.args_off = put(tab.data(), tab.size() * sizeof(tab[0]));
            p.start_off = put(start.data(), start.size() * sizeof(uint32_t));
            pend.push_back(p);
        };
        auto add_s = [&](int pass, bool xaxis, const std::pair<int, std::vector<size_t>>& g) {
            std::vector<SArgs<R, R>> tab;
            std::vector<uint32_t> start;
            uint64_t tiles = 0;
            int smode = S_PLAIN;
            const int W = fast_s_width(sizeof(R));
            for (size_t i : g.second) {
                const Fft3d<R>& f = *fields[i];
                const IterArgs<R>& ia = args[i];
                auto a = sargs<R, R>(f, ia.spec, xaxis);
                a.st = ia.st;
                a.sign = pass == 1 ? -1 : +1;
                if (pass == 2) {
                    a.fmode = ia.fmode;
                    a.delta = ia.delta;
                    a.delta_dense = ia.delta_dense;
                    a.dscale = ia.dscale;
                    a.df = ia.df;
                    a.orig = ia.orig;
                    a.delta_ptw = ia.delta_ptw;
                    a.ptw_mask = ia.ptw_mask;
                    a.ptw_corr = ia.ptw_corr;
                    // one mode per launch: a field without frequency bounds still
                    // runs the clamp kernel, whose FM_NONE projection is a no-op
                    smode = S_CLAMP;
                }
                a.W = W;
                a.nct = (int)ceil_div((size_t)a.ncols, (size_t)W);
                a.dNct = FastDiv((uint32_t)a.nct);
                tab.push_back(a);
                start.push_back((uint32_t)tiles);
                tiles += (uint64_t)(xaxis ? f.ny : f.nx) * (uint64_t)a.nct;
            }
            FFCZ_REQUIRE(tiles < (1ull << 31), ERR_INVALID, "batch too large for one launch");
            Pending p{};
            p.g.pass = pass;
            p.g.L = g.first;
            p.g.smode = smode;
            p.g.grid = (unsigned)tiles;
            p.g.n = (int)tab.size();
            p.args_off = put(tab.data(), tab.size() * sizeof(tab[0]));
            p.start_off = put(start.data(), start.size() * sizeof(uint32_t));
            pend.push_back(p);
        };
        // the fused iteration form: the forward z pass of the initial error
        // once (the prologue), then [y fwd, x fused, y inv, z fused] per
        // iteration
        for (auto& g : zgroups) add_z(0, g);
        for (auto& g : ygroups) add_s(1, false, g);
        for (auto& g : xgroups) add_s(2, true, g);
        for (auto& g : ygroups) add_s(3, false, g);
        for (auto& g : zgroups) add_z(4, g);
    }
    std::vector<FinEntry<R>> fin;
    for (size_t i : batched) {
        const Fft3d<R>& f = *fields[i];
        fin.push_back({args[i].vpart, ceil_div(f.rows, (size_t)fast_z_rows(f.Lz, f.half, fast_zb_var(f.Lz))), args[i].st});
    }
    size_t fin_off = 0;
    if (!fin.empty()) fin_off = put(fin.data(), fin.size() * sizeof(fin[0]));
    std::vector<DevState*> sts;
    for (auto& a : args) sts.push_back(a.st);
    const size_t sts_off = put(sts.data(), sts.size() * sizeof(DevState*));
    const size_t done_off = put(std::vector<DevState>(sts.size()).data(), sts.size() * sizeof(DevState));
    if (host.size() > bp.dcap) {
        if (bp.dmem) FFCZ_HIP(hipFree(bp.dmem));
        bp.dmem = nullptr;
        bp.dcap = 0;
        FFCZ_HIP(hipMalloc(&bp.dmem, host.size()));
        bp.dcap = host.size();
    }
    FFCZ_HIP(hipMemcpy(bp.dmem, host.data(), host.size(), hipMemcpyHostToDevice));
    for (auto& p : pend) {
        p.g.args = bp.dmem + p.args_off;
        p.g.start = reinterpret_cast<const uint32_t*>(bp.dmem + p.start_off);
        (p.g.pass == 0 ? bp.pro : bp.groups).push_back(p.g);
    }
    bp.nfin = (int)fin.size();
    bp.fin = fin.empty() ? nullptr : bp.dmem + fin_off;
    bp.sts = reinterpret_cast<DevState**>(bp.dmem + sts_off);
    bp.gathered = reinterpret_cast<DevState*>(bp.dmem + done_off);
}

template <typename R>
bool batch_fusable(const Fft3d<R>& f) {
    return batchable(f);
}

template <typename R>
void batch_prologue(const BatchPlan<R>& bp, hipStream_t s) {
    if constexpr (sizeof(R) == 4) {
        for (const auto& g : bp.pro) {
            BatchTab<ZArgs<R, R>> t{static_cast<const ZArgs<R, R>*>(g.args), g.start, g.n};
            FFCZ_REQUIRE(fast_zfwd_batch(g.L, t, g.grid, s), ERR_STATE, "batched pass without a kernel for its length");
            FFCZ_HIP(hipGetLastError());
        }
    }
}

template <typename R>
void batch_iteration(const BatchPlan<R>& bp, hipStream_t s) {
    if constexpr (sizeof(R) == 4) {
        for (const auto& g : bp.groups) {
            bool ok;
            if (g.pass == 4) {
                BatchTab<ZArgs<R, R>> t{static_cast<const ZArgs<R, R>*>(g.args), g.start, g.n};
                ok = fast_zz_batch(g.L, t, g.grid, s);
            } else {
                BatchTab<SArgs<R, R>> t{static_cast<const SArgs<R, R>*>(g.args), g.start, g.n};
                ok = fast_strided_batch(g.smode, g.L, t, g.grid, s);
            }
            FFCZ_REQUIRE(ok, ERR_STATE, "batched pass without a kernel for its length");
            FFCZ_HIP(hipGetLastError());
        }
        if (bp.nfin) {
            hipLaunchKernelGGL((k_finalize_batch<R>), dim3(bp.nfin), dim3(1024), 0, s,
                               static_cast<const FinEntry<R>*>(bp.fin));
            FFCZ_HIP(hipGetLastError());
        }
    }
    for (const auto& fb : bp.fallback) iteration(*fb.first, fb.second, s);
}

template <typename R>
void batch_gather_states(const BatchPlan<R>& bp, DevState* out, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_states, dim3(1), dim3(256), 0, s, (DevState* const*)bp.sts, bp.nfields,
                       bp.gathered);
    FFCZ_HIP(hipGetLastError());
    FFCZ_HIP(hipMemcpyAsync(out, bp.gathered, bp.nfields * sizeof(DevState), hipMemcpyDeviceToHost, s));
}

template <typename R>
void batch_reset_states(const BatchPlan<R>& bp, int max_iter, double tol, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_states, dim3(1), dim3(256), 0, s, (DevState* const*)bp.sts, bp.nfields,
                       (unsigned)(max_iter < 0 ? 0 : max_iter), tol);
    FFCZ_HIP(hipGetLastError());
}

template <typename T>
static const double* cert_df_field(const Fft3d<double>& f, const CertArgs<T>& c, hipStream_t s) {
    // C2R(deq df)/N in FP64: x pass loads the split accumulator
    if (!c.have_df) return nullptr;
    auto x = sargs<double, T>(f, c.work_spec, true);
    x.sign = +1;
    x.dfsrc = c.df;
    x.qf = c.qf;
    launch_strided(f, x, S_LOAD_DF, true, s);
    auto y = sargs<double, T>(f, c.work_spec, false);
    y.sign = +1;
    launch_strided(f, y, S_PLAIN, false, s);
    auto z = zargs<double, T>(f);
    z.spec = c.work_spec;
    z.out = c.work_real;
    z.scale = 1.0 / (double)f.N;
    launch_zinv<double, T, Z_PLAIN>(f, z, s);
    return c.work_real;
}

template <typename T>
static ZArgs<double, T> cert_zargs(const Fft3d<double>& f, const CertArgs<T>& c, const double* dfi) {
    auto z = zargs<double, T>(f);
    z.spec = c.work_spec;
    z.e0 = c.e0;
    z.cds = c.ds;
    z.dfi = dfi;
    z.qs = c.qs;
    z.ceps = c.eps;
    z.stats = c.stats;
    return z;
}

template <typename T>
static void cert_xcheck(const Fft3d<double>& f, const CertArgs<T>& c, hipStream_t s) {
    auto x = sargs<double, T>(f, c.work_spec, true);
    x.cdelta = c.delta;
    x.cdelta_dense = c.delta_dense;
    x.stats = c.stats;
    launch_strided(f, x, S_CERT_CHECK, true, s);
}

template <typename T>
void certify_pair(const Fft3d<double>& f, const CertArgs<T>& a, const CertArgs<T>& b, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_stats2, dim3(1), dim3(64), 0, s, a.stats, b.stats);
    FFCZ_HIP(hipGetLastError());
    const double* dfa = cert_df_field(f, a, s);
    const double* dfb = cert_df_field(f, b, s);
    const ZArgs<double, T> za = cert_zargs(f, a, dfa), zb = cert_zargs(f, b, dfb);
    if (!(f.cz_pair && fast_fft_enabled() && f.zf_var == 0 && fast_cert_zfwd_pair(za, zb, s))) {
        launch_zfwd<double, T, Z_CERT>(f, za, s);
        launch_zfwd<double, T, Z_CERT>(f, zb, s);
    }
    FFCZ_HIP(hipGetLastError());
    if (b.work_spec == a.work_spec + f.spec_elems()) {
        auto y = sargs<double, T>(f, a.work_spec, false);
        y.sign = -1;
        launch_strided(f, y, S_PLAIN, false, s, 2 * f.nx);  // both spectra: 2 nx planes
    } else {
        for (const CertArgs<T>* c : {&a, &b}) {
            auto y = sargs<double, T>(f, c->work_spec, false);
            y.sign = -1;
            launch_strided(f, y, S_PLAIN, false, s);
        }
    }
    // the checks are independent (each field's own statistics); the second
    // spectrum goes first, being the last one the y launch wrote
    cert_xcheck(f, b, s);
    cert_xcheck(f, a, s);
}

// ===========================================================================
// batched certification
// ===========================================================================
void DevBuffer::ensure(size_t n) {
    if (n <= cap) return;
    release();
    FFCZ_HIP(hipMalloc(&p, n));
    cap = n;
}
void DevBuffer::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

__global__ void k_reset_stats_many(CertStats* const* st, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) *st[i] = CertStats{};
}

__global__ void k_gather_stats(CertStats* const* st, int n, CertStats* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = *st[i];
}

template <typename T>
void certify_batch(const std::vector<const Fft3d<double>*>& fs, const std::vector<CertArgs<T>>& cs,
                   DevBuffer& tables, hipStream_t s, CertStats* host_out) {
    FFCZ_REQUIRE(fs.size() == cs.size(), ERR_INVALID, "certification batch mismatch");
    std::vector<size_t> bat;
    std::vector<size_t> rest;
    for (size_t i = 0; i < fs.size(); ++i) {
        const Fft3d<double>& f = *fs[i];
        const bool ok = std::is_same<T, float>::value && fast_fft_enabled() && f.half &&
                        fast_z_rows(f.Lz, f.half, 0) > 0 && fast_s_supported((int)f.ny) &&
                        fast_s_supported((int)f.nx);
        (ok ? bat : rest).push_back(i);
    }
    std::vector<unsigned char> host;
    auto put = [&](const void* p, size_t bytes) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        std::memcpy(host.data() + off, p, bytes);
        return off;
    };
    // every field's stats: reset by the batched passes' first launch, gathered after the last
    std::vector<CertStats*> all_st;
    for (const CertArgs<T>& c : cs) all_st.push_back(c.stats);
    const size_t all_off = put(all_st.data(), all_st.size() * sizeof(CertStats*));
    const size_t out_off = (host.size() + 255) & ~(size_t)255;
    host.resize(out_off + all_st.size() * sizeof(CertStats));
    bool uploaded = false;
    if constexpr (std::is_same<T, float>::value) {
        if (!bat.empty()) {
            struct G {
                int kind;  // 0 z fwd cert, 1 z inv plain, 2 strided
                int smode, L;
                unsigned grid;
                size_t args_off, start_off;
                int n;
            };
            std::vector<G> gs;
            const int W = fast_s_width(sizeof(double), 0);
            // one launch per axis length over the fields in `which`; make(i) builds field i's arguments
            auto add_z = [&](int kind, const std::vector<size_t>& which, auto make) {
                std::vector<std::pair<int, std::vector<size_t>>> byL;
                for (size_t i : which) {
                    const int L = fs[i]->Lz;
                    auto it = std::find_if(byL.begin(), byL.end(), [&](auto& g) { return g.first == L; });
                    if (it == byL.end()) byL.push_back({L, {i}});
                    else it->second.push_back(i);
                }
                for (auto& g : byL) {
                    std::vector<ZArgs<double, T>> tab;
                    std::vector<uint32_t> start;
                    uint64_t tiles = 0;
                    for (size_t i : g.second) {
                        tab.push_back(make(i));
                        start.push_back((uint32_t)tiles);
                        tiles += ceil_div(fs[i]->rows, (size_t)fast_z_rows(fs[i]->Lz, true, fast_zb_var(fs[i]->Lz)));
                    }
                    FFCZ_REQUIRE(tiles < (1ull << 31), ERR_INVALID, "batch too large for one launch");
                    gs.push_back({kind, 0, g.first, (unsigned)tiles, put(tab.data(), tab.size() * sizeof(tab[0])),
                                  put(start.data(), start.size() * sizeof(uint32_t)), (int)tab.size()});
                }
            };
            auto add_s = [&](int smode, bool xaxis, const std::vector<size_t>& which, auto make) {
                std::vector<std::pair<int, std::vector<size_t>>> byL;
                for (size_t i : which) {
                    const int L = (int)(xaxis ? fs[i]->nx : fs[i]->ny);
                    auto it = std::find_if(byL.begin(), byL.end(), [&](auto& g) { return g.first == L; });
                    if (it == byL.end()) byL.push_back({L, {i}});
                    else it->second.push_back(i);
                }
                for (auto& g : byL) {
                    std::vector<SArgs<double, T>> tab;
                    std::vector<uint32_t> start;
                    uint64_t tiles = 0;
                    for (size_t i : g.second) {
                        SArgs<double, T> a = make(i);
                        a.W = W;
                        a.nct = (int)ceil_div((size_t)a.ncols, (size_t)W);
                        a.dNct = FastDiv((uint32_t)a.nct);
                        tab.push_back(a);
                        start.push_back((uint32_t)tiles);
                        tiles += (uint64_t)(xaxis ? fs[i]->ny : fs[i]->nx) * (uint64_t)a.nct;
                    }
                    FFCZ_REQUIRE(tiles < (1ull << 31), ERR_INVALID, "batch too large for one launch");
                    gs.push_back({2, smode, g.first, (unsigned)tiles, put(tab.data(), tab.size() * sizeof(tab[0])),
                                  put(start.data(), start.size() * sizeof(uint32_t)), (int)tab.size()});
                }
            };
            std::vector<size_t> withdf;
            for (size_t i : bat)
                if (cs[i].have_df) withdf.push_back(i);
            // C2R(deq df)/N of the fields with frequency edits: x (loads df), y, z inverse
            add_s(S_LOAD_DF, true, withdf, [&](size_t i) {
                auto x = sargs<double, T>(*fs[i], cs[i].work_spec, true);
                x.sign = +1;
                x.dfsrc = cs[i].df;
                x.qf = cs[i].qf;
                return x;
            });
            add_s(S_PLAIN, false, withdf, [&](size_t i) {
                auto y = sargs<double, T>(*fs[i], cs[i].work_spec, false);
                y.sign = +1;
                return y;
            });
            add_z(1, withdf, [&](size_t i) {
                auto z = zargs<double, T>(*fs[i]);
                z.spec = cs[i].work_spec;
                z.out = cs[i].work_real;
                z.scale = 1.0 / (double)fs[i]->N;
                return z;
            });
            // forward FP64 transform of x_hat - x with the spatial check, then the frequency check
            add_z(0, bat, [&](size_t i) {
                auto z = zargs<double, T>(*fs[i]);
                z.spec = cs[i].work_spec;
                z.e0 = cs[i].e0;
                z.cds = cs[i].ds;
                z.dfi = cs[i].have_df ? cs[i].work_real : nullptr;
                z.qs = cs[i].qs;
                z.ceps = cs[i].eps;
                z.stats = cs[i].stats;
                return z;
            });
            add_s(S_PLAIN, false, bat, [&](size_t i) {
                auto y = sargs<double, T>(*fs[i], cs[i].work_spec, false);
                y.sign = -1;
                return y;
            });
            add_s(S_CERT_CHECK, true, bat, [&](size_t i) {
                auto x = sargs<double, T>(*fs[i], cs[i].work_spec, true);
                x.cdelta = cs[i].delta;
                x.cdelta_dense = cs[i].delta_dense;
                x.stats = cs[i].stats;
                return x;
            });
            tables.ensure(host.size());
            FFCZ_HIP(hipMemcpyAsync(tables.p, host.data(), host.size(), hipMemcpyHostToDevice, s));
            FFCZ_HIP(hipStreamSynchronize(s));  // the host staging goes out of scope
            uploaded = true;
            hipLaunchKernelGGL(k_reset_stats_many, dim3(1), dim3(256), 0, s,
                               (CertStats* const*)(tables.p + all_off), (int)all_st.size());
            FFCZ_HIP(hipGetLastError());
            for (const G& g : gs) {
                bool ok;
                if (g.kind == 2) {
                    BatchTab<SArgs<double, T>> t{(const SArgs<double, T>*)(tables.p + g.args_off),
                                                 (const uint32_t*)(tables.p + g.start_off), g.n};
                    ok = fast_cert_strided_batch(g.smode, g.L, t, g.grid, s);
                } else {
                    BatchTab<ZArgs<double, T>> t{(const ZArgs<double, T>*)(tables.p + g.args_off),
                                                 (const uint32_t*)(tables.p + g.start_off), g.n};
                    ok = g.kind == 0 ? fast_cert_zfwd_batch(g.L, t, g.grid, s) : fast_cert_zinv_batch(g.L, t, g.grid, s);
                }
                FFCZ_REQUIRE(ok, ERR_STATE, "batched certification pass without a kernel for its length");
                FFCZ_HIP(hipGetLastError());
            }
        }
    }
    for (size_t i : rest) certify<T>(*fs[i], cs[i], s);
    if (host_out && !cs.empty()) {
        if (!uploaded) {
            // no batched field: the stats pointers go up now
            tables.ensure(host.size());
            FFCZ_HIP(hipMemcpyAsync(tables.p, host.data(), host.size(), hipMemcpyHostToDevice, s));
            FFCZ_HIP(hipStreamSynchronize(s));
        }
        CertStats* dout = (CertStats*)(tables.p + out_off);
        hipLaunchKernelGGL(k_gather_stats, dim3(1), dim3(256), 0, s, (CertStats* const*)(tables.p + all_off),
                           (int)all_st.size(), dout);
        FFCZ_HIP(hipGetLastError());
        FFCZ_HIP(hipMemcpyAsync(host_out, dout, all_st.size() * sizeof(CertStats), hipMemcpyDeviceToHost, s));
    }
}

template struct Fft3d<float>;
template struct Fft3d<double>;
template void fft_r2c<float>(const Fft3d<float>&, const float*, cplx<float>*, hipStream_t);
template void fft_r2c<double>(const Fft3d<double>&, const double*, cplx<double>*, hipStream_t);
template void fft_c2r<float>(const Fft3d<float>&, cplx<float>*, float*, float, hipStream_t);
template void fft_c2r<double>(const Fft3d<double>&, cplx<double>*, double*, double, hipStream_t);
template void spec_to_compact<float>(const Fft3d<float>&, const cplx<float>*, cplx<float>*, hipStream_t);
template void spec_to_compact<double>(const Fft3d<double>&, const cplx<double>*, cplx<double>*, hipStream_t);
template void compact_to_spec<float>(const Fft3d<float>&, const cplx<float>*, cplx<float>*, hipStream_t);
template void compact_to_spec<double>(const Fft3d<double>&, const cplx<double>*, cplx<double>*, hipStream_t);
template void reset_state<float>(DevState*, int, double, hipStream_t);
template void reset_state<double>(DevState*, int, double, hipStream_t);
template void iteration_kernel<float>(const Fft3d<float>&, const IterArgs<float>&, hipStream_t, int);
template void iteration_kernel<double>(const Fft3d<double>&, const IterArgs<double>&, hipStream_t, int);
template void iteration<float>(const Fft3d<float>&, const IterArgs<float>&, hipStream_t, EventLog*);
template void iteration<double>(const Fft3d<double>&, const IterArgs<double>&, hipStream_t, EventLog*);
template bool fused_iteration_available<float>(const Fft3d<float>&);
template bool fused_iteration_available<double>(const Fft3d<double>&);
template size_t zz_blocks<float>(const Fft3d<float>&);
template size_t zz_blocks<double>(const Fft3d<double>&);
template size_t zz_log_entries<float>(const Fft3d<float>&);
template size_t zz_log_entries<double>(const Fft3d<double>&);
template size_t zz_blocks_max<float>(const Fft3d<float>&);
template size_t zz_blocks_max<double>(const Fft3d<double>&);
template void iteration_fused<float>(const Fft3d<float>&, const IterArgs<float>&, bool, hipStream_t, EventLog*);
template void iteration_fused<double>(const Fft3d<double>&, const IterArgs<double>&, bool, hipStream_t, EventLog*);
template void zz_apply<float>(const Fft3d<float>&, const IterArgs<float>&, int, hipStream_t);
template void zz_apply<double>(const Fft3d<double>&, const IterArgs<double>&, int, hipStream_t);
template void zz_recover_e<float>(const Fft3d<float>&, const IterArgs<float>&, int, hipStream_t);
template void zz_recover_e<double>(const Fft3d<double>&, const IterArgs<double>&, int, hipStream_t);
template void epilogue_clamp<float>(const Fft3d<float>&, const IterArgs<float>&, hipStream_t);
template void epilogue_clamp<double>(const Fft3d<double>&, const IterArgs<double>&, hipStream_t);
template void certify<float>(const Fft3d<double>&, const CertArgs<float>&, hipStream_t);
template void certify<double>(const Fft3d<double>&, const CertArgs<double>&, hipStream_t);
template void certify_pair<float>(const Fft3d<double>&, const CertArgs<float>&, const CertArgs<float>&, hipStream_t);
template void certify_pair<double>(const Fft3d<double>&, const CertArgs<double>&, const CertArgs<double>&,
                                   hipStream_t);
template void certify_batch<float>(const std::vector<const Fft3d<double>*>&, const std::vector<CertArgs<float>>&,
                                   DevBuffer&, hipStream_t, CertStats*);
template void certify_batch<double>(const std::vector<const Fft3d<double>*>&, const std::vector<CertArgs<double>>&,
                                    DevBuffer&, hipStream_t, CertStats*);
template void batch_plan_build<float>(BatchPlan<float>&, const std::vector<const Fft3d<float>*>&,
                                      const std::vector<IterArgs<float>>&);
template void batch_plan_build<double>(BatchPlan<double>&, const std::vector<const Fft3d<double>*>&,
                                       const std::vector<IterArgs<double>>&);
template void batch_plan_free<float>(BatchPlan<float>&);
template void batch_plan_free<double>(BatchPlan<double>&);
template bool batch_fusable<float>(const Fft3d<float>&);
template bool batch_fusable<double>(const Fft3d<double>&);
template void batch_prologue<float>(const BatchPlan<float>&, hipStream_t);
template void batch_prologue<double>(const BatchPlan<double>&, hipStream_t);
template void batch_iteration<float>(const BatchPlan<float>&, hipStream_t);
template void batch_iteration<double>(const BatchPlan<double>&, hipStream_t);
template void batch_gather_states<float>(const BatchPlan<float>&, DevState*, hipStream_t);
template void batch_gather_states<double>(const BatchPlan<double>&, DevState*, hipStream_t);
template void batch_reset_states<float>(const BatchPlan<float>&, int, double, hipStream_t);
template void batch_reset_states<double>(const BatchPlan<double>&, int, double, hipStream_t);

}  // namespace ffcz
