// Length-specialised pass kernels (fft_fast_impl.hpp).  Each returns false
// when the axis length has no specialisation, and the caller then runs the
// generic run-time-planned kernels of fft3d.hip.
#pragma once

#include <atomic>

#include "fft_kernels_common.hpp"

namespace ffcz {

template <typename R, typename T>
bool fast_zfwd(int mode, const ZArgs<R, T>& a, hipStream_t s);
template <typename R, typename T>
bool fast_zinv(int mode, const ZArgs<R, T>& a, hipStream_t s);
template <typename R, typename T>
bool fast_strided(int mode, const SArgs<R, T>& a, size_t outer, hipStream_t s);

// specialisations defined by fft_fast_f32.hip / fft_fast_f64.hip / fft_fast_c32.hip
#define FFCZ_FAST_DECL(R, T)                                                                  \
    template <>                                                                               \
    bool fast_zfwd<R, T>(int mode, const ZArgs<R, T>& a, hipStream_t s);                      \
    template <>                                                                               \
    bool fast_zinv<R, T>(int mode, const ZArgs<R, T>& a, hipStream_t s);                      \
    template <>                                                                               \
    bool fast_strided<R, T>(int mode, const SArgs<R, T>& a, size_t outer, hipStream_t s);
FFCZ_FAST_DECL(float, float)
FFCZ_FAST_DECL(double, double)
FFCZ_FAST_DECL(double, float)
#undef FFCZ_FAST_DECL

// MFMA split-precision variants of the float passes (fft_fast_mfma.hip):
// false when a.prec is PREC_VALU or the length has no radix-16 kernel.
bool mfma_zfwd(int mode, const ZArgs<float, float>& a, hipStream_t s);
bool mfma_zinv(int mode, const ZArgs<float, float>& a, hipStream_t s);
bool mfma_strided(int mode, const SArgs<float, float>& a, size_t outer, hipStream_t s);

// Batched launches of the vector-ALU pass kernels (float): one launch over
// the tiles of every field in the table, all of axis length L (z: the
// complex half length).  False when L has no specialised kernel.
bool fast_zfwd_batch(int Lz, const BatchTab<ZArgs<float, float>>& t, unsigned grid, hipStream_t s);
bool fast_zinv_batch(int Lz, const BatchTab<ZArgs<float, float>>& t, unsigned grid, hipStream_t s);
bool fast_strided_batch(int mode, int L, const BatchTab<SArgs<float, float>>& t, unsigned grid, hipStream_t s);
// the fused z pass of the fused iteration form (tile variant 0) over every field in the table
bool fast_zz_batch(int Lz, const BatchTab<ZArgs<float, float>>& t, unsigned grid, hipStream_t s);
// Batched FP64 certification passes over float fields (fft_fast_c32.hip):
// z forward in Z_CERT mode, z inverse in Z_PLAIN mode, y/x in S_PLAIN,
// S_CERT_CHECK or S_LOAD_DF mode.
bool fast_cert_zfwd_batch(int Lz, const BatchTab<ZArgs<double, float>>& t, unsigned grid, hipStream_t s);
bool fast_cert_zinv_batch(int Lz, const BatchTab<ZArgs<double, float>>& t, unsigned grid, hipStream_t s);
bool fast_cert_strided_batch(int mode, int L, const BatchTab<SArgs<double, float>>& t, unsigned grid,
                             hipStream_t s);
// Both certification z passes (unquantised a, quantised b) in one launch;
// false when the length has no specialised kernel.
bool fast_cert_zfwd_pair(const ZArgs<double, float>& a, const ZArgs<double, float>& b, hipStream_t s);
bool fast_cert_zfwd_pair(const ZArgs<double, double>& a, const ZArgs<double, double>& b, hipStream_t s);
// whether a y/x axis length has a specialised kernel, and the columns per
// tile of those kernels for a real type of `bytes` bytes
bool fast_s_supported(int L);
int fast_s_width(size_t bytes, int var = 0);
// a y/x length with a length-specialised kernel whose reversed radix order is
// a different plan (the tile variant 2 of the iteration's vector-ALU passes)
bool fast_s_rev_supported(int L);

// Process-wide tuning knobs (C ABI: ffcz_set_tuning).
struct Tuning {
    std::atomic<int> force_generic{0};  // 1: run the generic kernels only
    std::atomic<long> slab_mb{0};       // slab size of the z/y interleave (<= 0: whole field)
    Tuning() = default;
    Tuning(const Tuning& o) : force_generic(o.force_generic.load()), slab_mb(o.slab_mb.load()) {}
};
Tuning& tuning();
bool fast_fft_enabled();
// rows per workgroup of the specialised z kernels for half length Lz and tile
// variant var (0: the generic kernels run)
int fast_z_rows(int Lz, bool half, int var = 0);
// the z tile variant the batched launches use for a half length
int fast_zb_var(int Lz);

}  // namespace ffcz
