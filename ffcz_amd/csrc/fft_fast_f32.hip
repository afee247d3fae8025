// Length-specialised passes for R=float, T=float.
#include <cstdlib>

#include "fft_fast_impl.hpp"

namespace ffcz {

template <>
bool fast_zfwd<float, float>(int mode, const ZArgs<float, float>& a, hipStream_t s) {
    if (a.prec != PREC_VALU && mfma_zfwd(mode, a, s)) return true;
    return zfwd_dispatch<float, float, (1u << Z_PLAIN) | (1u << Z_ITER)>(mode, a, s);
}
template <>
bool fast_zinv<float, float>(int mode, const ZArgs<float, float>& a, hipStream_t s) {
    if (a.prec != PREC_VALU && mfma_zinv(mode, a, s)) return true;
    return zinv_dispatch<float, float, (1u << Z_PLAIN) | (1u << Z_ITER) | (1u << Z_FUSED)>(mode, a, s);
}
template <>
bool fast_strided<float, float>(int mode, const SArgs<float, float>& a, size_t outer, hipStream_t s) {
    if (a.prec != PREC_VALU && mfma_strided(mode, a, outer, s)) return true;
    return strided_dispatch<float, float, (1u << S_PLAIN) | (1u << S_CLAMP) | (1u << S_CLAMP_PASS)>(mode, a, outer, s);
}

bool fast_zfwd_batch(int Lz, const BatchTab<ZArgs<float, float>>& t, unsigned grid, hipStream_t s) {
    return zfwd_batch_dispatch<float, float, Z_ITER>(Lz, t, grid, s);
}
bool fast_zinv_batch(int Lz, const BatchTab<ZArgs<float, float>>& t, unsigned grid, hipStream_t s) {
    return zinv_batch_dispatch<float, float, Z_ITER>(Lz, t, grid, s);
}
bool fast_zz_batch(int Lz, const BatchTab<ZArgs<float, float>>& t, unsigned grid, hipStream_t s) {
    return zz_batch_dispatch<float>(Lz, t, grid, s);
}
bool fast_strided_batch(int mode, int L, const BatchTab<SArgs<float, float>>& t, unsigned grid, hipStream_t s) {
    if (mode == S_CLAMP) return strided_batch_dispatch<float, float, S_CLAMP>(L, t, grid, s);
    if (mode == S_PLAIN) return strided_batch_dispatch<float, float, S_PLAIN>(L, t, grid, s);
    return false;
}
bool fast_s_supported(int L) {
    switch (L) {
#define FFCZ_SS(Lv) case Lv:
        FFCZ_FAST_S_LENGTHS(FFCZ_SS)
#undef FFCZ_SS
        return true;
        default:
            return false;
    }
}
bool fast_s_rev_supported(int L) { return fast_s_supported(L) && pen::rev_differs(L); }
int fast_s_width(size_t bytes, int var) { return bytes == 4 ? s_width<float>(var) : s_width<double>(var); }

bool fast_fft_enabled() { return tuning().force_generic.load() == 0; }

int fast_z_rows(int Lz, bool half, int var) {
    if (!half || !fast_fft_enabled()) return 0;
    switch (Lz) {
#define FFCZ_ZR(Lv) \
    case Lv:        \
        return z_rows(Lv, var);
        FFCZ_FAST_Z_LENGTHS(FFCZ_ZR)
#undef FFCZ_ZR
        default:
            return 0;
    }
}

int fast_zb_var(int Lz) {
    switch (Lz) {
#define FFCZ_ZB(Lv) \
    case Lv:        \
        return zb_var(Lv);
        FFCZ_FAST_Z_LENGTHS(FFCZ_ZB)
#undef FFCZ_ZB
        default:
            return 0;
    }
}

Tuning& tuning() {
    // defaults from the environment: FFCZ_GENERIC_FFT=1, FFCZ_SLAB_MB=<MiB>
    static Tuning t = [] {
        Tuning v;
        const char* g = std::getenv("FFCZ_GENERIC_FFT");
        v.force_generic = (g && g[0] == '1') ? 1 : 0;
        const char* m = std::getenv("FFCZ_SLAB_MB");
        v.slab_mb = m ? std::atol(m) : 0L;
        return v;
    }();
    return t;
}

}  // namespace ffcz
