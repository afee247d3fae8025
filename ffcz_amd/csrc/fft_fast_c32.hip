// Length-specialised passes for R=double, T=float.
#include <cstdlib>

#include "fft_fast_impl.hpp"

namespace ffcz {

template <>
bool fast_zfwd<double, float>(int mode, const ZArgs<double, float>& a, hipStream_t s) {
    return zfwd_dispatch<double, float, (1u << Z_CERT)>(mode, a, s);
}
template <>
bool fast_zinv<double, float>(int mode, const ZArgs<double, float>& a, hipStream_t s) {
    return zinv_dispatch<double, float, (1u << Z_PLAIN)>(mode, a, s);
}
template <>
bool fast_strided<double, float>(int mode, const SArgs<double, float>& a, size_t outer, hipStream_t s) {
    return strided_dispatch<double, float, (1u << S_PLAIN) | (1u << S_CERT_CHECK) | (1u << S_LOAD_DF)>(mode, a, outer, s);
}

bool fast_cert_zfwd_pair(const ZArgs<double, float>& a, const ZArgs<double, float>& b, hipStream_t s) {
    return zfwd_pair_dispatch<double, float, Z_CERT>(a, b, s);
}

// batched certification passes (FP64 transforms of float edits)
bool fast_cert_zfwd_batch(int Lz, const BatchTab<ZArgs<double, float>>& t, unsigned grid, hipStream_t s) {
    return zfwd_batch_dispatch<double, float, Z_CERT>(Lz, t, grid, s);
}
bool fast_cert_zinv_batch(int Lz, const BatchTab<ZArgs<double, float>>& t, unsigned grid, hipStream_t s) {
    return zinv_batch_dispatch<double, float, Z_PLAIN>(Lz, t, grid, s);
}
bool fast_cert_strided_batch(int mode, int L, const BatchTab<SArgs<double, float>>& t, unsigned grid,
                             hipStream_t s) {
    if (mode == S_PLAIN) return strided_batch_dispatch<double, float, S_PLAIN>(L, t, grid, s);
    if (mode == S_CERT_CHECK) return strided_batch_dispatch<double, float, S_CERT_CHECK>(L, t, grid, s);
    if (mode == S_LOAD_DF) return strided_batch_dispatch<double, float, S_LOAD_DF>(L, t, grid, s);
    return false;
}

}  // namespace ffcz
