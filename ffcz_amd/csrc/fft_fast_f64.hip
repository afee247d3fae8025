// Length-specialised passes for R=double, T=double.
#include <cstdlib>

#include "fft_fast_impl.hpp"

namespace ffcz {

template <>
bool fast_zfwd<double, double>(int mode, const ZArgs<double, double>& a, hipStream_t s) {
    return zfwd_dispatch<double, double, (1u << Z_PLAIN) | (1u << Z_ITER) | (1u << Z_CERT)>(mode, a, s);
}
template <>
bool fast_zinv<double, double>(int mode, const ZArgs<double, double>& a, hipStream_t s) {
    return zinv_dispatch<double, double, (1u << Z_PLAIN) | (1u << Z_ITER) | (1u << Z_FUSED)>(mode, a, s);
}
template <>
bool fast_strided<double, double>(int mode, const SArgs<double, double>& a, size_t outer, hipStream_t s) {
    return strided_dispatch<double, double, (1u << S_PLAIN) | (1u << S_CLAMP) | (1u << S_CERT_CHECK) | (1u << S_LOAD_DF) |
                                            (1u << S_CLAMP_PASS)>(mode, a, outer, s);
}

bool fast_cert_zfwd_pair(const ZArgs<double, double>& a, const ZArgs<double, double>& b, hipStream_t s) {
    return zfwd_pair_dispatch<double, double, Z_CERT>(a, b, s);
}

}  // namespace ffcz
