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

}  // namespace ffcz
