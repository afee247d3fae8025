// MFMA split-precision (BF16x3, scaled FP16x3) variants of the float passes.
#include "fft_fast_impl.hpp"

namespace ffcz {

bool mfma_zfwd(int mode, const ZArgs<float, float>& a, hipStream_t s) {
    if (a.prec == PREC_BF16X3) return mfma_zfwd_dispatch<PREC_BF16X3>(mode, a, s);
    if (a.prec == PREC_FP16X3) return mfma_zfwd_dispatch<PREC_FP16X3>(mode, a, s);
    return false;
}
bool mfma_zinv(int mode, const ZArgs<float, float>& a, hipStream_t s) {
    if (a.prec == PREC_BF16X3) return mfma_zinv_dispatch<PREC_BF16X3>(mode, a, s);
    if (a.prec == PREC_FP16X3) return mfma_zinv_dispatch<PREC_FP16X3>(mode, a, s);
    return false;
}
bool mfma_strided(int mode, const SArgs<float, float>& a, size_t outer, hipStream_t s) {
    if (a.prec == PREC_BF16X3) return mfma_strided_dispatch<PREC_BF16X3>(mode, a, outer, s);
    if (a.prec == PREC_FP16X3) return mfma_strided_dispatch<PREC_FP16X3>(mode, a, outer, s);
    return false;
}

}  // namespace ffcz
