#pragma once

#include "ffcz_common.hpp"

namespace ffcz {

template <typename T>
void launch_project_spatial(T* e, T eps, T* ds, size_t n, hipStream_t s);
template <typename T>
void launch_project_freq_abs(cplx<T>* E, T d, T* df, size_t F, hipStream_t s);
template <typename T>
double convergence_metric(const T* e, T eps, size_t n, hipStream_t s);
template <typename T>
void launch_quant_dequant(const T* v, T* out, T mn, T mx, size_t n, T thr, hipStream_t s);

}  // namespace ffcz
