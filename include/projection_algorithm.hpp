// projection_algorithm.hpp - C++ drop-in for the reference's
// ProjectionSolver<T> (projection_algorithm.cuh:238-333), header-only, over the
// C ABI of ffcz.h.  Same public methods, argument meaning and behaviour:
//  * solve* runs the projection loop and then the extraction
//    (projection_algorithm_impl.cuh:813, 857, 901);
//  * solve* before initialize* throws std::runtime_error("Solver not initialized")
//    (projection_algorithm_impl.cuh:778-781);
//  * repeated solve* calls keep accumulating edits until reset_edit_accumulators.
// Differences: errors throw std::runtime_error instead of exit(1)
// (projection_algorithm.cuh:11-31); the class is move-only; flag packing is
// LSB-first over all flags; iteration count and violation have getters.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ffcz.h"

namespace ffcz_detail {
inline void check(int status) {
    if (status != FFCZ_OK) {
        const char* m = ffcz_last_error();
        throw std::runtime_error(m ? std::string(m) : std::string("ffcz error ") + std::to_string(status));
    }
}
}  // namespace ffcz_detail

template <typename T>
class ProjectionSolver {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                  "ProjectionSolver<T> is defined for float and double");
    static constexpr bool F32 = std::is_same<T, float>::value;

   public:
    ProjectionSolver(size_t Nx, size_t Ny, size_t Nz) {
        ffcz_detail::check(ffcz_create(F32 ? FFCZ_F32 : FFCZ_F64, Nx, Ny, Nz, &h_));
    }
    ~ProjectionSolver() {
        if (h_) ffcz_destroy(h_);
    }
    ProjectionSolver(const ProjectionSolver&) = delete;
    ProjectionSolver& operator=(const ProjectionSolver&) = delete;
    ProjectionSolver(ProjectionSolver&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    ProjectionSolver& operator=(ProjectionSolver&& o) noexcept {
        if (this != &o) {
            if (h_) ffcz_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }

    void initialize(const T* h_orig_data, const size_t* h_freq_indices, const T* h_delta_k, T spatial_epsilon,
                    size_t num_freq_constraints) {
        if constexpr (F32)
            ffcz_detail::check(ffcz_init_direct_f32(h_, h_orig_data, h_freq_indices, h_delta_k, spatial_epsilon,
                                                    num_freq_constraints));
        else
            ffcz_detail::check(ffcz_init_direct_f64(h_, h_orig_data, h_freq_indices, h_delta_k, spatial_epsilon,
                                                    num_freq_constraints));
    }
    void initialize_all_abs(const T* h_orig_data, T spatial_epsilon, T delta) {
        if constexpr (F32) ffcz_detail::check(ffcz_init_all_abs_f32(h_, h_orig_data, spatial_epsilon, delta));
        else ffcz_detail::check(ffcz_init_all_abs_f64(h_, h_orig_data, spatial_epsilon, delta));
    }
    void initialize_all_ptw(const T* h_orig_data, T spatial_epsilon, T delta_ptw) {
        if constexpr (F32) ffcz_detail::check(ffcz_init_all_ptw_f32(h_, h_orig_data, spatial_epsilon, delta_ptw));
        else ffcz_detail::check(ffcz_init_all_ptw_f64(h_, h_orig_data, spatial_epsilon, delta_ptw));
    }
    void initialize_partial_ptw(const T* h_orig_data, const size_t* h_freq_indices, const T delta_ptw,
                                T spatial_epsilon, size_t num_freq_constraints) {
        if constexpr (F32)
            ffcz_detail::check(ffcz_init_partial_ptw_f32(h_, h_orig_data, h_freq_indices, delta_ptw, spatial_epsilon,
                                                         num_freq_constraints));
        else
            ffcz_detail::check(ffcz_init_partial_ptw_f64(h_, h_orig_data, h_freq_indices, delta_ptw, spatial_epsilon,
                                                         num_freq_constraints));
    }

    void solve(T* h_spatial_error, int max_iterations, T convergence_tolerance) {
        if constexpr (F32) ffcz_detail::check(ffcz_solve_f32(h_, h_spatial_error, max_iterations, convergence_tolerance));
        else ffcz_detail::check(ffcz_solve_f64(h_, h_spatial_error, max_iterations, convergence_tolerance));
    }
    void solve_all_abs(T* h_spatial_error, int max_iterations, T convergence_tolerance) {
        if constexpr (F32)
            ffcz_detail::check(ffcz_solve_all_abs_f32(h_, h_spatial_error, max_iterations, convergence_tolerance));
        else ffcz_detail::check(ffcz_solve_all_abs_f64(h_, h_spatial_error, max_iterations, convergence_tolerance));
    }
    void solve_all_ptw(T* h_spatial_error, int max_iterations, T convergence_tolerance) {
        if constexpr (F32)
            ffcz_detail::check(ffcz_solve_all_ptw_f32(h_, h_spatial_error, max_iterations, convergence_tolerance));
        else ffcz_detail::check(ffcz_solve_all_ptw_f64(h_, h_spatial_error, max_iterations, convergence_tolerance));
    }

    void get_results(T* h_spatial_error) {
        size_t n = 0;
        ffcz_detail::check(ffcz_get_results(h_, nullptr, 0, &n));
        ffcz_detail::check(ffcz_get_results(h_, h_spatial_error, n, &n));
    }

    void extract_compact_representation() { ffcz_detail::check(ffcz_extract(h_)); }
    std::vector<bool> get_freq_flags() { return bools(&ffcz_get_freq_flags); }
    std::vector<bool> get_spatial_flags() { return bools(&ffcz_get_spatial_flags); }
    std::vector<uint8_t> get_freq_flags_pack() { return vec<uint8_t>(&ffcz_get_freq_flags_pack); }
    std::vector<uint8_t> get_spatial_flags_pack() { return vec<uint8_t>(&ffcz_get_spatial_flags_pack); }
    std::vector<uint16_t> get_active_freq_quant_edit_distances() { return vec<uint16_t>(&ffcz_get_active_freq_quant); }
    std::vector<uint16_t> get_active_spatial_quant_edit_distances() {
        return vec<uint16_t>(&ffcz_get_active_spatial_quant);
    }
    std::vector<T> get_active_freq_edit_distances() { return vecv(&ffcz_get_active_freq_edits); }
    std::vector<T> get_active_spatial_edit_distances() { return vecv(&ffcz_get_active_spatial_edits); }
    std::vector<T> get_freq_edits_full() { return vecv(&ffcz_get_freq_edits_full); }
    std::vector<T> get_spatial_edits_full() { return vecv(&ffcz_get_spatial_edits_full); }

    T get_min_freq_edit() { return range(0); }
    T get_max_freq_edit() { return range(1); }
    T get_min_spatial_edit() { return range(2); }
    T get_max_spatial_edit() { return range(3); }

    void reset_edit_accumulators() { ffcz_detail::check(ffcz_reset_edit_accumulators(h_)); }

    // additions
    int iteration_count() const {
        int it = 0;
        ffcz_detail::check(ffcz_get_status(h_, &it, nullptr, nullptr));
        return it;
    }
    T convergence_metric() const {
        double v = 0;
        ffcz_detail::check(ffcz_get_status(h_, nullptr, &v, nullptr));
        return (T)v;
    }
    ffcz_cert certify(bool quantized) {
        ffcz_cert c{};
        ffcz_detail::check(ffcz_certify(h_, quantized ? 1 : 0, &c));
        return c;
    }
    void set_margins(double spatial, double freq) { ffcz_detail::check(ffcz_set_margins(h_, spatial, freq)); }
    void set_quantizer(int kind, double freq_step, double spatial_step) {
        ffcz_detail::check(ffcz_set_quantizer(h_, kind, freq_step, spatial_step));
    }
    ffcz_solver* handle() const { return h_; }

   private:
    template <typename U, typename Fn>
    std::vector<U> vec(Fn fn) {
        size_t n = 0;
        ffcz_detail::check(fn(h_, nullptr, 0, &n));
        std::vector<U> v(n);
        if (n) ffcz_detail::check(fn(h_, v.data(), n, &n));
        return v;
    }
    template <typename Fn>
    std::vector<T> vecv(Fn fn) {
        size_t n = 0;
        ffcz_detail::check(fn(h_, nullptr, 0, &n));
        std::vector<T> v(n);
        if (n) ffcz_detail::check(fn(h_, (void*)v.data(), n, &n));
        return v;
    }
    template <typename Fn>
    std::vector<bool> bools(Fn fn) {
        std::vector<uint8_t> b = vec<uint8_t>(fn);
        return std::vector<bool>(b.begin(), b.end());
    }
    T range(int which) {
        double v[4] = {0, 0, 0, 0};
        ffcz_detail::check(ffcz_get_edit_range(h_, &v[0], &v[1], &v[2], &v[3]));
        return (T)v[which];
    }
    ffcz_solver* h_ = nullptr;
};
