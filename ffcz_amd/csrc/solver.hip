// ProjectionSolver orchestration on one GPU: device buffers, the projection
// loop with device-side termination, extraction and FP64 certification.
#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>

#include "fft_fast.hpp"
#include "plan_cache.hpp"
#include "solver.hpp"

namespace ffcz {

// the reference's split layout [Re(F) | Im(F)] (cu:164,179) of the solver's
// (Re, Im)-per-bin frequency edits
template <typename T>
__global__ void k_df_split(const cplx<T>* dfc, T* df, size_t F) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < F; i += (size_t)gridDim.x * blockDim.x) {
        const cplx<T> v = dfc[i];
        df[i] = v.x;
        df[i + F] = v.y;
    }
}

template <typename T>
__global__ void k_zero(T* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = T(0);
}
template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// dense per-bin bound: min over duplicate indices; flags an out-of-range index
template <typename T>
__global__ void k_scatter_delta(T* dense, size_t F, const size_t* idx, const T* dk, size_t K, unsigned int* bad) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        const size_t j = idx[i];
        const T d = dk[i];
        if (j >= F || !(d >= T(0))) {
            atomicOr(bad, 1u);
            continue;
        }
        // non-negative values order like their bit patterns
        if constexpr (sizeof(T) == 4) atomicMin((unsigned int*)&dense[j], __float_as_uint((float)d));
        else atomicMin((unsigned long long*)&dense[j], (unsigned long long)__double_as_longlong((double)d));
    }
}
static unsigned gridn(size_t n) {
    size_t g = ceil_div(n ? n : 1, (size_t)256);
    return (unsigned)std::min<size_t>(g, 1u << 16);
}

template <typename T>
static void dalloc(T** p, size_t n) {
    FFCZ_REQUIRE(n > 0, ERR_INVALID, "zero-size allocation");
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess)
        throw Error(ERR_ALLOC, "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
                                   hipGetErrorString(e));
}

template <typename T>
Solver<T>::Solver(size_t nx, size_t ny, size_t nz) {
    f_.create(nx, ny, nz);
    try {
        FFCZ_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
        dalloc(&x_, f_.N);
        dalloc(&e_, f_.N);
        dalloc(&e0_, f_.N);
        dalloc(&ds_, f_.N);
        dalloc(&df_, 2 * f_.F);
        dalloc(&dfc_, f_.F);
        dalloc(&spec_, f_.spec_elems());
        dalloc(&st_, 1);
        dalloc(&vpart_, f_.rows);
        hipLaunchKernelGGL((k_zero<T>), dim3(gridn(f_.rows)), dim3(256), 0, s_, vpart_, f_.rows);
        reset_edit_accumulators();  // projection_algorithm_impl.cuh:60
        reset_state<T>(st_, 0, 0.0, s_);
        FFCZ_HIP(hipStreamSynchronize(s_));
    } catch (...) {
        release();
        throw;
    }
}

template <typename T>
Solver<T>::~Solver() {
    release();
}

template <typename T>
void Solver<T>::release() {
    if (s_) (void)hipStreamSynchronize(s_);
    release_graph();
    void* ptrs[] = {x_, e_, e0_, ds_, df_, spec_, st_, vpart_, delta_dense_, orig_, partial_idx_, fpack_, spack_,
                    fcomp_, scomp_, fquant_, squant_, xscratch_, cert_real_, cert_spec_, cstats_, xglob_,
                    cert_real2_, zlog_off_, zlog_v_, zlog_cnt_, dfc_, ptw_mask_};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (xhost_) (void)hipHostFree(xhost_);
    if (chost_) (void)hipHostFree(chost_);
    if (sthost_) (void)hipHostFree(sthost_);
    sthost_ = nullptr;
    xglob_ = nullptr;
    xhost_ = nullptr;
    chost_ = nullptr;
    cert_pending_ = false;
    xpending_ = false;
    x_ = e_ = e0_ = ds_ = df_ = vpart_ = nullptr;
    e0_ext_ = nullptr;
    spec_ = nullptr;
    st_ = nullptr;
    delta_dense_ = nullptr;
    orig_ = nullptr;
    partial_idx_ = nullptr;
    ptw_mask_ = nullptr;
    ptw_mask_corr_ = -1;
    fpack_ = spack_ = nullptr;
    fcomp_ = scomp_ = nullptr;
    fquant_ = squant_ = nullptr;
    xscratch_ = nullptr;
    cert_real_ = nullptr;
    cert_real2_ = nullptr;
    cert_spec_ = nullptr;
    cstats_ = nullptr;
    zlog_off_ = zlog_cnt_ = nullptr;
    dfc_ = nullptr;
    zlog_v_ = nullptr;
    if (f64_) f64_->destroy();
    f64_.reset();
    f_.destroy();
    if (s_) (void)hipStreamDestroy(s_);
    s_ = nullptr;
}

template <typename T>
void Solver<T>::upload_x(const T* x) {
    FFCZ_REQUIRE(x != nullptr, ERR_INVALID, "null original data");
    // initialize* leaves the previous solve's iterate in place (the ptw
    // modes reuse the spectrum rows it may still be held in)
    materialize_e();
    FFCZ_HIP(hipMemcpyAsync(x_, x, f_.N * sizeof(T), hipMemcpyDefault, s_));
}

template <typename T>
void Solver<T>::reset_edit_accumulators() {
    // df: only when it may hold edits; ds: deferred (materialize_ds, or the
    // next solve's first spatial clamp stores it whole)
    if (df_touched_) {
        hipLaunchKernelGGL((k_zero<T>), dim3(gridn(2 * f_.F)), dim3(256), 0, s_, reinterpret_cast<T*>(dfc_),
                           2 * f_.F);
        df_split_stale_ = true;
        FFCZ_HIP(hipGetLastError());
        df_touched_ = false;
    }
    ds_fresh_ = true;
}

template <typename T>
void Solver<T>::accumulators_clobbered() {
    materialize_ds();
    df_touched_ = true;
    df_split_stale_ = true;
}

template <typename T>
void Solver<T>::materialize_df_split(hipStream_t s) const {
    if (!df_split_stale_) return;
    hipLaunchKernelGGL((k_df_split<T>), dim3(gridn(f_.F)), dim3(256), 0, s ? s : s_, (const cplx<T>*)dfc_, df_, f_.F);
    FFCZ_HIP(hipGetLastError());
    df_split_stale_ = false;
}

template <typename T>
void Solver<T>::materialize_ds(hipStream_t s) const {
    if (!ds_fresh_) return;
    hipLaunchKernelGGL((k_zero<T>), dim3(gridn(f_.N)), dim3(256), 0, s ? s : s_, ds_, f_.N);
    FFCZ_HIP(hipGetLastError());
    ds_fresh_ = false;
}

template <typename T>
void Solver<T>::init_all_abs(const T* x, T eps, T delta) {
    // projection_algorithm_impl.cuh:639-668 (its F(x) is not needed here)
    upload_x(x);
    eps_ = eps;
    delta_ = delta;
    init_mode_ = MODE_ABS;
    solve_mode_ = -1;
    initialized_ = true;
    FFCZ_HIP(hipStreamSynchronize(s_));
}

template <typename T>
void Solver<T>::init_direct(const T* x, const size_t* idx, const T* dk, T eps, size_t K) {
    // projection_algorithm_impl.cuh:598-637; the constraint list becomes a
    // dense per-bin bound (+inf where unconstrained) read by the fused pass.
    FFCZ_REQUIRE(K == 0 || (idx && dk), ERR_INVALID, "null constraint arrays");
    upload_x(x);
    eps_ = eps;
    if (!delta_dense_) dalloc(&delta_dense_, f_.F);
    hipLaunchKernelGGL((k_fill<T>), dim3(gridn(f_.F)), dim3(256), 0, s_, delta_dense_, f_.F,
                       std::numeric_limits<T>::infinity());
    if (K > 0) {
        size_t* didx = nullptr;
        T* ddk = nullptr;
        unsigned int* bad = nullptr;
        dalloc(&didx, K);
        dalloc(&ddk, K);
        dalloc(&bad, 1);
        unsigned int hbad = 0;
        try {
            FFCZ_HIP(hipMemcpyAsync(didx, idx, K * sizeof(size_t), hipMemcpyDefault, s_));
            FFCZ_HIP(hipMemcpyAsync(ddk, dk, K * sizeof(T), hipMemcpyDefault, s_));
            FFCZ_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned int), s_));
            hipLaunchKernelGGL((k_scatter_delta<T>), dim3(gridn(K)), dim3(256), 0, s_, delta_dense_, f_.F, didx, ddk, K, bad);
            FFCZ_HIP(hipGetLastError());
            FFCZ_HIP(hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, s_));
            FFCZ_HIP(hipStreamSynchronize(s_));
        } catch (...) {
            (void)hipFree(didx);
            (void)hipFree(ddk);
            (void)hipFree(bad);
            throw;
        }
        (void)hipFree(didx);
        (void)hipFree(ddk);
        (void)hipFree(bad);
        FFCZ_REQUIRE(hbad == 0, ERR_INVALID, "frequency constraint index >= F or negative/NaN bound");
    }
    init_mode_ = MODE_DIRECT;
    solve_mode_ = -1;
    initialized_ = true;
    FFCZ_HIP(hipStreamSynchronize(s_));
}

template <typename T>
void Solver<T>::init_all_ptw(const T* x, T eps, T delta_ptw) {
    // projection_algorithm_impl.cuh:670-699: F(x) is needed by the ptw bound.
    upload_x(x);
    eps_ = eps;
    delta_ptw_ = delta_ptw;
    if (!orig_) dalloc(&orig_, f_.F);
    fft_r2c(f_, x_, spec_, s_);
    spec_to_compact(f_, spec_, orig_, s_);
    init_mode_ = MODE_PTW;
    solve_mode_ = -1;
    initialized_ = true;
    FFCZ_HIP(hipStreamSynchronize(s_));
}

template <typename T>
void Solver<T>::init_partial_ptw(const T* x, const size_t* idx, T delta_ptw, T eps, size_t K) {
    // projection_algorithm_impl.cuh:701-738: stores F(x) and the index list,
    // consumed by the partial ptw projection (projection_algorithm.cu:261-333).
    std::vector<size_t> h(K);
    if (K) {
        FFCZ_REQUIRE(idx != nullptr, ERR_INVALID, "null index array");
        FFCZ_HIP(hipMemcpy(h.data(), idx, K * sizeof(size_t), hipMemcpyDefault));
        for (size_t j = 0; j < K; ++j)
            FFCZ_REQUIRE(h[j] < f_.F, ERR_INVALID, "frequency index out of range (>= nx*ny*(nz/2+1))");
    }
    init_all_ptw(x, eps, delta_ptw);
    init_mode_ = MODE_PARTIAL_PTW;
    solve_mode_ = -1;
    if (partial_idx_) (void)hipFree(partial_idx_);
    partial_idx_ = nullptr;
    partial_K_ = K;
    ptw_mask_corr_ = -1;
    if (K) {
        dalloc(&partial_idx_, K);
        FFCZ_HIP(hipMemcpyAsync(partial_idx_, h.data(), K * sizeof(size_t), hipMemcpyHostToDevice, s_));
        FFCZ_HIP(hipStreamSynchronize(s_));
    }
}

// The per-bin rule of the partial ptw projection.  Reference semantics
// (projection_algorithm.cu:261-333, one thread per constraint): constraint 0
// takes the DC rule on bin 0 whatever idx[0] holds, constraint j >= 1 the
// ratio rule on bin idx[j] (a later entry naming bin 0 is left to the DC rule:
// in the reference both would race on bin 0).  Corrected semantics: the
// listed bins take the ratio rule.
template <typename T>
void Solver<T>::build_ptw_mask() {
    if (ptw_mask_corr_ == (ptw_corr_ ? 1 : 0)) return;
    std::vector<size_t> h(partial_K_);
    if (partial_K_)
        FFCZ_HIP(hipMemcpy(h.data(), partial_idx_, partial_K_ * sizeof(size_t), hipMemcpyDeviceToHost));
    std::vector<uint8_t> m(f_.F, PTW_NONE);
    for (size_t j = 0; j < partial_K_; ++j) m[h[j]] = PTW_RATIO;
    if (!ptw_corr_ && partial_K_) m[0] = PTW_DC;
    if (!ptw_mask_) dalloc(&ptw_mask_, f_.F);
    FFCZ_HIP(hipMemcpyAsync(ptw_mask_, m.data(), f_.F, hipMemcpyHostToDevice, s_));
    FFCZ_HIP(hipStreamSynchronize(s_));
    ptw_mask_corr_ = ptw_corr_ ? 1 : 0;
}

template <typename T>
IterArgs<T> Solver<T>::iter_args() const {
    IterArgs<T> a{};
    a.e = e_;
    a.e_init = e_src_ ? e_src_ : e0_;
    a.e_copy = (e_src_ && !e0_ext_) ? e0_ : nullptr;
    a.ds_fresh = ds_fresh_ ? 1 : 0;
    a.ds = ds_;
    a.df = dfc_;
    a.spec = spec_;
    a.eps = (T)((double)eps_ * (1.0 - margin_s_));
    if (margin_s_ == 0.0) a.eps = eps_;
    const int mode = solve_mode();
    a.fmode = mode == MODE_ABS ? FM_ABS : (mode == MODE_DIRECT ? FM_DENSE : FM_PTW);
    a.ptw_mask = mode == MODE_PARTIAL_PTW ? ptw_mask_ : nullptr;
    a.ptw_corr = ptw_corr_ ? 1 : 0;
    a.delta = margin_f_ == 0.0 ? delta_ : (T)((double)delta_ * (1.0 - margin_f_));
    a.dscale = margin_f_ == 0.0 ? T(1) : (T)(1.0 - margin_f_);
    a.delta_dense = delta_dense_;
    a.orig = orig_;
    a.delta_ptw = delta_ptw_;
    a.norm = T(1) / static_cast<T>(f_.N);  // projection_algorithm_impl.cuh:242
    a.st = st_;
    a.vpart = vpart_;
    a.log_off = zlog_off_;
    a.log_v = zlog_v_;
    a.log_cnt = zlog_cnt_;
    return a;
}

template <typename T>
void Solver<T>::ensure_zz_log() {
    if (zlog_off_) return;
    const size_t n = zz_log_entries(f_), nb = zz_blocks_max(f_);
    FFCZ_REQUIRE(n > 0 && nb > 0, ERR_INVALID, "no fused z pass for this shape");
    dalloc(&zlog_off_, n);
    dalloc(&zlog_v_, n);
    dalloc(&zlog_cnt_, nb);
}

template <typename T>
void Solver<T>::run_iteration(const IterArgs<T>& a, bool fused, bool first, EventLog* log) {
    if (fused) iteration_fused(f_, a, first, s_, log);
    else iteration(f_, a, s_, log);
}

template <typename T>
void Solver<T>::materialize_e() {
    if (!e_lazy_) return;
    zz_recover_e(f_, iter_args(), fused_var_, s_);
    e_lazy_ = false;
}

template <typename T>
void Solver<T>::prepare_solve(int mode, const T* e0, int max_iter, T tol, hipStream_t s, bool reset) {
    FFCZ_REQUIRE(initialized_, ERR_STATE, "Solver not initialized");  // impl:778-781
    // solve_all_ptw also runs after initialize_partial_ptw, which computes
    // everything it needs (projection_algorithm_impl.cuh:701-738, 860-902)
    FFCZ_REQUIRE(mode == init_mode_ || (mode == MODE_PTW && init_mode_ == MODE_PARTIAL_PTW), ERR_STATE,
                 "solve mode does not match the initialize* call (direct<->initialize, abs<->initialize_all_abs, "
                 "ptw<->initialize_all_ptw or initialize_partial_ptw, partial ptw<->initialize_partial_ptw)");
    FFCZ_REQUIRE(e0 != nullptr, ERR_INVALID, "null spatial error");
    solve_mode_ = mode;
    if (mode == MODE_PARTIAL_PTW) build_ptw_mask();
    // e0_ keeps the initial error for certification.  The first iteration
    // reads the initial error directly (its inverse pass writes the iterate
    // e_); a device-resident e0 is copied into e0_ by that same pass.
    e_src_ = nullptr;
    e0_ext_ = nullptr;
    fused_ = false;
    e_lazy_ = false;
    hipPointerAttribute_t attr{};
    if (max_iter > 0 && hipPointerGetAttributes(&attr, e0) == hipSuccess &&
        (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged))
        e_src_ = e0;
    else
        (void)hipGetLastError();  // a host pointer is not an error here
    // a borrowed device e0 stays where it is: no copy for the certifications
    if (e_src_ && borrow_e0_) e0_ext_ = e0;
    if (!e_src_) FFCZ_HIP(hipMemcpyAsync(e0_, e0, f_.N * sizeof(T), hipMemcpyDefault, s));
    if (max_iter <= 0) FFCZ_HIP(hipMemcpyAsync(e_, e0_, f_.N * sizeof(T), hipMemcpyDeviceToDevice, s));
    if (reset) reset_state<T>(st_, max_iter, (double)tol, s);
}

// The fields of two iteration argument sets that an iteration after a
// solve's first reads (the graph holds them as kernel arguments).  The
// initial error and its copy are read by the first iteration only, which
// runs eagerly, so a caller's new initial-error buffer needs no new capture.
template <typename R>
static bool same_iter_args(const IterArgs<R>& a, const IterArgs<R>& b) {
    return a.e == b.e && a.ds_fresh == b.ds_fresh && a.ds == b.ds &&
           a.df == b.df && a.spec == b.spec && a.eps == b.eps && a.fmode == b.fmode && a.delta == b.delta &&
           a.delta_dense == b.delta_dense && a.dscale == b.dscale && a.orig == b.orig &&
           a.delta_ptw == b.delta_ptw && a.ptw_mask == b.ptw_mask && a.ptw_corr == b.ptw_corr && a.norm == b.norm &&
           a.st == b.st && a.vpart == b.vpart && a.log_off == b.log_off && a.log_v == b.log_v &&
           a.log_cnt == b.log_cnt;
}

template <typename T>
void Solver<T>::release_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    gexec_ = nullptr;
}

template <typename T>
void Solver<T>::set_graph(bool on) {
    graph_on_ = on;
    if (!on) {
        FFCZ_HIP(hipStreamSynchronize(s_));
        release_graph();
    }
}

// The graph of one iteration after the first (run_iteration with first =
// false): kernels only, no memset or copy nodes.
template <typename T>
bool Solver<T>::ensure_graph(const IterArgs<T>& a, bool fused) {
    const int key[8] = {fused ? 1 : 0, f_.prec, f_.zf_var, f_.zi_var, f_.y_var, f_.x_var, tuning().force_generic.load(),
                        (int)tuning().slab_mb.load()};
    if (gexec_ && same_iter_args(a, gargs_) && std::equal(key, key + 8, gkey_)) return true;
    // work queued earlier may still run the graph replaced here
    FFCZ_HIP(hipStreamSynchronize(s_));
    release_graph();
    hipGraph_t g = nullptr;
    FFCZ_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
        run_iteration(a, fused, false, nullptr);
    } catch (...) {
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    FFCZ_HIP(hipStreamEndCapture(s_, &g));
    const hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) gexec_ = nullptr;
    FFCZ_HIP(e);
    gargs_ = a;
    std::copy(key, key + 8, gkey_);
    return true;
}

template <typename T>
void Solver<T>::solve_loop(int mode, const T* e0, int max_iter, T tol) {
    const bool fused = use_fused();
    if (fused) ensure_zz_log();
    prepare_solve(mode, e0, max_iter, tol, s_);
    fused_ = fused;
    fused_var_ = f_.zi_var;
    const IterArgs<T> a = iter_args();
    const bool graph = graph_on_ && !pass_timing_ && max_iter > 1 && ensure_graph(a, fused);
    // iterations go out in chunks of 2, 4, 8, 16 between host checks; the
    // kernels of an iteration past convergence return at once
    int launched = 0, chunk = 2;
    DevState h{};
    bool known = false;
    if (pass_timing_) {
        plog_->used = 0;
        chunk_first_.clear();
    }
    while (launched < max_iter) {
        const int k = std::min(chunk, max_iter - launched);
        if (pass_timing_) chunk_first_.push_back(launched + 1);
        for (int i = 0; i < k; ++i) {
            const bool first = launched + i == 0;
            if (graph && !first) FFCZ_HIP(hipGraphLaunch(gexec_, s_));
            else run_iteration(a, fused, first, pass_timing_ ? plog_.get() : nullptr);
        }
        launched += k;
        h = read_state(s_);
        known = true;
        if (h.done) break;
        chunk = std::min(chunk * 2, 16);
    }
    if (pass_timing_ && known) account_pass_times((int)h.iter);
    finish_solve(s_, known && h.done ? &h : nullptr);
}

// The device state, through a pinned host copy (a pageable target would
// stage the copy), after everything queued on s.
template <typename T>
DevState Solver<T>::read_state(hipStream_t s) {
    if (!sthost_) FFCZ_HIP(hipHostMalloc((void**)&sthost_, sizeof(DevState), hipHostMallocDefault));
    FFCZ_HIP(hipMemcpyAsync(sthost_, st_, sizeof(DevState), hipMemcpyDeviceToHost, s));
    FFCZ_HIP(hipStreamSynchronize(s));
    return *sthost_;
}

template <typename T>
void Solver<T>::set_pass_timing(bool enable) {
    pass_timing_ = enable;
    if (enable && !plog_) plog_.reset(new EventLog());
    for (double& v : pass_ms_) v = 0.0;
    for (long& c : pass_cnt_) c = 0;
    pass_n_ = 0;
}

template <typename T>
void Solver<T>::account_pass_times(int iterations) {
    // the log holds, per launched iteration, a start mark (-1) and one mark
    // after each launch; intervals belong to the launch that ends them
    FFCZ_HIP(hipStreamSynchronize(s_));
    int it = 0;
    bool first_launch = false;
    for (size_t i = 0; i < plog_->used; ++i) {
        if (plog_->kinds[i] < 0) {
            ++it;
            first_launch = std::find(chunk_first_.begin(), chunk_first_.end(), it) != chunk_first_.end();
            continue;
        }
        const bool skip = first_launch || it > iterations || i == 0;
        first_launch = false;
        if (skip) continue;
        float t = 0;
        FFCZ_HIP(hipEventElapsedTime(&t, plog_->pool[i - 1], plog_->pool[i]));
        pass_ms_[plog_->kinds[i]] += t;
        pass_cnt_[plog_->kinds[i]] += 1;
    }
    pass_n_ += iterations;
}

template <typename T>
void Solver<T>::pass_counts(long* n6) const {
    for (int k = 0; k < 6; ++k) n6[k] = pass_cnt_[k];
}

template <typename T>
long Solver<T>::pass_timing(float* ms6) const {
    for (int k = 0; k < 6; ++k) ms6[k] = pass_cnt_[k] ? (float)(pass_ms_[k] / (double)pass_cnt_[k]) : 0.0f;
    return pass_n_;
}

template <typename T>
void Solver<T>::finish_solve(hipStream_t s, const DevState* known) {
    const IterArgs<T> a = iter_args();
    if (fused_) {
        // the last z fused pass clamped: kept after an unconverged loop (the
        // reference's final spatial projection, impl:851), its deferred
        // clamps then added to ds; a converged loop drops them (it stops
        // before that projection, impl:845)
        if (!known || (!known->converged && known->iter > 0)) zz_apply(f_, a, fused_var_, s);
    } else if (!known || (known->done && !known->converged && known->iter > 0)) {
        // the clamp applies after an unconverged loop only: a known final
        // state that says otherwise needs no launch
        epilogue_clamp(f_, a, s);
    }
    DevState h{};
    if (known) {
        h = *known;  // the epilogue leaves the state as it is
    } else {
        h = read_state(s);
    }
    e_src_ = nullptr;
    if (ds_fresh_) {
        // the first spatial clamp stored ds whole: the forward z pass of
        // iteration 1, or the epilogue after one unconverged iteration
        // (the fused form: the z fused pass of iteration 1, deferred clamps as 0)
        const bool written = fused_ ? h.iter >= 1 : (h.iter >= 2 || (h.iter == 1 && !h.converged));
        ds_fresh_ = false;
        if (!written) {
            hipLaunchKernelGGL((k_zero<T>), dim3(gridn(f_.N)), dim3(256), 0, s, ds_, f_.N);
            FFCZ_HIP(hipGetLastError());
        }
    }
    e_lazy_ = fused_ && h.iter > 0;
    iters_ = (int)h.iter;
    viol_ = h.last_viol;
    converged_ = h.converged != 0;
    df_touched_ = df_touched_ || h.freq_clamped != 0;
    if (h.freq_clamped) df_split_stale_ = true;
    solved_ = true;
    extracted_ = false;
}

template <typename T>
void Solver<T>::set_quantizer(int kind, double fstep, double sstep) {
    FFCZ_REQUIRE(kind == Q_REF16 || kind == Q_STEP, ERR_INVALID, "unknown quantizer kind");
    if (kind == Q_STEP) FFCZ_REQUIRE(fstep > 0 && sstep > 0, ERR_INVALID, "quantizer steps must be > 0");
    qkind_ = kind;
    qstep_f_ = fstep;
    qstep_s_ = sstep;
    extracted_ = false;
}

template <typename T>
QuantDesc Solver<T>::freq_quant() const {
    QuantDesc q{};
    q.kind = qkind_;
    q.threshold = 1e-7;  // projection_algorithm_impl.cuh:389
    q.step = qstep_f_;
    q.inv_step = qstep_f_ > 0 ? 1.0 / qstep_f_ : 0.0;
    q.vmin = xf_.vmin;
    q.vmax = xf_.vmax;
    q.dminmax = xglob_ ? &xglob_[0].vmin : nullptr;
    return q;
}
template <typename T>
QuantDesc Solver<T>::spatial_quant() const {
    QuantDesc q{};
    q.kind = qkind_;
    q.threshold = 1e-7;  // projection_algorithm_impl.cuh:398
    q.step = qstep_s_;
    q.inv_step = qstep_s_ > 0 ? 1.0 / qstep_s_ : 0.0;
    q.vmin = xs_.vmin;
    q.vmax = xs_.vmax;
    q.dminmax = xglob_ ? &xglob_[1].vmin : nullptr;
    return q;
}

template <typename T>
void Solver<T>::ensure_extract_buffers() {
    if (fpack_) return;
    dalloc(&fpack_, ceil_div(2 * f_.F, (size_t)64) * 8);
    dalloc(&spack_, ceil_div(f_.N, (size_t)64) * 8);
    dalloc(&fcomp_, 2 * f_.F);
    dalloc(&scomp_, f_.N);
    dalloc(&fquant_, 2 * f_.F);
    dalloc(&squant_, f_.N);
    const size_t sb = std::max(extract_scratch_bytes(2 * f_.F), extract_scratch_bytes(f_.N));
    unsigned char* p = nullptr;
    dalloc(&p, sb);
    xscratch_ = p;
    dalloc(&xglob_, 2);
    FFCZ_HIP(hipHostMalloc((void**)&xhost_, 2 * sizeof(ExtractGlobals), hipHostMallocDefault));
}

template <typename T>
void Solver<T>::extract() {
    // extract_compact_representation, projection_algorithm_impl.cuh:371-456
    ensure_extract_buffers();
    sync_extract();  // a previous extraction's results are still being copied
    materialize_ds();
    QuantDesc qf = freq_quant(), qs = spatial_quant();
    qf.dminmax = qs.dminmax = nullptr;  // the kernels compute the range themselves
    if (df_touched_) {
        materialize_df_split();
        extract_edits_async<T>(df_, 2 * f_.F, qf, fpack_, fcomp_, fquant_, xscratch_, &xglob_[0], nullptr, s_);
    } else {
        // no frequency clamp fired since df was cleared, so df holds +0 only:
        // what the full pass would find (an all-zero flag pack, no active bin,
        // the range [0, 0]) is written without reading df
        extract_zeros_async(fpack_, 2 * f_.F, &xglob_[0], s_);
    }
    extract_edits_async<T>(ds_, f_.N, qs, spack_, scomp_, squant_, xscratch_, &xglob_[1], nullptr, s_);
    // both arrays' counts and ranges reach the host in one copy
    FFCZ_HIP(hipMemcpyAsync(xhost_, xglob_, 2 * sizeof(ExtractGlobals), hipMemcpyDeviceToHost, s_));
    xpending_ = true;
    extracted_ = true;
}

template <typename T>
void Solver<T>::extract_segments(std::vector<ExtractSeg<T>>& out, hipStream_t s) {
    ensure_extract_buffers();
    sync_extract();
    materialize_ds(s);
    QuantDesc qf = freq_quant(), qs = spatial_quant();
    qf.dminmax = qs.dminmax = nullptr;
    // an untouched df (+0 only, as in extract()) goes in with no tiles: the
    // batch's finalisation then reports count 0 and the range [0, 0], and
    // its flag pack is cleared here on the batch stream
    const bool dfz = !df_touched_;
    if (dfz) extract_zeros_async(fpack_, 2 * f_.F, &xglob_[0], s);
    else materialize_df_split(s);
    out.push_back(ExtractSeg<T>{df_, dfz ? 0 : 2 * f_.F, qf, (unsigned int*)fpack_, fcomp_, fquant_, &xglob_[0], 0u});
    out.push_back(ExtractSeg<T>{ds_, f_.N, qs, (unsigned int*)spack_, scomp_, squant_, &xglob_[1], 0u});
    extracted_ = false;  // until the results are taken over
}

template <typename T>
void Solver<T>::extract_adopt(const ExtractGlobals& hf, const ExtractGlobals& hs) {
    xhost_[0] = hf;
    xhost_[1] = hs;
    xpending_ = false;
    extracted_ = true;
    take_extract_results(hf, hs);
}

template <typename T>
void Solver<T>::sync_extract() const {
    if (!xpending_) return;
    FFCZ_HIP(hipStreamSynchronize(s_));
    xpending_ = false;
    take_extract_results(xhost_[0], xhost_[1]);
}

template <typename T>
void Solver<T>::take_extract_results(const ExtractGlobals& hf, const ExtractGlobals& hs) const {
    xf_ = ExtractResult{(size_t)hf.count, hf.vmin, hf.vmax};
    xs_ = ExtractResult{(size_t)hs.count, hs.vmin, hs.vmax};
    if ((hf.overflow | hs.overflow) & 2u) {
        const_cast<Solver<T>*>(this)->extracted_ = false;
        throw Error(ERR_HIP, "extraction look-back did not complete");
    }
    if (hf.overflow || hs.overflow) {
        const_cast<Solver<T>*>(this)->extracted_ = false;
        throw Error(ERR_RANGE, "quantised edit does not fit a 16-bit symbol (raise the quantiser step)");
    }
}

template <typename T>
void Solver<T>::ensure_cert_buffers() {
    if (f64_) return;
    f64_.reset(new Fft3d<double>());
    f64_->create(f_.nx, f_.ny, f_.nz);
    dalloc(&cert_real_, f_.N);
    dalloc(&cert_spec_, 2 * f64_->spec_elems());
    dalloc(&cstats_, 2);
    FFCZ_HIP(hipHostMalloc((void**)&chost_, 2 * sizeof(CertStats), hipHostMallocDefault));
}

template <typename T>
CertResult Solver<T>::certify(bool quantized) {
    certify_enqueue(quantized);
    return certify_collect(quantized);
}

template <typename T>
CertArgs<T> Solver<T>::certify_prepare(bool quantized) {
    FFCZ_REQUIRE(solved_, ERR_STATE, "certify requires a completed solve");
    FFCZ_REQUIRE(init_mode_ == MODE_ABS || init_mode_ == MODE_DIRECT, ERR_INVALID,
                 "certification is defined for the absolute and per-bin bound modes");
    if (quantized) FFCZ_REQUIRE(extracted_, ERR_STATE, "certify(quantized) requires extraction first");
    ensure_cert_buffers();
    return cert_args(quantized);
}

template <typename T>
void Solver<T>::certify_enqueue(bool quantized) {
    FFCZ_REQUIRE(!cert_pending_, ERR_STATE, "the queued certification's results have not been read");
    const CertArgs<T> c = certify_prepare(quantized);
    ffcz::certify<T>(*f64_, c, s_);
    certify_copy_stats(quantized, s_);
}

template <typename T>
void Solver<T>::certify_both(CertResult* out) {
    FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null result");
    certify_both_enqueue();
    certify_both_results(out);
}

template <typename T>
void Solver<T>::certify_both_enqueue() {
    FFCZ_REQUIRE(!cert_pending_, ERR_STATE, "the queued certification's results have not been read");
    CertArgs<T> a = certify_prepare(false);
    CertArgs<T> b = certify_prepare(true);
    b.work_spec = cert_spec_ + f64_->spec_elems();
    if (b.have_df) {
        if (!cert_real2_) dalloc(&cert_real2_, f_.N);
        b.work_real = cert_real2_;
    }
    ffcz::certify_pair<T>(*f64_, a, b, s_);
    // both results in one copy (cstats_ and chost_ hold unquantised, quantised)
    FFCZ_HIP(hipMemcpyAsync(chost_, cstats_, 2 * sizeof(CertStats), hipMemcpyDeviceToHost, s_));
    cert_pending_ = true;
}

template <typename T>
void Solver<T>::certify_both_results(CertResult* out) {
    FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null result");
    FFCZ_REQUIRE(cert_pending_, ERR_STATE, "no certification queued");
    FFCZ_HIP(hipStreamSynchronize(s_));
    cert_pending_ = false;
    out[0] = certify_read(false);
    out[1] = certify_read(true);
}

template <typename T>
void Solver<T>::certify_copy_stats(bool quantized, hipStream_t s) {
    FFCZ_HIP(hipMemcpyAsync(chost_ + (quantized ? 1 : 0), cstats_ + (quantized ? 1 : 0), sizeof(CertStats),
                            hipMemcpyDeviceToHost, s));
}

template <typename T>
CertArgs<T> Solver<T>::cert_args(bool quantized) const {
    materialize_ds();
    CertArgs<T> c{};
    c.nx = f_.nx;
    c.ny = f_.ny;
    c.nz = f_.nz;
    c.e0 = e0_ext_ ? e0_ext_ : e0_;
    c.ds = ds_;
    c.df = dfc_;
    if (quantized) {
        // the dequantisers read the extraction's range on the device
        c.qs = spatial_quant();
        c.qf = freq_quant();
    } else {
        c.qs.kind = Q_NONE;
        c.qf.kind = Q_NONE;
    }
    // the frequency accumulator is non-zero only if some clamp fired
    c.have_df = df_touched_;
    c.eps = (double)eps_;
    c.delta = (double)delta_;
    c.delta_dense = init_mode_ == MODE_DIRECT ? delta_dense_ : nullptr;
    c.work_real = cert_real_;
    c.work_spec = cert_spec_;
    c.stats = cstats_ + (quantized ? 1 : 0);
    return c;
}

template <typename T>
CertResult Solver<T>::certify_collect(bool quantized) {
    FFCZ_REQUIRE(chost_ != nullptr, ERR_STATE, "no certification enqueued");
    FFCZ_HIP(hipStreamSynchronize(s_));
    return certify_read(quantized);
}

template <typename T>
CertResult Solver<T>::certify_adopt(bool quantized, const CertStats& st) {
    FFCZ_REQUIRE(chost_ != nullptr, ERR_STATE, "no certification enqueued");
    chost_[quantized ? 1 : 0] = st;
    return certify_read(quantized);
}

template <typename T>
CertResult Solver<T>::certify_read(bool quantized) {
    FFCZ_REQUIRE(chost_ != nullptr, ERR_STATE, "no certification enqueued");
    if (quantized) sync_extract();  // a quantiser overflow voids the quantised check
    const CertStats hs = chost_[quantized ? 1 : 0];
    CertResult r;
    r.spatial_violations = hs.spatial_viol;
    r.freq_violations = hs.freq_viol;
    r.spatial_max_excess = hs.spatial_max_bits ? __builtin_bit_cast(double, hs.spatial_max_bits) : 0.0;
    r.freq_max_excess = hs.freq_max_bits ? __builtin_bit_cast(double, hs.freq_max_bits) : 0.0;
    r.ok = (hs.spatial_viol == 0 && hs.freq_viol == 0) ? 1 : 0;
    return r;
}

template <typename T>
void Solver<T>::get_results(T* e) {
    // projection_algorithm_impl.cuh:904-910
    materialize_e();
    FFCZ_HIP(hipMemcpyAsync(e, e_, f_.N * sizeof(T), hipMemcpyDefault, s_));
    FFCZ_HIP(hipStreamSynchronize(s_));
}

template <typename T>
void Solver<T>::time_iteration(int reps, float* ms5) {
    // ms5[0..4]: the five passes; ms5[5]: a whole iteration
    FFCZ_REQUIRE(solved_, ERR_STATE, "time_iteration requires a completed solve");
    FFCZ_REQUIRE(reps >= 1, ERR_INVALID, "reps must be >= 1");
    materialize_e();
    accumulators_clobbered();
    const bool fused = use_fused();
    if (fused) ensure_zz_log();
    reset_state<T>(st_, 1 << 30, -1.0, s_);  // tol < 0: never converges
    const IterArgs<T> a = iter_args();
    // passes: events between the launches of real iterations (an interval
    // is attributed to the launch that ends it); the fused form's iterations
    // after the first (no forward z pass of the initial error)
    float acc[6] = {0, 0, 0, 0, 0, 0};  // [5]: violation reduction
    EventLog log;
    run_iteration(a, fused, false, nullptr);  // warm-up
    for (int r = 0; r < reps; ++r) {
        log.used = 0;
        run_iteration(a, fused, false, &log);
        FFCZ_REQUIRE(log.used >= 2, ERR_HIP, "event creation failed");
        FFCZ_HIP(hipEventSynchronize(log.pool[log.used - 1]));
        for (size_t i = 1; i < log.used; ++i) {
            float t = 0;
            FFCZ_HIP(hipEventElapsedTime(&t, log.pool[i - 1], log.pool[i]));
            acc[log.kinds[i]] += t;
        }
    }
    for (int k = 0; k < 5; ++k) ms5[k] = acc[k] / (float)reps;
    // whole iterations (five passes + the violation reduction), no events
    // between the launches
    hipEvent_t e0 = log.pool[0], e1 = log.pool[1];
    FFCZ_HIP(hipEventRecord(e0, s_));
    for (int r = 0; r < reps; ++r) run_iteration(a, fused, false, nullptr);
    FFCZ_HIP(hipEventRecord(e1, s_));
    FFCZ_HIP(hipEventSynchronize(e1));
    float t = 0;
    FFCZ_HIP(hipEventElapsedTime(&t, e0, e1));
    ms5[5] = t / (float)reps;
    reset_state<T>(st_, 0, 0.0, s_);
    FFCZ_HIP(hipStreamSynchronize(s_));
}

template <typename T>
void Solver<T>::set_fft_precision(int prec) {
    FFCZ_REQUIRE(prec >= PREC_VALU && prec <= PREC_FP16X3, ERR_INVALID, "unknown FFT precision variant");
    FFCZ_REQUIRE(prec == PREC_VALU || sizeof(T) == 4, ERR_INVALID,
                 "the MFMA split-precision FFT variants are FP32 transforms; FP64 solvers use the vector ALU");
    f_.prec = prec;
}

// ---------------------------------------------------------------------------
// autotuner: tile variant of every pass kernel, chosen by timing each
// candidate on this field's own buffers; plans are kept per device, type,
// shape, FFT variant and bound mode, for the process and on disk
// (plan_cache.hpp)
// ---------------------------------------------------------------------------
namespace {
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        FFCZ_HIP(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
        if (hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess) {
            (void)hipEventDestroy(a);
            throw Error(ERR_HIP, "event creation failed");
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    template <class F>
    float time(hipStream_t s, int reps, F&& fn) {
        fn();  // warm-up
        FFCZ_HIP(hipEventRecord(a, s));
        for (int r = 0; r < reps; ++r) fn();
        FFCZ_HIP(hipEventRecord(b, s));
        FFCZ_HIP(hipEventSynchronize(b));
        float t = 0;
        FFCZ_HIP(hipEventElapsedTime(&t, a, b));
        return t / (float)reps;
    }
};
}  // namespace

template <typename T>
TileVars Solver<T>::tile_vars() const {
    TileVars v{};
    v.zf = f_.zf_var;
    v.zi = f_.zi_var;
    v.y = f_.y_var;
    v.x = f_.x_var;
    if (f64_) {
        v.czf = f64_->zf_var;
        v.cy = f64_->y_var;
        v.cx = f64_->x_var;
        v.cpair = f64_->cz_pair ? 1 : 0;
    }
    return v;
}

template <typename T>
void Solver<T>::set_tile_vars(const TileVars& v) {
    const int* p = &v.zf;
    for (int i = 0; i < 7; ++i)
        FFCZ_REQUIRE(p[i] == 0 || p[i] == 1 || ((i == 2 || i == 3) && p[i] == 2), ERR_INVALID,
                     "tile variants are 0 or 1 (the iteration's y and x passes also 2: radices in reversed order)");
    f_.zf_var = v.zf;
    f_.zi_var = v.zi;
    f_.y_var = v.y;
    f_.x_var = v.x;
    ensure_cert_buffers();
    f64_->zf_var = v.czf;
    f64_->y_var = v.cy;
    f64_->x_var = v.cx;
    f64_->cz_pair = v.cpair != 0;
}

template <typename T>
TileVars Solver<T>::autotune(int reps, bool use_cache, float* ms_out) {
    FFCZ_REQUIRE(initialized_, ERR_STATE, "autotune requires an initialized solver");
    FFCZ_REQUIRE(reps >= 1, ERR_INVALID, "reps must be >= 1");
    const PlanKey key{sizeof(T), f_.nx, f_.ny, f_.nz, f_.prec, init_mode_};
    if (use_cache && !ms_out) {
        TileVars cached{};
        const int src = plan_lookup(key, cached);
        if (src != TUNE_NONE) {
            set_tile_vars(cached);
            tune_source_ = src;
            return cached;
        }
    }
    ensure_cert_buffers();
    // the tuner runs whole iterations on this solver's buffers (the passes
    // share caches, so each is timed in its real sequence): a zero initial
    // error (no violations, as in a converged field), zeroed accumulators and
    // a state that never stops the kernels
    materialize_e();
    e_lazy_ = false;
    FFCZ_HIP(hipMemsetAsync(e0_, 0, f_.N * sizeof(T), s_));
    reset_edit_accumulators();
    materialize_ds();
    reset_state<T>(st_, 1 << 30, -1.0, s_);
    f_.zf_var = f_.y_var = f_.x_var = f_.zi_var = 0;
    const bool fused = use_fused();
    if (fused) ensure_zz_log();
    const IterArgs<T> a = iter_args();
    EventPair ev;
    float ms[9] = {};
    // each timing starts from iteration 0: no fused pass replays deferred
    // clamps logged under another tile variant
    auto t_iter = [&] {
        reset_state<T>(st_, 1 << 30, -1.0, s_);
        return ev.time(s_, reps, [&] { run_iteration(a, fused, false, nullptr); });
    };
    float best = t_iter();
    ms[0] = best;
    // coordinate descent over the four iteration pass kernels; an alternative
    // must win clearly, so timing noise keeps the default.  The MFMA passes
    // run the full-width tiles only; in their iteration just the fused z pass
    // (a vector-ALU kernel) has tile variants.
    // The y and x passes of a vector-ALU iteration also try their length's
    // radices in reversed order (variant 2) where that is a different plan.
    int* fv[4] = {&f_.zf_var, &f_.y_var, &f_.x_var, &f_.zi_var};
    const size_t plen[4] = {0, f_.ny, f_.nx, 0};
    for (int p = 0; p < 4; ++p) {
        if (f_.prec != PREC_VALU && !(p == 3 && fused)) {
            ms[1 + p] = best;
            continue;
        }
        float tmin = 0.f;
        for (int cand = 1; cand <= 2; ++cand) {
            if (cand == 2 && !((p == 1 || p == 2) && fast_s_rev_supported((int)plen[p]))) continue;
            const int keep = *fv[p];
            *fv[p] = cand;
            const float t = t_iter();
            tmin = (cand == 1 || t < tmin) ? t : tmin;
            if (t < 0.97f * best) best = t;
            else *fv[p] = keep;
        }
        ms[1 + p] = tmin;
    }
    ms[5] = best;
    reset_state<T>(st_, 0, 0.0, s_);
    // certification passes: the paired FP64 check of the unquantised and
    // quantised fields, as certify_both runs it, for each combination of the
    // z, y and x variants (combos 0-7, the two forward z transforms in one
    // launch where the z variant is 0) and for z variant 0 with one z launch
    // per field (combos 8-11)
    CertArgs<T> c = cert_args(false);
    c.have_df = false;
    CertArgs<T> c2 = cert_args(true);
    c2.have_df = false;
    c2.work_spec = cert_spec_ + f64_->spec_elems();
    auto set_combo = [&](int combo) {
        const int k = combo < 8 ? combo : (combo - 8) << 1;
        f64_->zf_var = k & 1;
        f64_->y_var = (k >> 1) & 1;
        f64_->x_var = (k >> 2) & 1;
        f64_->cz_pair = combo < 8;
    };
    float cbest = 0;
    int cb = 0;
    const bool cert_defined = init_mode_ == MODE_ABS || init_mode_ == MODE_DIRECT;
    for (int combo = 0; combo < (cert_defined ? 12 : 0); ++combo) {
        set_combo(combo);
        const float t = ev.time(s_, reps, [&] { ffcz::certify_pair<T>(*f64_, c, c2, s_); });
        if (combo == 0) ms[6] = t;
        if (combo == 0 || t < 0.97f * cbest) {
            cbest = t;
            cb = combo;
        }
    }
    set_combo(cb);
    if (cert_defined) ms[8] = ev.time(s_, reps, [&] { ffcz::certify<T>(*f64_, c, s_); });
    accumulators_clobbered();
    reset_edit_accumulators();
    FFCZ_HIP(hipStreamSynchronize(s_));
    solved_ = false;
    extracted_ = false;
    if (ms_out) {
        ms[7] = cbest;
        for (int i = 0; i < 9; ++i) ms_out[i] = ms[i];
    }
    const TileVars v = tile_vars();
    plan_store(key, v);
    tune_source_ = TUNE_TIMED;
    return v;
}

template class Solver<float>;
template class Solver<double>;

}  // namespace ffcz
