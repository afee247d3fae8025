// ProjectionSolver state on one GPU (reference: class ProjectionSolver,
// projection_algorithm.cuh:238-333, projection_algorithm_impl.cuh:1-910).
#pragma once

#include <memory>
#include <vector>

#include "extract.hpp"
#include "fft3d.hpp"

namespace ffcz {

// MODE_PARTIAL_PTW: the pointwise power-spectrum bound on the bins listed by
// initialize_partial_ptw only (the reference's partial ptw kernel).
enum SolveMode : int { MODE_DIRECT = 0, MODE_ABS = 1, MODE_PTW = 2, MODE_PARTIAL_PTW = 3 };

// Tile variants (0/1; the y and x passes also 2, their radices reversed) of
// the iteration's z forward, z inverse, y and x pass
// kernels and of the FP64 certification's z, y and x passes.
struct TileVars {
    int zf, zi, y, x, czf, cy, cx;
    int cpair = 1;  // paired certification's z launch (Fft3d::cz_pair)
};

struct CertResult {
    unsigned long long spatial_violations = 0;
    unsigned long long freq_violations = 0;
    double spatial_max_excess = 0;
    double freq_max_excess = 0;
    int ok = 0;
};

template <typename T>
class Solver {
   public:
    Solver(size_t nx, size_t ny, size_t nz);
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    // initialize* (projection_algorithm_impl.cuh:598-738).  Pointers may be
    // host or device memory.
    void init_direct(const T* x, const size_t* idx, const T* delta_k, T eps, size_t K);
    void init_all_abs(const T* x, T eps, T delta);
    void init_all_ptw(const T* x, T eps, T delta_ptw);
    void init_partial_ptw(const T* x, const size_t* idx, T delta_ptw, T eps, size_t K);
    // Which pointwise bound the ptw modes apply: the reference's arithmetic
    // (default) or the corrected one (power ratio within [1-d, 1+d], Re and Im
    // clamped relative to |X|, DC held to the same rule).
    void set_ptw_semantics(bool corrected) { ptw_corr_ = corrected; }
    bool ptw_corrected() const { return ptw_corr_; }

    // Projection loop only (impl:791-808 / 835-852 / 879-896).
    void solve_loop(int mode, const T* e0, int max_iter, T tol);
    // the loop in pieces, for a caller that launches the iterations itself
    // (batches): checks + initial copy + state reset (left to the caller
    // when reset is false), then the epilogue and the read-back of the final
    // state, both on stream s
    void prepare_solve(int mode, const T* e0, int max_iter, T tol, hipStream_t s, bool reset = true);
    // known: the final state when the caller has already read it (no extra wait)
    void finish_solve(hipStream_t s, const DevState* known = nullptr);
    // after prepare_solve, for a caller running the fused iteration form
    // itself (batches): this solve's passes use tile variant var
    void set_fused_run(int var) {
        ensure_zz_log();
        fused_ = true;
        fused_var_ = var;
    }
    // iterations run outside a solve (timing) leave the accumulators undefined
    void accumulators_clobbered();
    // solve / solve_all_abs / solve_all_ptw: loop + extraction (impl:772-902).
    void solve(int mode, const T* e0, int max_iter, T tol) {
        solve_loop(mode, e0, max_iter, tol);
        extract();
    }
    void extract();
    // A batch's extraction: the field's two edit arrays (frequency, then
    // spatial) as segments of one launch on stream s, and that launch's
    // results taken over afterwards.
    void extract_segments(std::vector<ExtractSeg<T>>& out, hipStream_t s);
    void extract_adopt(const ExtractGlobals& hf, const ExtractGlobals& hs);
    void reset_edit_accumulators();
    CertResult certify(bool quantized);
    // both certifications (unquantised [0], quantised [1]) in one paired
    // sequence; requires the extraction
    void certify_both(CertResult* out2);
    // the same sequence queued without waiting, its results read (waiting for
    // them) by certify_both_results; work queued in between (the next field's
    // solve) runs after it on the solver's stream
    void certify_both_enqueue();
    void certify_both_results(CertResult* out2);
    // mean ms per launch of the five iteration passes, then per whole iteration
    void time_iteration(int reps, float* ms6);
    // Per-pass timing inside real solves: HIP events between the launches of
    // every iteration the loop runs (fence-free), averaged over the iterations
    // that did work.  enable resets the averages.
    void set_pass_timing(bool enable);
    // ms[6]: mean ms per launch of z fwd, y fwd, x fused, y inv, z inv,
    // finalize; returns the number of iterations timed.  The first launch of
    // each chunk is left out: the device waits for the host there, and that
    // wait would count as kernel time.
    long pass_timing(float* ms6) const;
    // n6[6]: how many launches of each of those passes were timed
    void pass_counts(long* n6) const;

    // margins: the loop iterates against eps*(1-ms) and delta*(1-mf); the
    // certification always checks the bounds given to initialize*.
    void set_margins(double ms, double mf) {
        margin_s_ = ms;
        margin_f_ = mf;
    }
    void set_quantizer(int kind, double freq_step, double spatial_step);

    // getters (copies into caller memory; host or device)
    void get_results(T* e);
    // the iterate (recovered first when the last solve ran the fused form)
    const T* d_e() { return materialize_e(), e_; }
    // the frequency edits in the reference's split layout [Re | Im]
    const T* d_df() const { return materialize_df_split(), df_; }
    const T* d_ds() const { return materialize_ds(), ds_; }
    size_t N() const { return f_.N; }
    size_t F() const { return f_.F; }
    size_t nx() const { return f_.nx; }
    size_t ny() const { return f_.ny; }
    size_t nz() const { return f_.nz; }
    int iterations() const { return iters_; }
    int init_mode() const { return init_mode_; }
    // the mode of the last solve (the initialize* call's before any)
    int solve_mode() const { return solve_mode_ >= 0 ? solve_mode_ : init_mode_; }
    T eps() const { return eps_; }
    T delta() const { return init_mode_ == MODE_PTW || init_mode_ == MODE_PARTIAL_PTW ? delta_ptw_ : delta_; }
    double violation() const { return viol_; }
    bool converged() const { return converged_; }
    bool extracted() const { return extracted_; }
    // extraction results (wait for a pending extraction first)
    size_t n_active_freq() const { return sync_extract(), xf_.count; }
    size_t n_active_spatial() const { return sync_extract(), xs_.count; }
    double freq_min() const { return sync_extract(), xf_.vmin; }
    double freq_max() const { return sync_extract(), xf_.vmax; }
    double spatial_min() const { return sync_extract(), xs_.vmin; }
    double spatial_max() const { return sync_extract(), xs_.vmax; }
    // extract() only enqueues; this waits for it and checks the quantiser range
    void sync_extract() const;
    // certify() in two halves: enqueue the FP64 passes (slot = quantized), then
    // wait for them and read the result; several solvers can be in flight
    void certify_enqueue(bool quantized);
    CertResult certify_collect(bool quantized);
    // pieces for a caller that queues the FP64 passes of many solvers itself
    // (batches): checks and arguments; the FP64 plan; the copy of the
    // counters to the host on stream s; reading them once s has finished
    CertArgs<T> certify_prepare(bool quantized);
    const Fft3d<double>& cert_fft() const { return *f64_; }
    void certify_copy_stats(bool quantized, hipStream_t s);
    CertResult certify_read(bool quantized);
    // a batch's certification: the stats its launch gathered for this field
    CertResult certify_adopt(bool quantized, const CertStats& st);
    // Tile variant of each pass kernel (iteration and FP64 certification).
    // autotune times every candidate on this field's buffers (after
    // initialize*; it clobbers the iterate and resets the edit accumulators)
    // and keeps the fastest; results are cached per shape and FFT variant.
    // ms (optional, 9 floats): whole-iteration times with the default
    // variants, with the z forward, y, x and z inverse pass switched in turn
    // (coordinate descent), the chosen iteration; then the certification with
    // the default and the chosen variants; the last slot is 0.
    TileVars autotune(int reps, bool use_cache, float* ms = nullptr);
    // where the last autotune's plan came from (TuneSource: 0 none yet,
    // 1 timed, 2 this process's plans, 3 the plan file)
    int tune_source() const { return tune_source_; }
    TileVars tile_vars() const;
    void set_tile_vars(const TileVars& v);
    const uint8_t* d_freq_pack() const { return fpack_; }
    const uint8_t* d_spatial_pack() const { return spack_; }
    const T* d_freq_compact() const { return fcomp_; }
    const T* d_spatial_compact() const { return scomp_; }
    const uint16_t* d_freq_quant() const { return fquant_; }
    const uint16_t* d_spatial_quant() const { return squant_; }
    hipStream_t stream() const { return s_; }
    // per-kernel timing of the iteration (for roofline reporting)
    const Fft3d<T>& fft() const { return f_; }
    // FftPrec of the iteration's pass kernels (float solvers only)
    void set_fft_precision(int prec);
    int fft_precision() const { return f_.prec; }
    IterArgs<T> iter_args() const;
    QuantDesc freq_quant() const;
    QuantDesc spatial_quant() const;

    // Graph replay of the iteration in solve_loop (default on): the loop's
    // iterations after the first are one HIP graph launch each, captured once
    // and captured again whenever the iteration's arguments, tile variants or
    // FFT variant change.  Per-pass timing runs the launches eagerly.
    void set_graph(bool on);
    // borrow=true: a device-resident initial error passed to a solve is read
    // in place by the certifications that follow, instead of being copied
    // (the caller keeps it unchanged until the next solve)
    void set_borrow_e0(bool borrow) { borrow_e0_ = borrow; }
    bool graph() const { return graph_on_; }

   private:
    CertArgs<T> cert_args(bool quantized) const;
    void account_pass_times(int iterations);
    bool ensure_graph(const IterArgs<T>& a, bool fused);
    void release_graph();
    hipGraphExec_t gexec_ = nullptr;
    IterArgs<T> gargs_{};
    int gkey_[8] = {};
    bool graph_on_ = true;
    bool pass_timing_ = false;
    std::unique_ptr<EventLog> plog_;
    std::vector<int> chunk_first_;  // iterations (1-based) that start a chunk
    double pass_ms_[6] = {};
    long pass_cnt_[6] = {};
    long pass_n_ = 0;
    void release();
    void upload_x(const T* x);
    void ensure_extract_buffers();
    void ensure_cert_buffers();

    Fft3d<T> f_;
    hipStream_t s_ = nullptr;
    T* x_ = nullptr;
    T* e_ = nullptr;
    T* e0_ = nullptr;
    T* df_ = nullptr;           // frequency edits, reference layout [Re | Im] (derived from dfc_)
    cplx<T>* dfc_ = nullptr;    // frequency edits, one (Re, Im) pair per bin: what the passes update
    mutable bool df_split_stale_ = false;  // df_ behind dfc_
    void materialize_df_split(hipStream_t s = nullptr) const;
    T* ds_ = nullptr;
    cplx<T>* spec_ = nullptr;
    DevState* st_ = nullptr;
    DevState* sthost_ = nullptr;  // pinned host copy of st_ (the loop's checks)
    DevState read_state(hipStream_t s);
    T* vpart_ = nullptr;  // per-workgroup violation partials of the inverse z pass
    T* delta_dense_ = nullptr;   // direct mode (and certification of it)
    cplx<T>* orig_ = nullptr;    // F(x), compact layout (ptw modes)
    size_t* partial_idx_ = nullptr;
    size_t partial_K_ = 0;
    // per-bin PtwRule of the partial ptw mode for the semantics ptw_mask_corr_
    // (built at the first solve that needs it)
    uint8_t* ptw_mask_ = nullptr;
    int ptw_mask_corr_ = -1;
    bool ptw_corr_ = false;
    void build_ptw_mask();
    // extraction outputs
    uint8_t* fpack_ = nullptr;
    uint8_t* spack_ = nullptr;
    T* fcomp_ = nullptr;
    T* scomp_ = nullptr;
    uint16_t* fquant_ = nullptr;
    uint16_t* squant_ = nullptr;
    void* xscratch_ = nullptr;
    ExtractGlobals* xglob_ = nullptr;  // device results of the two extractions (freq, spatial)
    ExtractGlobals* xhost_ = nullptr;  // pinned host copies
    mutable ExtractResult xf_, xs_;
    mutable bool xpending_ = false;
    bool extracted_ = false;
    // certification
    std::unique_ptr<Fft3d<double>> f64_;
    double* cert_real_ = nullptr;
    cplx<double>* cert_spec_ = nullptr;   // two FP64 spectra (the paired certification uses both)
    double* cert_real2_ = nullptr;         // second C2R work field (paired, frequency edits present)
    CertStats* cstats_ = nullptr;  // [2]: unquantised, quantised
    CertStats* chost_ = nullptr;   // pinned host copies
    bool cert_pending_ = false;    // a queued certification's results not yet read
    // parameters
    bool initialized_ = false;
    int init_mode_ = -1;
    int solve_mode_ = -1;  // the running (or last) solve's mode; -1 before the first solve after an initialize*
    T eps_ = 0, delta_ = 0, delta_ptw_ = 0;
    double margin_s_ = 0, margin_f_ = 0;
    int qkind_ = Q_REF16;
    double qstep_f_ = 0, qstep_s_ = 0;
    int iters_ = 0;
    double viol_ = 0;
    bool converged_ = false;
    bool solved_ = false;
    bool df_touched_ = true;   // df may be non-zero (a clamp fired, or contents unknown)
    // ds has not been written since the accumulators were reset: zeroed
    // lazily, or stored whole by the next solve's first spatial clamp
    mutable bool ds_fresh_ = false;
    const T* e_src_ = nullptr;  // device initial error read by the running solve's first iteration
    bool borrow_e0_ = false;
    const T* e0_ext_ = nullptr;  // the borrowed initial error of the last solve (null: e0_ holds it)
    void materialize_ds(hipStream_t s = nullptr) const;
    // Fused iteration form (iteration_fused): its log of deferred clamps, whether the last
    // solve ran it, and the iterate it leaves in the spectrum rows until asked for
    bool use_fused() const { return fused_iteration_available(f_); }
    void ensure_zz_log();
    void materialize_e();
    void run_iteration(const IterArgs<T>& a, bool fused, bool first, EventLog* log);
    uint32_t* zlog_off_ = nullptr;
    T* zlog_v_ = nullptr;
    uint32_t* zlog_cnt_ = nullptr;
    bool fused_ = false;
    int fused_var_ = 0;  // the z tile variant of the last fused loop (its log layout)
    int tune_source_ = 0;  // TuneSource of the last autotune
    bool e_lazy_ = false;
    void take_extract_results(const ExtractGlobals& hf, const ExtractGlobals& hs) const;
};

}  // namespace ffcz
