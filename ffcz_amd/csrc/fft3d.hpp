// 3D real FFT (R2C / C2R) for row-major Nx x Ny x Nz fields, z fastest, built
// from three LDS-tiled pencil passes.  The half spectrum is stored padded:
// bin (kx, ky, kz), kz <= Nz/2, lives at ((kx*Ny)+ky)*Mp + kz with the row
// pitch Mp a multiple of 16 elements (128 B for FP32), so every y/x tile load
// is an aligned full-line segment.  The reference's compact layout (pitch
// Nh = Nz/2+1, as cuFFT writes it) is used only for the edit accumulators.
#pragma once

#include <vector>

#include "fft_core.hpp"

namespace ffcz {

// Loop bookkeeping shared by the iteration kernels (device memory).  The last
// workgroup of the inverse z pass finalises an iteration: it publishes the max
// spatial violation, bumps ``iter`` and sets ``done`` on convergence or when
// ``iter`` reaches ``max_iter``; every iteration kernel returns immediately
// once ``done`` is set, so extra enqueued iterations are no-ops.
struct DevState {
    unsigned int done;
    unsigned int converged;
    unsigned int iter;
    unsigned int max_iter;
    unsigned int blocks_done;
    unsigned int freq_clamped;  // some frequency clamp fired during this solve
    unsigned long long viol_bits;  // max violation of the running iteration (order-preserving bits)
    double tol;
    double last_viol;
    float tol_f;
    unsigned int pad1;
};

// Spatial-bound statistics of a certification pass (FP64).
struct CertStats {
    unsigned long long spatial_viol;
    unsigned long long freq_viol;
    unsigned long long spatial_max_bits;  // max excess |d|-eps (ordbits of double)
    unsigned long long freq_max_bits;     // max excess max(|Re|,|Im|)-delta
};

// Quantiser descriptors (applied on the fly when reconstructing edits).
enum QuantKind : int { Q_NONE = 0, Q_REF16 = 1, Q_STEP = 2 };
struct QuantDesc {
    int kind;
    double vmin, vmax;  // Q_REF16 (cast to T inside the kernels)
    const double* dminmax;  // Q_REF16: {vmin, vmax} in device memory, used instead when set
    double step;        // Q_STEP
    double inv_step;    // Q_STEP: 1/step (0: every quotient by division)
    double threshold;   // Q_REF16 flag threshold (1e-7 in the reference)
};

template <typename R>
struct Fft3d {
    size_t nx = 0, ny = 0, nz = 0;
    size_t N = 0;       // nx*ny*nz
    size_t Nh = 0;      // nz/2+1 bins along z
    size_t F = 0;       // nx*ny*Nh
    size_t Mp = 0;      // padded z pitch of the internal half spectrum
    size_t rows = 0;    // nx*ny
    bool half = true;   // nz even: half-length complex trick
    int Lz = 1;         // complex length of the z transform (nz/2 or nz)
    FftPlan pz{}, py{}, px{};
    int Pz = 1, Wy = 1, Wx = 1;       // pencils per z tile, columns per y/x tile
    int prec = 0;                     // FftPrec of the FP32 pass kernels (0: vector ALU)
    // tile variant of each pass's length-specialised kernel (autotuned; the
    // MFMA variants and batched launches use variant 0)
    int zf_var = 0, zi_var = 0, y_var = 0, x_var = 0;
    // paired certification: both fields' forward z transforms in one launch
    // (tile variant 0), else one launch each (autotuned)
    bool cz_pair = true;
    int zinv_var() const { return prec == 0 ? zi_var : 0; }
    cplx<R>* d_tz = nullptr;   // W_Lz^m
    cplx<R>* d_tz2 = nullptr;  // W_nz^k, k <= nz/2 (R2C post / C2R pre)
    cplx<R>* d_ty = nullptr;
    cplx<R>* d_tx = nullptr;
    size_t spec_elems() const { return rows * Mp; }

    void create(size_t nx_, size_t ny_, size_t nz_);
    void destroy();
};

// ---- plain transforms (initialisation, tests) ------------------------------
// forward R2C: in[N] (real R) -> spec (padded half spectrum)
template <typename R>
void fft_r2c(const Fft3d<R>& f, const R* in, cplx<R>* spec, hipStream_t s);
// inverse C2R, unnormalised times ``scale``: spec is overwritten (y/x passes
// run in place) and out[N] receives the real field.
template <typename R>
void fft_c2r(const Fft3d<R>& f, cplx<R>* spec, R* out, R scale, hipStream_t s);
// padded <-> compact (pitch Nh) layout conversion
template <typename R>
void spec_to_compact(const Fft3d<R>& f, const cplx<R>* spec, cplx<R>* compact, hipStream_t s);
template <typename R>
void compact_to_spec(const Fft3d<R>& f, const cplx<R>* compact, cplx<R>* spec, hipStream_t s);

// ---- iteration passes ------------------------------------------------------
enum FreqMode : int { FM_ABS = 0, FM_DENSE = 1, FM_PTW = 2, FM_NONE = 3 };
// Per-bin rules of the pointwise power-spectrum bound (FM_PTW): unconstrained,
// the power-ratio rule, the DC rule.
enum PtwRule : uint8_t { PTW_NONE = 0, PTW_RATIO = 1, PTW_DC = 2 };

template <typename R>
struct IterArgs {
    R* e;                 // spatial iterate [N]
    const R* e_init;      // initial error [N], read by the first iteration instead of e (optional)
    R* e_copy;            // the first iteration stores a copy of e_init here (optional)
    int ds_fresh;         // ds not yet written since the accumulators were reset: the first
                          // spatial clamp (forward z pass of iteration 1, or the epilogue)
                          // stores every element instead of adding to it
    R* ds;                // spatial edit accumulator [N]
    cplx<R>* df;          // frequency edit accumulator [F], one (Re, Im) pair per bin (compact order)
    cplx<R>* spec;        // padded half spectrum work buffer
    R eps;                // spatial bound used by the iteration
    int fmode;            // FreqMode
    R delta;              // FM_ABS bound
    const R* delta_dense; // FM_DENSE per-bin bounds [F] (inf = unconstrained)
    R dscale;             // FM_DENSE bound multiplier (1 = the given bounds)
    const cplx<R>* orig;  // FM_PTW: F(x) compact layout [F]
    R delta_ptw;
    const uint8_t* ptw_mask;  // FM_PTW: per-bin rule [F] (PtwRule), null = every bin (DC by the DC rule)
    int ptw_corr;         // FM_PTW: corrected power-spectrum bound (not the reference's)
    R norm;               // T(1)/T(N)
    DevState* st;
    R* vpart;             // [rows] per-workgroup violation partials (zero-initialised)
    // log of the fused z pass's deferred clamps (iteration_fused):
    // zz_log_entries() entries of each kind, zz_blocks_max() counts
    uint32_t* log_off;
    R* log_v;
    uint32_t* log_cnt;
};

// forward z (with the deferred spatial clamp of the previous iteration fused
// into the load), forward y, fused x forward+clamp+inverse, inverse y,
// inverse z (+normalise, violation max, iteration bookkeeping).
// Optional per-launch event log (kind: 0 z fwd, 1 y fwd, 2 x fused,
// 3 y inv, 4 z inv; -1 start).  Events come from a caller-owned pool.
struct EventLog {
    std::vector<hipEvent_t> pool;
    std::vector<int> kinds;
    size_t used = 0;
    void mark(int kind, hipStream_t s) {
        if (used >= pool.size()) {
            hipEvent_t e;
            // no system-scope fence: recording must not flush the caches
            // between the kernels it separates
            if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
            pool.push_back(e);
        }
        (void)hipEventRecord(pool[used], s);
        if (kinds.size() <= used) kinds.resize(used + 1);
        kinds[used++] = kind;
    }
    ~EventLog() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};
template <typename R>
void iteration(const Fft3d<R>& f, const IterArgs<R>& a, hipStream_t s, EventLog* log = nullptr);
// One of the iteration's passes over the whole field (kind as in EventLog),
// for timing the kernels one at a time.
template <typename R>
void iteration_kernel(const Fft3d<R>& f, const IterArgs<R>& a, hipStream_t s, int kind);
// The fused iteration form: [y fwd, x fused, y inv, z fused, finalise], the
// z fused pass running the inverse z transform of one iteration, its spatial
// violation, the (speculative) spatial clamp and the forward z transform of
// the next iteration in place on the half spectrum (k2_zz: clamps a
// converged pass could make are deferred through a log); the iterate e is
// not written.  The first iteration of a solve starts with the forward z
// pass of the initial error.  Available where the z axis has a
// length-specialised half-length kernel and no slab interleave is set.
template <typename R>
bool fused_iteration_available(const Fft3d<R>& f);
// workgroups of the fused z pass (its tile variant; the larger of the two
// variants' counts), and the log entries (of each kind) either variant needs
template <typename R>
size_t zz_blocks(const Fft3d<R>& f);
template <typename R>
size_t zz_blocks_max(const Fft3d<R>& f);
template <typename R>
size_t zz_log_entries(const Fft3d<R>& f);
template <typename R>
void iteration_fused(const Fft3d<R>& f, const IterArgs<R>& a, bool first, hipStream_t s, EventLog* log = nullptr);
// After a fused loop: the last pass's deferred clamps (those with an excess
// below tol, which a converged loop drops) go into ds when the loop stopped
// at max_iter without converging; otherwise nothing is done.
// var: the z tile variant the loop ran with (the log's layout)
template <typename R>
void zz_apply(const Fft3d<R>& f, const IterArgs<R>& a, int var, hipStream_t s);
// The iterate after a fused loop: the inverse z transform of the rows the
// last z fused pass wrote (times 1/nz: the clamped iterate) into a.e, and,
// after a converged loop, the logged elements' unclamped values.
template <typename R>
void zz_recover_e(const Fft3d<R>& f, const IterArgs<R>& a, int var, hipStream_t s);

// Final spatial clamp when the loop stopped at max_iter without convergence.
template <typename R>
void epilogue_clamp(const Fft3d<R>& f, const IterArgs<R>& a, hipStream_t s);
template <typename R>
void reset_state(DevState* st, int max_iter, double tol, hipStream_t s);

// ---- one iteration over a batch of fields ----------------------------------
// Fields whose passes all have length-specialised kernels (batch_fusable) run
// the fused iteration form, each pass as one launch per axis length over
// every such field (tile -> field by a prefix table in device memory): the
// forward z pass of the initial error once (batch_prologue), then per
// iteration [y fwd, x fused, y inv, z fused] and one finalising launch with a
// workgroup per field.  The other fields (FP64, MFMA variants, lengths
// without a specialised kernel, slab interleave) run the five-pass
// iteration() one by one inside the same sequence.  A field whose state is
// done skips its tiles, so a batch keeps iterating until every field has
// converged.
template <typename R>
struct BatchPlan {
    struct Group {
        int pass;        // 0 z fwd, 1 y fwd, 2 x fused, 3 y inv, 4 z inv
        int L;           // axis length (z: complex half length)
        int smode;       // strided passes: S_PLAIN / S_CLAMP
        unsigned grid;   // total workgroups
        const void* args;
        const uint32_t* start;
        int n;
    };
    std::vector<Group> groups;   // the iteration's passes
    std::vector<Group> pro;      // the prologue: forward z of the initial errors
    std::vector<std::pair<const Fft3d<R>*, IterArgs<R>>> fallback;
    const void* fin = nullptr;   // finalize table of the batched fields
    int nfin = 0;
    DevState** sts = nullptr;    // every field's state
    DevState* gathered = nullptr;  // gathered copies of the states
    size_t nfields = 0;
    unsigned char* dmem = nullptr;
    size_t dcap = 0;
};
// Builds (or rebuilds) the plan; its device tables are reused when they fit,
// so no work that reads the previous plan may still be pending.
template <typename R>
void batch_plan_build(BatchPlan<R>& bp, const std::vector<const Fft3d<R>*>& fields,
                      const std::vector<IterArgs<R>>& args);
template <typename R>
void batch_plan_free(BatchPlan<R>& bp);
template <typename R>
bool batch_fusable(const Fft3d<R>& f);
template <typename R>
void batch_prologue(const BatchPlan<R>& bp, hipStream_t s);
template <typename R>
void batch_iteration(const BatchPlan<R>& bp, hipStream_t s);
// copies every field's state into host memory `out` (asynchronous on s)
template <typename R>
void batch_gather_states(const BatchPlan<R>& bp, DevState* out, hipStream_t s);
// every field's state reset for a solve of at most max_iter iterations (one launch)
template <typename R>
void batch_reset_states(const BatchPlan<R>& bp, int max_iter, double tol, hipStream_t s);

// ---- FP64 certification ----------------------------------------------------
template <typename T>
struct CertArgs {
    size_t nx, ny, nz;
    const T* e0;          // initial error x' - x [N]
    const T* ds;          // spatial edits [N]
    const cplx<T>* df;    // frequency edits [F], one (Re, Im) pair per bin (compact order)
    QuantDesc qs, qf;     // dequantisation applied to ds / df
    bool have_df;         // any frequency edit present
    double eps;
    double delta;         // uniform bound, or
    const T* delta_dense; // per-bin bounds [F] (nullptr: uniform)
    double* work_real;    // [N] doubles (C2R of df)
    cplx<double>* work_spec;  // padded FP64 spectrum
    CertStats* stats;
};
template <typename T>
void certify(const Fft3d<double>& f64, const CertArgs<T>& a, hipStream_t s);

// Both certifications of one field (unquantised a, quantised b) in one
// sequence: one forward z launch for both (the second read of x'-x and the
// spatial edits comes from cache), one y launch over both spectra when b's
// spectrum directly follows a's, then the two frequency checks.
template <typename T>
void certify_pair(const Fft3d<double>& f64, const CertArgs<T>& a, const CertArgs<T>& b, hipStream_t s);

// A device buffer that only grows (tables of batched launches).
struct DevBuffer {
    unsigned char* p = nullptr;
    size_t cap = 0;
    void ensure(size_t n);
    void release();
};
// The certification of many fields with one launch per pass and axis length
// (float fields with length-specialised kernels; the others run certify()
// one by one on the same stream).  Results land in each field's stats and,
// when host_out is set (pinned, one entry per field), are gathered and copied
// there behind the passes.
template <typename T>
void certify_batch(const std::vector<const Fft3d<double>*>& f64s, const std::vector<CertArgs<T>>& args,
                   DevBuffer& tables, hipStream_t s, CertStats* host_out = nullptr);

}  // namespace ffcz
