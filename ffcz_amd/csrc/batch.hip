// Batched projection loop (batch.hpp).
#include <algorithm>
#include <exception>

#include "batch.hpp"
#include "fft_fast.hpp"

namespace ffcz {

template <typename T>
Batch<T>::Batch(std::vector<Solver<T>*> fields) : f_(std::move(fields)) {
    FFCZ_REQUIRE(!f_.empty(), ERR_INVALID, "a batch needs at least one field");
    for (auto* p : f_) FFCZ_REQUIRE(p != nullptr, ERR_INVALID, "null field");
    FFCZ_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    FFCZ_HIP(hipHostMalloc((void**)&hstate_, f_.size() * sizeof(DevState), hipHostMallocDefault));
    FFCZ_HIP(hipHostMalloc((void**)&cert_host_, f_.size() * sizeof(CertStats), hipHostMallocDefault));
}

template <typename T>
Batch<T>::~Batch() {
    release_graph();
    batch_plan_free(plan_);
    cert_tables_.release();
    xwork_.release();
    if (hstate_) (void)hipHostFree(hstate_);
    if (cert_host_) (void)hipHostFree(cert_host_);
    if (s_) (void)hipStreamDestroy(s_);
}

template <typename T>
void Batch<T>::release_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    gexec_ = nullptr;
}

template <typename T>
void Batch<T>::solve(int mode, const T* const* e0, int max_iter, T tol, bool use_graph) {
    FFCZ_REQUIRE(e0 != nullptr, ERR_INVALID, "null initial-error table");
    // the fields' own streams may still run their initialisation (a query
    // first: most are idle, and waiting on an idle stream still costs a
    // round trip per field)
    for (auto* f : f_) {
        const hipError_t q = hipStreamQuery(f->stream());
        if (q == hipErrorNotReady) FFCZ_HIP(hipStreamSynchronize(f->stream()));
        else FFCZ_HIP(q);
    }
    for (size_t i = 0; i < f_.size(); ++i) f_[i]->prepare_solve(mode, e0[i], max_iter, tol, s_, false);
    // the batched fields run the fused iteration form (its batched launches
    // use the batch's z tile variant, which sets the log layout); the others
    // the five-pass form
    for (auto* f : f_)
        if (batch_fusable(f->fft())) f->set_fused_run(fast_zb_var(f->fft().Lz));
    build(use_graph && max_iter > 0);
    batch_reset_states(plan_, max_iter, (double)tol, s_);
    if (max_iter > 0) batch_prologue(plan_, s_);
    int launched = 0, chunk = 2;
    while (launched < max_iter) {
        const int k = std::min(chunk, max_iter - launched);
        for (int i = 0; i < k; ++i) {
            if (gexec_) FFCZ_HIP(hipGraphLaunch(gexec_, s_));
            else batch_iteration(plan_, s_);
        }
        launched += k;
        batch_gather_states(plan_, hstate_, s_);
        FFCZ_HIP(hipStreamSynchronize(s_));
        bool all = true;
        for (size_t i = 0; i < f_.size(); ++i) all = all && hstate_[i].done != 0;
        if (all) break;
        chunk = std::min(chunk * 2, 16);
    }
    launched_ = launched;
    graph_ = use_graph;
    // every state was read at the last check (the epilogue leaves them as they are)
    for (size_t i = 0; i < f_.size(); ++i) f_[i]->finish_solve(s_, max_iter > 0 ? &hstate_[i] : nullptr);
}

template <typename T>
void Batch<T>::build(bool use_graph) {
    // earlier work on the batch stream may still read the plan and graph replaced here
    FFCZ_HIP(hipStreamSynchronize(s_));
    std::vector<const Fft3d<T>*> ffts;
    std::vector<IterArgs<T>> args;
    for (auto* f : f_) {
        ffts.push_back(&f->fft());
        args.push_back(f->iter_args());
    }
    batch_plan_build(plan_, ffts, args);
    // The captured iteration reads every batched field's arguments from the
    // plan's device tables, which the build above rewrote in place.  The graph
    // is captured again only when its launches change (axis lengths, grids,
    // table addresses) or when fields run outside the batched kernels (their
    // arguments are kernel parameters inside the graph).
    std::vector<uintptr_t> sig;
    for (const auto& g : plan_.groups)
        sig.insert(sig.end(), {(uintptr_t)g.pass, (uintptr_t)g.L, (uintptr_t)g.smode, (uintptr_t)g.grid,
                               (uintptr_t)g.args, (uintptr_t)g.start, (uintptr_t)g.n});
    sig.insert(sig.end(), {(uintptr_t)plan_.fin, (uintptr_t)plan_.nfin});
    if (use_graph && gexec_ && plan_.fallback.empty() && sig == gsig_) return;
    release_graph();
    gsig_.clear();
    if (use_graph) {
        hipGraph_t g = nullptr;
        FFCZ_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        try {
            batch_iteration(plan_, s_);
        } catch (...) {
            (void)hipStreamEndCapture(s_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        FFCZ_HIP(hipStreamEndCapture(s_, &g));
        const hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FFCZ_HIP(e);
        gsig_ = std::move(sig);
    }
}

template <typename T>
float Batch<T>::time_iteration(int reps) {
    FFCZ_REQUIRE(plan_.nfields == f_.size(), ERR_STATE, "time_iteration requires a completed batch solve");
    FFCZ_REQUIRE(reps >= 1, ERR_INVALID, "reps must be >= 1");
    // the solve's plan may point at the callers' initial errors: rebuild it
    // over the fields' own buffers
    for (auto* f : f_) {
        f->accumulators_clobbered();
        FFCZ_HIP(hipStreamSynchronize(f->stream()));
    }
    build(graph_);
    for (auto* f : f_) reset_state<T>(f->iter_args().st, 1 << 30, -1.0, s_);  // tol < 0: never converges
    auto one = [&] {
        if (gexec_) FFCZ_HIP(hipGraphLaunch(gexec_, s_));
        else batch_iteration(plan_, s_);
    };
    one();  // warm-up
    hipEvent_t e0 = nullptr, e1 = nullptr;
    FFCZ_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
    float t = 0;
    try {
        FFCZ_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
        FFCZ_HIP(hipEventRecord(e0, s_));
        for (int r = 0; r < reps; ++r) one();
        FFCZ_HIP(hipEventRecord(e1, s_));
        FFCZ_HIP(hipEventSynchronize(e1));
        FFCZ_HIP(hipEventElapsedTime(&t, e0, e1));
    } catch (...) {
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto* f : f_) reset_state<T>(f->iter_args().st, 0, 0.0, s_);
    FFCZ_HIP(hipStreamSynchronize(s_));
    return t / (float)reps;
}

template <typename T>
void Batch<T>::extract_all() {
    // every field's two edit arrays in one launch of each extraction kernel
    // on the batch stream, after the fields' own streams have drained
    std::vector<ExtractSeg<T>> segs;
    for (auto* f : f_) FFCZ_HIP(hipStreamSynchronize(f->stream()));
    for (auto* f : f_) f->extract_segments(segs, s_);
    const ExtractGlobals* h = extract_edits_batch<T>(segs, xwork_, s_);
    FFCZ_HIP(hipStreamSynchronize(s_));
    // every field takes its results; the first field's error is reported
    std::exception_ptr first;
    for (size_t i = 0; i < f_.size(); ++i) {
        try {
            f_[i]->extract_adopt(h[2 * i], h[2 * i + 1]);
        } catch (...) {
            if (!first) first = std::current_exception();
        }
    }
    if (first) std::rethrow_exception(first);
}

template <typename T>
void Batch<T>::certify_all(bool quantized, CertResult* out) {
    FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null result array");
    // every field's FP64 passes as one launch per pass and axis length on the
    // batch stream, after the fields' own streams (extraction) have drained
    std::vector<const Fft3d<double>*> ffts;
    std::vector<CertArgs<T>> args;
    for (auto* f : f_) args.push_back(f->certify_prepare(quantized));
    for (auto* f : f_) {
        FFCZ_HIP(hipStreamSynchronize(f->stream()));
        ffts.push_back(&f->cert_fft());
    }
    certify_batch<T>(ffts, args, cert_tables_, s_, cert_host_);
    FFCZ_HIP(hipStreamSynchronize(s_));
    for (size_t i = 0; i < f_.size(); ++i) out[i] = f_[i]->certify_adopt(quantized, cert_host_[i]);
}

template class Batch<float>;
template class Batch<double>;

}  // namespace ffcz
