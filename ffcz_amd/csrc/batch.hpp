// Many fields of uneven shapes corrected together.
//
// The fields are ordinary solvers (each initialised, extracted and certified
// on its own); Batch::solve runs their projection loops in lock step: every
// iteration is one launch per pass and axis length over all fields
// (fft3d.hpp BatchPlan), each field gated by its own convergence state on the
// device, and the iteration is captured in a HIP graph that is replayed
// between the host's convergence checks (and kept for later solves whose
// launches are the same).
#pragma once

#include <cstdint>
#include <vector>

#include "solver.hpp"

namespace ffcz {

template <typename T>
class Batch {
   public:
    explicit Batch(std::vector<Solver<T>*> fields);
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    size_t size() const { return f_.size(); }
    Solver<T>& field(size_t i) { return *f_.at(i); }
    // solve*(e0[i], max_iter, tol) on every field (projection_algorithm_impl.cuh:772-858
    // per field), without the trailing extraction.  use_graph: replay a captured
    // iteration instead of launching its kernels one by one.
    void solve(int mode, const T* const* e0, int max_iter, T tol, bool use_graph);
    // iterations issued by the last solve (chunked; at least the slowest field's count)
    int launched() const { return launched_; }
    // mean ms of one whole batched iteration (every field iterating), timed
    // with HIP events on the batch stream after a completed solve; the
    // accumulators keep changing, so callers reset them afterwards
    float time_iteration(int reps);
    // extraction of every field: one launch of each extraction kernel over
    // all fields' edit arrays on the batch stream, then one wait
    void extract_all();
    // certification of every field (one launch per pass and axis length);
    // out[size()]
    void certify_all(bool quantized, CertResult* out);

   private:
    void release_graph();
    // batched plan (and graph) over the fields' current iteration arguments
    void build(bool use_graph);
    std::vector<Solver<T>*> f_;
    hipStream_t s_ = nullptr;
    BatchPlan<T> plan_;
    hipGraphExec_t gexec_ = nullptr;
    std::vector<uintptr_t> gsig_;  // the launches gexec_ was captured with
    DevState* hstate_ = nullptr;  // pinned host copies of the fields' states
    CertStats* cert_host_ = nullptr;  // pinned host copies of the fields' certification stats
    int launched_ = 0;
    bool graph_ = false;
    DevBuffer cert_tables_;
    ExtractBatchWork xwork_;
};

}  // namespace ffcz
