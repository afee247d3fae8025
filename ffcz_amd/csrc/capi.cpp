// C ABI over the HIP solver (include/ffcz.h).  Exceptions never cross the
// boundary: every entry point returns a status and records a thread-local
// message for ffcz_last_error().
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ffcz.h"
#include "batch.hpp"
#include "container.hpp"
#include "extract.hpp"
#include "fft_fast.hpp"
#include "huffman.hpp"
#include "kernels.hpp"
#include "plan_cache.hpp"
#include "solver.hpp"

using namespace ffcz;

static thread_local std::string g_err;

struct ffcz_solver {
    int dtype;
    std::unique_ptr<Solver<float>> f;
    std::unique_ptr<Solver<double>> d;
    bool borrowed = false;  // a field of a batch: freed with the batch
};

struct ffcz_batch {
    int dtype;
    std::vector<std::unique_ptr<ffcz_solver>> fields;
    std::unique_ptr<Batch<float>> bf;
    std::unique_ptr<Batch<double>> bd;
};

struct ffcz_coded {
    HuffmanCoded h;
    std::vector<uint8_t> zdata;
    bool zstd = false;
    ~ffcz_coded() { recycle_bitstream(std::move(h.bitstream)); }
};

struct ffcz_blob {
    std::vector<uint8_t> bytes;
};

template <typename F>
static int guard(F&& fn) {
    try {
        fn();
        return FFCZ_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return FFCZ_ERR_ALLOC;
    } catch (const std::exception& e) {
        g_err = e.what();
        return FFCZ_ERR_INVALID;
    } catch (...) {
        g_err = "unknown error";
        return FFCZ_ERR_INVALID;
    }
}

#define NEED(h) FFCZ_REQUIRE((h) != nullptr, ERR_INVALID, "null solver handle")

template <typename T>
static Solver<T>& sv(ffcz_solver* h);
template <>
Solver<float>& sv<float>(ffcz_solver* h) {
    NEED(h);
    FFCZ_REQUIRE(h->dtype == FFCZ_F32, ERR_INVALID, "solver is not float32");
    return *h->f;
}
template <>
Solver<double>& sv<double>(ffcz_solver* h) {
    NEED(h);
    FFCZ_REQUIRE(h->dtype == FFCZ_F64, ERR_INVALID, "solver is not float64");
    return *h->d;
}

// dispatch on the solver dtype
template <typename Fn>
static void with(ffcz_solver* h, Fn&& fn) {
    NEED(h);
    if (h->dtype == FFCZ_F32) fn(*h->f);
    else fn(*h->d);
}

static hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// copy device data out to a host-or-device buffer, two-call size pattern
static void copy_out(void* out, const void* src, size_t bytes, size_t elems, size_t cap, size_t* len,
                     hipStream_t s) {
    if (len) *len = elems;
    if (!out) return;
    FFCZ_REQUIRE(cap >= elems, ERR_RANGE, "output capacity too small");
    if (bytes) {
        FFCZ_HIP(hipMemcpyAsync(out, src, bytes, hipMemcpyDefault, s));
        FFCZ_HIP(hipStreamSynchronize(s));
    }
}

template <typename S>
static void need_extracted(const S& s) {
    FFCZ_REQUIRE(s.extracted(), ERR_STATE, "extract_compact_representation has not run");
}

template <typename S>
static void encode(const S* data, size_t n, int level, bool zstd, ffcz_coded** out) {
    FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null output");
    FFCZ_REQUIRE(n > 0 && data != nullptr, ERR_INVALID, "Empty data");  // HuffmanZSTDCoder.cuh:250-253
    hipPointerAttribute_t attr{};
    const bool on_device = hipPointerGetAttributes(&attr, data) == hipSuccess && attr.type == hipMemoryTypeDevice;
    (void)hipGetLastError();
    std::unique_ptr<ffcz_coded> c(new ffcz_coded());
    if (on_device) {
        c->h = huffman_encode_device<S>(data, n, nullptr);
    } else {
        S* d = nullptr;
        FFCZ_HIP(hipMalloc(&d, n * sizeof(S)));
        try {
            FFCZ_HIP(hipMemcpy(d, data, n * sizeof(S), hipMemcpyHostToDevice));
            c->h = huffman_encode_device<S>(d, n, nullptr);
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        (void)hipFree(d);
    }
    if (zstd) {
        c->zdata = zstd_compress(c->h.bitstream.data(), c->h.bitstream.size(), level);
        c->zstd = true;
    }
    *out = c.release();
}

template <typename S>
static void decompress(const uint8_t* comp, size_t clen, const uint32_t* sym, const uint32_t* code,
                       const int32_t* len, size_t nsym, size_t ne, size_t bb, S* out, size_t* out_len, bool zstd) {
    FFCZ_REQUIRE(out != nullptr && (nsym == 0 || (sym && code && len)), ERR_INVALID, "null argument");
    HuffmanTable t;
    t.sym.assign(sym, sym + nsym);
    t.code.assign(code, code + nsym);
    t.len.assign(len, len + nsym);
    std::vector<uint8_t> bits = zstd ? zstd_decompress(comp, clen, bb) : std::vector<uint8_t>(comp, comp + clen);
    std::vector<S> r = huffman_decode<S>(bits.data(), bits.size(), t, ne);
    std::memcpy(out, r.data(), r.size() * sizeof(S));
    if (out_len) *out_len = r.size();
}

template <typename T>
static void fft_entry(size_t nx, size_t ny, size_t nz, const void* in, void* out, hipStream_t s, bool fwd,
                      int prec = PREC_VALU) {
    FFCZ_REQUIRE(prec >= PREC_VALU && prec <= PREC_FP16X3, ERR_INVALID, "unknown FFT variant");
    FFCZ_REQUIRE(prec == PREC_VALU || sizeof(T) == 4, ERR_INVALID, "the MFMA FFT variants are FP32 transforms");
    Fft3d<T> f;
    f.create(nx, ny, nz);
    f.prec = prec;
    cplx<T>* spec = nullptr;
    try {
        FFCZ_HIP(hipMalloc(&spec, f.spec_elems() * sizeof(cplx<T>)));
        if (fwd) {
            fft_r2c<T>(f, (const T*)in, spec, s);
            spec_to_compact<T>(f, spec, (cplx<T>*)out, s);
        } else {
            compact_to_spec<T>(f, (const cplx<T>*)in, spec, s);
            fft_c2r<T>(f, spec, (T*)out, T(1), s);
        }
        FFCZ_HIP(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipFree(spec);
        f.destroy();
        throw;
    }
    (void)hipFree(spec);
    f.destroy();
}

static void unpack_flags(const uint8_t* dpack, size_t n, uint8_t* out, size_t cap, size_t* len, hipStream_t s) {
    if (len) *len = n;
    if (!out) return;
    FFCZ_REQUIRE(cap >= n, ERR_RANGE, "output capacity too small");
    std::vector<uint8_t> p((n + 7) / 8);
    if (!p.empty()) {
        FFCZ_HIP(hipMemcpyAsync(p.data(), dpack, p.size(), hipMemcpyDeviceToHost, s));
        FFCZ_HIP(hipStreamSynchronize(s));
    }
    std::vector<uint8_t> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (p[i >> 3] >> (i & 7)) & 1;
    FFCZ_HIP(hipMemcpy(out, tmp.data(), n, hipMemcpyDefault));
}

extern "C" {

const char* ffcz_last_error(void) { return g_err.c_str(); }
int ffcz_abi_version(void) { return 1; }

int ffcz_create(int dtype, size_t nx, size_t ny, size_t nz, ffcz_solver** out) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null output handle");
        FFCZ_REQUIRE(dtype == FFCZ_F32 || dtype == FFCZ_F64, ERR_INVALID, "dtype must be FFCZ_F32 or FFCZ_F64");
        std::unique_ptr<ffcz_solver> h(new ffcz_solver());
        h->dtype = dtype;
        if (dtype == FFCZ_F32) h->f.reset(new Solver<float>(nx, ny, nz));
        else h->d.reset(new Solver<double>(nx, ny, nz));
        *out = h.release();
    });
}

int ffcz_destroy(ffcz_solver* h) {
    return guard([&] {
        FFCZ_REQUIRE(h == nullptr || !h->borrowed, ERR_INVALID, "a batch field is destroyed with its batch");
        delete h;
    });
}

int ffcz_batch_create(int dtype, size_t nfields, const size_t* dims, ffcz_batch** out) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr && dims != nullptr, ERR_INVALID, "null argument");
        FFCZ_REQUIRE(nfields >= 1, ERR_INVALID, "a batch needs at least one field");
        FFCZ_REQUIRE(dtype == FFCZ_F32 || dtype == FFCZ_F64, ERR_INVALID, "dtype must be FFCZ_F32 or FFCZ_F64");
        std::unique_ptr<ffcz_batch> b(new ffcz_batch());
        b->dtype = dtype;
        std::vector<Solver<float>*> pf;
        std::vector<Solver<double>*> pd;
        for (size_t i = 0; i < nfields; ++i) {
            std::unique_ptr<ffcz_solver> h(new ffcz_solver());
            h->dtype = dtype;
            h->borrowed = true;
            if (dtype == FFCZ_F32) {
                h->f.reset(new Solver<float>(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]));
                pf.push_back(h->f.get());
            } else {
                h->d.reset(new Solver<double>(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]));
                pd.push_back(h->d.get());
            }
            b->fields.push_back(std::move(h));
        }
        if (dtype == FFCZ_F32) b->bf.reset(new Batch<float>(pf));
        else b->bd.reset(new Batch<double>(pd));
        *out = b.release();
    });
}
int ffcz_batch_destroy(ffcz_batch* b) {
    return guard([&] { delete b; });
}
int ffcz_batch_size(ffcz_batch* b, size_t* n) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr && n != nullptr, ERR_INVALID, "null argument");
        *n = b->fields.size();
    });
}
int ffcz_batch_field(ffcz_batch* b, size_t i, ffcz_solver** out) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr && out != nullptr, ERR_INVALID, "null argument");
        FFCZ_REQUIRE(i < b->fields.size(), ERR_RANGE, "field index out of range");
        *out = b->fields[i].get();
    });
}
int ffcz_batch_solve(ffcz_batch* b, int mode, const void* const* e0, int max_iter, double tol, int use_graph) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr, ERR_INVALID, "null batch handle");
        if (b->dtype == FFCZ_F32)
            b->bf->solve(mode, reinterpret_cast<const float* const*>(e0), max_iter, (float)tol, use_graph != 0);
        else
            b->bd->solve(mode, reinterpret_cast<const double* const*>(e0), max_iter, tol, use_graph != 0);
    });
}
int ffcz_batch_extract(ffcz_batch* b) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr, ERR_INVALID, "null batch handle");
        if (b->dtype == FFCZ_F32) b->bf->extract_all();
        else b->bd->extract_all();
    });
}
int ffcz_batch_certify(ffcz_batch* b, int quantized, ffcz_cert* out, size_t cap) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr && out != nullptr, ERR_INVALID, "null argument");
        FFCZ_REQUIRE(cap >= b->fields.size(), ERR_RANGE, "result capacity below the field count");
        std::vector<CertResult> r(b->fields.size());
        if (b->dtype == FFCZ_F32) b->bf->certify_all(quantized != 0, r.data());
        else b->bd->certify_all(quantized != 0, r.data());
        for (size_t i = 0; i < r.size(); ++i) {
            out[i].spatial_violations = r[i].spatial_violations;
            out[i].freq_violations = r[i].freq_violations;
            out[i].spatial_max_excess = r[i].spatial_max_excess;
            out[i].freq_max_excess = r[i].freq_max_excess;
            out[i].ok = r[i].ok;
        }
    });
}
int ffcz_batch_time_iteration(ffcz_batch* b, int reps, float* ms) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr && ms != nullptr, ERR_INVALID, "null argument");
        *ms = b->dtype == FFCZ_F32 ? b->bf->time_iteration(reps) : b->bd->time_iteration(reps);
    });
}
int ffcz_batch_status(ffcz_batch* b, int* iterations, double* violations, int* converged, size_t cap,
                      int* launched) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr, ERR_INVALID, "null batch handle");
        const size_t n = std::min(cap, b->fields.size());
        for (size_t i = 0; i < n; ++i) {
            with(b->fields[i].get(), [&](auto& s) {
                if (iterations) iterations[i] = s.iterations();
                if (violations) violations[i] = (double)s.violation();
                if (converged) converged[i] = s.converged() ? 1 : 0;
            });
        }
        if (launched) *launched = b->dtype == FFCZ_F32 ? b->bf->launched() : b->bd->launched();
    });
}

#define INIT_PAIR(name, call)                                                                       \
    int ffcz_##name##_f32 INIT_ARGS(float) { return guard([&] { sv<float>(h).call; }); }        \
    int ffcz_##name##_f64 INIT_ARGS(double) { return guard([&] { sv<double>(h).call; }); }

#define INIT_ARGS(T) (ffcz_solver * h, const T* x, T eps, T delta)
INIT_PAIR(init_all_abs, init_all_abs(x, eps, delta))
#undef INIT_ARGS
#define INIT_ARGS(T) (ffcz_solver * h, const T* x, const size_t* idx, const T* delta_k, T eps, size_t K)
INIT_PAIR(init_direct, init_direct(x, idx, delta_k, eps, K))
#undef INIT_ARGS
#define INIT_ARGS(T) (ffcz_solver * h, const T* x, T eps, T delta_ptw)
INIT_PAIR(init_all_ptw, init_all_ptw(x, eps, delta_ptw))
#undef INIT_ARGS
#define INIT_ARGS(T) (ffcz_solver * h, const T* x, const size_t* idx, T delta_ptw, T eps, size_t K)
INIT_PAIR(init_partial_ptw, init_partial_ptw(x, idx, delta_ptw, eps, K))
#undef INIT_ARGS
#define INIT_ARGS(T) (ffcz_solver * h, const T* e0, int max_iterations, T tol)
INIT_PAIR(solve, solve(MODE_DIRECT, e0, max_iterations, tol))
INIT_PAIR(solve_all_abs, solve(MODE_ABS, e0, max_iterations, tol))
INIT_PAIR(solve_all_ptw, solve(MODE_PTW, e0, max_iterations, tol))
INIT_PAIR(solve_partial_ptw, solve(MODE_PARTIAL_PTW, e0, max_iterations, tol))
#undef INIT_ARGS

int ffcz_set_ptw_semantics(ffcz_solver* h, int corrected) {
    return guard([&] {
        FFCZ_REQUIRE(corrected == 0 || corrected == 1, ERR_INVALID, "ptw semantics: 0 reference, 1 corrected");
        with(h, [&](auto& s) { s.set_ptw_semantics(corrected == 1); });
    });
}

int ffcz_solve_loop(ffcz_solver* h, int mode, const void* e0, int max_iterations, double tol) {
    return guard([&] {
        NEED(h);
        FFCZ_REQUIRE(mode >= 0 && mode <= 3, ERR_INVALID, "unknown solve mode");
        if (h->dtype == FFCZ_F32) h->f->solve_loop(mode, (const float*)e0, max_iterations, (float)tol);
        else h->d->solve_loop(mode, (const double*)e0, max_iterations, tol);
    });
}

int ffcz_get_status(ffcz_solver* h, int* iterations, double* violation, int* converged) {
    return guard([&] {
        with(h, [&](auto& s) {
            if (iterations) *iterations = s.iterations();
            if (violation) *violation = s.violation();
            if (converged) *converged = s.converged() ? 1 : 0;
        });
    });
}

int ffcz_extract(ffcz_solver* h) {
    return guard([&] { with(h, [&](auto& s) { s.extract(); }); });
}
int ffcz_reset_edit_accumulators(ffcz_solver* h) {
    return guard([&] { with(h, [&](auto& s) { s.reset_edit_accumulators(); }); });
}

int ffcz_get_results(ffcz_solver* h, void* e_out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            using T = std::remove_reference_t<decltype(*s.d_e())>;
            copy_out(e_out, s.d_e(), s.N() * sizeof(T), s.N(), cap, len, s.stream());
        });
    });
}



int ffcz_get_freq_flags(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            unpack_flags(s.d_freq_pack(), 2 * s.F(), out, cap, len, s.stream());
        });
    });
}
int ffcz_get_spatial_flags(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            unpack_flags(s.d_spatial_pack(), s.N(), out, cap, len, s.stream());
        });
    });
}
int ffcz_get_freq_flags_pack(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            const size_t nb = (2 * s.F() + 7) / 8;
            copy_out(out, s.d_freq_pack(), nb, nb, cap, len, s.stream());
        });
    });
}
int ffcz_get_spatial_flags_pack(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            const size_t nb = (s.N() + 7) / 8;
            copy_out(out, s.d_spatial_pack(), nb, nb, cap, len, s.stream());
        });
    });
}
int ffcz_get_active_freq_quant(ffcz_solver* h, uint16_t* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            const size_t n = s.n_active_freq();
            copy_out(out, s.d_freq_quant(), n * 2, n, cap, len, s.stream());
        });
    });
}
int ffcz_get_active_spatial_quant(ffcz_solver* h, uint16_t* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            const size_t n = s.n_active_spatial();
            copy_out(out, s.d_spatial_quant(), n * 2, n, cap, len, s.stream());
        });
    });
}
int ffcz_get_active_freq_edits(ffcz_solver* h, void* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            using T = std::remove_reference_t<decltype(*s.d_e())>;
            const size_t n = s.n_active_freq();
            copy_out(out, s.d_freq_compact(), n * sizeof(T), n, cap, len, s.stream());
        });
    });
}
int ffcz_get_active_spatial_edits(ffcz_solver* h, void* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            need_extracted(s);
            using T = std::remove_reference_t<decltype(*s.d_e())>;
            const size_t n = s.n_active_spatial();
            copy_out(out, s.d_spatial_compact(), n * sizeof(T), n, cap, len, s.stream());
        });
    });
}
int ffcz_get_freq_edits_full(ffcz_solver* h, void* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            using T = std::remove_reference_t<decltype(*s.d_e())>;
            copy_out(out, s.d_df(), 2 * s.F() * sizeof(T), 2 * s.F(), cap, len, s.stream());
        });
    });
}
int ffcz_get_spatial_edits_full(ffcz_solver* h, void* out, size_t cap, size_t* len) {
    return guard([&] {
        with(h, [&](auto& s) {
            using T = std::remove_reference_t<decltype(*s.d_e())>;
            copy_out(out, s.d_ds(), s.N() * sizeof(T), s.N(), cap, len, s.stream());
        });
    });
}
int ffcz_get_edit_range(ffcz_solver* h, double* fmin, double* fmax, double* smin, double* smax) {
    return guard([&] {
        with(h, [&](auto& s) {
            if (fmin) *fmin = s.freq_min();
            if (fmax) *fmax = s.freq_max();
            if (smin) *smin = s.spatial_min();
            if (smax) *smax = s.spatial_max();
        });
    });
}

int ffcz_set_margins(ffcz_solver* h, double sm, double fm) {
    return guard([&] {
        FFCZ_REQUIRE(sm >= 0 && sm < 1 && fm >= 0 && fm < 1, ERR_INVALID, "margins must lie in [0, 1)");
        with(h, [&](auto& s) { s.set_margins(sm, fm); });
    });
}
static void cert_out(const CertResult* r, ffcz_cert* out) {
    for (int i = 0; i < 2; ++i) {
        out[i].spatial_violations = r[i].spatial_violations;
        out[i].freq_violations = r[i].freq_violations;
        out[i].spatial_max_excess = r[i].spatial_max_excess;
        out[i].freq_max_excess = r[i].freq_max_excess;
        out[i].ok = r[i].ok;
    }
}
int ffcz_certify_both(ffcz_solver* h, ffcz_cert* out) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null result");
        with(h, [&](auto& s) {
            CertResult r[2];
            s.certify_both(r);
            cert_out(r, out);
        });
    });
}
int ffcz_certify_both_enqueue(ffcz_solver* h) {
    return guard([&] { with(h, [&](auto& s) { s.certify_both_enqueue(); }); });
}
int ffcz_certify_both_results(ffcz_solver* h, ffcz_cert* out) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null result");
        with(h, [&](auto& s) {
            CertResult r[2];
            s.certify_both_results(r);
            cert_out(r, out);
        });
    });
}
int ffcz_set_fft_precision(ffcz_solver* h, int variant) {
    return guard([&] { with(h, [&](auto& s) { s.set_fft_precision(variant); }); });
}
int ffcz_get_fft_precision(ffcz_solver* h, int* variant) {
    return guard([&] {
        FFCZ_REQUIRE(variant != nullptr, ERR_INVALID, "null output");
        with(h, [&](auto& s) { *variant = s.fft_precision(); });
    });
}
static void vars_out(const TileVars& v, int* out) {
    if (!out) return;
    const int a[7] = {v.zf, v.zi, v.y, v.x, v.czf, v.cy, v.cx};
    std::memcpy(out, a, sizeof(a));
}
int ffcz_autotune(ffcz_solver* h, int reps, int use_cache, int* vars, float* ms) {
    return guard([&] { with(h, [&](auto& s) { vars_out(s.autotune(reps, use_cache != 0, ms), vars); }); });
}
int ffcz_get_tune_source(ffcz_solver* h, int* source) {
    return guard([&] {
        FFCZ_REQUIRE(source != nullptr, ERR_INVALID, "null output");
        with(h, [&](auto& s) { *source = s.tune_source(); });
    });
}
int ffcz_set_plan_cache(const char* path) {
    return guard([&] { set_plan_cache_path(path); });
}
int ffcz_get_plan_cache(char* buf, size_t cap, size_t* len) {
    return guard([&] {
        FFCZ_REQUIRE(len != nullptr, ERR_INVALID, "null length");
        const std::string p = plan_cache_path();
        *len = p.size();
        if (buf && cap > 0) {
            const size_t n = std::min(cap - 1, p.size());
            std::memcpy(buf, p.data(), n);
            buf[n] = '\0';
        }
    });
}
int ffcz_get_tile_vars(ffcz_solver* h, int* vars) {
    return guard([&] {
        FFCZ_REQUIRE(vars != nullptr, ERR_INVALID, "null output");
        with(h, [&](auto& s) { vars_out(s.tile_vars(), vars); });
    });
}
int ffcz_set_tile_vars(ffcz_solver* h, const int* vars) {
    return guard([&] {
        FFCZ_REQUIRE(vars != nullptr, ERR_INVALID, "null input");
        with(h, [&](auto& s) {
            TileVars v = s.tile_vars();  // keeps the paired certification's choice
            int* p = &v.zf;
            for (int i = 0; i < 7; ++i) p[i] = vars[i];
            s.set_tile_vars(v);
        });
    });
}
int ffcz_set_cert_pair(ffcz_solver* h, int paired) {
    return guard([&] {
        FFCZ_REQUIRE(paired == 0 || paired == 1, ERR_INVALID, "paired is 0 or 1");
        with(h, [&](auto& s) {
            TileVars v = s.tile_vars();
            v.cpair = paired;
            s.set_tile_vars(v);
        });
    });
}
int ffcz_set_quantizer(ffcz_solver* h, int kind, double fstep, double sstep) {
    return guard([&] { with(h, [&](auto& s) { s.set_quantizer(kind, fstep, sstep); }); });
}
int ffcz_certify(ffcz_solver* h, int quantized, ffcz_cert* out) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null result");
        with(h, [&](auto& s) {
            CertResult r = s.certify(quantized != 0);
            out->spatial_violations = r.spatial_violations;
            out->freq_violations = r.freq_violations;
            out->spatial_max_excess = r.spatial_max_excess;
            out->freq_max_excess = r.freq_max_excess;
            out->ok = r.ok;
        });
    });
}
int ffcz_set_tuning(int force_generic, long slab_mb) {
    return guard([&] {
        if (force_generic >= 0) tuning().force_generic = force_generic ? 1 : 0;
        tuning().slab_mb = slab_mb;
    });
}
int ffcz_get_tuning(int* force_generic, long* slab_mb) {
    return guard([&] {
        if (force_generic) *force_generic = tuning().force_generic.load();
        if (slab_mb) *slab_mb = tuning().slab_mb.load();
    });
}
int ffcz_get_stream(ffcz_solver* h, void** stream) {
    return guard([&] {
        FFCZ_REQUIRE(stream != nullptr, ERR_INVALID, "null output");
        with(h, [&](auto& s) { *stream = (void*)s.stream(); });
    });
}
int ffcz_set_graph(ffcz_solver* h, int enable) {
    return guard([&] { with(h, [&](auto& s) { s.set_graph(enable != 0); }); });
}
int ffcz_set_borrow_e0(ffcz_solver* h, int borrow) {
    return guard([&] { with(h, [&](auto& s) { s.set_borrow_e0(borrow != 0); }); });
}
int ffcz_set_pass_timing(ffcz_solver* h, int enable) {
    return guard([&] { with(h, [&](auto& s) { s.set_pass_timing(enable != 0); }); });
}
int ffcz_get_pass_timing(ffcz_solver* h, float* ms6, long* iterations) {
    return guard([&] {
        FFCZ_REQUIRE(ms6 != nullptr, ERR_INVALID, "null output");
        with(h, [&](auto& s) {
            const long n = s.pass_timing(ms6);
            if (iterations) *iterations = n;
        });
    });
}
int ffcz_get_pass_counts(ffcz_solver* h, long* n6) {
    return guard([&] {
        FFCZ_REQUIRE(n6 != nullptr, ERR_INVALID, "null output");
        with(h, [&](auto& s) { s.pass_counts(n6); });
    });
}
int ffcz_time_iteration(ffcz_solver* h, int reps, float* ms5) {
    return guard([&] {
        FFCZ_REQUIRE(ms5 != nullptr, ERR_INVALID, "null output");
        with(h, [&](auto& s) { s.time_iteration(reps, ms5); });
    });
}

// ---- entropy coding --------------------------------------------------------

int ffcz_huffman_zstd_compress_u16(const uint16_t* data, size_t n, int level, ffcz_coded** out) {
    return guard([&] { encode<uint16_t>(data, n, level, true, out); });
}
int ffcz_huffman_zstd_compress_u8(const uint8_t* data, size_t n, int level, ffcz_coded** out) {
    return guard([&] { encode<uint8_t>(data, n, level, true, out); });
}
int ffcz_huffman_encode_u16(const uint16_t* data, size_t n, ffcz_coded** out) {
    return guard([&] { encode<uint16_t>(data, n, 0, false, out); });
}
int ffcz_huffman_encode_u8(const uint8_t* data, size_t n, ffcz_coded** out) {
    return guard([&] { encode<uint8_t>(data, n, 0, false, out); });
}
int ffcz_coded_info(const ffcz_coded* c, size_t* cb, size_t* bb, size_t* ne, size_t* ns, unsigned long long* nbits) {
    return guard([&] {
        FFCZ_REQUIRE(c != nullptr, ERR_INVALID, "null coded handle");
        if (cb) *cb = c->zstd ? c->zdata.size() : c->h.bitstream.size();
        if (bb) *bb = c->h.bitstream.size();
        if (ne) *ne = c->h.nelem;
        if (ns) *ns = c->h.table.sym.size();
        if (nbits) *nbits = c->h.nbits;
    });
}
int ffcz_coded_device_ms(const ffcz_coded* c, float* ms) {
    return guard([&] {
        FFCZ_REQUIRE(c != nullptr && ms != nullptr, ERR_INVALID, "null argument");
        *ms = c->h.device_ms;
    });
}
int ffcz_coded_data(const ffcz_coded* c, uint8_t* out, size_t cap) {
    return guard([&] {
        FFCZ_REQUIRE(c != nullptr && out != nullptr, ERR_INVALID, "null argument");
        const auto& v = c->zstd ? c->zdata : c->h.bitstream;
        FFCZ_REQUIRE(cap >= v.size(), ERR_RANGE, "output capacity too small");
        if (!v.empty()) std::memcpy(out, v.data(), v.size());
    });
}
int ffcz_coded_bitstream(const ffcz_coded* c, uint8_t* out, size_t cap) {
    return guard([&] {
        FFCZ_REQUIRE(c != nullptr && out != nullptr, ERR_INVALID, "null argument");
        const auto& v = c->h.bitstream;
        FFCZ_REQUIRE(cap >= v.size(), ERR_RANGE, "output capacity too small");
        if (!v.empty()) std::memcpy(out, v.data(), v.size());
    });
}
int ffcz_coded_table(const ffcz_coded* c, uint32_t* symbols, uint32_t* codes, int32_t* lengths, size_t cap) {
    return guard([&] {
        FFCZ_REQUIRE(c != nullptr, ERR_INVALID, "null coded handle");
        const auto& t = c->h.table;
        FFCZ_REQUIRE(cap >= t.sym.size(), ERR_RANGE, "output capacity too small");
        for (size_t i = 0; i < t.sym.size(); ++i) {
            if (symbols) symbols[i] = t.sym[i];
            if (codes) codes[i] = t.code[i];
            if (lengths) lengths[i] = t.len[i];
        }
    });
}
void ffcz_coded_free(ffcz_coded* c) { delete c; }

int ffcz_container_pack(ffcz_solver* h, int zstd_level, ffcz_blob** out) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null output");
        auto b = std::make_unique<ffcz_blob>();
        with(h, [&](auto& s) { b->bytes = container_pack(s, zstd_level); });
        *out = b.release();
    });
}
int ffcz_blob_size(const ffcz_blob* b, size_t* len) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr && len != nullptr, ERR_INVALID, "null argument");
        *len = b->bytes.size();
    });
}
int ffcz_blob_copy(const ffcz_blob* b, uint8_t* out, size_t cap) {
    return guard([&] {
        FFCZ_REQUIRE(b != nullptr && out != nullptr, ERR_INVALID, "null argument");
        FFCZ_REQUIRE(cap >= b->bytes.size(), ERR_RANGE, "output capacity too small");
        std::memcpy(out, b->bytes.data(), b->bytes.size());
    });
}
void ffcz_blob_free(ffcz_blob* b) { delete b; }
int ffcz_container_info(const uint8_t* buf, size_t len, int* dtype, size_t* dims, double* eps, double* delta,
                        int* quantizer, size_t* n_active) {
    return guard([&] {
        const ContainerInfo in = container_info(buf, len);
        if (dtype) *dtype = in.dtype;
        if (dims) {
            dims[0] = in.nx;
            dims[1] = in.ny;
            dims[2] = in.nz;
        }
        if (eps) *eps = in.eps;
        if (delta) *delta = in.delta;
        if (quantizer) *quantizer = in.qkind;
        if (n_active) {
            n_active[0] = in.n_freq;
            n_active[1] = in.n_spatial;
        }
    });
}
int ffcz_container_decode(const uint8_t* buf, size_t len, const void* x_prime, void* x_hat, void* stream) {
    return guard([&] { container_decode(buf, len, x_prime, x_hat, as_stream(stream)); });
}
int ffcz_base_compress(int dtype, const void* x, size_t nx, size_t ny, size_t nz, double eps, int zstd_level,
                       void* x_prime, ffcz_blob** out, void* stream) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null output");
        auto b = std::make_unique<ffcz_blob>();
        if (dtype == FFCZ_F32)
            b->bytes = base_compress<float>((const float*)x, nx, ny, nz, eps, zstd_level, (float*)x_prime,
                                            as_stream(stream));
        else if (dtype == FFCZ_F64)
            b->bytes = base_compress<double>((const double*)x, nx, ny, nz, eps, zstd_level, (double*)x_prime,
                                             as_stream(stream));
        else throw Error(ERR_INVALID, "bad dtype");
        *out = b.release();
    });
}
int ffcz_base_info(const uint8_t* buf, size_t len, int* dtype, size_t* dims, double* eps, size_t* n_outliers) {
    return guard([&] {
        const ContainerInfo in = base_info(buf, len);
        if (dtype) *dtype = in.dtype;
        if (dims) {
            dims[0] = in.nx;
            dims[1] = in.ny;
            dims[2] = in.nz;
        }
        if (eps) *eps = in.eps;
        if (n_outliers) *n_outliers = in.n_spatial;
    });
}
int ffcz_base_decompress(const uint8_t* buf, size_t len, void* x_prime, void* stream) {
    return guard([&] { base_decompress(buf, len, x_prime, as_stream(stream)); });
}

int ffcz_huffman_zstd_decompress_u16(const uint8_t* comp, size_t clen, const uint32_t* sym, const uint32_t* code,
                                     const int32_t* len, size_t nsym, size_t ne, size_t bb, uint16_t* out,
                                     size_t* out_len) {
    return guard([&] { decompress<uint16_t>(comp, clen, sym, code, len, nsym, ne, bb, out, out_len, true); });
}
int ffcz_huffman_zstd_decompress_u8(const uint8_t* comp, size_t clen, const uint32_t* sym, const uint32_t* code,
                                    const int32_t* len, size_t nsym, size_t ne, size_t bb, uint8_t* out,
                                    size_t* out_len) {
    return guard([&] { decompress<uint8_t>(comp, clen, sym, code, len, nsym, ne, bb, out, out_len, true); });
}

// ---- per-kernel entries ----------------------------------------------------

int ffcz_fft_r2c(int dtype, size_t nx, size_t ny, size_t nz, const void* in, void* out, void* stream) {
    return guard([&] {
        if (dtype == FFCZ_F32) fft_entry<float>(nx, ny, nz, in, out, as_stream(stream), true);
        else if (dtype == FFCZ_F64) fft_entry<double>(nx, ny, nz, in, out, as_stream(stream), true);
        else throw Error(ERR_INVALID, "bad dtype");
    });
}
int ffcz_fft_c2r(int dtype, size_t nx, size_t ny, size_t nz, const void* in, void* out, void* stream) {
    return guard([&] {
        if (dtype == FFCZ_F32) fft_entry<float>(nx, ny, nz, in, out, as_stream(stream), false);
        else if (dtype == FFCZ_F64) fft_entry<double>(nx, ny, nz, in, out, as_stream(stream), false);
        else throw Error(ERR_INVALID, "bad dtype");
    });
}
int ffcz_fft_r2c_prec(int dtype, int prec, size_t nx, size_t ny, size_t nz, const void* in, void* out,
                      void* stream) {
    return guard([&] {
        if (dtype == FFCZ_F32) fft_entry<float>(nx, ny, nz, in, out, as_stream(stream), true, prec);
        else if (dtype == FFCZ_F64) fft_entry<double>(nx, ny, nz, in, out, as_stream(stream), true, prec);
        else throw Error(ERR_INVALID, "bad dtype");
    });
}
int ffcz_fft_c2r_prec(int dtype, int prec, size_t nx, size_t ny, size_t nz, const void* in, void* out,
                      void* stream) {
    return guard([&] {
        if (dtype == FFCZ_F32) fft_entry<float>(nx, ny, nz, in, out, as_stream(stream), false, prec);
        else if (dtype == FFCZ_F64) fft_entry<double>(nx, ny, nz, in, out, as_stream(stream), false, prec);
        else throw Error(ERR_INVALID, "bad dtype");
    });
}
int ffcz_k_project_spatial(int dtype, void* e, double eps, void* ds, size_t n, void* stream) {
    return guard([&] {
        if (dtype == FFCZ_F32) launch_project_spatial<float>((float*)e, (float)eps, (float*)ds, n, as_stream(stream));
        else launch_project_spatial<double>((double*)e, eps, (double*)ds, n, as_stream(stream));
        FFCZ_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}
int ffcz_k_project_freq_abs(int dtype, void* spec, double delta, void* df, size_t F, void* stream) {
    return guard([&] {
        if (dtype == FFCZ_F32)
            launch_project_freq_abs<float>((cplx<float>*)spec, (float)delta, (float*)df, F, as_stream(stream));
        else launch_project_freq_abs<double>((cplx<double>*)spec, delta, (double*)df, F, as_stream(stream));
        FFCZ_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}
int ffcz_k_convergence(int dtype, const void* e, double eps, size_t n, double* out) {
    return guard([&] {
        FFCZ_REQUIRE(out != nullptr, ERR_INVALID, "null output");
        if (dtype == FFCZ_F32) *out = convergence_metric<float>((const float*)e, (float)eps, n, nullptr);
        else *out = convergence_metric<double>((const double*)e, eps, n, nullptr);
    });
}
int ffcz_k_quantize_dequantize(int dtype, const void* v, void* out, double vmin, double vmax, size_t n,
                               double threshold, void* stream) {
    return guard([&] {
        if (dtype == FFCZ_F32)
            launch_quant_dequant<float>((const float*)v, (float*)out, (float)vmin, (float)vmax, n, (float)threshold,
                                        as_stream(stream));
        else launch_quant_dequant<double>((const double*)v, (double*)out, vmin, vmax, n, threshold, as_stream(stream));
        FFCZ_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}
int ffcz_k_extract(int dtype, const void* v, size_t n, int kind, double step, uint8_t* pack, void* compact,
                   uint16_t* quant, size_t* count, double* vmin, double* vmax, void* stream) {
    return guard([&] {
        FFCZ_REQUIRE(kind == Q_REF16 || kind == Q_STEP, ERR_INVALID, "unknown quantizer kind");
        QuantDesc q{};
        q.kind = kind;
        q.step = step;
        q.inv_step = step > 0 ? 1.0 / step : 0.0;
        q.threshold = 1e-7;
        void* scratch = nullptr;
        FFCZ_HIP(hipMalloc(&scratch, extract_scratch_bytes(n)));
        ExtractResult r;
        try {
            if (dtype == FFCZ_F32)
                r = extract_edits<float>((const float*)v, n, q, pack, (float*)compact, quant, scratch, as_stream(stream));
            else
                r = extract_edits<double>((const double*)v, n, q, pack, (double*)compact, quant, scratch,
                                          as_stream(stream));
        } catch (...) {
            (void)hipFree(scratch);
            throw;
        }
        (void)hipFree(scratch);
        if (count) *count = r.count;
        if (vmin) *vmin = r.vmin;
        if (vmax) *vmax = r.vmax;
    });
}

int ffcz_k_extract_noalloc(int dtype, const void* v, size_t n, int kind, double step, uint8_t* pack, void* compact,
                           uint16_t* quant, size_t* count, double* vmin, double* vmax, void* stream) {
    return guard([&] {
        FFCZ_REQUIRE(n <= EXTRACT_NOALLOC_MAX, ERR_CAPACITY,
                     "array larger than the allocation-free extraction holds (2^30 elements)");
        FFCZ_REQUIRE(kind == Q_REF16 || kind == Q_STEP, ERR_INVALID, "unknown quantizer kind");
        FFCZ_REQUIRE(dtype == FFCZ_F32 || dtype == FFCZ_F64, ERR_INVALID, "unknown dtype");
        QuantDesc q{};
        q.kind = kind;
        q.step = step;
        q.inv_step = step > 0 ? 1.0 / step : 0.0;
        q.threshold = 1e-7;
        const ExtractResult r =
            dtype == FFCZ_F32
                ? extract_edits_noalloc<float>((const float*)v, n, q, pack, (float*)compact, quant, as_stream(stream))
                : extract_edits_noalloc<double>((const double*)v, n, q, pack, (double*)compact, quant,
                                                as_stream(stream));
        if (count) *count = r.count;
        if (vmin) *vmin = r.vmin;
        if (vmax) *vmax = r.vmax;
    });
}

}  // extern "C"
