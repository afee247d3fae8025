/*
 * ffcz.h - C ABI of the MI355X FFCz alternating-projection post-processor.
 *
 * Plain C: opaque handles, pointers, sizes and int status codes; no HIP or
 * torch types.  Data pointers may be host or device memory (the library uses
 * unified addressing to tell them apart) unless stated otherwise.
 *
 * Each entry point names the reference interface it replaces
 * (FastFourierCorrectionz/FFCz, file:line).  The C++ drop-in class
 * ProjectionSolver<T> in projection_algorithm.hpp is written over these.
 *
 * Errors: every function returns FFCZ_OK (0) or a status code; the message of
 * the last failure on the calling thread is ffcz_last_error().  Nothing calls
 * exit() (the reference exits the process on CUDA/cuFFT errors,
 * projection_algorithm.cuh:11-31).
 */
#ifndef FFCZ_H
#define FFCZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FFCZ_OK 0
#define FFCZ_ERR_INVALID 1
#define FFCZ_ERR_STATE 2
#define FFCZ_ERR_HIP 3
#define FFCZ_ERR_RANGE 4
#define FFCZ_ERR_ALLOC 5
#define FFCZ_ERR_ZSTD 6
#define FFCZ_ERR_CAPACITY 8

/* element types (the reference instantiates float and double,
 * projection_algorithm.cuh:39-53) */
#define FFCZ_F32 0
#define FFCZ_F64 1

/* solve modes */
#define FFCZ_MODE_DIRECT 0 /* solve()         projection_algorithm_impl.cuh:772-814 */
#define FFCZ_MODE_ABS 1    /* solve_all_abs() projection_algorithm_impl.cuh:816-858 */
#define FFCZ_MODE_PTW 2    /* solve_all_ptw() projection_algorithm_impl.cuh:860-902 */
#define FFCZ_MODE_PARTIAL_PTW 3 /* project_onto_frequency_polytope_partial_ptw() impl:303-333 in the loop */

/* edit quantisers */
#define FFCZ_QUANT_REF16 1 /* projection_algorithm.cu:447-455 (default) */
#define FFCZ_QUANT_STEP 2  /* round(v/step), zigzag 16-bit symbols */

typedef struct ffcz_solver ffcz_solver;
typedef struct ffcz_coded ffcz_coded;
typedef struct ffcz_blob ffcz_blob;

typedef struct ffcz_cert {
    unsigned long long spatial_violations; /* elements with |x_hat - x| > eps  */
    unsigned long long freq_violations;    /* bins with |Re| or |Im| of F(x_hat)-F(x) > delta */
    double spatial_max_excess;
    double freq_max_excess;
    int ok; /* 1 if both bounds hold everywhere (FP64) */
} ffcz_cert;

const char* ffcz_last_error(void);
int ffcz_abi_version(void);

/* ---- solver lifecycle ----------------------------------------------------
 * ProjectionSolver(Nx,Ny,Nz)  projection_algorithm_impl.cuh:4-18 (+ :27-61, :109-122)
 * ~ProjectionSolver()         projection_algorithm_impl.cuh:20-25            */
int ffcz_create(int dtype, size_t nx, size_t ny, size_t nz, ffcz_solver** out);
int ffcz_destroy(ffcz_solver* h);

/* initialize_all_abs   projection_algorithm_impl.cuh:639-668 */
int ffcz_init_all_abs_f32(ffcz_solver* h, const float* x, float eps, float delta);
int ffcz_init_all_abs_f64(ffcz_solver* h, const double* x, double eps, double delta);
/* initialize           projection_algorithm_impl.cuh:598-637 */
int ffcz_init_direct_f32(ffcz_solver* h, const float* x, const size_t* idx, const float* delta_k, float eps, size_t K);
int ffcz_init_direct_f64(ffcz_solver* h, const double* x, const size_t* idx, const double* delta_k, double eps,
                         size_t K);
/* initialize_all_ptw   projection_algorithm_impl.cuh:670-699 */
int ffcz_init_all_ptw_f32(ffcz_solver* h, const float* x, float eps, float delta_ptw);
int ffcz_init_all_ptw_f64(ffcz_solver* h, const double* x, double eps, double delta_ptw);
/* initialize_partial_ptw projection_algorithm_impl.cuh:701-738 */
int ffcz_init_partial_ptw_f32(ffcz_solver* h, const float* x, const size_t* idx, float delta_ptw, float eps, size_t K);
int ffcz_init_partial_ptw_f64(ffcz_solver* h, const double* x, const size_t* idx, double delta_ptw, double eps,
                              size_t K);

/* solve / solve_all_abs / solve_all_ptw: projection loop + extraction
 * projection_algorithm_impl.cuh:772-902.  e0 = x' - x (read only). */
int ffcz_solve_f32(ffcz_solver* h, const float* e0, int max_iterations, float tol);
int ffcz_solve_f64(ffcz_solver* h, const double* e0, int max_iterations, double tol);
int ffcz_solve_all_abs_f32(ffcz_solver* h, const float* e0, int max_iterations, float tol);
int ffcz_solve_all_abs_f64(ffcz_solver* h, const double* e0, int max_iterations, double tol);
int ffcz_solve_all_ptw_f32(ffcz_solver* h, const float* e0, int max_iterations, float tol);
int ffcz_solve_all_ptw_f64(ffcz_solver* h, const double* e0, int max_iterations, double tol);
/* the same loop with the partial ptw projection over initialize_partial_ptw's
 * index list (projection_algorithm_impl.cuh:303-333, projection_algorithm.cu:
 * 261-333); the reference stores the list (impl:701-738) but has no solve
 * that launches it.  Constraint 0 is the DC rule on bin 0, constraint j >= 1
 * the ratio rule on bin idx[j]. */
int ffcz_solve_partial_ptw_f32(ffcz_solver* h, const float* e0, int max_iterations, float tol);
int ffcz_solve_partial_ptw_f64(ffcz_solver* h, const double* e0, int max_iterations, double tol);
/* the loop alone (impl:791-808 / 835-852 / 879-896); e0 has the solver dtype */
int ffcz_solve_loop(ffcz_solver* h, int mode, const void* e0, int max_iterations, double tol);
/* iteration count and last violation: private results_ in the reference
 * (projection_algorithm.cuh:102-103), printed at impl:810/854/898 */
int ffcz_get_status(ffcz_solver* h, int* iterations, double* violation, int* converged);

/* extract_compact_representation  projection_algorithm_impl.cuh:371-456 */
int ffcz_extract(ffcz_solver* h);
/* reset_edit_accumulators  projection_algorithm_impl.cuh:133-147 */
int ffcz_reset_edit_accumulators(ffcz_solver* h);

/* ---- getters (two-call size pattern: out=NULL returns the length) --------
 * get_results                  impl:904-910  (N elements of the solver dtype) */
int ffcz_get_results(ffcz_solver* h, void* e_out, size_t cap, size_t* len);
/* get_freq_flags / get_spatial_flags            impl:459-485 (one byte per flag) */
int ffcz_get_freq_flags(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len);
int ffcz_get_spatial_flags(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len);
/* get_freq_flags_pack / get_spatial_flags_pack  impl:487-505 (LSB-first) */
int ffcz_get_freq_flags_pack(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len);
int ffcz_get_spatial_flags_pack(ffcz_solver* h, uint8_t* out, size_t cap, size_t* len);
/* get_active_*_quant_edit_distances             impl:507-529 */
int ffcz_get_active_freq_quant(ffcz_solver* h, uint16_t* out, size_t cap, size_t* len);
int ffcz_get_active_spatial_quant(ffcz_solver* h, uint16_t* out, size_t cap, size_t* len);
/* get_active_*_edit_distances                   impl:531-553 (solver dtype) */
int ffcz_get_active_freq_edits(ffcz_solver* h, void* out, size_t cap, size_t* len);
int ffcz_get_active_spatial_edits(ffcz_solver* h, void* out, size_t cap, size_t* len);
/* get_*_edits_full                              impl:555-571 (solver dtype) */
int ffcz_get_freq_edits_full(ffcz_solver* h, void* out, size_t cap, size_t* len);
int ffcz_get_spatial_edits_full(ffcz_solver* h, void* out, size_t cap, size_t* len);
/* get_{min,max}_{freq,spatial}_edit             impl:573-595 */
int ffcz_get_edit_range(ffcz_solver* h, double* fmin, double* fmax, double* smin, double* smax);

/* ---- additions (no counterpart in the reference) --------------------------
 * loop bounds shrunk by (1-margin) for certified output; quantiser choice;
 * FP64 certification of x_hat = x' + edits (quantized=1: dequantised edits). */
int ffcz_set_margins(ffcz_solver* h, double spatial_margin, double freq_margin);
/* the bound the ptw modes apply: 0 the reference's arithmetic (default;
 * projection_algorithm.cu:199-258), 1 the corrected bound: power ratio
 * |X+E|^2/|X|^2 within [1-d, 1+d] on every constrained bin, DC included,
 * Re and Im clamped to +-d'|X| with d' = min(1-sqrt(1-d), sqrt(1+d)-1)/sqrt(2) */
int ffcz_set_ptw_semantics(ffcz_solver* h, int corrected);
int ffcz_set_quantizer(ffcz_solver* h, int kind, double freq_step, double spatial_step);
int ffcz_certify(ffcz_solver* h, int quantized, ffcz_cert* out);
/* both certifications after extraction, out[0] unquantised and out[1]
 * quantised, as one paired FP64 sequence (one y launch over both spectra;
 * the two forward z transforms in one launch or two, as ffcz_autotune chose) */
int ffcz_certify_both(ffcz_solver* h, ffcz_cert* out2);
/* ffcz_certify_both in two halves, for callers that keep the GPU busy while
 * they read results: _enqueue queues the sequence and the copy of its results
 * without waiting; _results waits for them and fills out2 as ffcz_certify_both
 * does.  Work queued in between on the solver (the next field's reset and
 * solve) runs after the sequence; a second certification before the results
 * are read is an error (one result buffer).  No counterpart in the reference,
 * which has no certification. */
int ffcz_certify_both_enqueue(ffcz_solver* h);
int ffcz_certify_both_results(ffcz_solver* h, ffcz_cert* out2);
/* ---- batches: many fields of uneven shapes corrected together ----------
 * No counterpart in the reference (its driver, projection_example.cu:157-265,
 * runs one field per ProjectionSolver).  Each field is a borrowed solver
 * handle, from ffcz_batch_field: initialise it, and after the batch solve
 * extract, read and certify it, with the single-field calls; it is freed by
 * ffcz_batch_destroy, never by ffcz_destroy.  ffcz_batch_solve runs every
 * field's projection loop (solve/solve_all_abs/solve_all_ptw semantics per
 * field, mode as for ffcz_solve_loop) with one launch per pass and axis
 * length over all fields and per-field convergence flags on the device;
 * use_graph=1 replays the iteration as a captured HIP graph.  dims holds
 * (nx, ny, nz) per field; e0 one initial error (host or device) per field. */
typedef struct ffcz_batch ffcz_batch;
int ffcz_batch_create(int dtype, size_t nfields, const size_t* dims, ffcz_batch** out);
int ffcz_batch_destroy(ffcz_batch* b);
int ffcz_batch_size(ffcz_batch* b, size_t* nfields);
int ffcz_batch_field(ffcz_batch* b, size_t i, ffcz_solver** field);
int ffcz_batch_solve(ffcz_batch* b, int mode, const void* const* e0, int max_iter, double tol, int use_graph);
/* per-field iteration counts, final violations and convergence flags (any
 * may be NULL; cap entries at most); launched = iterations the batch issued
 * (in chunks between host checks; fields already converged skip them) */
int ffcz_batch_status(ffcz_batch* b, int* iterations, double* violations, int* converged, size_t cap,
                      int* launched);
/* extract_compact_representation, and certify(quantized), of every field of
 * the batch, with one launch of each kernel (extraction) or of each pass and
 * axis length (certification) over all fields, and one wait.  A field whose
 * quantised edits overflow keeps no extraction; the others keep theirs and
 * the first error is returned.  out receives one result per field
 * (cap >= field count). */
int ffcz_batch_extract(ffcz_batch* b);
int ffcz_batch_certify(ffcz_batch* b, int quantized, ffcz_cert* out, size_t cap);
/* mean ms of one whole batched iteration (HIP events on the batch stream),
 * after a completed batch solve; the edit accumulators change: reset them */
int ffcz_batch_time_iteration(ffcz_batch* b, int reps, float* ms);

/* FFT variant of the projection iteration (an addition; the reference runs
 * cuFFT at projection_algorithm_impl.cuh:205-214, 228-238):
 * FFCZ_FFT_VALU   FP32 vector-ALU Stockham passes (default)
 * FFCZ_FFT_BF16X3 radix-16 stages as DFT-matrix GEMMs on MFMA, split BF16x3
 * FFCZ_FFT_FP16X3 the same in split FP16x3 with per-butterfly power-of-two scaling
 * Lengths without a radix-16 factor keep the vector-ALU kernels.  Float
 * solvers only; certification always runs in FP64 on the vector ALU. */
#define FFCZ_FFT_VALU 0
#define FFCZ_FFT_BF16X3 1
#define FFCZ_FFT_FP16X3 2
int ffcz_set_fft_precision(ffcz_solver* h, int variant);
/* FFT plan autotuner (an addition; the reference takes cuFFT's plans,
 * projection_algorithm_impl.cuh:109-131).  Every pass kernel of the
 * iteration and of the FP64 certification has two tile variants (y/x:
 * 128- or 64-byte column tiles; z: up to 256 or 128 threads per workgroup),
 * and the iteration's vector-ALU y and x passes a third, 2: 128-byte tiles
 * with the length's radices in reversed order, where that is a different
 * plan (another rounding: results agree within the precision, not bit for
 * bit); ffcz_autotune times each on this field's buffers and keeps the
 * fastest.
 * Call it after initialize* and before solving: it overwrites the iterate
 * and resets the edit accumulators.  use_cache=1 reuses an earlier plan for
 * the same device, shape, type, FFT variant and bound mode (this process's,
 * or the plan file's; see ffcz_set_plan_cache) without timing.  vars[7] (optional)
 * receives the choice: iteration z fwd, z inv, y, x; certification z, y, x.
 * The iteration is tuned by coordinate descent over whole iterations (the
 * passes share caches); the certification over its 8 combinations, timed as
 * the paired sequence of ffcz_certify_both, and, with z variant 0, also with
 * one z launch per field instead of one for both (ffcz_set_cert_pair).  ms[9]
 * (optional; forces a fresh tuning): iteration ms with the defaults, with z
 * fwd, y, x, z inv switched in turn (a pass the FFT variant runs without tile
 * variants, as the MFMA passes, repeats the time before it), with the choice; paired certification ms
 * with the defaults and with the choice; single certification (ffcz_certify)
 * ms with the choice. */
int ffcz_autotune(ffcz_solver* h, int reps, int use_cache, int* vars, float* ms);
/* Plans persist across processes (no counterpart in the reference, which
 * re-plans with cuFFT in every constructor, projection_algorithm_impl.cuh:
 * 109-131).  A plan is keyed by device model (name, architecture, compute
 * units), library build, type, shape, FFT variant and bound mode; the file
 * is FFCZ_PLAN_CACHE, else $XDG_CACHE_HOME/ffcz/plans.txt, else
 * ~/.cache/ffcz/plans.txt, read once per process and merged under a file lock
 * after each tuning.  ffcz_set_plan_cache(path) replaces the path (NULL or ""
 * keeps plans in the process only); ffcz_get_plan_cache returns it ("" when
 * off), with the two-call size pattern.  ffcz_get_tune_source reports where
 * the solver's last ffcz_autotune plan came from: 0 none yet, 1 timed now,
 * 2 this process's plans, 3 the plan file. */
int ffcz_set_plan_cache(const char* path);
int ffcz_get_plan_cache(char* buf, size_t cap, size_t* len);
int ffcz_get_tune_source(ffcz_solver* h, int* source);
int ffcz_get_tile_vars(ffcz_solver* h, int* vars);
int ffcz_set_tile_vars(ffcz_solver* h, const int* vars);
/* ffcz_certify_both's forward z transforms: paired=1 in one launch for both
 * fields (when the certification z variant is 0), 0 in one launch each;
 * ffcz_autotune sets it, ffcz_set_tile_vars keeps it. */
int ffcz_set_cert_pair(ffcz_solver* h, int paired);
int ffcz_get_fft_precision(ffcz_solver* h, int* variant);
/* process-wide tuning: force_generic=1 runs the run-time-planned FFT kernels
 * instead of the length-specialised ones; slab_mb sets the slab size of the
 * z/y pass interleave (<= 0: whole field).  Negative force_generic leaves it. */
int ffcz_set_tuning(int force_generic, long slab_mb);
int ffcz_get_tuning(int* force_generic, long* slab_mb);
/* the solver's HIP stream (hipStream_t) */
int ffcz_get_stream(ffcz_solver* h, void** stream);
/* times the projection iteration with HIP events on the solver stream, after
 * a completed solve (the accumulators keep changing: reset them afterwards).
 * ms[6] = mean ms per launch of the z forward, y forward, fused x, y inverse
 * and z inverse passes (each launched `reps` times back to back), and mean ms
 * per whole iteration over `reps` back-to-back iterations */
int ffcz_time_iteration(ffcz_solver* h, int reps, float* ms6);
/* per-pass timing inside real solves: with enable=1 every later solve records
 * fence-free HIP events between its launches; ms[6] = mean ms per launch of
 * the z fwd, y fwd, x fused, y inv, z inv passes and the violation reduction
 * over the iterations that did work (iterations receives their count).  The
 * first launch after each host check is left out (its interval includes the
 * device waiting for the host).  Enabling resets the averages. */
int ffcz_set_pass_timing(ffcz_solver* h, int enable);
/* Graph replay of the projection loop (an addition; the reference launches
 * every kernel of every iteration, projection_algorithm_impl.cuh:835-852).
 * enable=1 (the default): each iteration after a solve's first is one HIP
 * graph launch on the solver's stream, captured on first use and again when
 * the iteration's buffers, bounds, mode, tile variants or FFT variant change.
 * Per-pass timing (above) launches eagerly.  Results are identical either way. */
int ffcz_set_graph(ffcz_solver* h, int enable);
/* borrow=1: a device-resident initial error passed to a solve is not copied;
 * the certifications after that solve read it in place, so the caller keeps
 * it allocated and unchanged until the last of them has run (for a queued
 * certification: until ffcz_certify_both_results returns) or the next solve
 * starts (the copy is one N-element write in the solve's first pass).  0, the default, copies it as the
 * reference copies its host input (projection_algorithm_impl.cuh:816-858);
 * host arrays are always copied. */
int ffcz_set_borrow_e0(ffcz_solver* h, int borrow);
int ffcz_get_pass_timing(ffcz_solver* h, float* ms6, long* iterations);
/* n6[6]: the number of timed launches behind each of those six means */
int ffcz_get_pass_counts(ffcz_solver* h, long* n6);

/* ---- entropy coding: SimpleCudaHuffmanZstdCompressor<T>
 * compress    HuffmanZSTDCoder.cuh:248-376 (symbols: host or device memory)
 * decompress  HuffmanZSTDCoder.cuh:378-450                                   */
int ffcz_huffman_zstd_compress_u16(const uint16_t* data, size_t n, int zstd_level, ffcz_coded** out);
int ffcz_huffman_zstd_compress_u8(const uint8_t* data, size_t n, int zstd_level, ffcz_coded** out);
/* Huffman only (no ZSTD): the pre-ZSTD bitstream of the same encoder */
int ffcz_huffman_encode_u16(const uint16_t* data, size_t n, ffcz_coded** out);
int ffcz_huffman_encode_u8(const uint8_t* data, size_t n, ffcz_coded** out);
int ffcz_coded_info(const ffcz_coded* c, size_t* compressed_bytes, size_t* bitstream_bytes, size_t* num_elements,
                    size_t* num_symbols, unsigned long long* num_bits);
int ffcz_coded_data(const ffcz_coded* c, uint8_t* out, size_t cap);      /* ZSTD output (or bitstream) */
int ffcz_coded_bitstream(const ffcz_coded* c, uint8_t* out, size_t cap); /* pre-ZSTD bytes */
/* code table sorted by symbol: CompressionResult::code_table, HuffmanZSTDCoder.cuh:243 */
int ffcz_coded_table(const ffcz_coded* c, uint32_t* symbols, uint32_t* codes, int32_t* lengths, size_t cap);
/* the encoder's device time for this result (histogram, codebook, lengths,
 * scan and emission, HIP events; the host transfers and, for alphabets of
 * more than 2048 symbols, the host codebook are not counted) */
int ffcz_coded_device_ms(const ffcz_coded* c, float* ms);
void ffcz_coded_free(ffcz_coded* c);
int ffcz_huffman_zstd_decompress_u16(const uint8_t* comp, size_t comp_len, const uint32_t* symbols,
                                     const uint32_t* codes, const int32_t* lengths, size_t nsym,
                                     size_t num_elements, size_t bitstream_bytes, uint16_t* out, size_t* out_len);
int ffcz_huffman_zstd_decompress_u8(const uint8_t* comp, size_t comp_len, const uint32_t* symbols,
                                    const uint32_t* codes, const int32_t* lengths, size_t nsym, size_t num_elements,
                                    size_t bitstream_bytes, uint8_t* out, size_t* out_len);

/* ---- container: the coded edits with everything a decoder needs ---------
 * No counterpart in the reference, whose code tables live in memory only and
 * which writes neither shape, type, bounds nor quantiser range
 * (HuffmanZSTDCoder.cuh:240-246, projection_example.cu:202-246).  The four streams are the ones the reference driver codes,
 * each Huffman-coded bit-exactly as HuffmanZSTDCoder.cuh:48-357 and
 * ZSTD-compressed (zstd_level <= 0: stored raw).  Layout: container.hip. */
int ffcz_container_pack(ffcz_solver* h, int zstd_level, ffcz_blob** out); /* after ffcz_extract */
int ffcz_blob_size(const ffcz_blob* b, size_t* len);
int ffcz_blob_copy(const ffcz_blob* b, uint8_t* out, size_t cap);
void ffcz_blob_free(ffcz_blob* b);
/* header fields: dims[3], n_active[2] (frequency, spatial); delta is NaN for per-bin bounds */
int ffcz_container_info(const uint8_t* buf, size_t len, int* dtype, size_t* dims, double* eps, double* delta,
                        int* quantizer, size_t* n_active);
/* x_hat = x' + C2R(deq df)/N + deq(ds) on the GPU (the reconstruction of the
 * reference's commented get_results, projection_algorithm_impl.cuh:912-985);
 * x_prime and x_hat: N elements of the container's dtype, host or device */
int ffcz_container_decode(const uint8_t* buf, size_t len, const void* x_prime, void* x_hat, void* stream);

/* ---- base compressor producing x' on the GPU ----------------------------
 * No counterpart in the reference, which takes x' from an external base
 * compressor through a file (projection_example.cu:164-176).  Dual-quantisation Lorenzo: x' = round(x/(2 eps)) * 2 eps
 * goes to x_prime (host or device, N elements, |x' - x| <= eps up to the
 * rounding of the product), and the blob codes the 3D Lorenzo residuals of
 * the integers (Huffman + ZSTD).  x' can feed ffcz_init_* / ffcz_solve_*
 * device to device. */
int ffcz_base_compress(int dtype, const void* x, size_t nx, size_t ny, size_t nz, double eps, int zstd_level,
                       void* x_prime, ffcz_blob** out, void* stream);
/* header fields: dims[3], the outlier count (residuals beyond 16 bits) */
int ffcz_base_info(const uint8_t* buf, size_t len, int* dtype, size_t* dims, double* eps, size_t* n_outliers);
int ffcz_base_decompress(const uint8_t* buf, size_t len, void* x_prime, void* stream);

/* ---- per-kernel entries on device pointers (stream may be NULL) ---------
 * 3D transforms with cuFFT's layout (compact half spectrum, interleaved
 * complex, unnormalised): cufftExecR2C/D2Z and C2R/Z2D at
 * projection_algorithm_impl.cuh:114-120, 155-161, 181-187 */
int ffcz_fft_r2c(int dtype, size_t nx, size_t ny, size_t nz, const void* in, void* out_compact, void* stream);
int ffcz_fft_c2r(int dtype, size_t nx, size_t ny, size_t nz, const void* in_compact, void* out, void* stream);
/* the same transforms with an FFT variant (FFCZ_FFT_*; dtype must be f32 for the MFMA ones) */
int ffcz_fft_r2c_prec(int dtype, int variant, size_t nx, size_t ny, size_t nz, const void* in, void* out_compact,
                      void* stream);
int ffcz_fft_c2r_prec(int dtype, int variant, size_t nx, size_t ny, size_t nz, const void* in_compact, void* out,
                      void* stream);
/* project_spatial_constraints_direct   projection_algorithm.cu:335-361 */
int ffcz_k_project_spatial(int dtype, void* e, double eps, void* ds, size_t n, void* stream);
/* project_frequency_constraints_all_abs projection_algorithm.cu:147-188 (compact spectrum, F bins) */
int ffcz_k_project_freq_abs(int dtype, void* spec, double delta, void* df, size_t F, void* stream);
/* compute_convergence_metric + host max  projection_algorithm.cu:574-613, impl:741-770 */
int ffcz_k_convergence(int dtype, const void* e, double eps, size_t n, double* out);
/* quantize_and_dequantize              projection_algorithm.cu:615-644 */
int ffcz_k_quantize_dequantize(int dtype, const void* v, void* out, double vmin, double vmax, size_t n,
                               double threshold, void* stream);
/* extract_active_flags + compact_and_quantize_active_elements + pack_bools
 * projection_algorithm.cu:363-550 on an arbitrary device array (kind: FFCZ_QUANT_*) */
int ffcz_k_extract(int dtype, const void* v, size_t n, int kind, double step, uint8_t* pack, void* compact,
                   uint16_t* quant, size_t* count, double* vmin, double* vmax, void* stream);

/* ffcz_k_extract without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and results as ffcz_k_extract; the look-back
 * status words, tile ranges, ticket and results live in a fixed device array
 * of the library (about 3 MB per GPU, taken at the first call on each device
 * and held until the process ends), so n is at most FFCZ_EXTRACT_NOALLOC_MAX
 * (FFCZ_ERR_CAPACITY beyond).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other threads, other streams) race.  The
 * caller passes every input and output; the call waits for its stream. */
#define FFCZ_EXTRACT_NOALLOC_MAX ((size_t)1 << 30)
int ffcz_k_extract_noalloc(int dtype, const void* v, size_t n, int kind, double step, uint8_t* pack, void* compact,
                           uint16_t* quant, size_t* count, double* vmin, double* vmax, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FFCZ_H */
