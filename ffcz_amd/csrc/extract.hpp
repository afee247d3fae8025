// Edit extraction: flags, min/max, stable compaction, quantisation and LSB-first
// flag packing (reference: extract_compact_representation,
// projection_algorithm_impl.cuh:371-456, and its helpers
// projection_algorithm.cu:363-550).
#pragma once

#include <vector>

#include "fft3d.hpp"

namespace ffcz {

constexpr int EXTRACT_TILE = 8192;     // elements per workgroup tile (8 per thread x 4 rounds)

struct ExtractResult {
    size_t count = 0;
    double vmin = 0, vmax = 0;
};

// Device-side results of one extraction (vmin, vmax adjacent: the
// dequantisers read them through QuantDesc::dminmax).
struct ExtractGlobals {
    unsigned long long count;
    double vmin, vmax;
    unsigned int overflow;  // bit 0: a step-quantised symbol exceeds 16 bits; bit 1: look-back stalled
    unsigned int pad;
};

// Device scratch needed by extract_edits for n elements.
size_t extract_scratch_bytes(size_t n);

// v[n] -> pack[ceil(n/8) rounded up to 8 bytes], compact[count], quant[count].
//   q.kind == Q_REF16: flag = |v| > threshold, min/max over the full array,
//     q = n>0.5 ? ceil(n*65535) : floor(n*65535) with n = (v-min)/(max-min).
//   q.kind == Q_STEP:  r = llround(v/step); flag = r != 0; quant = zigzag(r)
//     (ERR_RANGE if a symbol does not fit 16 bits).
// Returns count (synchronises the stream once).
template <typename T>
ExtractResult extract_edits(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                            void* scratch, hipStream_t s);
// The same without a host synchronisation: results land in dglob (device,
// kept until the next extraction into it) and, unless hglob is null, are
// copied asynchronously to hglob (pinned host memory) behind the kernels on
// stream s.
template <typename T>
void extract_edits_async(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                         void* scratch, ExtractGlobals* dglob, ExtractGlobals* hglob, hipStream_t s);
// What extract_edits_async gives for n values that are all +0 (no flag set:
// a zero flag pack of ceil(n/64) 64-bit words, count 0, range [0, 0]), in
// one launch that reads no values.
void extract_zeros_async(uint8_t* pack, size_t n, ExtractGlobals* dglob, hipStream_t s);

// One edit array of a batched extraction, with its quantiser and outputs.
template <typename T>
struct ExtractSeg {
    const T* v;
    size_t n;
    QuantDesc q;
    unsigned int* pack;  // the flag bytes, written four at a time
    T* compact;
    uint16_t* quant;
    ExtractGlobals* g;   // device results (zeroed first)
    unsigned int tile0;  // first tile in the batch's numbering (set by extract_edits_batch)
};

// Device scratch and pinned staging of batched extractions, grown as needed
// and reused (a call waits for the previous one's results to be copied out).
struct ExtractBatchWork {
    unsigned char* dev = nullptr;
    size_t dev_cap = 0;
    unsigned char* host = nullptr;
    size_t host_cap = 0;
    hipEvent_t done = nullptr;
    void release();
};

// Every segment extracted as by extract_edits_async, with one launch of each
// kernel for the whole batch on stream s.  Returns pinned host copies of the
// segments' results, in segment order, valid once s has drained (and until
// the next call with the same workspace).
template <typename T>
const ExtractGlobals* extract_edits_batch(std::vector<ExtractSeg<T>>& segs, ExtractBatchWork& w, hipStream_t s);

// extract_edits without any device allocation: the look-back status words,
// tile ranges, ticket and results live in a fixed device array of this
// library (one per GPU, taken at the first call on each device), which holds
// arrays of up to EXTRACT_NOALLOC_MAX elements (ERR_CAPACITY beyond).  Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU race.
constexpr size_t EXTRACT_NOALLOC_MAX = size_t(1) << 30;
template <typename T>
ExtractResult extract_edits_noalloc(const T* v, size_t n, QuantDesc q, uint8_t* pack, T* compact, uint16_t* quant,
                                    hipStream_t s);

// Dense dequantised copy (projection_algorithm.cu:615-644 for Q_REF16).
template <typename T>
void dequantize_dense(const T* v, size_t n, QuantDesc q, T* out, hipStream_t s);

}  // namespace ffcz
