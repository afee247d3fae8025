// Self-describing compressed form of one solve's edits, and its decoder.
//
// The reference keeps the Huffman code tables in memory only and writes no
// shape, type, bounds or quantiser range anywhere (HuffmanZSTDCoder.cuh:240-246,
// projection_example.cu:202-246), so its four coded streams cannot be decoded
// on their own.  The container carries all of it; the decoder rebuilds
// x_hat = x' + C2R(deq df)/N + deq(ds) on the GPU (the reconstruction the
// reference sketches in its commented get_results, projection_algorithm_impl.cuh:912-985).
#pragma once

#include <vector>

#include "solver.hpp"

namespace ffcz {

struct ContainerInfo {
    int dtype = 0;   // 0 float, 1 double
    int mode = 0;    // solve mode of the edits (informational)
    int qkind = 0;   // Q_REF16 or Q_STEP
    size_t nx = 0, ny = 0, nz = 0;
    double eps = 0, delta = 0;  // delta: NaN for per-bin bounds
    size_t n_freq = 0, n_spatial = 0;  // active (coded) edits
};

// The solver's extracted edits as a container (Huffman streams ZSTD-compressed
// at zstd_level, or stored raw when zstd_level <= 0).
template <typename T>
std::vector<uint8_t> container_pack(Solver<T>& s, int zstd_level);

// Header fields; throws ERR_INVALID on a malformed header.
ContainerInfo container_info(const uint8_t* buf, size_t len);

// x_hat = x' + C2R(deq df)/N + deq(ds) in the container's type.  x_prime and
// x_hat hold N elements (host or device memory); every stream is validated
// before any device work.
void container_decode(const uint8_t* buf, size_t len, const void* x_prime, void* x_hat, hipStream_t s);

// The base compressor that produces x' for the solver (the reference takes
// x' from an external compressor): dual-quantisation Lorenzo.  x' = round(x/(2 eps)) * 2 eps is
// written to x_prime (host or device, N elements), and the returned stream
// codes the 3D Lorenzo residuals of the integers (Huffman + ZSTD, outliers
// beyond 16 bits listed).  Layout: container.hip.
template <typename T>
std::vector<uint8_t> base_compress(const T* x, size_t nx, size_t ny, size_t nz, double eps, int zstd_level,
                                   T* x_prime, hipStream_t s);
// header fields of a base stream (n_spatial: its outlier count)
ContainerInfo base_info(const uint8_t* buf, size_t len);
// x' from a base stream (host or device, N elements of its dtype)
void base_decompress(const uint8_t* buf, size_t len, void* x_prime, hipStream_t s);

}  // namespace ffcz
