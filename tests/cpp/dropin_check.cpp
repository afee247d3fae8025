// Drop-in check: the reference driver's workflow (projection_example.cu:157-265)
// written against include/projection_algorithm.hpp and
// include/huffman_zstd_coder.hpp, on a small synthetic field.  Prints one line
// of key=value results; exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "huffman_zstd_coder.hpp"
#include "projection_algorithm.hpp"

template <typename T>
static int run(size_t Nx, size_t Ny, size_t Nz) {
    const size_t n = Nx * Ny * Nz;
    std::mt19937 rng(42);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<T> x(n), e(n);
    const T eps = T(1e-3);
    for (size_t i = 0; i < n; ++i) {
        x[i] = T(u(rng));
        e[i] = T((u(rng) * 3.2 - 1.6) * 1e-3);  // x' - x, spreading past eps
    }
    ProjectionSolver<T> solver(Nx, Ny, Nz);
    bool threw = false;
    try {
        solver.solve_all_abs(e.data(), 10, T(1e-7));
    } catch (const std::runtime_error&) {
        threw = true;  // "Solver not initialized"
    }
    solver.initialize_all_abs(x.data(), eps, T(0.02));
    solver.solve_all_abs(e.data(), 100, T(1e-7));
    auto fq = solver.get_active_freq_quant_edit_distances();
    auto sq = solver.get_active_spatial_quant_edit_distances();
    auto fp = solver.get_freq_flags_pack();
    auto sp = solver.get_spatial_flags_pack();
    size_t extra = 0;
    bool roundtrip = true;
    if (!fq.empty()) {
        auto c1 = cudaHuffmanZstdCompress(fq, 20);
        auto c2 = cudaHuffmanZstdCompress(fp, 20);
        extra += c1.compressed_data.size() + c2.compressed_data.size();
        roundtrip &= cudaHuffmanZstdDecompress(c1.compressed_data, c1.code_table, c1.num_elements,
                                               c1.original_bitstream_size) == fq;
        roundtrip &= cudaHuffmanZstdDecompress(c2.compressed_data, c2.code_table, c2.num_elements,
                                               c2.original_bitstream_size) == fp;
    }
    if (!sq.empty()) {
        auto c3 = cudaHuffmanZstdCompress(sq, 20);
        auto c4 = cudaHuffmanZstdCompress(sp, 20);
        extra += c3.compressed_data.size() + c4.compressed_data.size();
        roundtrip &= cudaHuffmanZstdDecompress(c3.compressed_data, c3.code_table, c3.num_elements,
                                               c3.original_bitstream_size) == sq;
    }
    std::vector<T> out(n);
    solver.get_results(out.data());
    double mx = 0;
    for (size_t i = 0; i < n; ++i) mx = std::fmax(mx, std::fabs((double)out[i]));
    std::printf("threw=%d iterations=%d active_freq=%zu active_spatial=%zu extra_bytes=%zu roundtrip=%d max_abs_e=%.9g\n",
                (int)threw, solver.iteration_count(), fq.size(), sq.size(), extra, (int)roundtrip, mx);
    return (threw && roundtrip && solver.iteration_count() > 0) ? 0 : 1;
}

int main() { return run<float>(8, 12, 16) | run<double>(6, 10, 12); }
