// matrix_generator.h -- seeded pattern generator (matrix_generator.h:10-34).
//
// generate_matrix<T> draws exactly as the reference does: std::mt19937(seed),
// bernoulli_distribution / uniform_real_distribution<double>(-1, 1) (ints:
// uniform_int_distribution(-1000, 1000)) in the same order, so a given
// (shape, sparsity, pattern, blocksize, seed) yields the same matrix.
#pragma once

#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "half_type.h"

namespace mg {

template <typename T>
using Matrix = std::vector<std::vector<T>>;

template <typename T>
Matrix<T> generate_dense_block(int rows, int cols, std::mt19937 &rng);

// Patterns: "random", "checkerboard", "diagonal", "blockdiagonal",
// "blockrandom", "pattern64by16", "pattern32by16".  Unknown names throw
// std::invalid_argument.
template <typename T>
Matrix<T> generate_matrix(int num_rows, int num_cols, double sparsity = 0.7,
                          const std::string &pattern = "random", int blocksize = 16,
                          unsigned int seed = 123u);

// ---- extensions (no reference counterpart) --------------------------------
enum class ValueDist { Uniform = 0, LogUniform = 1 };

struct GenSpec {
    std::string pattern = "random";
    double sparsity = 0.7;
    int bh = 16;                // block rows (square patterns use bh)
    int bw = 16;                // block cols for "blockrandom", "blockband", "blockrandom_rho"
    int band_halfwidth = 8;     // "blockband": block-columns either side of the diagonal
    double inblock_density = 1; // "blockrandom_rho": P(entry of a kept block is nonzero)
    ValueDist dist = ValueDist::Uniform;
    double lo = 1e-3, hi = 1e2; // LogUniform magnitude range
};

// Flat row-major float generation.  For the reference patterns with
// bh == bw == blocksize and Uniform values it equals generate_matrix<float>.
// Extra patterns: "blockband" (kept iff the block-column is within
// band_halfwidth of the block-columns crossed by the block-row's diagonal),
// "blockrandom_rho" (blockrandom whose kept blocks hold each entry with
// probability inblock_density).  LogUniform draws a sign then a magnitude
// lo*(hi/lo)^u.  Throws std::invalid_argument on an unknown pattern.
void generate_into(float *out, int rows, int cols, const GenSpec &spec, unsigned int seed);

// Same draws, rounded double -> float -> half (nvbench_main.cu:67-75).
void generate_half_into(half *out, int rows, int cols, const GenSpec &spec, unsigned int seed);

// The BCSR arrays (bh x bw blocks, float values) that BCSRMatrix would build
// from generate_half_into(rows, cols, spec, seed), produced one band of rows
// at a time without a dense matrix: same draws, a block stored iff one of its
// values is nonzero as a half.  Any pattern; throws std::invalid_argument for
// an unknown pattern or when rows % bh or cols % bw != 0.
struct BcsrArrays {
    std::vector<int32_t> hdr, idx;
    std::vector<float> data;
};
void generate_bcsr(int rows, int cols, const GenSpec &spec, unsigned int seed, int bh, int bw, BcsrArrays &out);

}  // namespace mg
