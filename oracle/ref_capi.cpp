// ref_capi.cpp -- C entry point over the reference's own generator.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference's matrix_generator.cpp, read where it lies; the result goes to
// oracle/_ref/ and is used by tests/ as the reference generator (the checker
// for the oracle's restatement and for the product generator).
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "matrix_generator.h"

namespace {
template <typename T>
int copy_out(int rows, int cols, double sparsity, const char *pattern, int blocksize,
             unsigned seed, double *out) {
    try {
        auto m = mg::generate_matrix<T>(rows, cols, sparsity, std::string(pattern), blocksize, seed);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j) out[(size_t)i * cols + j] = (double)m[i][j];
        return 0;
    } catch (const std::invalid_argument &) {
        return -1;
    }
}
}  // namespace

extern "C" int ref_generate_matrix(int rows, int cols, double sparsity, const char *pattern,
                                   int blocksize, unsigned seed, int elem_type, double *out) {
    switch (elem_type) {
        case 0: return copy_out<float>(rows, cols, sparsity, pattern, blocksize, seed, out);
        case 1: return copy_out<double>(rows, cols, sparsity, pattern, blocksize, seed, out);
        case 2: return copy_out<int>(rows, cols, sparsity, pattern, blocksize, seed, out);
        case 3: return copy_out<long long>(rows, cols, sparsity, pattern, blocksize, seed, out);
        default: return -2;
    }
}

// Float-only variant writing float directly (halves the memory of large cases).
extern "C" int ref_generate_matrix_f32(int rows, int cols, double sparsity, const char *pattern,
                                       int blocksize, unsigned seed, float *out) {
    try {
        auto m = mg::generate_matrix<float>(rows, cols, sparsity, std::string(pattern), blocksize, seed);
        for (int i = 0; i < rows; ++i) std::memcpy(out + (size_t)i * cols, m[i].data(), sizeof(float) * cols);
        return 0;
    } catch (const std::invalid_argument &) {
        return -1;
    }
}
