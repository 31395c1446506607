// capi_host.cpp -- C ABI over the host-side generator
// (mg::generate_matrix<T>, matrix_generator.cpp:28-140, and its extensions).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "bcsr_mi355x.h"
#include "matrix_generator.h"

namespace {

template <typename T>
int copy_generated(int32_t rows, int32_t cols, double sparsity, const char *pattern, int32_t blocksize,
                   uint32_t seed, double *out) {
    auto m = mg::generate_matrix<T>(rows, cols, sparsity, std::string(pattern), blocksize, seed);
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) out[static_cast<size_t>(i) * cols + j] = static_cast<double>(m[i][j]);
    return BCSR_OK;
}

int make_spec(double sparsity, const char *pattern, int32_t bh, int32_t bw, int32_t band, double rho,
              int32_t dist, double lo, double hi, mg::GenSpec &spec) {
    if (!pattern || bh <= 0 || bw <= 0 || band < 0 || (dist != 0 && dist != 1)) return BCSR_ERR_INVALID_ARG;
    if (dist == 1 && !(lo > 0.0 && hi >= lo)) return BCSR_ERR_INVALID_ARG;
    spec.pattern = pattern;
    spec.sparsity = sparsity;
    spec.bh = bh;
    spec.bw = bw;
    spec.band_halfwidth = band;
    spec.inblock_density = rho;
    spec.dist = dist ? mg::ValueDist::LogUniform : mg::ValueDist::Uniform;
    spec.lo = lo;
    spec.hi = hi;
    return BCSR_OK;
}

}  // namespace

extern "C" int mg_generate_matrix(int32_t rows, int32_t cols, double sparsity, const char *pattern,
                                  int32_t blocksize, uint32_t seed, int32_t elem_type, double *out) {
    if (!pattern || !out || rows < 0 || cols < 0 || blocksize <= 0) return BCSR_ERR_INVALID_ARG;
    try {
        switch (elem_type) {
            case 0: return copy_generated<float>(rows, cols, sparsity, pattern, blocksize, seed, out);
            case 1: return copy_generated<double>(rows, cols, sparsity, pattern, blocksize, seed, out);
            case 2: return copy_generated<int>(rows, cols, sparsity, pattern, blocksize, seed, out);
            case 3: return copy_generated<long long>(rows, cols, sparsity, pattern, blocksize, seed, out);
            default: return BCSR_ERR_INVALID_ARG;
        }
    } catch (const std::invalid_argument &) {
        return BCSR_ERR_UNKNOWN_PATTERN;
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
}

extern "C" int mg_generate_f32(int32_t rows, int32_t cols, double sparsity, const char *pattern, int32_t bh,
                               int32_t bw, int32_t band_halfwidth, double inblock_density, int32_t value_dist,
                               double lo, double hi, uint32_t seed, float *out) {
    if (!out || rows < 0 || cols < 0) return BCSR_ERR_INVALID_ARG;
    mg::GenSpec spec;
    int rc = make_spec(sparsity, pattern, bh, bw, band_halfwidth, inblock_density, value_dist, lo, hi, spec);
    if (rc != BCSR_OK) return rc;
    try {
        mg::generate_into(out, rows, cols, spec, seed);
    } catch (const std::invalid_argument &) {
        return BCSR_ERR_UNKNOWN_PATTERN;
    }
    return BCSR_OK;
}

extern "C" int mg_generate_f16(int32_t rows, int32_t cols, double sparsity, const char *pattern, int32_t bh,
                               int32_t bw, int32_t band_halfwidth, double inblock_density, int32_t value_dist,
                               double lo, double hi, uint32_t seed, void *out) {
    if (!out || rows < 0 || cols < 0) return BCSR_ERR_INVALID_ARG;
    mg::GenSpec spec;
    int rc = make_spec(sparsity, pattern, bh, bw, band_halfwidth, inblock_density, value_dist, lo, hi, spec);
    if (rc != BCSR_OK) return rc;
    try {
        mg::generate_half_into(static_cast<half *>(out), rows, cols, spec, seed);
    } catch (const std::invalid_argument &) {
        return BCSR_ERR_UNKNOWN_PATTERN;
    }
    return BCSR_OK;
}

extern "C" int mg_generate_bcsr(int32_t rows, int32_t cols, double sparsity, const char *pattern, int32_t bh,
                                int32_t bw, int32_t band_halfwidth, double inblock_density, int32_t value_dist,
                                double lo, double hi, uint32_t seed, int32_t *hdr, int64_t *nnzb, int32_t **idx,
                                float **data) {
    if (!hdr || !nnzb || !idx || !data || rows < 0 || cols < 0) return BCSR_ERR_INVALID_ARG;
    if (bh <= 0 || bw <= 0) return BCSR_ERR_BLOCK_SIZE;
    if (rows % bh || cols % bw) return BCSR_ERR_SHAPE;
    *idx = nullptr;
    *data = nullptr;
    mg::GenSpec spec;
    int rc = make_spec(sparsity, pattern, bh, bw, band_halfwidth, inblock_density, value_dist, lo, hi, spec);
    if (rc != BCSR_OK) return rc;
    mg::BcsrArrays a;
    try {
        mg::generate_bcsr(rows, cols, spec, seed, bh, bw, a);
    } catch (const std::invalid_argument &) {
        return BCSR_ERR_UNKNOWN_PATTERN;
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
    const size_t n = a.idx.size();
    std::copy(a.hdr.begin(), a.hdr.end(), hdr);
    *nnzb = static_cast<int64_t>(n);
    *idx = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * (n ? n : 1)));
    *data = static_cast<float *>(std::malloc(sizeof(float) * (a.data.size() ? a.data.size() : 1)));
    if (!*idx || !*data) {
        std::free(*idx);
        std::free(*data);
        *idx = nullptr;
        *data = nullptr;
        return BCSR_ERR_OUT_OF_MEMORY;
    }
    std::copy(a.idx.begin(), a.idx.end(), *idx);
    std::copy(a.data.begin(), a.data.end(), *data);
    return BCSR_OK;
}

extern "C" void mg_free(void *p) { std::free(p); }
