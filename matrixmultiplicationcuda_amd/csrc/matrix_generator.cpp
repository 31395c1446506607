// matrix_generator.cpp -- seeded pattern generator.
//
// Draw-for-draw the same sequence as the reference (matrix_generator.cpp:10-140):
// one std::mt19937(seed); for every block/element the pattern visits in
// row-major order, a bernoulli_distribution decision where the pattern has
// one, then the values of a kept block drawn row-major over its clipped
// extent.  Reals are uniform_real_distribution<double>(-1, 1) narrowed with
// static_cast<T>; ints are uniform_int_distribution(-1000, 1000).  Keeping
// the same std distributions in the same order is what makes the matrices
// identical for the same seed (libstdc++ on both sides).
#include "matrix_generator.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <type_traits>

namespace mg {

namespace {

template <typename T>
T draw_value(std::mt19937 &rng, const GenSpec *spec) {
    if constexpr (std::is_integral_v<T>) {
        std::uniform_int_distribution<long long> dist(-1000, 1000);
        return static_cast<T>(dist(rng));
    } else {
        if (spec && spec->dist == ValueDist::LogUniform) {
            double u_sign = std::generate_canonical<double, std::numeric_limits<double>::digits>(rng);
            double u_mag = std::generate_canonical<double, std::numeric_limits<double>::digits>(rng);
            double sgn = u_sign < 0.5 ? -1.0 : 1.0;
            return static_cast<T>(sgn * (spec->lo * std::pow(spec->hi / spec->lo, u_mag)));
        }
        std::uniform_real_distribution<double> dist(-1.0, 1.0);
        return static_cast<T>(dist(rng));
    }
}

// Visits the pattern in the reference's draw order and hands every generated
// value to sink(row, col, value).  Entries never visited stay zero.
template <typename T, typename Sink>
void generate_core(int rows, int cols, const GenSpec &spec, unsigned int seed, Sink &&sink) {
    std::mt19937 rng(seed);
    const std::string &pattern = spec.pattern;
    const GenSpec *sp = &spec;

    auto fill_block = [&](int r0, int c0, int bh, int bw) {
        int r1 = std::min(r0 + bh, rows);
        int c1 = std::min(c0 + bw, cols);
        for (int i = r0; i < r1; ++i)
            for (int j = c0; j < c1; ++j) sink(i, j, draw_value<T>(rng, sp));
    };

    if (pattern == "random") {
        std::bernoulli_distribution keep(1.0 - spec.sparsity);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j)
                if (keep(rng)) {
                    if constexpr (std::is_integral_v<T>) {
                        std::uniform_int_distribution<int> d(-1000, 1000);
                        sink(i, j, static_cast<T>(d(rng)));
                    } else {
                        sink(i, j, draw_value<T>(rng, sp));
                    }
                }
    } else if (pattern == "checkerboard") {
        int bs = spec.bh;
        int nbi = (rows + bs - 1) / bs, nbj = (cols + bs - 1) / bs;
        for (int i = 0; i < nbi; ++i)
            for (int j = 0; j < nbj; ++j)
                if (((i % 2) ^ (j % 2)) == 0) fill_block(i * bs, j * bs, bs, bs);
    } else if (pattern == "diagonal") {
        int m = std::min(rows, cols);
        for (int i = 0; i < m; ++i) sink(i, i, draw_value<T>(rng, sp));
    } else if (pattern == "blockdiagonal") {
        int bs = spec.bh;
        int n = std::min(rows, cols) / bs + 1;
        for (int i = 0; i < n; ++i) fill_block(i * bs, i * bs, bs, bs);
    } else if (pattern == "blockrandom") {
        int bh = spec.bh, bw = spec.bw;
        int nbi = (rows + bh - 1) / bh, nbj = (cols + bw - 1) / bw;
        std::bernoulli_distribution keep(1.0 - spec.sparsity);
        for (int i = 0; i < nbi; ++i)
            for (int j = 0; j < nbj; ++j)
                if (keep(rng)) fill_block(i * bh, j * bw, bh, bw);
    } else if (pattern == "pattern64by16" || pattern == "pattern32by16") {
        const int brow = pattern == "pattern64by16" ? 64 : 32, bcol = 16;
        int nbi = (rows + brow - 1) / brow, nbj = (cols + bcol - 1) / bcol;
        std::bernoulli_distribution keep(1.0 - spec.sparsity);
        for (int i = 0; i < nbi; ++i)
            for (int j = 0; j < nbj; ++j)
                if (keep(rng)) fill_block(i * brow, j * bcol, brow, bcol);
    } else if (pattern == "blockband") {
        int bh = spec.bh, bw = spec.bw, w = spec.band_halfwidth;
        int nbi = (rows + bh - 1) / bh, nbj = (cols + bw - 1) / bw;
        for (int i = 0; i < nbi; ++i) {
            int d0 = (i * bh) / bw, d1 = (i * bh + bh - 1) / bw;
            for (int j = std::max(0, d0 - w); j < nbj && j <= d1 + w; ++j)
                fill_block(i * bh, j * bw, bh, bw);
        }
    } else if (pattern == "blockrandom_rho") {
        int bh = spec.bh, bw = spec.bw;
        int nbi = (rows + bh - 1) / bh, nbj = (cols + bw - 1) / bw;
        std::bernoulli_distribution keep(1.0 - spec.sparsity);
        std::bernoulli_distribution entry(spec.inblock_density);
        for (int bi = 0; bi < nbi; ++bi)
            for (int bj = 0; bj < nbj; ++bj)
                if (keep(rng)) {
                    int r1 = std::min(bi * bh + bh, rows), c1 = std::min(bj * bw + bw, cols);
                    for (int i = bi * bh; i < r1; ++i)
                        for (int j = bj * bw; j < c1; ++j)
                            if (entry(rng)) sink(i, j, draw_value<T>(rng, sp));
                }
    } else {
        throw std::invalid_argument("Unknown pattern: " + pattern);
    }
}

}  // namespace

template <typename T>
Matrix<T> generate_dense_block(int rows, int cols, std::mt19937 &rng) {
    Matrix<T> block(rows, std::vector<T>(cols));
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) block[i][j] = draw_value<T>(rng, nullptr);
    return block;
}

template <typename T>
Matrix<T> generate_matrix(int num_rows, int num_cols, double sparsity, const std::string &pattern,
                          int blocksize, unsigned int seed) {
    GenSpec spec;
    spec.pattern = pattern;
    spec.sparsity = sparsity;
    spec.bh = spec.bw = blocksize;
    Matrix<T> mat(num_rows, std::vector<T>(num_cols));
    generate_core<T>(num_rows, num_cols, spec, seed, [&](int i, int j, T v) { mat[i][j] = v; });
    return mat;
}

void generate_into(float *out, int rows, int cols, const GenSpec &spec, unsigned int seed) {
    std::memset(out, 0, sizeof(float) * static_cast<size_t>(rows) * static_cast<size_t>(cols));
    generate_core<float>(rows, cols, spec, seed,
                         [&](int i, int j, float v) { out[static_cast<size_t>(i) * cols + j] = v; });
}

void generate_half_into(half *out, int rows, int cols, const GenSpec &spec, unsigned int seed) {
    std::memset(out, 0, sizeof(half) * static_cast<size_t>(rows) * static_cast<size_t>(cols));
    generate_core<float>(rows, cols, spec, seed, [&](int i, int j, float v) {
        out[static_cast<size_t>(i) * cols + j] = float2half_rn(v);
    });
}

void generate_bcsr(int rows, int cols, const GenSpec &spec, unsigned int seed, int bh, int bw, BcsrArrays &out) {
    if (bh <= 0 || bw <= 0 || rows % bh || cols % bw)
        throw std::invalid_argument("generate_bcsr: matrix is not a whole number of blocks");
    // Rows of one pattern block: every pattern visits its row bands in order
    // (a band = the rows of one pattern block-row), so once a value lands in
    // a later band the earlier ones are final.  The band buffer spans
    // lcm(pattern rows, bh) rows so that it always holds whole BCSR block-rows.
    const std::string &p = spec.pattern;
    int prow = 1;
    if (p == "blockrandom" || p == "blockrandom_rho" || p == "blockband" || p == "checkerboard" ||
        p == "blockdiagonal")
        prow = spec.bh;
    else if (p == "pattern64by16")
        prow = 64;
    else if (p == "pattern32by16")
        prow = 32;
    else if (p != "random" && p != "diagonal")
        throw std::invalid_argument("Unknown pattern: " + p);
    if (prow <= 0) throw std::invalid_argument("generate_bcsr: bad pattern block size");
    const int64_t g = std::gcd(static_cast<int64_t>(prow), static_cast<int64_t>(bh));
    const int64_t band64 = static_cast<int64_t>(prow) / g * bh;
    if (band64 > (int64_t(1) << 24)) throw std::invalid_argument("generate_bcsr: pattern band too tall");
    const int band = static_cast<int>(std::min<int64_t>(band64, std::max(rows, 1)));
    const int nbr = rows / bh, nbc = cols / bw;
    const size_t blk = static_cast<size_t>(bh) * bw;
    out.hdr.assign(nbr + 1, 0);
    out.idx.clear();
    out.data.clear();
    std::vector<float> buf(static_cast<size_t>(band) * cols, 0.0f);
    int band_row0 = 0;  // first row held by buf
    auto flush = [&]() {
        const int r1 = std::min(band_row0 + band, rows);
        for (int bi = band_row0 / bh; bi < r1 / bh; ++bi)
            for (int bj = 0; bj < nbc; ++bj) {
                bool nz = false;
                for (int i = 0; i < bh && !nz; ++i)
                    for (int j = 0; j < bw; ++j)
                        if (float2half_rn(buf[static_cast<size_t>(bi * bh + i - band_row0) * cols + bj * bw + j]) !=
                            static_cast<half>(0.0f)) {
                            nz = true;
                            break;
                        }
                if (!nz) continue;
                out.idx.push_back(bj);
                for (int i = 0; i < bh; ++i) {
                    const float *src = &buf[static_cast<size_t>(bi * bh + i - band_row0) * cols + bj * bw];
                    out.data.insert(out.data.end(), src, src + bw);
                }
                out.hdr[bi + 1]++;
            }
        std::fill(buf.begin(), buf.end(), 0.0f);
    };
    generate_core<float>(rows, cols, spec, seed, [&](int i, int j, float v) {
        while (i >= band_row0 + band) {
            flush();
            band_row0 += band;
        }
        buf[static_cast<size_t>(i - band_row0) * cols + j] = v;
    });
    while (band_row0 < rows) {
        flush();
        band_row0 += band;
    }
    for (int r = 0; r < nbr; ++r) out.hdr[r + 1] += out.hdr[r];
    (void)blk;
}

template Matrix<float> generate_matrix<float>(int, int, double, const std::string &, int, unsigned int);
template Matrix<double> generate_matrix<double>(int, int, double, const std::string &, int, unsigned int);
template Matrix<int> generate_matrix<int>(int, int, double, const std::string &, int, unsigned int);
template Matrix<long long> generate_matrix<long long>(int, int, double, const std::string &, int,
                                                      unsigned int);
template Matrix<float> generate_dense_block<float>(int, int, std::mt19937 &);
template Matrix<double> generate_dense_block<double>(int, int, std::mt19937 &);
template Matrix<int> generate_dense_block<int>(int, int, std::mt19937 &);
template Matrix<long long> generate_dense_block<long long>(int, int, std::mt19937 &);

}  // namespace mg
