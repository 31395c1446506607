// matrix.cpp -- dense fp16 container, ".mat" text I/O and the reference's
// miscellaneous helpers (Matrix.cu:6-46, miscutil.cpp:14-87).
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "Matrix.h"
#include "miscutil.h"

Matrix::Matrix(int rows_, int cols_) : rows(rows_), cols(cols_), nonZeros(0) {
    if (rows_ < 0 || cols_ < 0) throw std::invalid_argument("Matrix: negative dimension");
    size_t n = static_cast<size_t>(rows_) * static_cast<size_t>(cols_);
    data = static_cast<half *>(std::calloc(n ? n : 1, sizeof(half)));
    if (!data) throw std::bad_alloc();
}

Matrix::Matrix(const std::string &filename) {
    std::ifstream in(filename);
    if (!in) throw std::runtime_error("Error opening file: " + filename);
    long long r = -1, c = -1;
    if (!(in >> r >> c) || r < 0 || c < 0 || r > (1LL << 31) - 1 || c > (1LL << 31) - 1)
        throw std::runtime_error("Malformed .mat header in " + filename);
    rows = static_cast<int>(r);
    cols = static_cast<int>(c);
    size_t n = static_cast<size_t>(rows) * static_cast<size_t>(cols);
    data = static_cast<half *>(std::calloc(n ? n : 1, sizeof(half)));
    if (!data) throw std::bad_alloc();
    for (size_t i = 0; i < n; ++i) {
        float v;
        if (!(in >> v)) {
            std::free(data);
            data = nullptr;
            throw std::runtime_error("Malformed .mat body in " + filename + ": expected " +
                                     std::to_string(n) + " values, got " + std::to_string(i));
        }
        if (v != 0.0f) nonZeros++;
        data[i] = float2half_rn(v);
    }
}

Matrix::~Matrix() { std::free(data); }

void Matrix::print() const {
    std::cout << rows << ' ' << cols << "\n===\n";
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) std::cout << half2float(data[static_cast<size_t>(i) * cols + j]) << " ";
        std::cout << "\n";
    }
}

void Matrix::save(const std::string &filename) const {
    std::ofstream out(filename);
    if (!out) throw std::runtime_error("Cannot write " + filename);
    out << rows << ' ' << cols << '\n';
    out.precision(9);
    for (int i = 0; i < rows; ++i) {
        for (int j = 0; j < cols; ++j) {
            if (j) out << ' ';
            out << half2float(data[static_cast<size_t>(i) * cols + j]);
        }
        out << '\n';
    }
    if (!out) throw std::runtime_error("Write failed: " + filename);
}

// miscutil.cpp:14-20 prints from index -1; this prints the n*n entries.
void printMatrix(const float *M, unsigned int n) {
    std::cout << "Matrix size: " << n << std::endl;
    for (size_t i = 0; i < static_cast<size_t>(n) * n; i++) std::cout << M[i] << " ";
    std::cout << std::endl;
}

// Exact equality of two n x n matrices (miscutil.cpp:22-33 with its
// CHECK_CORRECTNESS switch always on).
bool checkMatrix(const float *A, const float *B, unsigned int n) {
    for (size_t i = 0; i < static_cast<size_t>(n) * n; i++)
        if (A[i] != B[i]) {
            std::cout << "Value at i = " << i << " mismatch " << A[i] << "!=" << B[i] << '\n';
            return false;
        }
    return true;
}

double rmse(const float *A, const float *B, unsigned int n) {
    double sum = 0.0;
    const size_t total = static_cast<size_t>(n) * n;
    for (size_t i = 0; i < total; i++) {
        const double d = static_cast<double>(A[i]) - static_cast<double>(B[i]);
        sum += d * d;
    }
    return std::sqrt(sum / static_cast<double>(total));
}

float maxdiff(const float *A, const float *B, unsigned int n) {
    float m = 0.0f;
    for (size_t i = 0; i < static_cast<size_t>(n) * n; i++) m = std::fmax(m, std::fabs(A[i] - B[i]));
    return m;
}

float avgrelerr(const float *A, const float *B, unsigned int n) {
    double sum = 0.0;
    const size_t total = static_cast<size_t>(n) * n;
    for (size_t i = 0; i < total; i++) {
        const double den = std::fabs(static_cast<double>(B[i]));
        const double num = std::fabs(static_cast<double>(A[i]) - static_cast<double>(B[i]));
        sum += den > 1e-12 ? num / den : num;
    }
    return static_cast<float>(sum / static_cast<double>(total));
}

float blockDensity(const Matrix &matrix, int i, int j, int blockSizeRow, int blockSizeCol) {
    int nz = 0;
    for (int i1 = i; i1 < i + blockSizeRow; i1++)
        for (int j1 = j; j1 < j + blockSizeCol; j1++)
            if (matrix.data[static_cast<size_t>(i1) * matrix.cols + j1] != static_cast<half>(0.0f)) nz++;
    return static_cast<float>(nz) / static_cast<float>(blockSizeRow * blockSizeCol);
}

// ---- C ABI over the ".mat" reader/writer ----------------------------------
#include <cstring>

#include "bcsr_mi355x.h"

extern "C" int mat_read_header(const char *path, int32_t *rows, int32_t *cols) {
    if (!path || !rows || !cols) return BCSR_ERR_INVALID_ARG;
    std::ifstream in(path);
    long long r = -1, c = -1;
    if (!in || !(in >> r >> c) || r < 0 || c < 0 || r > INT32_MAX || c > INT32_MAX) return BCSR_ERR_INVALID_ARG;
    *rows = static_cast<int32_t>(r);
    *cols = static_cast<int32_t>(c);
    return BCSR_OK;
}

extern "C" int mat_read(const char *path, int32_t rows, int32_t cols, void *out_f16, int32_t *nonzeros) {
    if (!path || !out_f16) return BCSR_ERR_INVALID_ARG;
    try {
        Matrix m{std::string(path)};
        if (m.rows != rows || m.cols != cols) return BCSR_ERR_SHAPE;
        std::memcpy(out_f16, m.data, sizeof(half) * static_cast<size_t>(rows) * cols);
        if (nonzeros) *nonzeros = m.nonZeros;
        return BCSR_OK;
    } catch (const std::runtime_error &) {
        return BCSR_ERR_INVALID_ARG;
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
}

extern "C" int mat_write(const char *path, int32_t rows, int32_t cols, const void *data_f16) {
    if (!path || !data_f16 || rows < 0 || cols < 0) return BCSR_ERR_INVALID_ARG;
    try {
        Matrix m(rows, cols);
        std::memcpy(m.data, data_f16, sizeof(half) * static_cast<size_t>(rows) * cols);
        m.save(path);
        return BCSR_OK;
    } catch (const std::runtime_error &) {
        return BCSR_ERR_INVALID_ARG;
    }
}
