// formats.cpp -- host construction of the BCSR, CSR and hybrid formats
// (formats/BCSRMatrix.cu:22-104, formats/CSRMatrix.cu:13-40,
// formats/HCSRMatrix.cpp:10-36).  The device side of copyToDevice lives with
// the plan builder (bcsr_plan.hip).
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <stdexcept>
#include <string>

#include "BCSRMatrix.h"
#include "CSRMatrix.h"
#include "HCSRMatrix.h"
#include "bcsr_mi355x.h"
#include "miscutil.h"

namespace {

bool block_nonzero(const half *A, int cols, int r0, int c0, int bh, int bw) {
    for (int x = 0; x < bh; ++x) {
        const half *row = A + static_cast<size_t>(r0 + x) * cols + c0;
        for (int y = 0; y < bw; ++y)
            if (row[y] != static_cast<half>(0.0f)) return true;
    }
    return false;
}

void check_block_shape(int rows, int cols, int bh, int bw) {
    if (bh < 16 || bw < 16)
        throw std::invalid_argument("Block sizes must be at least 16x16 for MFMA tiles. Got " +
                                    std::to_string(bh) + "x" + std::to_string(bw));
    if (bh % 16 != 0 || bw % 16 != 0)
        throw std::invalid_argument("Block sizes must be multiples of 16 for MFMA tiles. Got " +
                                    std::to_string(bh) + "x" + std::to_string(bw));
    if (rows % bh != 0 || cols % bw != 0)
        throw std::invalid_argument("Matrix " + std::to_string(rows) + "x" + std::to_string(cols) +
                                    " is not a whole number of " + std::to_string(bh) + "x" +
                                    std::to_string(bw) + " blocks");
}

}  // namespace

// Two scans as in formats/BCSRMatrix.cu:36-71: count stored blocks per
// block-row, then copy each stored block row-major.
BCSRMatrix::BCSRMatrix(const Matrix &matrix, int bsr, int bsc) : blockSizeRow(bsr), blockSizeCol(bsc) {
    check_block_shape(matrix.rows, matrix.cols, bsr, bsc);
    blockRows = matrix.rows / bsr;
    blockCols = matrix.cols / bsc;
    hdr = static_cast<int *>(std::malloc(sizeof(int) * (blockRows + 1)));
    if (!hdr) throw std::bad_alloc();
    int64_t nnzb = 0;
    if (bcsr_host_count(matrix.data, matrix.rows, matrix.cols, bsr, bsc, hdr, &nnzb) != BCSR_OK)
        throw std::invalid_argument("BCSRMatrix: block count failed");
    size_t blk = static_cast<size_t>(bsr) * bsc;
    idx = static_cast<int *>(std::malloc(sizeof(int) * (nnzb ? nnzb : 1)));
    data = static_cast<half *>(std::malloc(sizeof(half) * blk * (nnzb ? nnzb : 1)));
    if (!idx || !data) throw std::bad_alloc();
    bcsr_host_fill(matrix.data, matrix.rows, matrix.cols, bsr, bsc, hdr, idx, data);
    nonZeros = static_cast<int>(nnzb * static_cast<int64_t>(blk));
}

BCSRMatrix::BCSRMatrix(int br, int bc, int bsr, int bsc, int *h, int *i, half *d)
    : hdr(h), idx(i), data(d), blockRows(br), blockCols(bc), blockSizeRow(bsr), blockSizeCol(bsc) {
    check_block_shape(br * bsr, bc * bsc, bsr, bsc);
    nonZeros = static_cast<int>(static_cast<int64_t>(hdr[blockRows]) * bsr * bsc);
}

BCSRMatrix::~BCSRMatrix() {
    std::free(hdr);
    std::free(idx);
    std::free(data);
}

void BCSRMatrix::print() const {
    std::cout << "Block size: " << blockSizeRow << "x" << blockSizeCol << std::endl;
    std::cout << "hdr:\n\t";
    for (int i = 0; i <= blockRows; i++) std::cout << hdr[i] << " ";
    std::cout << "\nidx:\n\t";
    for (int i = 0; i < hdr[blockRows]; i++) std::cout << idx[i] << " ";
    std::cout << "\ndata:\n\t";
    for (int i = 0; i < hdr[blockRows]; i++) {
        std::cout << "=== Block " << i << " ===\n";
        for (int j = 0; j < blockSizeRow; j++) {
            std::cout << '\t';
            for (int k = 0; k < blockSizeCol; k++)
                std::cout << half2float(data[static_cast<size_t>(i) * blockSizeRow * blockSizeCol +
                                             j * blockSizeCol + k])
                          << " ";
            std::cout << '\n';
        }
    }
}

CSRMatrix::CSRMatrix(const Matrix &matrix) {
    N = matrix.rows;
    hdr = static_cast<int *>(std::malloc(sizeof(int) * (matrix.rows + 1)));
    if (!hdr) throw std::bad_alloc();
    hdr[0] = 0;
    for (int i = 0; i < matrix.rows; i++) {
        hdr[i + 1] = hdr[i];
        for (int j = 0; j < matrix.cols; j++)
            if (matrix.data[static_cast<size_t>(i) * matrix.cols + j] != static_cast<half>(0.0f)) hdr[i + 1]++;
    }
    size_t nnz = static_cast<size_t>(hdr[matrix.rows]);
    idx = static_cast<int *>(std::malloc(sizeof(int) * (nnz ? nnz : 1)));
    data = static_cast<half *>(std::malloc(sizeof(half) * (nnz ? nnz : 1)));
    if (!idx || !data) throw std::bad_alloc();
    size_t k = 0, total = static_cast<size_t>(matrix.rows) * matrix.cols;
    for (size_t i = 0; i < total; i++)
        if (matrix.data[i] != static_cast<half>(0.0f)) {
            idx[k] = static_cast<int>(i % static_cast<size_t>(matrix.cols));
            data[k] = matrix.data[i];
            k++;
        }
}

CSRMatrix::~CSRMatrix() {
    std::free(hdr);
    std::free(idx);
    std::free(data);
}

// 16x16 blocks with density >= threshold are copied to the BCSR part, every
// other block to the CSR part (formats/HCSRMatrix.cpp:10-36).
HCSRMatrix::HCSRMatrix(const Matrix &matrix, float threshold) {
    const int bs = 16;
    check_block_shape(matrix.rows, matrix.cols, bs, bs);
    Matrix dense_part(matrix.rows, matrix.cols), sparse_part(matrix.rows, matrix.cols);
    for (int i = 0; i < matrix.rows; i += bs)
        for (int j = 0; j < matrix.cols; j += bs) {
            Matrix &dst = blockDensity(matrix, i, j) >= threshold ? dense_part : sparse_part;
            for (int i1 = i; i1 < i + bs; i1++)
                std::memcpy(dst.data + static_cast<size_t>(i1) * matrix.cols + j,
                            matrix.data + static_cast<size_t>(i1) * matrix.cols + j, sizeof(half) * bs);
        }
    bcsr = new BCSRMatrix(dense_part);
    csr = new CSRMatrix(sparse_part);
}

HCSRMatrix::~HCSRMatrix() {
    delete bcsr;
    delete csr;
}

// ---- C ABI: host-side BCSR construction --------------------------------------
extern "C" int bcsr_host_count(const void *dense_f16, int32_t rows, int32_t cols, int32_t bh, int32_t bw,
                               int32_t *hdr, int64_t *nnzb) {
    if (!dense_f16 || !hdr || rows < 0 || cols < 0) return BCSR_ERR_INVALID_ARG;
    if (bh < 16 || bw < 16 || bh % 16 || bw % 16) return BCSR_ERR_BLOCK_SIZE;
    if (rows % bh || cols % bw) return BCSR_ERR_SHAPE;
    const half *A = static_cast<const half *>(dense_f16);
    int nbr = rows / bh, nbc = cols / bw;
    hdr[0] = 0;
    for (int i = 0; i < nbr; ++i) {
        hdr[i + 1] = hdr[i];
        for (int j = 0; j < nbc; ++j)
            if (block_nonzero(A, cols, i * bh, j * bw, bh, bw)) hdr[i + 1]++;
    }
    if (nnzb) *nnzb = hdr[nbr];
    return BCSR_OK;
}

extern "C" int bcsr_host_fill(const void *dense_f16, int32_t rows, int32_t cols, int32_t bh, int32_t bw,
                              const int32_t *hdr, int32_t *idx, void *data_f16) {
    if (!dense_f16 || !hdr || !idx || !data_f16) return BCSR_ERR_INVALID_ARG;
    if (bh < 16 || bw < 16 || bh % 16 || bw % 16) return BCSR_ERR_BLOCK_SIZE;
    if (rows % bh || cols % bw) return BCSR_ERR_SHAPE;
    const half *A = static_cast<const half *>(dense_f16);
    half *data = static_cast<half *>(data_f16);
    int nbr = rows / bh, nbc = cols / bw;
    int64_t k = 0;
    for (int i = 0; i < nbr; ++i)
        for (int j = 0; j < nbc; ++j)
            if (block_nonzero(A, cols, i * bh, j * bw, bh, bw)) {
                if (k >= hdr[nbr]) return BCSR_ERR_INVALID_ARG;
                idx[k] = j;
                for (int x = 0; x < bh; ++x)
                    std::memcpy(data + (k * bh + x) * static_cast<int64_t>(bw),
                                A + static_cast<size_t>(i * bh + x) * cols + static_cast<size_t>(j) * bw,
                                sizeof(half) * bw);
                k++;
            }
    return k == hdr[nbr] ? BCSR_OK : BCSR_ERR_INVALID_ARG;
}
