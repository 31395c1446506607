// BCSRMatrix.h -- block CSR (replaces formats/BCSRMatrix.cuh:9-26).
//
// Host layout is the reference's: hdr[blockRows+1] prefix counts, idx[nnzb]
// ascending block-columns per block-row, data = nnzb row-major
// blockSizeRow x blockSizeCol blocks.  A block is stored iff some entry is
// nonzero as a half; zeros inside a stored block stay stored.
#pragma once
#include <cstdint>

#include "Matrix.h"

class BCSRMatrix {
public:
    int *hdr = nullptr;
    int *idx = nullptr;
    half *data = nullptr;
    int blockRows = 0;
    int blockCols = 0;
    int nonZeros = 0;  // stored entries = nnzb * blockSizeRow * blockSizeCol
    int blockSizeRow = 0;
    int blockSizeCol = 0;

    // Throws std::invalid_argument when a block size is below 16 or not a
    // multiple of 16 (formats/BCSRMatrix.cu:25-34) and -- where the reference
    // reads and writes out of bounds -- when the matrix is not a whole number
    // of blocks.
    explicit BCSRMatrix(const Matrix &matrix, int blockSizeRow = 16, int blockSizeCol = 16);
    // Adopts already-built arrays (malloc'd; freed by the destructor).
    BCSRMatrix(int blockRows, int blockCols, int blockSizeRow, int blockSizeCol, int *hdr,
               int *idx, half *data);
    ~BCSRMatrix();
    BCSRMatrix(const BCSRMatrix &) = delete;
    BCSRMatrix &operator=(const BCSRMatrix &) = delete;

    int64_t nnzb() const { return hdr[blockRows]; }
    void print() const;

    // Uploads the matrix in the device layout used by the MI355X kernels
    // (blocks repacked into MFMA operand order, 32-bit indices, block-rows
    // ordered by length) and registers it, so that the returned gpuHdr
    // identifies it to the sparseMatrixMulTensor* entry points.  gpuHdr holds
    // the reference prefix array, gpuIdx the block-column indices and gpuData
    // the repacked blocks.  Throws std::runtime_error on a HIP failure (the
    // reference prints and asserts, formats/BCSRMatrix.cu:15-19).  Free with
    // bcsr_release_device(gpuHdr).
    void copyToDevice(int **gpuHdr, int **gpuIdx, half **gpuData) const;
};
