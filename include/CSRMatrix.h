// CSRMatrix.h -- scalar CSR (replaces formats/CSRMatrix.cuh:13-22).
// The column index is flat_index % cols (the reference's "% rows",
// formats/CSRMatrix.cu:32, is only right for square matrices).
#pragma once
#include "Matrix.h"

class CSRMatrix {
public:
    int N;  // rows
    int *hdr = nullptr;
    int *idx = nullptr;
    half *data = nullptr;
    explicit CSRMatrix(const Matrix &matrix);
    ~CSRMatrix();
    CSRMatrix(const CSRMatrix &) = delete;
    CSRMatrix &operator=(const CSRMatrix &) = delete;
    int nnz() const { return hdr[N]; }
    // Plain cudaMalloc+memcpy equivalent (formats/CSRMatrix.cu:42-63); throws
    // std::runtime_error on a HIP failure.
    void copyToDevice(int **gpuHdr, int **gpuIdx, half **gpuData) const;
};
