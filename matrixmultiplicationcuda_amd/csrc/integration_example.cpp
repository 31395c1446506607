// integration_example.cpp -- the reference's NVBench call site of
// sparseMatrixMulTensor_v2 (nvbench_main.cu:393-402) rewritten for this
// library, as a complete program:
//
//   reference                                  here
//   prepare_buffers (nvbench_main.cu:139-175)  the same generator calls and seeds
//   BCSRMatrix(A, 16, 16).copyToDevice(...)    the same class and call
//   cudaMalloc / cudaMemcpy of B and C         hipMalloc / hipMemcpy
//   sparseMatrixMulTensor_v2<<<grid, block,    sparseMatrixMulTensor_v2(d_hdr, d_idx, d_data,
//       0, stream>>>(d_hdr, d_idx, d_data,         d_B, d_C, M, N)   -- null stream by default
//       d_B, d_C, M, N)
//
// usage: bcsr_integration_example out.bin [M K N sparsity]
// Writes C (M x N fp32, row-major) to out.bin and prints M K N nnzb.  The
// defaults are the reference's small case: 1024 x 1024 A with 10 % of its
// 16 x 16 blocks ("blockrandom", seed 123), B 1024 x 256 (seed 456).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "BCSRMatrix.h"
#include "Matrix.h"
#include "bcsr_mi355x.h"
#include "matrix_generator.h"

static int fail(const char *what, int rc) {
    std::fprintf(stderr, "%s: %s\n", what, bcsr_status_string(rc));
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.bin [M K N sparsity]\n", argv[0]);
        return 2;
    }
    const int M = argc > 2 ? std::atoi(argv[2]) : 1024;
    const int K = argc > 3 ? std::atoi(argv[3]) : 1024;
    const int N = argc > 4 ? std::atoi(argv[4]) : 256;
    const double sp = argc > 5 ? std::atof(argv[5]) : 0.9;

    // prepare_buffers: A and B from the generator, double -> float -> half
    Matrix A(M, K), B(K, N);
    const auto gA = mg::generate_matrix<float>(M, K, sp, "blockrandom", 16, 123);
    const auto gB = mg::generate_matrix<float>(K, N, 0.0, "random", 16, 456);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < K; ++j) A.data[static_cast<size_t>(i) * K + j] = float2half_rn(gA[i][j]);
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < N; ++j) B.data[static_cast<size_t>(i) * N + j] = float2half_rn(gB[i][j]);

    BCSRMatrix A_bcsr(A, 16, 16);
    int *d_hdr = nullptr, *d_idx = nullptr;
    half *d_data = nullptr;
    A_bcsr.copyToDevice(&d_hdr, &d_idx, &d_data);

    half *d_B = nullptr;
    float *d_C = nullptr;
    if (hipMalloc(&d_B, sizeof(half) * K * N) != hipSuccess || hipMalloc(&d_C, sizeof(float) * M * N) != hipSuccess ||
        hipMemcpy(d_B, B.data, sizeof(half) * K * N, hipMemcpyHostToDevice) != hipSuccess)
        return fail("hip", BCSR_ERR_HIP);

    // the reference's launch, without its launch configuration
    int rc = sparseMatrixMulTensor_v2(d_hdr, d_idx, d_data, d_B, d_C, M, N);
    if (rc != BCSR_OK) return fail("sparseMatrixMulTensor_v2", rc);

    std::vector<float> C(static_cast<size_t>(M) * N);
    if (hipMemcpy(C.data(), d_C, sizeof(float) * C.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return fail("hipMemcpy", BCSR_ERR_HIP);
    FILE *f = std::fopen(argv[1], "wb");
    if (!f || std::fwrite(C.data(), sizeof(float), C.size(), f) != C.size()) return fail("write", BCSR_ERR_INVALID_ARG);
    std::fclose(f);
    std::printf("%d %d %d %lld\n", M, K, N, static_cast<long long>(A_bcsr.nnzb()));

    (void)hipFree(d_B);
    (void)hipFree(d_C);
    rc = bcsr_release_device(d_hdr);  // the reference leaks these arrays
    return rc == BCSR_OK ? 0 : fail("bcsr_release_device", rc);
}
