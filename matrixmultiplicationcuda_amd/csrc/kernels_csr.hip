// kernels_csr.hip -- the scalar CSR product, the hybrid (HCSR) sequence and
// the dense fp16 GEMM comparison row.
//
// CSR x dense (replaces the reference's sparseMatrixMult1Co,
// src/cuda_kernels.cu:271-285, used by its hybrid path, main.cu:567-590):
// one wave per (row, 256-column tile); lane l owns columns tile + 4 l .. + 3
// and walks the row's nonzeros in CSR order, loading 8 bytes of B per
// nonzero and accumulating in fp32 (products of two fp16 values are exact in
// fp32, so the result is the CPU k-order sum bit for bit).  accumulate = 1
// adds the finished row sum to C (C + sum, the order of the reference's
// addMatrices after sparseMatrixMult1Co), which is how the hybrid path adds
// its CSR part onto the BCSR part without a second buffer; accumulate = 2
// starts the running sum from C (((C + a0 b0) + a1 b1) + ..., the order of
// the reference's in-place sparseMatrixMult1, src/cuda_kernels.cu:287-299).
//
// Dense GEMM: rocBLAS gemm_ex (fp16 inputs, fp32 output and accumulation),
// the comparison row the reference takes from cuBLAS (nvbench_main.cu:
// 1078-1180, main.cu:508-522).
#include <rocblas/rocblas.h>

#include <mutex>
#include <new>
#include <vector>

#include "bcsr_internal.h"

namespace {

constexpr int kCsrTile = 256;

__global__ __launch_bounds__(256) void csr_spmm(const int *__restrict__ hdr, const int *__restrict__ idx,
                                                const _Float16 *__restrict__ data, const _Float16 *__restrict__ B,
                                                float *__restrict__ C, int M, int N, int n_tiles, int accumulate) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int row = __builtin_amdgcn_readfirstlane(wave / n_tiles);
    const int tile = __builtin_amdgcn_readfirstlane(wave % n_tiles);
    if (row >= M) return;
    const int lane = threadIdx.x & 63;
    const int n0 = tile * kCsrTile + 4 * lane;
    const bool valid = n0 < N;  // N % 16 == 0: all four columns or none
    const int p0 = hdr[row], p1 = hdr[row + 1];
    f32x4 *cp = reinterpret_cast<f32x4 *>(C + static_cast<size_t>(row) * N + (valid ? n0 : 0));
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    if (accumulate == 2 && valid) {
        const f32x4 c = *cp;
        acc[0] = c[0];
        acc[1] = c[1];
        acc[2] = c[2];
        acc[3] = c[3];
    }
    const _Float16 *bcol = B + (valid ? n0 : 0);
    auto step = [&](float a, u32x2 b) {
        const float b0 = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(b.x & 0xffffu)));
        const float b1 = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(b.x >> 16)));
        const float b2 = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(b.y & 0xffffu)));
        const float b3 = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(b.y >> 16)));
        acc[0] = fmaf(a, b0, acc[0]);
        acc[1] = fmaf(a, b1, acc[1]);
        acc[2] = fmaf(a, b2, acc[2]);
        acc[3] = fmaf(a, b3, acc[3]);
    };
    // eight nonzeros per step: their index, value and B loads are all in
    // flight together, and the fmas still run in CSR order (bit-exact)
    constexpr int U = 8;
    int p = p0;
    for (; p + U <= p1; p += U) {
        int k[U];
        float a[U];
        u32x2 b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            k[u] = idx[p + u];
            a[u] = static_cast<float>(data[p + u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) b[u] = *reinterpret_cast<const u32x2 *>(bcol + static_cast<size_t>(k[u]) * N);
#pragma unroll
        for (int u = 0; u < U; ++u) step(a[u], b[u]);
    }
    for (; p < p1; ++p)
        step(static_cast<float>(data[p]), *reinterpret_cast<const u32x2 *>(bcol + static_cast<size_t>(idx[p]) * N));
    if (!valid) return;
    f32x4 out{acc[0], acc[1], acc[2], acc[3]};
    if (accumulate == 1) out += *cp;
    *cp = out;
}

// C[i] += P[i] over count floats (the reference's addMatrices,
// src/cuda_kernels.cu:357-365).
__global__ __launch_bounds__(256) void add_into(float *__restrict__ C, const float *__restrict__ P, size_t count) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        C[i] += P[i];
}

int check_dense_operands(const void *B, const float *C, int N) {
    if (N < 0) return BCSR_ERR_INVALID_ARG;
    if (N % 16) return BCSR_ERR_SHAPE;
    if ((reinterpret_cast<uintptr_t>(B) & 15u) || (reinterpret_cast<uintptr_t>(C) & 15u)) return BCSR_ERR_SHAPE;
    return BCSR_OK;
}

// one rocBLAS handle per device, created on first use
std::mutex g_blas_mu;
std::vector<rocblas_handle> g_blas;

int blas_handle(rocblas_handle &h) {
    int dev = 0;
    int rc = hip_status(hipGetDevice(&dev));
    if (rc != BCSR_OK) return rc;
    std::lock_guard<std::mutex> lk(g_blas_mu);
    try {
        if (static_cast<int>(g_blas.size()) <= dev) g_blas.resize(dev + 1, nullptr);
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
    if (!g_blas[dev] && rocblas_create_handle(&g_blas[dev]) != rocblas_status_success) {
        g_blas[dev] = nullptr;
        return BCSR_ERR_UNSUPPORTED;
    }
    h = g_blas[dev];
    return BCSR_OK;
}

}  // namespace

extern "C" int bcsr_csr_spmm(const int32_t *hdr, const int32_t *idx, const void *data_f16, const void *B, float *C,
                             int32_t M, int32_t K, int32_t N, int32_t accumulate, void *stream) {
    if (M < 0 || K < 0 || accumulate < 0 || accumulate > 2) return BCSR_ERR_INVALID_ARG;
    int rc = check_dense_operands(B, C, N);
    if (rc != BCSR_OK) return rc;
    if (!M || !N) return BCSR_OK;
    if (!hdr || !C || (K && !B)) return BCSR_ERR_INVALID_ARG;
    const int n_tiles = (N + kCsrTile - 1) / kCsrTile;
    const int64_t waves = static_cast<int64_t>(M) * n_tiles;
    if (waves > (static_cast<int64_t>(INT32_MAX) - 3)) return BCSR_ERR_UNSUPPORTED;
    const unsigned grid = static_cast<unsigned>((waves + 3) / 4);
    hipLaunchKernelGGL(csr_spmm, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), hdr, idx,
                       static_cast<const _Float16 *>(data_f16), static_cast<const _Float16 *>(B), C, M, N, n_tiles,
                       accumulate);
    return hip_status(hipGetLastError());
}

extern "C" int sparseMatrixMult1(const int *hdr, const int *idx, const void *data, const void *B, float *C,
                                 unsigned int n, void *stream) {
    if (n > static_cast<unsigned>(INT32_MAX)) return BCSR_ERR_INVALID_ARG;
    return bcsr_csr_spmm(hdr, idx, data, B, C, static_cast<int32_t>(n), static_cast<int32_t>(n),
                         static_cast<int32_t>(n), 2, stream);
}

extern "C" int bcsr_add_into(float *C, const float *P, int64_t count, void *stream) {
    if (count < 0 || (count && (!C || !P))) return BCSR_ERR_INVALID_ARG;
    if (!count) return BCSR_OK;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL(add_into, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, static_cast<hipStream_t>(stream), C,
                       P, static_cast<size_t>(count));
    return hip_status(hipGetLastError());
}

extern "C" int sparseMatrixMult1Co(const int *hdr, const int *idx, const void *data, const void *B, float *C,
                                   unsigned int n, void *stream) {
    if (n > static_cast<unsigned>(INT32_MAX)) return BCSR_ERR_INVALID_ARG;
    return bcsr_csr_spmm(hdr, idx, data, B, C, static_cast<int32_t>(n), static_cast<int32_t>(n),
                         static_cast<int32_t>(n), 0, stream);
}

extern "C" int addMatrices(float *C, const float *CPart, unsigned int n, void *stream) {
    return bcsr_add_into(C, CPart, static_cast<int64_t>(n) * n, stream);
}

extern "C" int bcsr_hybrid_spmm(const bcsr_plan *plan, const int32_t *csr_hdr, const int32_t *csr_idx,
                                const void *csr_data_f16, const void *B, float *C, int32_t N, void *stream) {
    if (!plan) return BCSR_ERR_INVALID_ARG;
    int rc = bcsr_spmm(plan, B, C, N, stream);
    if (rc != BCSR_OK) return rc;
    return bcsr_csr_spmm(csr_hdr, csr_idx, csr_data_f16, B, C, plan->info.rows, plan->info.cols, N, 1, stream);
}

extern "C" int bcsr_dense_gemm_f16(const void *A, const void *B, float *C, int32_t M, int32_t K, int32_t N,
                                   void *stream) {
    if (M < 0 || K < 0 || N < 0) return BCSR_ERR_INVALID_ARG;
    if (!M || !N) return BCSR_OK;
    if (!A || !B || !C) return BCSR_ERR_INVALID_ARG;
    rocblas_handle h = nullptr;
    int rc = blas_handle(h);
    if (rc != BCSR_OK) return rc;
    std::lock_guard<std::mutex> lk(g_blas_mu);  // the handle's stream is per call
    if (rocblas_set_stream(h, static_cast<hipStream_t>(stream)) != rocblas_status_success) return BCSR_ERR_HIP;
    const float alpha = 1.0f, beta = 0.0f;
    // row-major C (M x N) = A (M x K) . B (K x N)  ==  column-major C^T = B^T . A^T
    const rocblas_status st =
        rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &alpha, B, rocblas_datatype_f16_r, N,
                        A, rocblas_datatype_f16_r, K, &beta, C, rocblas_datatype_f32_r, N, C, rocblas_datatype_f32_r, N,
                        rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
    return st == rocblas_status_success ? BCSR_OK : BCSR_ERR_HIP;
}
