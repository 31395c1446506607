// legacy_driver.cpp -- the reference's argv-driven ".mat" runner (main.cu:352-611)
// on this library: reads A and B from ".mat" files, runs every product 20
// times, and prints one CSV line per product in the reference's schema
//     name,time,diff,maxrelativeerr,size,pattern,sparsity,matrixApath
// (main.cu:81-82): the median of the 20 timed runs in milliseconds, then
// maxdiff and avgrelerr (miscutil.cpp:47-74) against the dense product, which
// is the correctness reference as in the reference's driver (main.cu:427-428);
// here that product is rocBLAS's fp16 GEMM with fp32 accumulation.
//
// Usage: bcsr_legacy_driver A.mat B.mat [pattern] [sparsity]
// Unlike the reference it accepts rectangular A (M x K) and B (K x N) with
// M, K multiples of 16 and N a multiple of 16, and reports size as N.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "BCSRMatrix.h"
#include "CSRMatrix.h"
#include "HCSRMatrix.h"
#include "Matrix.h"
#include "bcsr_mi355x.h"
#include "miscutil.h"

namespace {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void bcsr_check(int rc, const char *what) {
    if (rc != BCSR_OK) throw std::runtime_error(std::string(what) + ": " + bcsr_status_string(rc));
}

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { hip_check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc"); }
    ~DevBuf() { (void)hipFree(p); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

struct Context {
    std::string a_path, pattern, sparsity;
    int M = 0, K = 0, N = 0;
    float *gpuC = nullptr;
    std::vector<float> host_c, correct;
};

// 20 timed runs (events around each), median printed with the error metrics.
void run20(Context &cx, const char *name, const std::function<void()> &call, bool is_reference = false) {
    hipEvent_t t1, t2;
    hip_check(hipEventCreate(&t1), "hipEventCreate");
    hip_check(hipEventCreate(&t2), "hipEventCreate");
    call();  // one untimed run (first-use costs: plan registry, library handles)
    hip_check(hipDeviceSynchronize(), "warm-up");
    float times[20];
    for (float &t : times) {
        hip_check(hipEventRecord(t1, nullptr), "hipEventRecord");
        call();
        hip_check(hipEventRecord(t2, nullptr), "hipEventRecord");
        hip_check(hipEventSynchronize(t2), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(&t, t1, t2), "hipEventElapsedTime");
    }
    (void)hipEventDestroy(t1);
    (void)hipEventDestroy(t2);
    std::sort(times, times + 20);
    const size_t count = static_cast<size_t>(cx.M) * cx.N;
    hip_check(hipMemcpy(cx.host_c.data(), cx.gpuC, sizeof(float) * count, hipMemcpyDeviceToHost), "copy C");
    if (is_reference) cx.correct = cx.host_c;
    // the metrics of miscutil take a square n; over M x N they are the same sums
    float md = 0.0f;
    double rel = 0.0;
    for (size_t i = 0; i < count; ++i) {
        md = std::max(md, std::fabs(cx.host_c[i] - cx.correct[i]));
        const double denom = std::fabs(static_cast<double>(cx.correct[i]));
        const double numer = std::fabs(static_cast<double>(cx.host_c[i]) - static_cast<double>(cx.correct[i]));
        rel += denom > 1e-12 ? numer / denom : numer;
    }
    std::printf("%s,%f,%lf,%lf,%u,%s,%s,%s\n", name, times[10], static_cast<double>(md),
                count ? rel / static_cast<double>(count) : 0.0, static_cast<unsigned>(cx.N), cx.pattern.c_str(),
                cx.sparsity.c_str(), cx.a_path.c_str());
    std::fflush(stdout);
    hip_check(hipMemset(cx.gpuC, 0, sizeof(float) * count), "clear C");
}

// fp32 copy of a fp16 BCSR's blocks (for the modes that quantise fp32 values)
std::vector<float> blocks_f32(const BCSRMatrix &b) {
    const size_t n = static_cast<size_t>(b.nnzb()) * b.blockSizeRow * b.blockSizeCol;
    std::vector<float> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = half2float(b.data[i]);
    return out;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s A.mat B.mat [pattern] [sparsity]\n", argv[0]);
        return 2;
    }
    try {
        Context cx;
        cx.a_path = argv[1];
        const std::string b_path = argv[2];
        cx.pattern = argc >= 4 ? argv[3] : "random";
        Matrix A(cx.a_path);
        Matrix B(b_path);
        std::cout << "Matrix A: " << A.rows << " x " << A.cols << ", nonZeros: " << A.nonZeros << std::endl;
        if (!A.rows || !A.cols || !B.rows || !B.cols) throw std::runtime_error("empty matrix");
        if (A.cols != B.rows) throw std::runtime_error("A.cols != B.rows");
        cx.M = A.rows;
        cx.K = A.cols;
        cx.N = B.cols;
        if (cx.M % 16 || cx.K % 16 || cx.N % 16) throw std::runtime_error("M, K and N must be multiples of 16");
        {
            std::ostringstream oss;
            oss << std::fixed << std::setprecision(4)
                << static_cast<double>(A.nonZeros) / (static_cast<double>(A.rows) * A.cols);
            cx.sparsity = argc >= 5 ? argv[4] : oss.str();
        }
        const size_t mk = static_cast<size_t>(cx.M) * cx.K, kn = static_cast<size_t>(cx.K) * cx.N;
        const size_t mn = static_cast<size_t>(cx.M) * cx.N;
        DevBuf dA(sizeof(half) * mk), dB(sizeof(half) * kn), dC(sizeof(float) * mn);
        hip_check(hipMemcpy(dA.p, A.data, sizeof(half) * mk, hipMemcpyHostToDevice), "copy A");
        hip_check(hipMemcpy(dB.p, B.data, sizeof(half) * kn, hipMemcpyHostToDevice), "copy B");
        cx.gpuC = static_cast<float *>(dC.p);
        cx.host_c.resize(mn);

        // dense product: the correctness reference of the remaining rows
        run20(cx, "rocBLAS GeMM", [&] {
            bcsr_check(bcsr_dense_gemm_f16(dA.p, dB.p, cx.gpuC, cx.M, cx.K, cx.N, nullptr), "dense");
        }, true);

        // CSR scalar product (the reference's "SpMM 1 Co")
        CSRMatrix csr(A);
        int *ch = nullptr, *ci = nullptr;
        half *cd = nullptr;
        csr.copyToDevice(&ch, &ci, &cd);
        run20(cx, "SpMM 1 Co", [&] {
            bcsr_check(bcsr_csr_spmm(ch, ci, cd, dB.p, cx.gpuC, cx.M, cx.K, cx.N, 0, nullptr), "csr");
        });

        // BCSR products through the reference's entry points
        BCSRMatrix bcsr(A);
        int *bh = nullptr, *bi = nullptr;
        half *bd = nullptr;
        bcsr.copyToDevice(&bh, &bi, &bd);
        run20(cx, "SpMM with Tensors", [&] {
            bcsr_check(sparseMatrixMulTensor(bh, bi, bd, dB.p, cx.gpuC, cx.M, cx.N, nullptr), "spmm");
        });
        if (cx.M == cx.N && cx.K == cx.N)
            run20(cx, "SpMM with Tensors Op", [&] {
                bcsr_check(sparseMatrixMulTensor1(bh, bi, bd, dB.p, cx.gpuC, cx.N, nullptr), "spmm1");
            });

        // vendor sparse library rows (the reference's cuSPARSE CSR, main.cu:530-563)
        {
            bcsr_sparse_ref *ref = nullptr;
            bcsr_check(bcsr_rocsparse_csr_create(ch, ci, cd, cx.M, cx.K, csr.nnz(), dB.p, cx.gpuC, cx.N, &ref),
                       "rocsparse csr");
            run20(cx, "rocSPARSE CSR", [&] { bcsr_check(bcsr_rocsparse_run(ref, nullptr), "rocsparse csr run"); });
            bcsr_rocsparse_destroy(ref);
            const size_t nnzb = static_cast<size_t>(bcsr.nnzb());
            DevBuf ph(sizeof(int) * (bcsr.blockRows + 1)), pi(sizeof(int) * (nnzb ? nnzb : 1)),
                pd(sizeof(half) * (nnzb ? nnzb : 1) * 256);
            hip_check(hipMemcpy(ph.p, bcsr.hdr, sizeof(int) * (bcsr.blockRows + 1), hipMemcpyHostToDevice), "hdr");
            hip_check(hipMemcpy(pi.p, bcsr.idx, sizeof(int) * nnzb, hipMemcpyHostToDevice), "idx");
            hip_check(hipMemcpy(pd.p, bcsr.data, sizeof(half) * nnzb * 256, hipMemcpyHostToDevice), "data");
            bcsr_check(bcsr_rocsparse_create(static_cast<int *>(ph.p), static_cast<int *>(pi.p), pd.p, bcsr.blockRows,
                                             bcsr.blockCols, static_cast<int64_t>(nnzb), 16, dB.p, cx.gpuC, cx.N, &ref),
                       "rocsparse bsr");
            run20(cx, "rocSPARSE BSR", [&] { bcsr_check(bcsr_rocsparse_run(ref, nullptr), "rocsparse bsr run"); });
            bcsr_rocsparse_destroy(ref);
        }

        // storage-precision modes of this library
        const std::vector<float> f32 = blocks_f32(bcsr);
        const struct {
            const char *name;
            int prec;
        } modes[] = {{"SpMM BF16", BCSR_PREC_BF16}, {"SpMM FP8 E4M3", BCSR_PREC_FP8_E4M3}, {"SpMM hi+lo", BCSR_PREC_F16_HILO}};
        for (const auto &m : modes) {
            bcsr_plan *plan = nullptr;
            bcsr_check(bcsr_plan_create_f32(bcsr.hdr, bcsr.idx, f32.data(), bcsr.blockRows, bcsr.blockCols, 16, 16,
                                            m.prec, &plan),
                       "plan");
            run20(cx, m.name, [&] { bcsr_check(bcsr_spmm(plan, dB.p, cx.gpuC, cx.N, nullptr), "spmm mode"); });
            bcsr_plan_destroy(plan);
        }

        // hybrid: dense 16 x 16 blocks in BCSR, the rest in CSR (main.cu:567-590)
        for (float threshold : {0.0f, 0.2f, 0.4f, 0.6f, 0.8f}) {
            HCSRMatrix hc(A, threshold);
            bcsr_plan *plan = nullptr;
            bcsr_check(bcsr_plan_create(hc.bcsr->hdr, hc.bcsr->idx, hc.bcsr->data, hc.bcsr->blockRows,
                                        hc.bcsr->blockCols, 16, 16, &plan),
                       "hybrid plan");
            int *hh = nullptr, *hi = nullptr;
            half *hd = nullptr;
            hc.csr->copyToDevice(&hh, &hi, &hd);
            std::ostringstream name;
            name << "Hybrid CSR " << std::fixed << std::setprecision(6) << threshold;
            run20(cx, name.str().c_str(), [&] {
                bcsr_check(bcsr_hybrid_spmm(plan, hh, hi, hd, dB.p, cx.gpuC, cx.N, nullptr), "hybrid");
            });
            bcsr_plan_destroy(plan);
            (void)hipFree(hh);
            (void)hipFree(hi);
            (void)hipFree(hd);
        }
        bcsr_release_device(bh);
        (void)hipFree(ch);
        (void)hipFree(ci);
        (void)hipFree(cd);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
