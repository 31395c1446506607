// Host-code memory and undefined-behaviour check (built by
// tests/test_sanitizers.py with AddressSanitizer + UBSan, CPU only): drives
// the host library (generator, .mat I/O, BCSR / CSR / HCSR construction, the
// generator's C ABI) and the oracle's C restatement over every pattern, odd
// shapes and the error paths, and prints "ok <checksum>".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "BCSRMatrix.h"
#include "CSRMatrix.h"
#include "HCSRMatrix.h"
#include "Matrix.h"
#include "bcsr_mi355x.h"
#include "matrix_generator.h"

extern "C" {
#include "oracle.h"
}

static const char *kPatterns[] = {"random",        "checkerboard",  "diagonal", "blockdiagonal", "blockrandom",
                                  "pattern64by16", "pattern32by16", "blockband", "blockrandom_rho"};

static bool reference_pattern(const std::string &p) { return p != "blockband" && p != "blockrandom_rho"; }

int main(int argc, char **argv) {
    const std::string tmp = argc > 1 ? argv[1] : "/tmp";
    double sum = 0.0;
    // ---- host library ------------------------------------------------------
    for (const char *p : kPatterns) {
        for (int rows : {1, 97, 128}) {
            for (int cols : {1, 131, 256}) {
                if (reference_pattern(p)) {
                    sum += mg::generate_matrix<float>(rows, cols, 0.7, p, 16, 7).size();
                    sum += mg::generate_matrix<int>(rows, cols, 0.7, p, 16, 7).size();
                    sum += mg::generate_matrix<double>(rows, cols, 0.7, p, 32, 7).size();
                }
                mg::GenSpec sp;
                sp.pattern = p;
                sp.sparsity = 0.6;
                sp.inblock_density = 0.5;
                sp.dist = mg::ValueDist::LogUniform;
                std::vector<float> f(static_cast<size_t>(rows) * cols);
                mg::generate_into(f.data(), rows, cols, sp, 3);
                Matrix M(rows, cols);
                mg::generate_half_into(M.data, rows, cols, sp, 3);
                if (rows % 16 || cols % 16) {
                    try {
                        BCSRMatrix bad(M, 16, 16);
                    } catch (const std::invalid_argument &) {
                        sum += 1;
                    }
                    continue;
                }
                for (int bh : {16, 32, 64}) {
                    if (rows % bh) continue;
                    mg::BcsrArrays arr;
                    mg::GenSpec s2 = sp;
                    s2.bh = bh;
                    mg::generate_bcsr(rows, cols, s2, 3, bh, 16, arr);
                    BCSRMatrix B(M, bh, 16);
                    sum += static_cast<double>(B.nnzb()) + arr.idx.size();
                }
                CSRMatrix C(M);
                for (float th : {0.0f, 0.3f, 1.0f}) {
                    HCSRMatrix H(M, th);
                    sum += H.csr->nnz() + H.bcsr->nnzb();
                }
                sum += C.nnz();
                const std::string path = tmp + "/sanitize.mat";
                M.save(path);
                Matrix L(path);
                sum += L.nonZeros;
            }
        }
    }
    try {
        Matrix bad16(64, 64);
        BCSRMatrix b(bad16, 8, 16);
    } catch (const std::invalid_argument &) {
        sum += 1;
    }
    try {
        (void)mg::generate_matrix<float>(8, 8, 0.5, "nosuch", 2, 1);
    } catch (const std::invalid_argument &) {
        sum += 1;
    }
    {
        const std::string path = tmp + "/sanitize_bad.mat";
        FILE *fp = std::fopen(path.c_str(), "w");
        std::fprintf(fp, "3 3\n1 2\n");
        std::fclose(fp);
        try {
            Matrix bad(path);
        } catch (const std::exception &) {
            sum += 1;
        }
    }
    {
        std::vector<int32_t> hdr(64 + 1);
        int64_t nnzb = 0;
        int32_t *idx = nullptr;
        float *data = nullptr;
        if (mg_generate_bcsr(1024, 512, 0.9, "blockrandom", 16, 16, 8, 1.0, 1, 1e-3, 1e2, 5, hdr.data(), &nnzb, &idx,
                             &data) == 0) {
            sum += static_cast<double>(nnzb);
            mg_free(idx);
            mg_free(data);
        }
        if (mg_generate_bcsr(72, 64, 0.5, "blockrandom", 16, 16, 8, 1.0, 0, 1e-3, 1e2, 5, hdr.data(), &nnzb, &idx,
                             &data) != 0)
            sum += 1;  // 72 rows: not whole blocks
    }
    // ---- oracle ------------------------------------------------------------
    for (const char *p : kPatterns) {
        const int rows = 96, cols = 80;
        std::vector<double> d(static_cast<size_t>(rows) * cols);
        if (reference_pattern(p))
            for (int t : {ORC_T_FLOAT, ORC_T_DOUBLE, ORC_T_INT, ORC_T_LONGLONG})
                sum += orc_generate_matrix(rows, cols, 0.7, p, 16, 9, t, d.data()) == 0;
        std::vector<float> fa(static_cast<size_t>(rows) * cols);
        sum += orc_generate_ex(rows, cols, 0.7, p, 16, 16, 9, ORC_VAL_LOGUNIFORM, 1e-3, 1e2, 2, 0.5, fa.data()) == 0;
        std::vector<uint16_t> A16(fa.size());
        orc_f32_to_f16(fa.data(), A16.data(), fa.size());
        std::vector<int32_t> hdr(rows / 16 + 1);
        const int64_t nnzb = orc_bcsr_count(A16.data(), rows, cols, 16, 16, hdr.data());
        std::vector<int32_t> idx(nnzb + 1);
        std::vector<uint16_t> data(static_cast<size_t>(nnzb) * 256 + 1);
        std::vector<float> data32(static_cast<size_t>(nnzb) * 256 + 1);
        orc_bcsr_fill(A16.data(), rows, cols, 16, 16, hdr.data(), idx.data(), data.data());
        orc_bcsr_fill_f32(fa.data(), rows, cols, 16, 16, hdr.data(), idx.data(), data32.data());
        std::vector<uint16_t> dense(A16.size());
        orc_bcsr_densify(hdr.data(), idx.data(), data.data(), rows / 16, cols / 16, 16, 16, dense.data());
        const int N = 48;
        std::vector<float> Bf(static_cast<size_t>(cols) * N);
        std::vector<uint16_t> B16(Bf.size());
        for (size_t i = 0; i < Bf.size(); ++i) Bf[i] = orc_h2f(orc_f2h(static_cast<float>(i % 7) - 3.0f));
        orc_f32_to_f16(Bf.data(), B16.data(), Bf.size());
        std::vector<float> C(static_cast<size_t>(rows) * N), Cd(C.size());
        orc_bcsr_spmm_f32(hdr.data(), idx.data(), data.data(), 16, 16, 0, rows / 16, Bf.data(), N, C.data(), 3);
        orc_matmul_dense_ref(A16.data(), B16.data(), rows, cols, N, Cd.data());
        std::vector<double> C64(C.size()), Cabs(C.size());
        orc_bcsr_spmm_f64(hdr.data(), idx.data(), data32.data(), 16, 16, 0, rows / 16, Bf.data(), N, C64.data(),
                          Cabs.data(), 2);
        std::vector<int32_t> chdr(rows + 1);
        const int64_t nnz = orc_csr_count(A16.data(), rows, cols, chdr.data());
        std::vector<int32_t> cidx(nnz + 1);
        std::vector<uint16_t> cdata(nnz + 1);
        orc_csr_fill(A16.data(), rows, cols, chdr.data(), cidx.data(), cdata.data());
        orc_csr_spmm_f32(chdr.data(), cidx.data(), cdata.data(), rows, Bf.data(), N, C.data());
        orc_csr_spmm_acc_f32(chdr.data(), cidx.data(), cdata.data(), rows, Bf.data(), N, C.data());
        std::vector<float> Aq(data32.size());
        orc_quant_e4m3_tiles(hdr.data(), idx.data(), data32.data(), rows / 16, 16, 16, Aq.data());
        for (size_t i = 0; i < C.size(); ++i) sum += C[i] + Cd[i] + C64[i] + Cabs[i];
    }
    {
        const int K = 48, N = 592;  // 3 column chunks, the last partial
        std::vector<float> B(static_cast<size_t>(K) * N);
        for (size_t i = 0; i < B.size(); ++i) B[i] = static_cast<float>((i * 37) % 101) - 50.0f;
        // one scale per (16-row slab, 256-column chunk), chunk rows padded to 16 bytes
        std::vector<uint8_t> q(B.size()), sc(static_cast<size_t>((N + 255) / 256) * ((K / 16 + 15) & ~15));
        orc_quant_e4m3_B(B.data(), K, N, q.data(), sc.data());
        for (uint8_t v : q) sum += v;
        for (uint8_t v : sc) sum += v;
        uint16_t hi, lo;
        orc_hilo_split(1.2345f, &hi, &lo);
        sum += hi + lo;
    }
    std::printf("ok %.6g\n", sum);
    return 0;
}
