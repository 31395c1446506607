/*
 * oracle.h -- CPU restatement of the reference's BCSR x dense path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded only by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  The product (matrixmultiplicationcuda_amd) never calls it.
 *
 * What each function restates (reference file:line) and what pins it:
 *
 *   orc_mt_*              std::mt19937 as used by matrix_generator.cpp:34
 *                         pinned: C++11 known answer (10000th output of the
 *                         default-seeded engine is 4123659995) and numpy's
 *                         independent MT19937 (RandomState legacy seeding).
 *   orc_canonical         libstdc++ generate_canonical<double,53> (the engine
 *                         adaptor behind uniform_real_distribution<double> and
 *                         bernoulli_distribution, matrix_generator.cpp:20,50)
 *                         pinned: oracle/_ref (the reference generator compiled
 *                         from matrix_generator.cpp) on every pattern.
 *   orc_uniform_int       libstdc++ uniform_int_distribution (Lemire downscale),
 *                         matrix_generator.cpp:15,54.  pinned: oracle/_ref int.
 *   orc_generate          mg::generate_matrix<T>, matrix_generator.cpp:28-140,
 *                         all seven reference patterns; pinned: oracle/_ref
 *                         bit-exact + committed golden hashes.
 *                         Extensions with no reference counterpart (parity
 *                         unpinned vs the reference; pinned by structural
 *                         invariants and distribution checks in tests):
 *                         "blockband", "blockrandom_rho", value dist "loguniform".
 *   orc_f2h / orc_h2f     __float2half / __half2float (RNE), used by
 *                         nvbench_main.cu:67-75 and Matrix.cu:30.
 *                         pinned: numpy float16 conversion over all halves and
 *                         a dense float sweep including ties and subnormals.
 *   orc_bcsr_count/fill   BCSRMatrix::BCSRMatrix, formats/BCSRMatrix.cu:22-74 with
 *                         blockDensity, miscutil.cpp:76-87.  pinned: invariants
 *                         that fix the format uniquely + densify round trip.
 *   orc_matmul_dense_ref  cpu_matmul_ref, nvbench_main.cu:78-91 (literal form).
 *   orc_bcsr_spmm_f32     the same product in BCSR order; bit-identical to the
 *                         dense form (fp16*fp16 is exact in fp32, skipped zero
 *                         terms add +0, block columns ascend).  pinned: exact
 *                         equality with orc_matmul_dense_ref, closed forms
 *                         (identity blocks, one-hot B), numpy float64 matmul.
 *   orc_bcsr_spmm_f64     FP64 oracle on the unquantised generator values.
 *                         pinned: numpy float64 matmul.
 *   orc_csr_*             CSRMatrix::CSRMatrix, formats/CSRMatrix.cu:13-40
 *                         (column index taken modulo cols; the reference's
 *                         "% rows" is only correct for square matrices).
 */
#ifndef BCSR_ORACLE_H
#define BCSR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t mt[624];
    int idx;
} orc_mt;

void orc_mt_seed(orc_mt *g, uint32_t seed);
uint32_t orc_mt_next(orc_mt *g);
/* Fill out[n] with consecutive engine outputs of mt19937(seed). */
void orc_mt_stream(uint32_t seed, uint32_t *out, size_t n);
double orc_canonical(orc_mt *g);
int orc_bernoulli(orc_mt *g, double p);
double orc_uniform(orc_mt *g, double a, double b);
int64_t orc_uniform_int(orc_mt *g, int64_t a, int64_t b);

uint16_t orc_f2h(float f);
float orc_h2f(uint16_t h);
void orc_f32_to_f16(const float *in, uint16_t *out, size_t n);
void orc_f16_to_f32(const uint16_t *in, float *out, size_t n);

/* Value distributions for generated nonzeros. */
enum { ORC_VAL_UNIFORM = 0, ORC_VAL_LOGUNIFORM = 1 };
/* Element types of the reference's generate_matrix<T> instantiations. */
enum { ORC_T_FLOAT = 0, ORC_T_DOUBLE = 1, ORC_T_INT = 2, ORC_T_LONGLONG = 3 };

/*
 * mg::generate_matrix<T>(rows, cols, sparsity, pattern, blocksize, seed),
 * written row-major as double into out[rows*cols] (a float/int result is
 * exactly representable in double).  elem_type selects T.
 * Returns 0, or -1 for an unknown pattern (the reference throws
 * std::invalid_argument, matrix_generator.cpp:135-137).
 */
int orc_generate_matrix(int rows, int cols, double sparsity, const char *pattern,
                        int blocksize, uint32_t seed, int elem_type, double *out);

/*
 * Extended generator (float only).  bh x bw block shape for the block
 * patterns, band_w = half-width in block-columns for "blockband",
 * rho = in-block element density for "blockrandom_rho", value_dist one of
 * ORC_VAL_*, with [lo, hi] the magnitude range for ORC_VAL_LOGUNIFORM.
 * Reference patterns reproduce orc_generate_matrix(..., ORC_T_FLOAT) when
 * bh == bw == blocksize and value_dist == ORC_VAL_UNIFORM.
 */
int orc_generate_ex(int rows, int cols, double sparsity, const char *pattern,
                    int bh, int bw, uint32_t seed, int value_dist, double lo,
                    double hi, int band_w, double rho, float *out);

/* BCSR construction from a dense row-major fp16 matrix. */
int64_t orc_bcsr_count(const uint16_t *A, int rows, int cols, int bh, int bw,
                       int32_t *hdr /* rows/bh + 1 */);
void orc_bcsr_fill(const uint16_t *A, int rows, int cols, int bh, int bw,
                   const int32_t *hdr, int32_t *idx, uint16_t *data);
/* Same structure, float payload (for the FP64 oracle on unquantised values). */
void orc_bcsr_fill_f32(const float *A, int rows, int cols, int bh, int bw,
                       const int32_t *hdr, int32_t *idx, float *data);
/* Densify: out[rows*cols] fp16 from BCSR. */
void orc_bcsr_densify(const int32_t *hdr, const int32_t *idx, const uint16_t *data,
                      int block_rows, int block_cols, int bh, int bw, uint16_t *out);

/* CSR construction (column index = flat index % cols). */
int64_t orc_csr_count(const uint16_t *A, int rows, int cols, int32_t *hdr);
void orc_csr_fill(const uint16_t *A, int rows, int cols, const int32_t *hdr,
                  int32_t *idx, uint16_t *data);

/* CSR x dense in CSR order, fp32 from +0 (the product sparseMatrixMult1Co
 * computes, src/cuda_kernels.cu:271-285, restated on the CPU); pinned by
 * exact equality with the dense cpu_matmul_ref loop on the densified CSR
 * matrix (fp16 products are exact in fp32, zero terms add +0, k ascends). */
void orc_csr_spmm_f32(const int32_t *hdr, const int32_t *idx, const uint16_t *data, int M, const float *Bf, int N,
                      float *C);
/* The same loop summing onto what C holds, the in-place running sum of
 * sparseMatrixMult1 (src/cuda_kernels.cu:287-299): c = ((c + a0 b0) + a1 b1)...;
 * pinned by C = 0 giving orc_csr_spmm_f32 and by a rounding case whose result
 * differs from c + (a0 b0 + a1 b1). */
void orc_csr_spmm_acc_f32(const int32_t *hdr, const int32_t *idx, const uint16_t *data, int M, const float *Bf,
                          int N, float *C);
/* cpu_matmul_ref (dense, literal loop order). */
void orc_matmul_dense_ref(const uint16_t *A, const uint16_t *B, int M, int K, int N,
                          float *C);
/*
 * BCSR-order fp32 product over block-rows [br0, br1) (C rows of those
 * block-rows are overwritten), nthreads >= 1 splits the block-rows.
 * Bf is B already widened to fp32 (K x N).
 */
void orc_bcsr_spmm_f32(const int32_t *hdr, const int32_t *idx, const uint16_t *data,
                       int bh, int bw, int br0, int br1, const float *Bf, int N,
                       float *C, int nthreads);
/* FP64 product and |A|.|B| bound on float-valued BCSR data (same structure). */
void orc_bcsr_spmm_f64(const int32_t *hdr, const int32_t *idx, const float *data,
                       int bh, int bw, int br0, int br1, const float *Bf, int N,
                       double *C, double *Cabs, int nthreads);

/* ---- FP8 E4M3 (OCP "FN": bias 7, max 448, no infinity) -----------------
 * Extension of the reference (which is fp16 only): the quantiser of the FP8
 * storage mode.  Round to nearest even, finite overflow saturates to +-448.
 * Scales are powers of two stored as E8M0 bytes (2^(e - 127)): the smallest
 * 2^k with absmax / 2^k <= 448, 127 for an all-zero group.
 * pinned (tests/test_oracle_precision.py): all 256 codes decode as PyTorch's
 * float8_e4m3fn; encoding equals PyTorch's RNE cast on a 200k-value sweep
 * plus every tie between neighbouring codes and the subnormal boundaries;
 * saturation, -0 and NaN as special cases; E8M0 choice against its defining
 * inequality; group and B-slab quantisers against PyTorch's cast of each
 * group divided by its own scale. */
uint8_t orc_f2e4m3(float f);
float orc_e4m32f(uint8_t b);
uint8_t orc_e8m0_for_absmax(float absmax);
/* Quantise n groups of `len` floats, one E8M0 scale per group. */
void orc_quant_e4m3_groups(const float *x, int64_t n, int64_t len, uint8_t *q, uint8_t *scale);
/* Quantise B (K x N fp32, row-major, K % 16 == 0): one scale per (16-row
 * slab s, 256-column chunk h), the E8M0 exponent of that 16 x 256 patch's
 * absmax, at scale[h * ((K / 16 + 15) & ~15) + s]; q row-major like B. */
void orc_quant_e4m3_B(const float *B, int K, int N, uint8_t *q, uint8_t *scale);
/* The FP8 mode's A: each 16-row slice of a block-row is a sequence of
 * 16 x 16 device blocks (stored blocks in order, each cut into bw / 16
 * column pieces); consecutive device blocks (0, 1), (2, 3), ... share one
 * E8M0 scale (a lone last block has its own).  out = the dequantised values
 * in the source layout (nnzb row-major bh x bw blocks). */
void orc_quant_e4m3_tiles(const int32_t *hdr, const int32_t *idx, const float *data, int block_rows, int bh,
                          int bw, float *out);

/* The per-block FP32-scale form of FP8 storage (measured against the E8M0
 * tile form above; the kernels do not use it): per block of `blk` floats,
 * s = absmax / 448 in fp32, out = E4M3(x / s) * s (all zeros for a zero
 * block).  pinned (tests/test_oracle_precision.py): equals
 * orc_quant_e4m3_groups' result wherever s is a power of two, the block's
 * largest element maps to +-448 * s, and every value stays within E4M3's
 * rounding bound (2^-4 relative above 2^-6 s, 2^-10 s below). */
void orc_quant_e4m3_blocks_f32scale(const float *data, int64_t nblk, int64_t blk, float *out);

/* ---- other storage modes (extensions) ---------------------------------
 * bf16: float -> bfloat16 round to nearest even.  hi+lo: hi = fp16(a),
 * lo = fp16((a - hi) * 2^11), a ~ hi + 2^-11 lo.  pinned
 * (tests/test_oracle_precision.py): bf16 equals PyTorch's bfloat16 cast on a
 * wide sweep; hi equals numpy's float16 cast, lo numpy's cast of the exact
 * residual, and hi + 2^-11 lo reconstructs a to 2^-21 relative. */
uint16_t orc_f2bf16(float f);
float orc_bf162f(uint16_t b);
void orc_hilo_split(float a, uint16_t *hi, uint16_t *lo);
/* orc_f2bf16 / orc_hilo_split applied to n values. */
void orc_f2bf16_n(const float *x, int64_t n, uint16_t *out);
void orc_hilo_split_n(const float *x, int64_t n, uint16_t *hi, uint16_t *lo);

#ifdef __cplusplus
}
#endif
#endif
