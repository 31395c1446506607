/*
 * bcsr_mi355x.h -- C ABI of the MI355X BCSR x dense engine.
 *
 * C = A_bcsr (M x K, bh x bw blocks) . B (K x N, fp16, row-major) -> fp32
 * row-major M x N, every element of C written (empty block-rows give zeros).
 *
 * All functions return a bcsr_status (0 = ok, negative = error); none throws
 * and none synchronises the stream it is given.  `stream` is a hipStream_t
 * (NULL = the null stream).  Device pointers are HIP device addresses.
 *
 * Two ways in:
 *  - plans: bcsr_plan_create* from host BCSR arrays, then bcsr_spmm*;
 *  - drop-in: BCSRMatrix::copyToDevice (or bcsr_copy_to_device) returns
 *    device hdr/idx/data arrays that the sparseMatrixMulTensor* entry points
 *    below accept.  Those replace the reference's __global__ kernels of the
 *    same names (src/cuda_kernels.cuh:39-114): same arguments plus a trailing
 *    stream, and the library picks the launch geometry itself.
 */
#ifndef BCSR_MI355X_H
#define BCSR_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum bcsr_status {
    BCSR_OK = 0,
    BCSR_ERR_INVALID_ARG = -1,    /* null pointer, negative size, bad enum */
    BCSR_ERR_BLOCK_SIZE = -2,     /* block size < 16, not a multiple of 16, or not the entry point's */
    BCSR_ERR_SHAPE = -3,          /* M/N/K inconsistent with the plan, N % 16 != 0, misaligned B or C */
    BCSR_ERR_HIP = -4,            /* a HIP runtime call failed (bcsr_last_hip_error) */
    BCSR_ERR_OUT_OF_MEMORY = -5,  /* device or host allocation failed */
    BCSR_ERR_NOT_REGISTERED = -6, /* device hdr pointer not produced by bcsr_copy_to_device */
    BCSR_ERR_UNSUPPORTED = -7,    /* precision/variant not available for this plan */
    BCSR_ERR_UNKNOWN_PATTERN = -8, /* generator pattern name not recognised */
    BCSR_ERR_SCRATCH_LIMIT = -9    /* bcsr_spmm_noalloc: the product needs more than its fixed scratch */
};

/* Storage precision of A on the device.  B is fp16 and C fp32 in every mode. */
enum bcsr_precision {
    BCSR_PREC_F16 = 0,
    BCSR_PREC_BF16 = 1,
    BCSR_PREC_FP8_E4M3 = 2, /* per-block power-of-two (E8M0) scale; B quantised per call */
    BCSR_PREC_F16_HILO = 3  /* hi + lo fp16 split, near-fp32 accuracy */
};

typedef struct bcsr_plan bcsr_plan;

typedef struct {
    int32_t tile_n;   /* output columns per wave: 0 = auto, else 32, 64, 128 or 256 (FP8 plans
                       * with 64-row device blocks: up to 128; 256 gives BCSR_ERR_UNSUPPORTED) */
    int32_t xcd_map;  /* -1 = auto, 0 = linear block order, 1 = column tiles pinned to XCDs,
                       * 2 = block-rows pinned to XCDs (all tiles of a row on one XCD),
                       * 3 = tile-major work cut into 8 equal runs, one per XCD (any tile count),
                       * 4 = 2 x 4 grid: XCD x takes half the column tiles (x & 1) and a
                       *     quarter of the block-rows (x >> 1); even tile counts only */
    int32_t variant;  /* -1 = auto; 0 plain, 1 double-buffered, 2 64-bit addressing (16-bit modes);
                       * 3 = panel kernel: B slabs staged in LDS and shared by 16 block-rows
                       * (fp16 / bf16, 16-row device blocks, 128-column tiles only; never
                       * chosen automatically) */
    int32_t waves_per_block; /* 0 = auto; 1, 2 or 4 waves (independent work items) per workgroup */
    int32_t reserved[4];
} bcsr_launch_cfg;

typedef struct {
    int32_t rows, cols;             /* M, K */
    int32_t bh, bw;                 /* block shape of the source BCSR */
    int32_t block_rows, block_cols; /* rows / bh, cols / bw */
    int32_t precision;              /* bcsr_precision */
    int32_t dev_bh;                 /* rows per device block-row (16, 32 or 64) */
    int64_t nnzb;                   /* stored bh x bw blocks */
    int64_t dev_pairs;              /* device block pairs (K = 32 MFMA steps), padding included */
    int64_t device_bytes;           /* device memory held by the plan */
} bcsr_plan_info;

const char *bcsr_status_string(int status);
/* hipError_t of the last failing HIP call made by this library on this thread. */
int bcsr_last_hip_error(void);

/* ---- plans ---------------------------------------------------------------
 * Host arrays in the reference layout: hdr[block_rows+1], idx[nnzb],
 * data = nnzb row-major bh x bw blocks.  bcsr_plan_create takes fp16 data
 * (uint16 bit patterns) and builds an fp16 plan; bcsr_plan_create_f32 takes
 * fp32 block values and quantises them to `precision`. */
int bcsr_plan_create(const int32_t *hdr, const int32_t *idx, const void *data_f16,
                     int32_t block_rows, int32_t block_cols, int32_t bh, int32_t bw,
                     bcsr_plan **out);
int bcsr_plan_create_f32(const int32_t *hdr, const int32_t *idx, const float *data,
                         int32_t block_rows, int32_t block_cols, int32_t bh, int32_t bw,
                         int32_t precision, bcsr_plan **out);
int bcsr_plan_destroy(bcsr_plan *plan);
int bcsr_plan_get_info(const bcsr_plan *plan, bcsr_plan_info *info);

/* C (device, M x N fp32) = plan . B (device, K x N fp16).  N % 16 == 0,
 * B and C 16-byte aligned.  cfg may be NULL (automatic choice).  The plan's
 * device must be the current one (BCSR_ERR_INVALID_ARG otherwise).  Calls of
 * the 16-bit modes only read the plan, so they may overlap on any streams.
 * An FP8 plan quantises B into a workspace of its own (allocated by the
 * first call, grown by a larger N, the outgrown ones kept until
 * bcsr_plan_destroy), so the caller orders calls of one FP8 plan (one
 * stream, or events); bcsr_quantize_b_e4m3 + bcsr_spmm_b8 take the caller's
 * buffers instead. */
int bcsr_spmm(const bcsr_plan *plan, const void *B, float *C, int32_t N, void *stream);
int bcsr_spmm_cfg(const bcsr_plan *plan, const void *B, float *C, int32_t N,
                  const bcsr_launch_cfg *cfg, void *stream);
/* reps back-to-back launches of the same product on one stream (timing). */
int bcsr_spmm_repeat(const bcsr_plan *plan, const void *B, float *C, int32_t N,
                     const bcsr_launch_cfg *cfg, int32_t reps, void *stream);
/* Average time of `reps` launches after `warmup` untimed ones, measured with
 * HIP events on `stream` (synchronises the stream). */
int bcsr_spmm_time(const bcsr_plan *plan, const void *B, float *C, int32_t N,
                   const bcsr_launch_cfg *cfg, int32_t warmup, int32_t reps, void *stream,
                   float *ms);
/* The FP8 mode's B quantiser, on its own (one kernel, one pass over B):
 * B (device, K x N fp16) -> B8 (device, K * N bytes), stored k-minor: byte
 * (s * N + n) * 16 + r holds E4M3(B[16 s + r][n] / scale(s, n / 256)) -- and
 * scales (device, bcsr_b8_scale_bytes(K, N) bytes): the E8M0 exponent of each
 * (16-row slab s, 256-column chunk h), the smallest power of two with
 * |B| / scale <= 448 over that slab's 16 x 256 patch, at byte
 * h * ((K / 16 + 15) & ~15) + s.  work is unused (may be NULL).  B, B8 and
 * scales 16-byte aligned, K % 16 == 0, N % 16 == 0. */
int bcsr_quantize_b_e4m3(const void *B, int32_t K, int32_t N, void *B8, void *scales, void *work,
                         void *stream);
/* Bytes of bcsr_quantize_b_e4m3's scales for a K x N B (0 for K or N < 0). */
int64_t bcsr_b8_scale_bytes(int32_t K, int32_t N);
/* The FP8 product on a B quantised beforehand by bcsr_quantize_b_e4m3 (K =
 * the plan's columns, same N): only the product kernel runs, so a caller that
 * multiplies one B by several FP8 plans, or the same B more than once,
 * quantises it once.  FP8 plans only (BCSR_ERR_UNSUPPORTED otherwise); B8
 * and C 16-byte aligned, N % 16 == 0; cfg may be NULL.  C is identical to
 * bcsr_spmm's on the B the quantiser was given.  The scales 16-byte aligned
 * (BCSR_ERR_SHAPE otherwise).  No reference counterpart. */
int bcsr_spmm_b8(const bcsr_plan *plan, const void *B8, const void *scales, float *C, int32_t N,
                 const bcsr_launch_cfg *cfg, void *stream);
/* bcsr_spmm_cfg that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The 16-bit
 * modes launch exactly as bcsr_spmm_cfg (variant 3 only once an earlier
 * launch built its schedule; BCSR_ERR_UNSUPPORTED before).  An FP8 plan
 * quantises B into a fixed per-GPU scratch of BCSR_NOALLOC_SCRATCH_BYTES
 * (static device memory each GPU holds until the process ends) instead of
 * the plan's workspace: align256(K * N) + bcsr_b8_scale_bytes(K, N) bytes
 * must fit (K * N up to 16 MiB, e.g. 16384 x 1024), else
 * BCSR_ERR_SCRATCH_LIMIT before any launch.  Calls on
 * one GPU share that scratch: calls that overlap on one GPU (unordered
 * streams) race.  Results equal bcsr_spmm_cfg's bit for bit. */
#define BCSR_NOALLOC_SCRATCH_BYTES ((16u << 20) + (64u << 10))
int bcsr_spmm_noalloc(const bcsr_plan *plan, const void *B, float *C, int32_t N,
                      const bcsr_launch_cfg *cfg, void *stream);
/* Name of the kernel a launch with this configuration runs (for profiling). */
int bcsr_spmm_kernel_name(const bcsr_plan *plan, int32_t N, const bcsr_launch_cfg *cfg,
                          char *buf, size_t buflen);

/* ---- batches: many products in one persistent launch -------------------------
 * No reference counterpart (the reference launches one kernel per product).
 * bcsr_batch_create binds count products C_i = plans[i] . B_i (N[i] columns;
 * B_i, C_i device pointers as for bcsr_spmm) and builds the device work queue;
 * all plans share one precision and the current device.  tile_n[i] is the
 * output tile width of problem i (0 or tile_n == NULL: automatic); widths the
 * persistent kernel is not compiled for give BCSR_ERR_UNSUPPORTED
 * (bcsr_batch_tile_supported tells).  bcsr_batch_run enqueues the batch
 * (counter reset + persistent kernel; FP8 first quantises every B);
 * bcsr_batch_run_graph captures that sequence into a hipGraph on its first
 * call and replays it (stream must not be NULL).  Results equal the
 * per-matrix launches with the same tile widths bit for bit.  The batch keeps
 * the plans' device arrays and the B_i, C_i pointers: plans, B and C must
 * outlive it.  Runs of one batch share its queue counter (and FP8
 * workspaces), so they must be ordered (one stream, or events). */
typedef struct bcsr_batch bcsr_batch;
typedef struct {
    int32_t waves_per_cu; /* persistent waves per CU: 0 = 8 (FP8) or 6 (16-bit modes) */
    int32_t bin_work;     /* work per queue item, in 16 x 16 x 32 MFMA-tile steps: 0 = auto */
    int32_t queues;       /* 0 = auto, 1 = one global queue, 2 = one queue per XCD (column tiles
                           * pinned to XCDs, idle waves take the other queues' items) */
    int32_t reserved[5];
} bcsr_batch_cfg;
typedef struct {
    int32_t count, items, grid, precision;
    double flops;         /* sum over problems of 2 * nnzb * bh * bw * N */
    int64_t device_bytes; /* queue, descriptors and FP8 workspaces */
} bcsr_batch_info;
int bcsr_batch_tile_supported(int32_t precision, int32_t dev_bh, int32_t tile_n);
int bcsr_batch_create(const bcsr_plan *const *plans, const void *const *B, float *const *C,
                      const int32_t *N, const int32_t *tile_n, int32_t count,
                      const bcsr_batch_cfg *cfg, bcsr_batch **out);
int bcsr_batch_run(bcsr_batch *batch, void *stream);
int bcsr_batch_run_graph(bcsr_batch *batch, void *stream);
/* Average time of one batch run (graph != 0: graph replay) over reps runs
 * after warmup untimed ones, HIP events on `stream`. */
int bcsr_batch_time(bcsr_batch *batch, int32_t graph, int32_t warmup, int32_t reps, void *stream,
                    float *ms);
int bcsr_batch_get_info(const bcsr_batch *batch, bcsr_batch_info *info);
int bcsr_batch_tiles(const bcsr_batch *batch, int32_t *tiles);
int bcsr_batch_destroy(bcsr_batch *batch);

/* ---- learned dispatch: the cost model's MLP ------------------------------------
 * No reference counterpart.  A 3-layer perceptron in -> hidden -> hidden ->
 * out (GELU after the two hidden layers, fp32) trained on the GPU by this
 * library's own forward, backward and Adam kernels.  Parameters are one flat
 * fp32 array W1[hidden][in] b1 W2[hidden][hidden] b2 W3[out][hidden] b3;
 * create initialises W ~ U(-1/sqrt(fan_in), +) from mt19937(seed), b = 0.
 * Activations X [n][in], targets Y and mask M [n][out] are device arrays;
 * the loss is sum M (P - Y)^2 * inv_count (inv_count = 1 / sum M). */
typedef struct bcsr_costmodel bcsr_costmodel;
int bcsr_cm_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint32_t seed,
                   bcsr_costmodel **out);
int bcsr_cm_destroy(bcsr_costmodel *cm);
int bcsr_cm_num_params(const bcsr_costmodel *cm, int32_t *n);
int bcsr_cm_set_params(bcsr_costmodel *cm, const float *host, int32_t n); /* resets Adam */
int bcsr_cm_get_params(const bcsr_costmodel *cm, float *host, int32_t n);
int bcsr_cm_get_grads(const bcsr_costmodel *cm, float *host, int32_t n);
/* P [n][out] = model(X).  n (rows) is at most 2^24 per call in forward,
 * backward and train (BCSR_ERR_INVALID_ARG above). */
int bcsr_cm_forward(bcsr_costmodel *cm, const float *X, int32_t n, float *P, void *stream);
/* Gradient of the loss at the current parameters (read with get_grads);
 * *loss (optional) = the loss, which synchronises the stream. */
int bcsr_cm_backward(bcsr_costmodel *cm, const float *X, const float *Y, const float *M, int32_t n,
                     float inv_count, float *loss, void *stream);
/* steps full-batch Adam steps; *loss (optional) = loss before the last step. */
int bcsr_cm_train(bcsr_costmodel *cm, const float *X, const float *Y, const float *M, int32_t n,
                  float inv_count, int32_t steps, float lr, float beta1, float beta2, float eps,
                  float *loss, void *stream);

/* ---- drop-in device arrays -------------------------------------------------
 * Replaces BCSRMatrix::copyToDevice (formats/BCSRMatrix.cu:106-132): builds an
 * fp16 plan from host arrays and returns its device arrays.  *gpu_hdr is the
 * reference prefix array, *gpu_idx the block-column indices, *gpu_data the
 * blocks repacked into MFMA operand order. */
int bcsr_copy_to_device(const int32_t *hdr, const int32_t *idx, const void *data_f16,
                        int32_t block_rows, int32_t block_cols, int32_t bh, int32_t bw,
                        int **gpu_hdr, int **gpu_idx, void **gpu_data);
/* Frees what bcsr_copy_to_device allocated. */
int bcsr_release_device(const int *gpu_hdr);
/* The plan behind a registered device hdr pointer (owned by the registry). */
int bcsr_plan_from_device(const int *gpu_hdr, const bcsr_plan **out);

/* The reference launches its kernels on a stream chosen at the call site
 * (<<<grid, block, smem, stream>>>, or the null stream when omitted); the
 * entry points below take it as their last argument, which C++ callers may
 * leave out (the null stream), so a reference call site only drops the
 * launch configuration. */
#ifdef __cplusplus
#define BCSR_STREAM_ARG void *stream = nullptr
#else
#define BCSR_STREAM_ARG void *stream
#endif

/* ---- the reference's entry points ------------------------------------------
 * hdr/idx/data must come from bcsr_copy_to_device (BCSR_ERR_NOT_REGISTERED
 * otherwise); the block shape named by the entry point must match the
 * registered matrix (BCSR_ERR_BLOCK_SIZE otherwise).  M is the row count of
 * A and C, N the column count of B and C; the "1" variants take one n for
 * square M = N = n.  colBlocksPerRow of the Adaptive kernels is accepted and
 * ignored (the library balances work itself). */
#define BCSR_ENTRY(name) \
    int name(const int *hdr, const int *idx, const void *data, const void *B, float *C, \
             unsigned int M, unsigned int N, BCSR_STREAM_ARG)
BCSR_ENTRY(sparseMatrixMulTensor);             /* src/cuda_kernels.cu:301 16x16 */
BCSR_ENTRY(sparseMatrixMulTensor_v2);          /* src/cuda_kernels.cu:370 16x16 */
BCSR_ENTRY(sparseMatrixMulTensor_v3);          /* src/cuda_kernels.cu:423 16x16 */
BCSR_ENTRY(sparseMatrixMulTensor_v1_improved); /* src/cuda_kernels.cu:495 16x16 */
BCSR_ENTRY(sparseMatrixMulTensor_option2_ldmatrix_sm80); /* src/cuda_kernels.cu:552 16x16 */
BCSR_ENTRY(sparseMatrixMulTensor64x16);        /* src/cuda_kernels.cu:618 64x16 */
BCSR_ENTRY(sparseMatrixMulTensor64x16_v2);     /* src/cuda_kernels.cu:939 64x16 */
BCSR_ENTRY(sparseMatrixMulTensor32x16);        /* src/cuda_kernels.cu:699 32x16 */
BCSR_ENTRY(sparseMatrixMulTensor32x16_shared); /* src/cuda_kernels.cu:768 32x16 */
BCSR_ENTRY(sparseMatrixMulTensor32x16_shared_vectorized);    /* src/cuda_kernels.cu:853 */
BCSR_ENTRY(sparseMatrixMulTensor32x16_v2);     /* src/cuda_kernels.cu:1001 32x16 */
BCSR_ENTRY(sparseMatrixMulTensor32x16_v2_shared_vectorized); /* src/cuda_kernels.cu:1056 */
#undef BCSR_ENTRY
/* src/cuda_kernels.cu:330, square n x n */
int sparseMatrixMulTensor1(const int *hdr, const int *idx, const void *data, const void *B,
                           float *C, unsigned int n, BCSR_STREAM_ARG);
/* src/cuda_kernels.cu:1171, 1227, 1293 */
int sparseMatrixMulTensorAdaptive(const int *hdr, const int *idx, const void *data, const void *B,
                                  float *C, unsigned int M, unsigned int N,
                                  unsigned int colBlocksPerRow, BCSR_STREAM_ARG);
int sparseMatrixMulTensorAdaptive32x16(const int *hdr, const int *idx, const void *data,
                                       const void *B, float *C, unsigned int M, unsigned int N,
                                       unsigned int colBlocksPerRow, BCSR_STREAM_ARG);
int sparseMatrixMulTensorAdaptive64x16(const int *hdr, const int *idx, const void *data,
                                       const void *B, float *C, unsigned int M, unsigned int N,
                                       unsigned int colBlocksPerRow, BCSR_STREAM_ARG);

/* ---- CSR, hybrid and dense comparison ----------------------------------------
 * bcsr_csr_spmm: C (M x N fp32) = A_csr . B (accumulate 0), C = C + (A_csr . B)
 * (accumulate 1: the finished row sum added to C) or the running sum started
 * from C (accumulate 2: ((C + a0 b0) + a1 b1) + ...); hdr[M + 1], idx[nnz]
 * (column indices < K), data fp16, all device arrays; fp32 accumulation in
 * CSR order.  B, C 16-byte aligned, N % 16 == 0.  Replaces sparseMatrixMult1Co
 * and sparseMatrixMult1 (src/cuda_kernels.cu:271-299). */
int bcsr_csr_spmm(const int32_t *hdr, const int32_t *idx, const void *data_f16, const void *B, float *C,
                  int32_t M, int32_t K, int32_t N, int32_t accumulate, void *stream);
/* src/cuda_kernels.cu:271 (square n x n, overwrites C). */
int sparseMatrixMult1Co(const int *hdr, const int *idx, const void *data, const void *B, float *C,
                        unsigned int n, BCSR_STREAM_ARG);
/* src/cuda_kernels.cu:287 (square n x n, C += A . B in place, summed from C). */
int sparseMatrixMult1(const int *hdr, const int *idx, const void *data, const void *B, float *C,
                      unsigned int n, BCSR_STREAM_ARG);
/* C[i] += P[i], count floats; addMatrices (src/cuda_kernels.cu:357) takes n x n. */
int bcsr_add_into(float *C, const float *P, int64_t count, void *stream);
int addMatrices(float *C, const float *CPart, unsigned int n, BCSR_STREAM_ARG);
/* The hybrid (HCSR) product of the reference's legacy driver (main.cu:567-590):
 * C = plan . B (the dense blocks) then C += csr . B (the sparse remainder),
 * without a second C buffer. */
int bcsr_hybrid_spmm(const bcsr_plan *plan, const int32_t *csr_hdr, const int32_t *csr_idx,
                     const void *csr_data_f16, const void *B, float *C, int32_t N, void *stream);
/* Dense comparison row: C (M x N fp32) = A (M x K fp16) . B (K x N fp16),
 * row-major, rocBLAS gemm_ex with fp32 accumulation (the reference's cuBLAS
 * rows, nvbench_main.cu:1078-1180). */
int bcsr_dense_gemm_f16(const void *A, const void *B, float *C, int32_t M, int32_t K, int32_t N,
                        void *stream);

/* Library comparison rows: rocSPARSE generic SpMM, fp32 compute, B (K x N
 * fp16 row-major) into C (M x N fp32 row-major), on the plain BCSR arrays
 * (device hdr[block_rows + 1], idx[nnzb], data: nnzb square block_dim blocks,
 * each row-major, fp16) or on CSR arrays (hdr[M + 1], idx[nnz], data[nnz]).
 * The reference's counterpart is its cuSPARSE CSR row (main.cu:530-563).
 * create does the descriptors, the work buffer and the analysis (blocking);
 * run computes on a stream. */
typedef struct bcsr_sparse_ref bcsr_sparse_ref;
int bcsr_rocsparse_create(const int32_t *hdr, const int32_t *idx, const void *data_f16,
                          int32_t block_rows, int32_t block_cols, int64_t nnzb,
                          int32_t block_dim, const void *B, float *C, int32_t N,
                          bcsr_sparse_ref **out);
int bcsr_rocsparse_csr_create(const int32_t *hdr, const int32_t *idx, const void *data_f16,
                              int32_t M, int32_t K, int64_t nnz, const void *B, float *C,
                              int32_t N, bcsr_sparse_ref **out);
int bcsr_rocsparse_run(bcsr_sparse_ref *ref, void *stream);
int bcsr_rocsparse_destroy(bcsr_sparse_ref *ref);

/* ---- device plumbing ---------------------------------------------------------
 * Thin wrappers over the HIP runtime this library is linked against, so that
 * callers without their own HIP bindings (tests, bench, scripting languages)
 * can allocate, copy, order and time work on the same runtime. */
int bcsr_device_count(int32_t *count);
int bcsr_set_device(int32_t device);
int bcsr_dev_malloc(void **ptr, size_t bytes);
int bcsr_dev_free(void *ptr);
int bcsr_memcpy_htod(void *dst, const void *src, size_t bytes);
int bcsr_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int bcsr_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int bcsr_memset_async(void *dst, int32_t byte_value, size_t bytes, void *stream);
/* Device-to-device copy by a plain streaming kernel (bcsr_stream_copy);
 * bytes and both pointers must be multiples of 16.  Also the known-size
 * reference for hardware byte counters. */
int bcsr_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);
int bcsr_stream_create(void **stream);
int bcsr_stream_destroy(void *stream);
int bcsr_stream_synchronize(void *stream);
int bcsr_device_synchronize(void);
int bcsr_event_create(void **event);
int bcsr_event_destroy(void *event);
int bcsr_event_record(void *event, void *stream);
int bcsr_event_elapsed_ms(void *start, void *end, float *ms);
/* HIP graphs of this library's launches (no reference counterpart: the
 * reference launches every kernel from the host).  bcsr_graph_begin starts
 * capturing `stream` (not NULL; thread-local capture mode); the library
 * calls that follow on it are recorded, not run; bcsr_graph_end stops and
 * instantiates, giving an executable graph that bcsr_graph_launch replays on
 * any stream.  The library's launches record kernel nodes only (no memset
 * nodes).  An FP8 plan call allocates its workspace when it needs a larger
 * one than the plan holds, which a capture cannot record
 * (BCSR_ERR_UNSUPPORTED): call the plan once at that N before capturing.
 * The captured workspace stays valid until bcsr_plan_destroy. */
int bcsr_graph_begin(void *stream);
int bcsr_graph_end(void *stream, void **graph_exec);
int bcsr_graph_launch(void *graph_exec, void *stream);
int bcsr_graph_destroy(void *graph_exec);

/* ---- host-side generator and formats ----------------------------------------
 * mg::generate_matrix<T> (matrix_generator.cpp:28-140) written row-major as
 * double; elem_type 0 float, 1 double, 2 int, 3 long long. */
int mg_generate_matrix(int32_t rows, int32_t cols, double sparsity, const char *pattern,
                       int32_t blocksize, uint32_t seed, int32_t elem_type, double *out);
/* Extended float generator (mg::generate_into): bh x bw blocks, band
 * half-width, in-block density, value_dist 0 uniform / 1 log-uniform on
 * [lo, hi].  _f16 rounds double -> float -> half. */
int mg_generate_f32(int32_t rows, int32_t cols, double sparsity, const char *pattern, int32_t bh,
                    int32_t bw, int32_t band_halfwidth, double inblock_density,
                    int32_t value_dist, double lo, double hi, uint32_t seed, float *out);
int mg_generate_f16(int32_t rows, int32_t cols, double sparsity, const char *pattern, int32_t bh,
                    int32_t bw, int32_t band_halfwidth, double inblock_density,
                    int32_t value_dist, double lo, double hi, uint32_t seed, void *out);
/* mg::generate_bcsr: the BCSR form (bh x bw blocks) of mg_generate_f16's
 * matrix, any pattern, generated one band of block-rows at a time (no dense
 * matrix).  Rejects rows % bh or cols % bw != 0.  hdr must hold
 * rows / bh + 1 entries; *idx and *data (nnzb and nnzb * bh * bw floats) are
 * allocated by the library and released with mg_free. */
int mg_generate_bcsr(int32_t rows, int32_t cols, double sparsity, const char *pattern, int32_t bh,
                     int32_t bw, int32_t band_halfwidth, double inblock_density, int32_t value_dist,
                     double lo, double hi, uint32_t seed, int32_t *hdr, int64_t *nnzb,
                     int32_t **idx, float **data);
void mg_free(void *p);
/* BCSRMatrix construction from dense fp16 (formats/BCSRMatrix.cu:22-74) in
 * two calls: count (fills hdr[rows/bh + 1], returns nnzb) then fill. */
int bcsr_host_count(const void *dense_f16, int32_t rows, int32_t cols, int32_t bh, int32_t bw,
                    int32_t *hdr, int64_t *nnzb);
int bcsr_host_fill(const void *dense_f16, int32_t rows, int32_t cols, int32_t bh, int32_t bw,
                   const int32_t *hdr, int32_t *idx, void *data_f16);
/* The ".mat" text format of Matrix(const string&) (Matrix.cu:14-32):
 * "rows cols" then rows*cols values.  mat_read_header reads the shape,
 * mat_read the fp16 values (and the count of values nonzero as floats),
 * mat_write writes the format back.  BCSR_ERR_INVALID_ARG on a missing or
 * malformed file. */
int mat_read_header(const char *path, int32_t *rows, int32_t *cols);
int mat_read(const char *path, int32_t rows, int32_t cols, void *out_f16, int32_t *nonzeros);
int mat_write(const char *path, int32_t rows, int32_t cols, const void *data_f16);

#ifdef __cplusplus
}
#endif
#endif
