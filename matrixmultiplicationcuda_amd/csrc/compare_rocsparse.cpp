// compare_rocsparse.cpp -- the vendor-library comparison rows: rocSPARSE's
// generic SpMM on a BSR matrix or a CSR matrix (fp16 A and B, fp32 C and
// compute), the library counterparts of this library's BCSR and CSR products
// (the reference times cuSPARSE CSR beside its kernels, main.cu:530-563).
// They are comparisons only: nothing in the library dispatches to them.
//
// The BSR matrix is the plain BCSR layout (prefix hdr, block-column idx,
// each square block row-major), so rocSPARSE reads it with
// rocsparse_direction_row.  Descriptors, the work buffer and the analysis are
// made once at creation; bcsr_rocsparse_run only computes.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cstdint>
#include <new>

#include "bcsr_mi355x.h"

struct bcsr_sparse_ref {
    rocsparse_handle handle = nullptr;
    rocsparse_spmat_descr A = nullptr;
    rocsparse_dnmat_descr B = nullptr, C = nullptr;
    rocsparse_spmm_alg alg = rocsparse_spmm_alg_default;
    void *buffer = nullptr;
    size_t buffer_bytes = 0;
};

namespace {

int sp_status(rocsparse_status s) {
    switch (s) {
    case rocsparse_status_success:
        return BCSR_OK;
    case rocsparse_status_memory_error:
        return BCSR_ERR_OUT_OF_MEMORY;
    case rocsparse_status_not_implemented:
        return BCSR_ERR_UNSUPPORTED;
    case rocsparse_status_invalid_size:
        return BCSR_ERR_SHAPE;
    default:
        return BCSR_ERR_HIP;
    }
}

void release(bcsr_sparse_ref *r) {
    if (!r) return;
    if (r->A) rocsparse_destroy_spmat_descr(r->A);
    if (r->B) rocsparse_destroy_dnmat_descr(r->B);
    if (r->C) rocsparse_destroy_dnmat_descr(r->C);
    if (r->buffer) (void)hipFree(r->buffer);
    if (r->handle) rocsparse_destroy_handle(r->handle);
    delete r;
}

int stage(bcsr_sparse_ref *r, rocsparse_spmm_stage st, size_t *bytes) {
    const float alpha = 1.0f, beta = 0.0f;
    return sp_status(rocsparse_spmm(r->handle, rocsparse_operation_none, rocsparse_operation_none, &alpha, r->A, r->B,
                                    &beta, r->C, rocsparse_datatype_f32_r, r->alg, st, bytes,
                                    r->buffer));
}

// r->handle and r->A exist; adds B, C, the buffer and the analysis
int finish(bcsr_sparse_ref *r, int64_t M, int64_t K, const void *B, float *C, int32_t N) {
    int rc = sp_status(rocsparse_create_dnmat_descr(&r->B, K, N, N, const_cast<void *>(B), rocsparse_datatype_f16_r,
                                                    rocsparse_order_row));
    if (rc == BCSR_OK)
        rc = sp_status(
            rocsparse_create_dnmat_descr(&r->C, M, N, N, C, rocsparse_datatype_f32_r, rocsparse_order_row));
    if (rc == BCSR_OK) rc = stage(r, rocsparse_spmm_stage_buffer_size, &r->buffer_bytes);
    if (rc == BCSR_OK && hipMalloc(&r->buffer, r->buffer_bytes ? r->buffer_bytes : 16) != hipSuccess) {
        r->buffer = nullptr;
        rc = BCSR_ERR_OUT_OF_MEMORY;
    }
    if (rc == BCSR_OK) rc = stage(r, rocsparse_spmm_stage_preprocess, &r->buffer_bytes);
    return rc;
}

}  // namespace

extern "C" int bcsr_rocsparse_create(const int32_t *hdr, const int32_t *idx, const void *data_f16, int32_t block_rows,
                                     int32_t block_cols, int64_t nnzb, int32_t block_dim, const void *B, float *C,
                                     int32_t N, bcsr_sparse_ref **out) {
    if (!out) return BCSR_ERR_INVALID_ARG;
    *out = nullptr;
    if (block_rows <= 0 || block_cols <= 0 || nnzb < 0 || N <= 0 || !hdr || !B || !C) return BCSR_ERR_INVALID_ARG;
    if (nnzb && (!idx || !data_f16)) return BCSR_ERR_INVALID_ARG;
    if (block_dim <= 0) return BCSR_ERR_BLOCK_SIZE;
    bcsr_sparse_ref *r = new (std::nothrow) bcsr_sparse_ref;
    if (!r) return BCSR_ERR_OUT_OF_MEMORY;
    r->alg = rocsparse_spmm_alg_bsr;
    const int64_t M = static_cast<int64_t>(block_rows) * block_dim, K = static_cast<int64_t>(block_cols) * block_dim;
    int rc = sp_status(rocsparse_create_handle(&r->handle));
    if (rc == BCSR_OK)
        rc = sp_status(rocsparse_create_bsr_descr(
            &r->A, block_rows, block_cols, nnzb, rocsparse_direction_row, block_dim, const_cast<int32_t *>(hdr),
            const_cast<int32_t *>(idx), const_cast<void *>(data_f16), rocsparse_indextype_i32, rocsparse_indextype_i32,
            rocsparse_index_base_zero, rocsparse_datatype_f16_r));
    if (rc == BCSR_OK) rc = finish(r, M, K, B, C, N);
    if (rc != BCSR_OK) {
        release(r);
        return rc;
    }
    *out = r;
    return BCSR_OK;
}

extern "C" int bcsr_rocsparse_csr_create(const int32_t *hdr, const int32_t *idx, const void *data_f16, int32_t M,
                                         int32_t K, int64_t nnz, const void *B, float *C, int32_t N,
                                         bcsr_sparse_ref **out) {
    if (!out) return BCSR_ERR_INVALID_ARG;
    *out = nullptr;
    if (M <= 0 || K <= 0 || nnz < 0 || N <= 0 || !hdr || !B || !C) return BCSR_ERR_INVALID_ARG;
    if (nnz && (!idx || !data_f16)) return BCSR_ERR_INVALID_ARG;
    bcsr_sparse_ref *r = new (std::nothrow) bcsr_sparse_ref;
    if (!r) return BCSR_ERR_OUT_OF_MEMORY;
    r->alg = rocsparse_spmm_alg_csr;
    int rc = sp_status(rocsparse_create_handle(&r->handle));
    if (rc == BCSR_OK)
        rc = sp_status(rocsparse_create_csr_descr(&r->A, M, K, nnz, const_cast<int32_t *>(hdr),
                                                  const_cast<int32_t *>(idx), const_cast<void *>(data_f16),
                                                  rocsparse_indextype_i32, rocsparse_indextype_i32,
                                                  rocsparse_index_base_zero, rocsparse_datatype_f16_r));
    if (rc == BCSR_OK) rc = finish(r, M, K, B, C, N);
    if (rc != BCSR_OK) {
        release(r);
        return rc;
    }
    *out = r;
    return BCSR_OK;
}

extern "C" int bcsr_rocsparse_run(bcsr_sparse_ref *r, void *stream) {
    if (!r) return BCSR_ERR_INVALID_ARG;
    int rc = sp_status(rocsparse_set_stream(r->handle, static_cast<hipStream_t>(stream)));
    if (rc != BCSR_OK) return rc;
    return stage(r, rocsparse_spmm_stage_compute, &r->buffer_bytes);
}

extern "C" int bcsr_rocsparse_destroy(bcsr_sparse_ref *r) {
    release(r);
    return BCSR_OK;
}
