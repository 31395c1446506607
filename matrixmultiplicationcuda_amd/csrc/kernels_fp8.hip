// kernels_fp8.hip -- FP8 E4M3 storage mode (no reference counterpart: the
// reference's kernels are fp16 only): the B quantiser and the per-matrix
// product kernels.  The device code is in fp8_device.h.
#include <algorithm>
#include <cstdio>
#include <new>

#include "fp8_device.h"

namespace {

using namespace fp8dev;

// B quantiser, grid (ceil(slabs / 2), chunks) x 128 threads: each slab's
// scale from its own absmax, then its quantised records.
__global__ __launch_bounds__(128) void quant_b_fp8(const _Float16 *__restrict__ B, uint8_t *__restrict__ B8,
                                                   uint8_t *__restrict__ scale, int N, int slabs) {
    quant_slab(B, B8, scale, N, slabs, blockIdx.x, blockIdx.y);
}

}  // namespace

size_t b8_scale_bytes(int K, int N) {
    return static_cast<size_t>(scale_stride(K / 16)) * static_cast<size_t>((N + kChunk - 1) / kChunk);
}

int enqueue_quant(const void *B, int K, int N, uint8_t *B8, uint8_t *scales, hipStream_t stream) {
    const int slabs = K / 16, chunks = (N + kChunk - 1) / kChunk;
    const dim3 grid((slabs + 1) / 2, chunks);
    hipLaunchKernelGGL(quant_b_fp8, grid, dim3(128), 0, stream, static_cast<const _Float16 *>(B), B8, scales, N,
                       slabs);
    return hip_status(hipGetLastError());
}

namespace {

template <int RS, int EPL>
__global__ __launch_bounds__(256) void bcsr_spmm_fp8(SpmmArgs a) {
    int tile, slot;
    if (!map_work(a, tile, slot)) return;
    fp8_work<RS, EPL>(a, a.row_order[slot], tile);
}

typedef void (*Kernel)(SpmmArgs);

template <int RS>
Kernel pick_epl(int epl) {
    switch (epl) {
        case 2: return bcsr_spmm_fp8<RS, 2>;
        case 4: return bcsr_spmm_fp8<RS, 4>;
        case 8: return bcsr_spmm_fp8<RS, 8>;
        case 16:
            if constexpr (fp8_shape_supported(RS, 16)) return bcsr_spmm_fp8<RS, 16>;
            return nullptr;
        default: return nullptr;
    }
}

Kernel pick_kernel(int rs, int epl) {
    switch (rs) {
        case 1: return pick_epl<1>(epl);
        case 2: return pick_epl<2>(epl);
        case 4: return pick_epl<4>(epl);
        default: return nullptr;
    }
}

}  // namespace

void fp8_kernel_name(const bcsr_plan *plan, const LaunchShape &ls, char *buf, size_t len) {
    std::snprintf(buf, len, "bcsr_spmm_fp8<%d, %d>", plan->rs, ls.epl);
}

int launch_fp8_product(const bcsr_plan *plan, const LaunchShape &ls, const uint8_t *B8, const uint8_t *bscale,
                       float *C, int N, hipStream_t stream) {
    Kernel k = pick_kernel(plan->rs, ls.epl);
    if (!k) return BCSR_ERR_UNSUPPORTED;
    SpmmArgs a{};
    a.grp_ptr = plan->d_grp_ptr;
    a.grp_col = plan->d_grp_col;
    a.vals = static_cast<const u32x4 *>(plan->d_vals);
    a.ascale = plan->d_ascale;
    a.row_order = ls.map_mode == 3 ? plan->d_mix_order : plan->d_row_order;
    a.B8 = B8;
    a.b8scale = bscale;
    a.C = C;
    a.N = N;
    a.n_tiles = ls.n_tiles;
    a.dev_rows = plan->dev_rows;
    a.map_mode = ls.map_mode;
    a.b_bytes = static_cast<int>(static_cast<size_t>(plan->info.cols) * N);
    a.b_chunks = (N + kChunk - 1) / kChunk;
    a.b_slabs = plan->info.cols / 16;
    a.a_bytes = static_cast<int>(std::min<int64_t>(plan->info.dev_pairs * plan->rs * 2048, INT32_MAX));
    a.wpb = ls.wpb;
    hipLaunchKernelGGL(k, dim3(ls.grid), dim3(64 * ls.wpb), 0, stream, a);
    return hip_status(hipGetLastError());
}

int launch_fp8(bcsr_plan *plan, const LaunchShape &ls, const void *B, float *C, int N, hipStream_t stream) {
    if (!pick_kernel(plan->rs, ls.epl)) return BCSR_ERR_UNSUPPORTED;
    const int K = plan->info.cols;
    const int slabs = K / 16;
    const size_t q_bytes = static_cast<size_t>(K) * N;
    const size_t off_s = (q_bytes + 255) / 256 * 256;
    const size_t need = off_s + b8_scale_bytes(K, N);
    std::lock_guard<std::mutex> lk(*plan->ws_mu);
    if (plan->ws_bytes < need) {
        // an allocation cannot be recorded into a graph under capture
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        int rc0 = hip_status(hipStreamIsCapturing(stream, &cap));
        if (rc0 != BCSR_OK) return rc0;
        if (cap != hipStreamCaptureStatusNone) return BCSR_ERR_UNSUPPORTED;
        // grow at least 2x, so a plan called with rising N keeps few buffers
        const size_t bytes = std::max(need, 2 * plan->ws_bytes);
        void *ws = nullptr;
        int rc = hip_status(hipMalloc(&ws, bytes));
        if (rc != BCSR_OK) return rc;
        if (plan->d_ws) {
            try {
                plan->ws_retired.push_back(plan->d_ws);
            } catch (const std::bad_alloc &) {
                (void)hipFree(ws);
                return BCSR_ERR_OUT_OF_MEMORY;
            }
        }
        plan->d_ws = ws;
        plan->ws_bytes = bytes;
    }
    uint8_t *B8 = static_cast<uint8_t *>(plan->d_ws);
    uint8_t *bscale = B8 + off_s;
    // K = 0: B has no rows to quantise; the product kernel writes C = 0
    // without reading B or its scales
    if (slabs > 0) {
        int rc = enqueue_quant(B, K, N, B8, bscale, stream);
        if (rc != BCSR_OK) return rc;
    }
    return launch_fp8_product(plan, ls, B8, bscale, C, N, stream);
}

extern "C" int bcsr_quantize_b_e4m3(const void *B, int32_t K, int32_t N, void *B8, void *scales, void *work,
                                    void *stream) {
    (void)work;  // the one-pass quantiser needs no scratch
    if (K < 0 || N < 0 || K % 16 || N % 16) return BCSR_ERR_SHAPE;
    if (!K || !N) return BCSR_OK;
    if (!B || !B8 || !scales) return BCSR_ERR_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(B8) & 15u) || (reinterpret_cast<uintptr_t>(B) & 15u) ||
        (reinterpret_cast<uintptr_t>(scales) & 15u))
        return BCSR_ERR_SHAPE;
    return enqueue_quant(B, K, N, static_cast<uint8_t *>(B8), static_cast<uint8_t *>(scales),
                         static_cast<hipStream_t>(stream));
}

extern "C" int64_t bcsr_b8_scale_bytes(int32_t K, int32_t N) {
    if (K < 0 || N < 0) return 0;
    return static_cast<int64_t>(b8_scale_bytes(K, N));
}
