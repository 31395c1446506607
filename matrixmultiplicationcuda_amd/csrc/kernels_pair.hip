// kernels_pair.hip -- BCSR x dense kernels on the 16-bit MFMA
// (v_mfma_f32_16x16x32_{f16,bf16}) for the pair layout: fp16 storage (the
// reference's sparseMatrixMulTensor* computation, src/cuda_kernels.cu), bf16
// storage, and hi+lo fp16 storage.  The device code is in pair_device.h.
#include <algorithm>
#include <cstdio>

#include "pair_device.h"

namespace {

using namespace pairdev;

template <int RS, int EPL, bool PF, int PREC>
__global__ __launch_bounds__(256) void bcsr_spmm_pair(SpmmArgs a) {
    int tile, slot;
    if (!map_work(a, tile, slot)) return;
    pair_work<RS, EPL, PF, PREC>(a, a.row_order[slot], tile);
}

// 64-bit global-load kernel (variant 2): any B size.
template <int RS, int EPL, int PREC>
__global__ __launch_bounds__(256) void bcsr_spmm_pair_global(SpmmArgs a) {
    int tile, slot;
    if (!map_work(a, tile, slot)) return;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int br = a.row_order[slot];
    const int n0 = tile * (16 * EPL) + EPL * j;
    const bool nvalid = n0 < a.N;
    const size_t N = static_cast<size_t>(a.N);
    const _Float16 *Bcol = a.B + (nvalid ? n0 : 0) + static_cast<size_t>((g & 1) * 8) * N;
    const bool second = g >= 2;
    constexpr int H = AOps<RS, PREC>::H;
    Acc<RS, EPL, PREC> acc;
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int s = 0; s < RS; ++s)
#pragma unroll
            for (int e = 0; e < EPL; ++e) acc.v[h][s][e] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int p0 = a.grp_ptr[br], p1 = a.grp_ptr[br + 1];
    const u32x4 *vals = a.vals + lane;
    for (int p = p0; p < p1; ++p) {
        const int c = second ? a.grp_col[2 * p + 1] : a.grp_col[2 * p];
        BRow<EPL> raw[8];
        load_patch_global<EPL>(Bcol + static_cast<size_t>(c) * 16 * N, N, raw);
        AOps<RS, PREC> av;
#pragma unroll
        for (int s = 0; s < RS; ++s)
#pragma unroll
            for (int h = 0; h < H; ++h) av.v[s][h] = vals[((static_cast<size_t>(p) * RS + s) * H + h) * 64];
        consume<RS, EPL, PREC>(acc, raw, av);
    }
    if (!nvalid) return;
    store_tile<RS, EPL, PREC>(acc, a.C, N, br, g, n0);
}

typedef void (*Kernel)(SpmmArgs);

// Instantiated shapes: 256-column tiles only with 16-row blocks, and the
// prefetching kernel only while its two load sets and accumulators fit the
// register file (RS * EPL * H <= 16).
template <int RS, int EPL, int PREC>
constexpr bool shape_ok(int variant) {
    constexpr int H = PREC == PREC_HILO ? 2 : 1;
    if (EPL == 16 && (RS != 1 || H != 1)) return false;
    if (variant == 1) return RS * EPL * H <= 16;
    return true;
}

template <int RS, int EPL, int PREC>
Kernel pick_variant(int variant) {
    if constexpr (EPL == 16 && (RS != 1 || PREC == PREC_HILO)) {
        return nullptr;
    } else {
        switch (variant) {
            case 0: return bcsr_spmm_pair<RS, EPL, false, PREC>;
            case 1:
                if constexpr (shape_ok<RS, EPL, PREC>(1)) return bcsr_spmm_pair<RS, EPL, true, PREC>;
                return nullptr;
            case 2: return bcsr_spmm_pair_global<RS, EPL, PREC>;
            default: return nullptr;
        }
    }
}

template <int RS, int PREC>
Kernel pick_epl(int epl, int variant) {
    switch (epl) {
        case 2: return pick_variant<RS, 2, PREC>(variant);
        case 4: return pick_variant<RS, 4, PREC>(variant);
        case 8: return pick_variant<RS, 8, PREC>(variant);
        case 16: return pick_variant<RS, 16, PREC>(variant);
        default: return nullptr;
    }
}

template <int PREC>
Kernel pick_rs(int rs, int epl, int variant) {
    switch (rs) {
        case 1: return pick_epl<1, PREC>(epl, variant);
        case 2: return pick_epl<2, PREC>(epl, variant);
        case 4: return pick_epl<4, PREC>(epl, variant);
        default: return nullptr;
    }
}

Kernel pick_kernel(int precision, int rs, int epl, int variant) {
    switch (precision) {
        case BCSR_PREC_F16: return pick_rs<PREC_F16>(rs, epl, variant);
        case BCSR_PREC_BF16: return pick_rs<PREC_BF16>(rs, epl, variant);
        case BCSR_PREC_F16_HILO: return pick_rs<PREC_HILO>(rs, epl, variant);
        default: return nullptr;
    }
}

const char *prec_tag(int precision) {
    switch (precision) {
        case BCSR_PREC_BF16: return "bf16";
        case BCSR_PREC_F16_HILO: return "hilo";
        default: return "f16";
    }
}

}  // namespace

bool pair_shape_supported(int precision, int rs, int epl, int variant) {
    return pick_kernel(precision, rs, epl, variant) != nullptr;
}

void pair_kernel_name(const bcsr_plan *plan, const LaunchShape &ls, char *buf, size_t len) {
    const char *tag = prec_tag(plan->info.precision);
    if (ls.variant == 2)
        std::snprintf(buf, len, "bcsr_spmm_pair_global<%d, %d, %s>", plan->rs, ls.epl, tag);
    else
        std::snprintf(buf, len, "bcsr_spmm_pair<%d, %d, %s, %s>", plan->rs, ls.epl, ls.variant == 1 ? "true" : "false",
                      tag);
}

int launch_pairs(const bcsr_plan *plan, const LaunchShape &ls, const void *B, float *C, int N, hipStream_t stream) {
    Kernel k = pick_kernel(plan->info.precision, plan->rs, ls.epl, ls.variant);
    if (!k) return BCSR_ERR_UNSUPPORTED;
    SpmmArgs a{};
    a.grp_ptr = plan->d_grp_ptr;
    a.grp_col = plan->d_grp_col;
    a.vals = static_cast<const u32x4 *>(plan->d_vals);
    a.row_order = ls.map_mode == 3 ? plan->d_mix_order : plan->d_row_order;
    a.B = static_cast<const _Float16 *>(B);
    a.C = C;
    a.N = N;
    a.n_tiles = ls.n_tiles;
    a.dev_rows = plan->dev_rows;
    a.map_mode = ls.map_mode;
    a.b_bytes = static_cast<int>(std::min<int64_t>(static_cast<int64_t>(plan->info.cols) * N * 2, INT32_MAX));
    a.wpb = ls.wpb;
    hipLaunchKernelGGL(k, dim3(ls.grid), dim3(64 * ls.wpb), 0, stream, a);
    return hip_status(hipGetLastError());
}
