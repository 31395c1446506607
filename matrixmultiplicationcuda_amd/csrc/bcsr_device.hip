// bcsr_device.hip -- device layout, gfx950 MFMA kernels and the C ABI of the
// BCSR x dense engine (replaces the sparseMatrixMulTensor* kernels of
// src/cuda_kernels.cu and BCSRMatrix::copyToDevice, formats/BCSRMatrix.cu:106-132).
//
// Device layout of a plan
//   A bh x bw BCSR matrix is viewed as dev_bh x 16 blocks (dev_bh = 64, 32 or
//   16, the largest that divides bh; bw is cut into 16-column pieces).  Each
//   device block-row's blocks (ascending block-column) are taken two at a
//   time: a pair is one K = 32 step of v_mfma_f32_16x16x32_f16.  An odd
//   block-row is padded with a zero block that repeats the last block-column.
//     pair_ptr[dev_rows + 1]  prefix count of pairs per device block-row
//     pair_col[2 * pairs]     the two 16-column block indices of each pair
//     vals                    pairs x (dev_bh / 16) x 64 lanes x 8 halves: lane
//                             l of row-subtile s holds A[16 s + l % 16][8 (l / 16) ..
//                             + 7] of the pair's 16 x 32 slice, i.e. exactly its
//                             MFMA A operand, so one 16-byte load per lane
//                             fetches it and a wave reads 1 KB contiguously.
//     row_order[dev_rows]     device block-rows by descending pair count.
//
// Kernel (one 64-lane wave per (device block-row, column tile of 16*EPL))
//   Lane l = 16 g + j loads, for each pair, 8 consecutive B rows (rows 8 (g % 2)
//   .. + 7 of the slab of block g / 2) at columns n0 = tile + EPL j .. + EPL - 1,
//   i.e. an 8 x EPL patch, and transposes it in registers into EPL MFMA B
//   operands (8 consecutive k at one column each).  MFMA number e of the pair
//   then computes output columns tile + EPL j + e for j = 0..15, so no LDS is
//   needed and every lane ends with EPL consecutive fp32 outputs per row:
//   the C tile is written with 16-byte stores, EPL * 4 bytes per lane per row.
//   B is read through L2: with xcd_map = 1 every column tile is pinned to one
//   XCD (workgroups are dispatched round-robin over the 8 XCDs), so each
//   XCD's L2 holds only its K x 16*EPL slice of B.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "BCSRMatrix.h"
#include "CSRMatrix.h"
#include "bcsr_mi355x.h"

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

namespace {

thread_local int g_last_hip_error = 0;

int hip_status(hipError_t e) {
    if (e == hipSuccess) return BCSR_OK;
    g_last_hip_error = static_cast<int>(e);
    return e == hipErrorOutOfMemory ? BCSR_ERR_OUT_OF_MEMORY : BCSR_ERR_HIP;
}

#define BCSR_HIP_TRY(expr)                              \
    do {                                                \
        int rc_ = hip_status(expr);                     \
        if (rc_ != BCSR_OK) return rc_;                 \
    } while (0)

// ---------------------------------------------------------------------------
// Kernel
// ---------------------------------------------------------------------------
struct SpmmArgs {
    const int32_t *pair_ptr;
    const int32_t *pair_col;
    const u32x4 *vals;
    const int32_t *row_order;
    const _Float16 *B;
    float *C;
    int N;
    int n_tiles;
    int dev_rows;
    int map_mode;
};

// Work item of workgroup `bid`: (column tile, slot in row_order).
__device__ __forceinline__ bool map_work(const SpmmArgs &a, int bid, int &tile, int &slot) {
    if (a.map_mode == 0) {
        tile = bid % a.n_tiles;
        slot = bid / a.n_tiles;
    } else {
        const int x = bid & 7, i = bid >> 3;
        if (a.n_tiles >= 8) {  // n_tiles % 8 == 0: XCD x owns tiles x, x + 8, ...
            const int per = a.n_tiles >> 3;
            tile = x + 8 * (i % per);
            slot = i / per;
        } else {  // n_tiles in {1, 2, 4}: 8 / n_tiles XCDs share each tile
            const int q = 8 / a.n_tiles;
            tile = x % a.n_tiles;
            slot = i * q + x / a.n_tiles;
        }
    }
    return slot < a.dev_rows;
}

// EPL consecutive halves of one B row.
template <int EPL>
struct BRow {
    unsigned int d[EPL / 2];
};

template <int EPL>
__device__ __forceinline__ BRow<EPL> load_brow(const _Float16 *p) {
    BRow<EPL> r;
    if constexpr (EPL == 2) {
        r.d[0] = *reinterpret_cast<const unsigned int *>(p);
    } else if constexpr (EPL == 4) {
        u32x2 v = *reinterpret_cast<const u32x2 *>(p);
        r.d[0] = v.x;
        r.d[1] = v.y;
    } else {
#pragma unroll
        for (int h = 0; h < EPL / 8; ++h) {
            u32x4 v = reinterpret_cast<const u32x4 *>(p)[h];
            r.d[4 * h + 0] = v.x;
            r.d[4 * h + 1] = v.y;
            r.d[4 * h + 2] = v.z;
            r.d[4 * h + 3] = v.w;
        }
    }
    return r;
}

template <int EPL>
__device__ __forceinline__ BRow<EPL> zero_brow() {
    BRow<EPL> r;
#pragma unroll
    for (int q = 0; q < EPL / 2; ++q) r.d[q] = 0u;
    return r;
}

// Column e of an 8 x EPL patch (rows r = 0..7) as an MFMA B operand: the
// 8 halves patch[0..7][e] packed two per dword.
template <int EPL>
__device__ __forceinline__ f16x8 patch_column(const BRow<EPL> (&raw)[8], int e) {
    u32x4 v;
    const int w = e >> 1;
    if (e & 1) {
        v.x = (raw[0].d[w] >> 16) | (raw[1].d[w] & 0xffff0000u);
        v.y = (raw[2].d[w] >> 16) | (raw[3].d[w] & 0xffff0000u);
        v.z = (raw[4].d[w] >> 16) | (raw[5].d[w] & 0xffff0000u);
        v.w = (raw[6].d[w] >> 16) | (raw[7].d[w] & 0xffff0000u);
    } else {
        v.x = (raw[0].d[w] & 0xffffu) | (raw[1].d[w] << 16);
        v.y = (raw[2].d[w] & 0xffffu) | (raw[3].d[w] << 16);
        v.z = (raw[4].d[w] & 0xffffu) | (raw[5].d[w] << 16);
        v.w = (raw[6].d[w] & 0xffffu) | (raw[7].d[w] << 16);
    }
    return __builtin_bit_cast(f16x8, v);
}

template <int RS, int EPL>
__global__ __launch_bounds__(64) void bcsr_spmm_f16(SpmmArgs a) {
    int tile, slot;
    if (!map_work(a, blockIdx.x, tile, slot)) return;
    const int lane = threadIdx.x;
    const int g = lane >> 4, j = lane & 15;
    const int br = a.row_order[slot];
    const int n0 = tile * (16 * EPL) + EPL * j;
    const bool nvalid = n0 < a.N;
    const size_t N = static_cast<size_t>(a.N);
    const int krow = (g & 1) * 8;
    const bool second = g >= 2;

    f32x4 acc[RS][EPL];
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
        for (int e = 0; e < EPL; ++e) acc[s][e] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int p0 = a.pair_ptr[br], p1 = a.pair_ptr[br + 1];
    const _Float16 *Bcol = a.B + n0;
    for (int p = p0; p < p1; ++p) {
        const int c0 = a.pair_col[2 * p], c1 = a.pair_col[2 * p + 1];
        const int c = second ? c1 : c0;
        const _Float16 *bp = Bcol + static_cast<size_t>(c * 16 + krow) * N;
        BRow<EPL> raw[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) raw[r] = nvalid ? load_brow<EPL>(bp + r * N) : zero_brow<EPL>();
        f16x8 av[RS];
#pragma unroll
        for (int s = 0; s < RS; ++s)
            av[s] = __builtin_bit_cast(f16x8, a.vals[(static_cast<size_t>(p) * RS + s) * 64 + lane]);
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            const f16x8 bv = patch_column<EPL>(raw, e);
#pragma unroll
            for (int s = 0; s < RS; ++s)
                acc[s][e] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[s], bv, acc[s][e], 0, 0, 0);
        }
    }

    if (!nvalid) return;
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const size_t row = static_cast<size_t>(br) * (16 * RS) + 16 * s + 4 * g + r;
            float *cp = a.C + row * N + n0;
            if constexpr (EPL == 2) {
                *reinterpret_cast<f32x2 *>(cp) = f32x2{acc[s][0][r], acc[s][1][r]};
            } else {
#pragma unroll
                for (int h = 0; h < EPL / 4; ++h)
                    reinterpret_cast<f32x4 *>(cp)[h] =
                        f32x4{acc[s][4 * h][r], acc[s][4 * h + 1][r], acc[s][4 * h + 2][r], acc[s][4 * h + 3][r]};
            }
        }
}

typedef void (*SpmmKernel)(SpmmArgs);

template <int RS>
SpmmKernel pick_epl(int epl) {
    switch (epl) {
        case 2: return bcsr_spmm_f16<RS, 2>;
        case 4: return bcsr_spmm_f16<RS, 4>;
        case 8: return bcsr_spmm_f16<RS, 8>;
        case 16: return bcsr_spmm_f16<RS, 16>;
        default: return nullptr;
    }
}

SpmmKernel pick_kernel(int rs, int epl) {
    switch (rs) {
        case 1: return pick_epl<1>(epl);
        case 2: return pick_epl<2>(epl);
        case 4: return pick_epl<4>(epl);
        default: return nullptr;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
struct bcsr_plan {
    bcsr_plan_info info{};
    int device = 0;
    int rs = 1;        // dev_bh / 16
    int dev_rows = 0;  // M / dev_bh
    void *dev_base = nullptr;
    int32_t *d_hdr = nullptr;
    int32_t *d_idx = nullptr;
    int32_t *d_pair_ptr = nullptr;
    int32_t *d_pair_col = nullptr;
    int32_t *d_row_order = nullptr;
    void *d_vals = nullptr;
    float *d_scales = nullptr;
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Source block values: fp16 bits or fp32.
struct SrcBlocks {
    const uint16_t *f16 = nullptr;
    const float *f32 = nullptr;
    uint16_t half_bits(int64_t i) const {
        if (f16) return f16[i];
        _Float16 h = static_cast<_Float16>(f32[i]);
        uint16_t b;
        std::memcpy(&b, &h, 2);
        return b;
    }
};

int validate_host(const int32_t *hdr, const int32_t *idx, int32_t block_rows, int32_t block_cols,
                  int32_t bh, int32_t bw) {
    if (!hdr || block_rows < 0 || block_cols < 0) return BCSR_ERR_INVALID_ARG;
    if (bh < 16 || bw < 16 || bh % 16 || bw % 16) return BCSR_ERR_BLOCK_SIZE;
    if (static_cast<int64_t>(block_rows) * bh > INT32_MAX || static_cast<int64_t>(block_cols) * bw > INT32_MAX)
        return BCSR_ERR_SHAPE;
    if (hdr[0] != 0) return BCSR_ERR_INVALID_ARG;
    for (int i = 0; i < block_rows; ++i) {
        if (hdr[i + 1] < hdr[i]) return BCSR_ERR_INVALID_ARG;
        for (int64_t k = hdr[i]; k < hdr[i + 1]; ++k) {
            if (!idx || idx[k] < 0 || idx[k] >= block_cols) return BCSR_ERR_INVALID_ARG;
            if (k > hdr[i] && idx[k] <= idx[k - 1]) return BCSR_ERR_INVALID_ARG;
        }
    }
    return BCSR_OK;
}

int build_plan(const int32_t *hdr, const int32_t *idx, const SrcBlocks &src, int32_t block_rows,
               int32_t block_cols, int32_t bh, int32_t bw, int32_t precision, bcsr_plan **out) {
    if (!out) return BCSR_ERR_INVALID_ARG;
    *out = nullptr;
    int rc = validate_host(hdr, idx, block_rows, block_cols, bh, bw);
    if (rc != BCSR_OK) return rc;
    if (precision != BCSR_PREC_F16) return BCSR_ERR_UNSUPPORTED;

    const int dev_bh = bh % 64 == 0 ? 64 : (bh % 32 == 0 ? 32 : 16);
    const int rs = dev_bh / 16;
    const int sub_rows = bh / dev_bh;    // device block-rows per source block-row
    const int pieces = bw / 16;          // 16-column pieces per source block
    const int dev_rows = block_rows * sub_rows;
    const int64_t nnzb = hdr[block_rows];

    // Pairs per device block-row: every source block of the row gives
    // `pieces` device blocks for each of its sub-rows.
    std::vector<int32_t> pair_ptr(dev_rows + 1, 0);
    for (int R = 0; R < block_rows; ++R) {
        int64_t nblk = static_cast<int64_t>(hdr[R + 1] - hdr[R]) * pieces;
        for (int t = 0; t < sub_rows; ++t) pair_ptr[R * sub_rows + t + 1] = static_cast<int32_t>((nblk + 1) / 2);
    }
    for (int r = 0; r < dev_rows; ++r) pair_ptr[r + 1] += pair_ptr[r];
    const int64_t pairs = pair_ptr[dev_rows];

    std::vector<int32_t> pair_col(2 * pairs);
    std::vector<uint16_t> vals(static_cast<size_t>(pairs) * rs * 64 * 8, 0);
    const int64_t blk_elems = static_cast<int64_t>(bh) * bw;
    for (int R = 0; R < block_rows; ++R) {
        const int64_t k0 = hdr[R], nsrc = hdr[R + 1] - hdr[R];
        const int64_t nblk = nsrc * pieces;
        for (int t = 0; t < sub_rows; ++t) {
            const int dr = R * sub_rows + t;
            int64_t p = pair_ptr[dr];
            for (int64_t b = 0; b < nblk; b += 2) {
                for (int which = 0; which < 2; ++which) {
                    int64_t db = b + which;
                    const bool pad = db >= nblk;
                    if (pad) db = nblk - 1;
                    const int64_t k = k0 + db / pieces;
                    const int q = static_cast<int>(db % pieces);
                    pair_col[2 * p + which] = idx[k] * pieces + q;
                    if (pad) continue;
                    for (int s = 0; s < rs; ++s)
                        for (int lane = 0; lane < 64; ++lane) {
                            const int g = lane >> 4, i = lane & 15;
                            if ((g >> 1) != which) continue;
                            const int row_in_blk = t * dev_bh + 16 * s + i;
                            for (int e = 0; e < 8; ++e) {
                                const int col_in_blk = q * 16 + (g & 1) * 8 + e;
                                vals[((static_cast<size_t>(p) * rs + s) * 64 + lane) * 8 + e] =
                                    src.half_bits(k * blk_elems + static_cast<int64_t>(row_in_blk) * bw + col_in_blk);
                            }
                        }
                }
                ++p;
            }
        }
    }

    std::vector<int32_t> row_order(dev_rows);
    for (int r = 0; r < dev_rows; ++r) row_order[r] = r;
    std::stable_sort(row_order.begin(), row_order.end(), [&](int x, int y) {
        return pair_ptr[x + 1] - pair_ptr[x] > pair_ptr[y + 1] - pair_ptr[y];
    });

    bcsr_plan *plan = new (std::nothrow) bcsr_plan();
    if (!plan) return BCSR_ERR_OUT_OF_MEMORY;
    const size_t A = 256;
    size_t off_hdr = 0;
    size_t off_idx = align_up(off_hdr + sizeof(int32_t) * (block_rows + 1), A);
    size_t off_pp = align_up(off_idx + sizeof(int32_t) * std::max<int64_t>(nnzb, 1), A);
    size_t off_pc = align_up(off_pp + sizeof(int32_t) * (dev_rows + 1), A);
    size_t off_ro = align_up(off_pc + sizeof(int32_t) * std::max<int64_t>(2 * pairs, 2), A);
    size_t off_v = align_up(off_ro + sizeof(int32_t) * std::max(dev_rows, 1), A);
    size_t total = align_up(off_v + sizeof(uint16_t) * std::max<size_t>(vals.size(), 8), A);

    int dev = 0;
    rc = hip_status(hipGetDevice(&dev));
    if (rc == BCSR_OK) rc = hip_status(hipMalloc(&plan->dev_base, total));
    if (rc != BCSR_OK) {
        delete plan;
        return rc;
    }
    char *base = static_cast<char *>(plan->dev_base);
    plan->device = dev;
    plan->d_hdr = reinterpret_cast<int32_t *>(base + off_hdr);
    plan->d_idx = reinterpret_cast<int32_t *>(base + off_idx);
    plan->d_pair_ptr = reinterpret_cast<int32_t *>(base + off_pp);
    plan->d_pair_col = reinterpret_cast<int32_t *>(base + off_pc);
    plan->d_row_order = reinterpret_cast<int32_t *>(base + off_ro);
    plan->d_vals = base + off_v;
    hipError_t e = hipMemcpy(plan->d_hdr, hdr, sizeof(int32_t) * (block_rows + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnzb) e = hipMemcpy(plan->d_idx, idx, sizeof(int32_t) * nnzb, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(plan->d_pair_ptr, pair_ptr.data(), sizeof(int32_t) * (dev_rows + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && pairs)
        e = hipMemcpy(plan->d_pair_col, pair_col.data(), sizeof(int32_t) * 2 * pairs, hipMemcpyHostToDevice);
    if (e == hipSuccess && dev_rows)
        e = hipMemcpy(plan->d_row_order, row_order.data(), sizeof(int32_t) * dev_rows, hipMemcpyHostToDevice);
    if (e == hipSuccess && pairs)
        e = hipMemcpy(plan->d_vals, vals.data(), sizeof(uint16_t) * vals.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rc = hip_status(e);
        (void)hipFree(plan->dev_base);
        delete plan;
        return rc;
    }

    plan->rs = rs;
    plan->dev_rows = dev_rows;
    bcsr_plan_info &inf = plan->info;
    inf.rows = block_rows * bh;
    inf.cols = block_cols * bw;
    inf.bh = bh;
    inf.bw = bw;
    inf.block_rows = block_rows;
    inf.block_cols = block_cols;
    inf.precision = precision;
    inf.dev_bh = dev_bh;
    inf.nnzb = nnzb;
    inf.dev_pairs = pairs;
    inf.device_bytes = static_cast<int64_t>(total);
    *out = plan;
    return BCSR_OK;
}

struct LaunchShape {
    int epl = 8;
    int map_mode = 0;
    int n_tiles = 1;
    int grid = 0;
};

int resolve_launch(const bcsr_plan *plan, int N, const bcsr_launch_cfg *cfg, LaunchShape &ls) {
    int tile_n = cfg ? cfg->tile_n : 0;
    if (tile_n == 0) tile_n = N >= 128 ? 128 : (N >= 64 ? 64 : 32);
    if (tile_n != 32 && tile_n != 64 && tile_n != 128 && tile_n != 256) return BCSR_ERR_INVALID_ARG;
    if (cfg && cfg->variant != 0) return BCSR_ERR_UNSUPPORTED;
    ls.epl = tile_n / 16;
    ls.n_tiles = (N + tile_n - 1) / tile_n;
    const bool xcd_ok = ls.n_tiles == 1 || ls.n_tiles == 2 || ls.n_tiles == 4 || ls.n_tiles % 8 == 0;
    int map = cfg ? cfg->xcd_map : -1;
    if (map < 0) map = xcd_ok ? 1 : 0;
    if (map == 1 && !xcd_ok) return BCSR_ERR_INVALID_ARG;
    if (map != 0 && map != 1) return BCSR_ERR_INVALID_ARG;
    ls.map_mode = map;
    const int rows = plan->dev_rows;
    if (map == 0) {
        ls.grid = ls.n_tiles * rows;
    } else if (ls.n_tiles >= 8) {
        ls.grid = ls.n_tiles * rows;
    } else {
        const int q = 8 / ls.n_tiles;
        ls.grid = 8 * ((rows + q - 1) / q);
    }
    return BCSR_OK;
}

int check_operands(const bcsr_plan *plan, const void *B, const float *C, int N) {
    if (!plan) return BCSR_ERR_INVALID_ARG;
    if (N < 0) return BCSR_ERR_INVALID_ARG;
    if (N % 16 != 0) return BCSR_ERR_SHAPE;
    if (N == 0 || plan->info.rows == 0) return BCSR_OK;
    if (!B || !C) return BCSR_ERR_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(B) | reinterpret_cast<uintptr_t>(C)) & 15u) return BCSR_ERR_SHAPE;
    return BCSR_OK;
}

int launch_spmm(const bcsr_plan *plan, const void *B, float *C, int N, const bcsr_launch_cfg *cfg,
                hipStream_t stream) {
    int rc = check_operands(plan, B, C, N);
    if (rc != BCSR_OK) return rc;
    if (N == 0 || plan->info.rows == 0) return BCSR_OK;
    LaunchShape ls;
    rc = resolve_launch(plan, N, cfg, ls);
    if (rc != BCSR_OK) return rc;
    SpmmKernel k = pick_kernel(plan->rs, ls.epl);
    if (!k) return BCSR_ERR_UNSUPPORTED;
    SpmmArgs a;
    a.pair_ptr = plan->d_pair_ptr;
    a.pair_col = plan->d_pair_col;
    a.vals = static_cast<const u32x4 *>(plan->d_vals);
    a.row_order = plan->d_row_order;
    a.B = static_cast<const _Float16 *>(B);
    a.C = C;
    a.N = N;
    a.n_tiles = ls.n_tiles;
    a.dev_rows = plan->dev_rows;
    a.map_mode = ls.map_mode;
    hipLaunchKernelGGL(k, dim3(ls.grid), dim3(64), 0, stream, a);
    return hip_status(hipGetLastError());
}

// Registry of plans made by bcsr_copy_to_device, keyed by their device hdr.
std::mutex g_reg_mu;
std::unordered_map<const int *, bcsr_plan *> &registry() {
    static std::unordered_map<const int *, bcsr_plan *> r;
    return r;
}

// bh > 0 additionally requires a bh x 16 matrix (the reference's kernels
// hard-code 16-column blocks).
int lookup(const int *hdr, int bh, bcsr_plan **plan) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = registry().find(hdr);
    if (it == registry().end()) return BCSR_ERR_NOT_REGISTERED;
    if (bh > 0 && (it->second->info.bh != bh || it->second->info.bw != 16)) return BCSR_ERR_BLOCK_SIZE;
    *plan = it->second;
    return BCSR_OK;
}

int entry(int bh, const int *hdr, const int *idx, const void *data, const void *B, float *C, unsigned M,
          unsigned N, void *stream) {
    bcsr_plan *plan = nullptr;
    int rc = lookup(hdr, bh, &plan);
    if (rc != BCSR_OK) return rc;
    if (idx != plan->d_idx || data != plan->d_vals) return BCSR_ERR_INVALID_ARG;
    if (M != static_cast<unsigned>(plan->info.rows) || N > static_cast<unsigned>(INT32_MAX)) return BCSR_ERR_SHAPE;
    return launch_spmm(plan, B, C, static_cast<int>(N), nullptr, static_cast<hipStream_t>(stream));
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *bcsr_status_string(int status) {
    switch (status) {
        case BCSR_OK: return "ok";
        case BCSR_ERR_INVALID_ARG: return "invalid argument";
        case BCSR_ERR_BLOCK_SIZE: return "unsupported block size";
        case BCSR_ERR_SHAPE: return "inconsistent shape or misaligned operand";
        case BCSR_ERR_HIP: return "HIP runtime error";
        case BCSR_ERR_OUT_OF_MEMORY: return "out of memory";
        case BCSR_ERR_NOT_REGISTERED: return "device hdr pointer not registered by copyToDevice";
        case BCSR_ERR_UNSUPPORTED: return "unsupported precision or variant";
        case BCSR_ERR_UNKNOWN_PATTERN: return "unknown pattern";
        default: return "unknown status";
    }
}

int bcsr_last_hip_error(void) { return g_last_hip_error; }

int bcsr_plan_create(const int32_t *hdr, const int32_t *idx, const void *data_f16, int32_t block_rows,
                     int32_t block_cols, int32_t bh, int32_t bw, bcsr_plan **out) {
    if (!data_f16 && hdr && block_rows > 0 && hdr[block_rows] > 0) return BCSR_ERR_INVALID_ARG;
    SrcBlocks src;
    src.f16 = static_cast<const uint16_t *>(data_f16);
    try {
        return build_plan(hdr, idx, src, block_rows, block_cols, bh, bw, BCSR_PREC_F16, out);
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
}

int bcsr_plan_create_f32(const int32_t *hdr, const int32_t *idx, const float *data, int32_t block_rows,
                         int32_t block_cols, int32_t bh, int32_t bw, int32_t precision, bcsr_plan **out) {
    if (!data && hdr && block_rows > 0 && hdr[block_rows] > 0) return BCSR_ERR_INVALID_ARG;
    if (precision < BCSR_PREC_F16 || precision > BCSR_PREC_F16_HILO) return BCSR_ERR_INVALID_ARG;
    SrcBlocks src;
    src.f32 = data;
    try {
        return build_plan(hdr, idx, src, block_rows, block_cols, bh, bw, precision, out);
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
}

int bcsr_plan_destroy(bcsr_plan *plan) {
    if (!plan) return BCSR_OK;
    int rc = hip_status(hipFree(plan->dev_base));
    delete plan;
    return rc;
}

int bcsr_plan_get_info(const bcsr_plan *plan, bcsr_plan_info *info) {
    if (!plan || !info) return BCSR_ERR_INVALID_ARG;
    *info = plan->info;
    return BCSR_OK;
}

int bcsr_spmm(const bcsr_plan *plan, const void *B, float *C, int32_t N, void *stream) {
    return launch_spmm(plan, B, C, N, nullptr, static_cast<hipStream_t>(stream));
}

int bcsr_spmm_cfg(const bcsr_plan *plan, const void *B, float *C, int32_t N, const bcsr_launch_cfg *cfg,
                  void *stream) {
    return launch_spmm(plan, B, C, N, cfg, static_cast<hipStream_t>(stream));
}

int bcsr_spmm_repeat(const bcsr_plan *plan, const void *B, float *C, int32_t N, const bcsr_launch_cfg *cfg,
                     int32_t reps, void *stream) {
    if (reps < 0) return BCSR_ERR_INVALID_ARG;
    for (int i = 0; i < reps; ++i) {
        int rc = launch_spmm(plan, B, C, N, cfg, static_cast<hipStream_t>(stream));
        if (rc != BCSR_OK) return rc;
    }
    return BCSR_OK;
}

int bcsr_spmm_kernel_name(const bcsr_plan *plan, int32_t N, const bcsr_launch_cfg *cfg, char *buf,
                          size_t buflen) {
    if (!plan || !buf || buflen == 0) return BCSR_ERR_INVALID_ARG;
    LaunchShape ls;
    int rc = resolve_launch(plan, N, cfg, ls);
    if (rc != BCSR_OK) return rc;
    std::snprintf(buf, buflen, "bcsr_spmm_f16<%d, %d>", plan->rs, ls.epl);
    return BCSR_OK;
}

int bcsr_copy_to_device(const int32_t *hdr, const int32_t *idx, const void *data_f16, int32_t block_rows,
                        int32_t block_cols, int32_t bh, int32_t bw, int **gpu_hdr, int **gpu_idx,
                        void **gpu_data) {
    if (!gpu_hdr || !gpu_idx || !gpu_data) return BCSR_ERR_INVALID_ARG;
    bcsr_plan *plan = nullptr;
    int rc = bcsr_plan_create(hdr, idx, data_f16, block_rows, block_cols, bh, bw, &plan);
    if (rc != BCSR_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        registry()[plan->d_hdr] = plan;
    }
    *gpu_hdr = plan->d_hdr;
    *gpu_idx = plan->d_idx;
    *gpu_data = plan->d_vals;
    return BCSR_OK;
}

int bcsr_release_device(const int *gpu_hdr) {
    bcsr_plan *plan = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        auto it = registry().find(gpu_hdr);
        if (it == registry().end()) return BCSR_ERR_NOT_REGISTERED;
        plan = it->second;
        registry().erase(it);
    }
    return bcsr_plan_destroy(plan);
}

int bcsr_plan_from_device(const int *gpu_hdr, const bcsr_plan **out) {
    if (!out) return BCSR_ERR_INVALID_ARG;
    bcsr_plan *plan = nullptr;
    int rc = lookup(gpu_hdr, 0, &plan);
    *out = plan;
    return rc;
}

#define BCSR_DEFINE_ENTRY(name, bh)                                                                  \
    int name(const int *hdr, const int *idx, const void *data, const void *B, float *C, unsigned int M, \
             unsigned int N, void *stream) {                                                          \
        return entry(bh, hdr, idx, data, B, C, M, N, stream);                                         \
    }
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor, 16)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor_v2, 16)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor_v3, 16)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor_v1_improved, 16)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor_option2_ldmatrix_sm80, 16)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor64x16, 64)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor64x16_v2, 64)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor32x16, 32)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor32x16_shared, 32)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor32x16_shared_vectorized, 32)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor32x16_v2, 32)
BCSR_DEFINE_ENTRY(sparseMatrixMulTensor32x16_v2_shared_vectorized, 32)
#undef BCSR_DEFINE_ENTRY

int sparseMatrixMulTensor1(const int *hdr, const int *idx, const void *data, const void *B, float *C,
                           unsigned int n, void *stream) {
    return entry(16, hdr, idx, data, B, C, n, n, stream);
}

int sparseMatrixMulTensorAdaptive(const int *hdr, const int *idx, const void *data, const void *B, float *C,
                                  unsigned int M, unsigned int N, unsigned int, void *stream) {
    return entry(16, hdr, idx, data, B, C, M, N, stream);
}

int sparseMatrixMulTensorAdaptive32x16(const int *hdr, const int *idx, const void *data, const void *B,
                                       float *C, unsigned int M, unsigned int N, unsigned int, void *stream) {
    return entry(32, hdr, idx, data, B, C, M, N, stream);
}

int sparseMatrixMulTensorAdaptive64x16(const int *hdr, const int *idx, const void *data, const void *B,
                                       float *C, unsigned int M, unsigned int N, unsigned int, void *stream) {
    return entry(64, hdr, idx, data, B, C, M, N, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device plumbing
// ---------------------------------------------------------------------------
extern "C" {

int bcsr_device_count(int32_t *count) {
    if (!count) return BCSR_ERR_INVALID_ARG;
    int n = 0;
    int rc = hip_status(hipGetDeviceCount(&n));
    *count = rc == BCSR_OK ? n : 0;
    return rc;
}

int bcsr_set_device(int32_t device) { return hip_status(hipSetDevice(device)); }

int bcsr_dev_malloc(void **ptr, size_t bytes) {
    if (!ptr) return BCSR_ERR_INVALID_ARG;
    *ptr = nullptr;
    return hip_status(hipMalloc(ptr, bytes ? bytes : 1));
}

int bcsr_dev_free(void *ptr) { return hip_status(hipFree(ptr)); }

int bcsr_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (!bytes) return BCSR_OK;
    if (!dst || !src) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int bcsr_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (!bytes) return BCSR_OK;
    if (!dst || !src) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int bcsr_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return BCSR_OK;
    if (!dst || !src) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
}

int bcsr_memset_async(void *dst, int32_t byte_value, size_t bytes, void *stream) {
    if (!bytes) return BCSR_OK;
    if (!dst) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemsetAsync(dst, byte_value, bytes, static_cast<hipStream_t>(stream)));
}

int bcsr_stream_create(void **stream) {
    if (!stream) return BCSR_ERR_INVALID_ARG;
    hipStream_t s = nullptr;
    int rc = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return rc;
}

int bcsr_stream_destroy(void *stream) { return hip_status(hipStreamDestroy(static_cast<hipStream_t>(stream))); }

int bcsr_stream_synchronize(void *stream) {
    return hip_status(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

int bcsr_device_synchronize(void) { return hip_status(hipDeviceSynchronize()); }

int bcsr_event_create(void **event) {
    if (!event) return BCSR_ERR_INVALID_ARG;
    hipEvent_t e = nullptr;
    int rc = hip_status(hipEventCreate(&e));
    *event = e;
    return rc;
}

int bcsr_event_destroy(void *event) { return hip_status(hipEventDestroy(static_cast<hipEvent_t>(event))); }

int bcsr_event_record(void *event, void *stream) {
    return hip_status(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
}

int bcsr_event_elapsed_ms(void *start, void *end, float *ms) {
    if (!ms) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// C++ class device members
// ---------------------------------------------------------------------------
void BCSRMatrix::copyToDevice(int **gpuHdr, int **gpuIdx, half **gpuData) const {
    void *d = nullptr;
    int rc = bcsr_copy_to_device(hdr, idx, data, blockRows, blockCols, blockSizeRow, blockSizeCol, gpuHdr,
                                 gpuIdx, &d);
    if (rc != BCSR_OK)
        throw std::runtime_error(std::string("BCSRMatrix::copyToDevice: ") + bcsr_status_string(rc));
    *gpuData = static_cast<half *>(d);
}

void CSRMatrix::copyToDevice(int **gpuHdr, int **gpuIdx, half **gpuData) const {
    const size_t nnz = static_cast<size_t>(hdr[N]);
    auto fail = [](const char *what, hipError_t e) {
        throw std::runtime_error(std::string("CSRMatrix::copyToDevice ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc(reinterpret_cast<void **>(gpuData), sizeof(half) * (nnz ? nnz : 1))) != hipSuccess)
        fail("malloc gpuData", e);
    if ((e = hipMalloc(reinterpret_cast<void **>(gpuHdr), sizeof(int) * (N + 1))) != hipSuccess)
        fail("malloc gpuHdr", e);
    if ((e = hipMalloc(reinterpret_cast<void **>(gpuIdx), sizeof(int) * (nnz ? nnz : 1))) != hipSuccess)
        fail("malloc gpuIdx", e);
    if (nnz && (e = hipMemcpy(*gpuData, data, sizeof(half) * nnz, hipMemcpyHostToDevice)) != hipSuccess)
        fail("gpuData", e);
    if ((e = hipMemcpy(*gpuHdr, hdr, sizeof(int) * (N + 1), hipMemcpyHostToDevice)) != hipSuccess)
        fail("gpuHdr", e);
    if (nnz && (e = hipMemcpy(*gpuIdx, idx, sizeof(int) * nnz, hipMemcpyHostToDevice)) != hipSuccess)
        fail("gpuIdx", e);
}
