// bcsr_plan.hip -- device layout of a BCSR matrix (the plan), the registry
// behind the reference-named entry points, and the C ABI.
//
// Replaces BCSRMatrix::copyToDevice (formats/BCSRMatrix.cu:106-132) and the
// launch sites of the sparseMatrixMulTensor* kernels (src/cuda_kernels.cuh:
// 39-114): the caller no longer picks grids, the plan does.  The layout is
// described in bcsr_internal.h; the kernels live in kernels_pair.hip (16-bit
// modes) and kernels_fp8.hip (FP8 mode).
#include <algorithm>
#include <cmath>
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
#include "bcsr_internal.h"

namespace {
thread_local int g_last_hip_error = 0;
}

// A failed HIP call also leaves its error as the runtime's "last error" until
// hipGetLastError reads it; the launches check hipGetLastError, so a failure
// already reported here (say an allocation that did not fit) would come back
// as the status of the next launch.  Reading it here clears it.
int hip_status(hipError_t e) {
    if (e == hipSuccess) return BCSR_OK;
    g_last_hip_error = static_cast<int>(e);
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? BCSR_ERR_OUT_OF_MEMORY : BCSR_ERR_HIP;
}

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

uint16_t f32_to_f16_bits(float f) {
    _Float16 h = static_cast<_Float16>(f);
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

float f16_bits_to_f32(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return static_cast<float>(h);
}

// float -> bfloat16, round to nearest even (NaN kept quiet).
uint16_t f32_to_bf16_bits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return static_cast<uint16_t>(x >> 16);
}

}  // namespace

// float -> FP8 E4M3 (OCP: bias 7, max 448, NaN = 0x7f, no infinity), round
// to nearest even; finite overflow saturates to +-448.
uint8_t bcsr_f32_to_e4m3(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint8_t sign = static_cast<uint8_t>((x >> 24) & 0x80u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return sign | 0x7fu;
    if (ax >= 0x43e00000u) return sign | 0x7eu;
    if (ax < 0x3c800000u) {  // below the smallest normal 2^-6: multiples of 2^-9
        const uint32_t e = ax >> 23;
        if (e < 116u) return sign;
        const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 141u - e;
        uint32_t r = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (r & 1u))) r++;
        return sign | static_cast<uint8_t>(r);
    }
    const uint32_t e = ax >> 23;
    uint32_t r = ((e - 120u) << 3) | ((ax >> 20) & 7u);
    const uint32_t rem = ax & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (r & 1u))) r++;
    return sign | static_cast<uint8_t>(std::min<uint32_t>(r, 0x7eu));
}

// E8M0 exponent of the smallest power of two s with absmax / s <= 448.
uint8_t bcsr_e8m0_for_absmax(float absmax) {
    if (!(absmax > 0.0f)) return 127;
    int k = -127;
    while (k < 127 && std::ldexp(448.0f, k) < absmax) k++;
    return static_cast<uint8_t>(k + 127);
}

namespace {

// Source block values, fp16 bits or fp32.
struct SrcBlocks {
    const uint16_t *f16 = nullptr;
    const float *f32 = nullptr;
    float value(int64_t i) const { return f16 ? f16_bits_to_f32(f16[i]) : f32[i]; }
    uint16_t half_bits(int64_t i) const { return f16 ? f16[i] : f32_to_f16_bits(f32[i]); }
};

int validate_host(const int32_t *hdr, const int32_t *idx, int32_t block_rows, int32_t block_cols, int32_t bh,
                  int32_t bw) {
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
    if (precision < BCSR_PREC_F16 || precision > BCSR_PREC_F16_HILO) return BCSR_ERR_INVALID_ARG;

    const bool fp8 = precision == BCSR_PREC_FP8_E4M3;
    const int G = fp8 ? 8 : 2;                                  // blocks per group
    const int H = precision == BCSR_PREC_F16_HILO ? 2 : 1;      // operand copies (hi, lo)
    const int dev_bh = bh % 64 == 0 ? 64 : (bh % 32 == 0 ? 32 : 16);
    const int rs = dev_bh / 16;
    const int sub_rows = bh / dev_bh;
    const int pieces = bw / 16;
    const int dev_rows = block_rows * sub_rows;
    const int64_t nnzb = hdr[block_rows];
    const int64_t blk_elems = static_cast<int64_t>(bh) * bw;

    // group offsets are 32-bit on the device
    int64_t total_groups = 0;
    for (int R = 0; R < block_rows; ++R)
        total_groups += (static_cast<int64_t>(hdr[R + 1] - hdr[R]) * pieces + G - 1) / G * sub_rows;
    if (total_groups > INT32_MAX) return BCSR_ERR_SHAPE;
    std::vector<int32_t> grp_ptr(dev_rows + 1, 0);
    for (int R = 0; R < block_rows; ++R) {
        const int64_t nblk = static_cast<int64_t>(hdr[R + 1] - hdr[R]) * pieces;
        for (int t = 0; t < sub_rows; ++t) grp_ptr[R * sub_rows + t + 1] = static_cast<int32_t>((nblk + G - 1) / G);
    }
    for (int r = 0; r < dev_rows; ++r) grp_ptr[r + 1] += grp_ptr[r];
    const int64_t groups = grp_ptr[dev_rows];
    // pairs: per row-subtile (and hi/lo copy) 64 lanes x 16 bytes;
    // octets: per row-subtile 64 lanes x 32 bytes
    const size_t group_bytes = fp8 ? static_cast<size_t>(rs) * 64 * 32 : static_cast<size_t>(rs) * H * 64 * 16;

    std::vector<int32_t> grp_col(static_cast<size_t>(G) * groups);
    std::vector<uint8_t> vals(static_cast<size_t>(groups) * group_bytes, 0);
    std::vector<uint32_t> ascale(fp8 ? static_cast<size_t>(groups) * rs : 0, 0);
    for (int R = 0; R < block_rows; ++R) {
        const int64_t k0 = hdr[R], nblk = static_cast<int64_t>(hdr[R + 1] - hdr[R]) * pieces;
        for (int t = 0; t < sub_rows; ++t) {
            int64_t p = grp_ptr[R * sub_rows + t];
            for (int64_t b = 0; b < nblk; b += G, ++p) {
                uint8_t *gv = vals.data() + static_cast<size_t>(p) * group_bytes;
                // source offset of element (row, column c) of device block db
                auto src_index = [&](int64_t db, int row, int c) {
                    const int64_t k = k0 + db / pieces;
                    const int q = static_cast<int>(db % pieces);
                    return k * blk_elems + static_cast<int64_t>(row) * bw + q * 16 + c;
                };
                for (int slot = 0; slot < G; ++slot) {
                    int64_t db = b + slot;
                    if (db >= nblk) db = nblk - 1;  // padding repeats the last column
                    const int64_t k = k0 + db / pieces;
                    grp_col[static_cast<size_t>(G) * p + slot] = idx[k] * pieces + static_cast<int>(db % pieces);
                }
                if (fp8) {
                    // slots (2 u, 2 u + 1) share one E8M0 scale per row-subtile;
                    // slot -> lane group slot & 3, byte half slot >> 2
                    for (int s = 0; s < rs; ++s)
                        for (int u = 0; u < 4; ++u) {
                            float amax = 0.0f;
                            for (int m = 0; m < 2; ++m) {
                                const int64_t db = b + 2 * u + m;
                                if (db >= nblk) continue;
                                for (int i = 0; i < 16; ++i)
                                    for (int c = 0; c < 16; ++c)
                                        amax = std::max(amax, std::fabs(src.value(src_index(db, t * dev_bh + 16 * s + i, c))));
                            }
                            const uint8_t se = bcsr_e8m0_for_absmax(amax);
                            ascale[static_cast<size_t>(p) * rs + s] |= static_cast<uint32_t>(se) << (8 * u);
                            const float inv = std::ldexp(1.0f, 127 - static_cast<int>(se));
                            for (int m = 0; m < 2; ++m) {
                                const int slot = 2 * u + m;
                                const int64_t db = b + slot;
                                if (db >= nblk) continue;
                                const int g = slot & 3, half = slot >> 2;
                                for (int i = 0; i < 16; ++i) {
                                    uint8_t *dst = gv + (static_cast<size_t>(s) * 64 + 16 * g + i) * 32 + 16 * half;
                                    for (int c = 0; c < 16; ++c)
                                        dst[c] = bcsr_f32_to_e4m3(src.value(src_index(db, t * dev_bh + 16 * s + i, c)) * inv);
                                }
                            }
                        }
                    continue;
                }
                for (int slot = 0; slot < G; ++slot) {
                    const int64_t db = b + slot;
                    if (db >= nblk) continue;
                    for (int s = 0; s < rs; ++s)
                        for (int lane = 0; lane < 64; ++lane) {
                            const int g = lane >> 4, i = lane & 15;
                            if ((g >> 1) != slot) continue;
                            const int row = t * dev_bh + 16 * s + i;
                            for (int h = 0; h < H; ++h) {
                                uint16_t *dst = reinterpret_cast<uint16_t *>(
                                    gv + ((static_cast<size_t>(s) * H + h) * 64 + lane) * 16);
                                for (int e = 0; e < 8; ++e) {
                                    const int64_t si = src_index(db, row, (g & 1) * 8 + e);
                                    if (precision == BCSR_PREC_F16) {
                                        dst[e] = src.half_bits(si);
                                    } else if (precision == BCSR_PREC_BF16) {
                                        dst[e] = f32_to_bf16_bits(src.value(si));
                                    } else {  // hi + lo: a = hi + 2^-11 lo
                                        const float a = src.value(si);
                                        const uint16_t hi = f32_to_f16_bits(a);
                                        dst[e] = h == 0 ? hi : f32_to_f16_bits((a - f16_bits_to_f32(hi)) * 2048.0f);
                                    }
                                }
                            }
                        }
                }
            }
        }
    }

    // the panel kernel (launch variant 3) runs 16-row device blocks in the
    // 16-bit single-operand modes; its schedule is built by its first launch
    const bool panel_ok = !fp8 && H == 1 && rs == 1 && groups * 1024 <= INT32_MAX && dev_rows > 0;

    std::vector<int32_t> row_order(dev_rows);
    for (int r = 0; r < dev_rows; ++r) row_order[r] = r;
    std::stable_sort(row_order.begin(), row_order.end(),
                     [&](int x, int y) { return grp_ptr[x + 1] - grp_ptr[x] > grp_ptr[y + 1] - grp_ptr[y]; });
    // the same rows taken from both ends in turn, 32 at a time (the CUs of an
    // XCD): any run of consecutive entries then holds about the average work,
    // and so do the entries 32 apart that land on one CU (chunked mapping)
    std::vector<int32_t> mix_order(dev_rows);
    for (int k = 0, lo = 0, hi = dev_rows - 1; k < dev_rows; ++k) mix_order[k] = row_order[(k / 32) % 2 ? hi-- : lo++];

    bcsr_plan *plan = new (std::nothrow) bcsr_plan();
    if (!plan) return BCSR_ERR_OUT_OF_MEMORY;
    const size_t A = 256;
    const size_t off_hdr = 0;
    const size_t off_idx = align_up(off_hdr + sizeof(int32_t) * (block_rows + 1), A);
    const size_t off_gp = align_up(off_idx + sizeof(int32_t) * std::max<int64_t>(nnzb, 1), A);
    const size_t off_gc = align_up(off_gp + sizeof(int32_t) * (dev_rows + 1), A);
    const size_t off_ro = align_up(off_gc + sizeof(int32_t) * std::max<size_t>(grp_col.size(), 4), A);
    const size_t off_mo = align_up(off_ro + sizeof(int32_t) * std::max(dev_rows, 1), A);
    const size_t off_as = align_up(off_mo + sizeof(int32_t) * std::max(dev_rows, 1), A);
    const size_t off_v = align_up(off_as + sizeof(uint32_t) * std::max<size_t>(ascale.size(), 1), A);
    const size_t total = align_up(off_v + std::max<size_t>(vals.size(), 16), A);

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
    plan->d_grp_ptr = reinterpret_cast<int32_t *>(base + off_gp);
    plan->d_grp_col = reinterpret_cast<int32_t *>(base + off_gc);
    plan->d_row_order = reinterpret_cast<int32_t *>(base + off_ro);
    plan->d_mix_order = reinterpret_cast<int32_t *>(base + off_mo);
    plan->d_ascale = fp8 ? reinterpret_cast<uint32_t *>(base + off_as) : nullptr;
    plan->d_vals = base + off_v;
    struct Copy {
        void *dst;
        const void *src;
        size_t bytes;
    } copies[] = {{plan->d_hdr, hdr, sizeof(int32_t) * (block_rows + 1)},
                  {plan->d_idx, idx, sizeof(int32_t) * static_cast<size_t>(nnzb)},
                  {plan->d_grp_ptr, grp_ptr.data(), sizeof(int32_t) * (dev_rows + 1)},
                  {plan->d_grp_col, grp_col.data(), sizeof(int32_t) * grp_col.size()},
                  {plan->d_row_order, row_order.data(), sizeof(int32_t) * dev_rows},
                  {plan->d_mix_order, mix_order.data(), sizeof(int32_t) * dev_rows},
                  {base + off_as, ascale.data(), sizeof(uint32_t) * ascale.size()},
                  {plan->d_vals, vals.data(), vals.size()}};
    for (const Copy &c : copies) {
        if (!c.bytes) continue;
        rc = hip_status(hipMemcpy(c.dst, c.src, c.bytes, hipMemcpyHostToDevice));
        if (rc != BCSR_OK) {
            (void)hipFree(plan->dev_base);
            delete plan;
            return rc;
        }
    }
    if (panel_ok) {
        plan->n_panels = (dev_rows + 15) / 16;
        plan->h_grp_col = grp_col;
        plan->h_nblk.resize(dev_rows);
        for (int r = 0; r < dev_rows; ++r)
            plan->h_nblk[r] = (hdr[r / sub_rows + 1] - hdr[r / sub_rows]) * pieces;
        plan->panel_mu = new std::mutex();
    }
    plan->rs = rs;
    plan->dev_rows = dev_rows;
    plan->group = G;
    plan->h_grp_ptr = std::move(grp_ptr);
    plan->h_row_order = std::move(row_order);
    if (fp8) plan->ws_mu = new std::mutex();
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
    inf.dev_pairs = groups;
    inf.device_bytes = static_cast<int64_t>(total);
    *out = plan;
    return BCSR_OK;
}

}  // namespace

int check_operands(const bcsr_plan *plan, const void *B, const float *C, int N) {
    if (!plan) return BCSR_ERR_INVALID_ARG;
    if (N < 0) return BCSR_ERR_INVALID_ARG;
    // the plan's arrays live on the device it was built on
    int dev = 0;
    const int rc = hip_status(hipGetDevice(&dev));
    if (rc != BCSR_OK) return rc;
    if (dev != plan->device) return BCSR_ERR_INVALID_ARG;
    if (N % 16 != 0) return BCSR_ERR_SHAPE;
    if (N == 0 || plan->info.rows == 0) return BCSR_OK;
    if (!B || !C) return BCSR_ERR_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(B) | reinterpret_cast<uintptr_t>(C)) & 15u) return BCSR_ERR_SHAPE;
    return BCSR_OK;
}

namespace {

int launch_spmm(const bcsr_plan *plan, const void *B, float *C, int N, const bcsr_launch_cfg *cfg,
                hipStream_t stream) {
    int rc = check_operands(plan, B, C, N);
    if (rc != BCSR_OK) return rc;
    if (N == 0 || plan->info.rows == 0) return BCSR_OK;
    LaunchShape ls;
    rc = resolve_launch(plan, N, cfg, ls);
    if (rc != BCSR_OK) return rc;
    if (plan->info.precision == BCSR_PREC_FP8_E4M3)
        return launch_fp8(const_cast<bcsr_plan *>(plan), ls, B, C, N, stream);
    if (ls.variant == 3) {
        rc = ensure_panel(const_cast<bcsr_plan *>(plan), stream);
        if (rc != BCSR_OK) return rc;
        return launch_panel(plan, ls, B, C, N, stream);
    }
    return launch_pairs(plan, ls, B, C, N, stream);
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

int resolve_launch(const bcsr_plan *plan, int N, const bcsr_launch_cfg *cfg, LaunchShape &ls) {
    const bool fp8 = plan->info.precision == BCSR_PREC_FP8_E4M3;
    // fraction of the bh x bw blocks that are stored
    const double density = static_cast<double>(plan->info.nnzb) /
                           std::max(1.0, static_cast<double>(plan->info.block_rows) * plan->info.block_cols);
    int tile_n = cfg ? cfg->tile_n : 0;
    bool few_items = false;  // set by the automatic tile choice
    if (tile_n == 0 && cfg && cfg->variant == 3) tile_n = 128;  // the panel kernel's only width
    if (tile_n == 0) {
        // fastest measured on MI355X at 16384 x 16384, 1-20% blocks:
        // 128-column tiles for 16/32-row device blocks; for 64-row blocks
        // 128 columns (plain kernel: the pipelined one does not fit) from
        // N = 512 on (5%: 35.6 vs 42.4 us, 20%: 89.5 vs 134 us) and 64
        // columns, pipelined, below (N = 256, 5%: 23.6 vs 30.0 us)
        tile_n = N >= 128 ? 128 : (N >= 64 ? 64 : 32);
        if (plan->rs == 4 && tile_n == 128 && N < 512) tile_n = 64;
        // 64-row blocks at N <= 128: 32-column tiles give twice the waves;
        // fp16 N = 64: 4.6 vs 5.2 us (1 %), 11.9 vs 13.1 (5 %), 32.5 vs 37.1
        // (20 %); N = 128 and FP8: 2-14 % faster
        if (plan->rs == 4 && N <= 128) tile_n = 32;
        // 16-row blocks in the 16-bit modes at 512 <= N <= 1024 with >= 8 %
        // blocks and >= 16384 rows: 256-column tiles (bench.py at
        // 16384^2, N = 512: 10 %: 81.3 vs 85.1 us, 20 %: 147.6 vs 152.9 us;
        // at 5 % 128 columns stay ahead, 49.5 vs 51.4 us).  Fewer rows give
        // too few work items for wider tiles (1280^2, N = 1280: 7.1 vs 5.5 us).
        const bool wide_ok = plan->rs <= 2 && plan->dev_rows * plan->rs >= 1024 && N >= 512 && N <= 1024;
        if (!fp8 && plan->rs == 1 && plan->info.precision != BCSR_PREC_F16_HILO && wide_ok && density >= 0.08)
            tile_n = 256;
        if (fp8) {
            // FP8, 32-row blocks at N = 128: 64-column tiles (18.1 vs 18.6 us
            // at 5 %, 28.3 vs 31.4 at 20 %); 16/32-row blocks in the range
            // above with >= 8 % blocks: 256-column tiles (whole calls at
            // 16384^2, 16-row blocks: 8 %, N = 512: 44.9 vs 46.9 us; 12 %:
            // 55.8 vs 60.7; N = 1024, 10 %: 86.4 vs 94.9; 32-row blocks,
            // N = 512, 10 %: 38.9 vs 41.3; at 5 % the widths are level)
            if (plan->rs == 2 && N > 64 && N <= 128) tile_n = 64;
            if (wide_ok && density >= 0.08) tile_n = 256;
        }
        // Products with few work items (device block-rows x column tiles
        // under kFewItems, about 2.5 per CU) leave CUs idle at these widths:
        // halve the tile until there are enough (or it is 32 wide).  Over the
        // 3500 generator matrices of the autotuner's timing database (512 to
        // 16384 rows and columns, 0.5-25 % blocks, N 32-512) this takes the
        // automatic choice's mean regret against exhaustive autotuning from
        // 15.2 % to 3.6 % (fp16) and from 6.4 % to 1.8 % (FP8); the 16384^2
        // products keep their widths.
        constexpr int64_t kFewItems = 640;
        few_items = static_cast<int64_t>(plan->dev_rows) * ((N + tile_n - 1) / tile_n) < kFewItems;
        while (tile_n > 32 && static_cast<int64_t>(plan->dev_rows) * ((N + tile_n - 1) / tile_n) < kFewItems)
            tile_n /= 2;
    }
    if (tile_n != 32 && tile_n != 64 && tile_n != 128 && tile_n != 256) return BCSR_ERR_INVALID_ARG;
    // automatic choice: the software-pipelined kernel where its register
    // footprint fits, else the plain one
    const bool auto_variant = !cfg || cfg->variant < 0;
    ls.variant = auto_variant ? 1 : cfg->variant;
    if (ls.variant == 3) {
        // the slab-streaming panel kernel: 128-column tiles, 1024 threads
        if (!panel_supported(plan) || tile_n != 128) return BCSR_ERR_UNSUPPORTED;
        if (static_cast<int64_t>(plan->info.cols) * N * 2 > INT32_MAX) return BCSR_ERR_UNSUPPORTED;
        ls.epl = 8;
        ls.n_tiles = (N + 127) / 128;
        const bool pin_ok = ls.n_tiles == 1 || ls.n_tiles == 2 || ls.n_tiles == 4 || ls.n_tiles % 8 == 0;
        int pmap = cfg ? cfg->xcd_map : -1;
        if (pmap < 0) pmap = pin_ok ? 1 : 0;
        if ((pmap == 1 && !pin_ok) || (pmap != 0 && pmap != 1)) return BCSR_ERR_INVALID_ARG;
        ls.map_mode = pmap;
        ls.wpb = 16;
        ls.grid = panel_grid(plan, ls.n_tiles, pmap);
        return BCSR_OK;
    }
    if (ls.variant > 2) return BCSR_ERR_UNSUPPORTED;
    if (!fp8 && !pair_shape_supported(plan->info.precision, plan->rs, tile_n / 16, ls.variant)) {
        if (!auto_variant || ls.variant != 1) return BCSR_ERR_UNSUPPORTED;
        ls.variant = 0;
        if (!pair_shape_supported(plan->info.precision, plan->rs, tile_n / 16, 0)) return BCSR_ERR_UNSUPPORTED;
    }
    if (fp8 && (ls.variant == 2 || !fp8_shape_supported(plan->rs, tile_n / 16))) return BCSR_ERR_UNSUPPORTED;
    // buffer offsets are 32-bit: larger B falls back to 64-bit global loads
    if (!fp8 && static_cast<int64_t>(plan->info.cols) * N * 2 > INT32_MAX) ls.variant = 2;
    if (fp8 && (static_cast<int64_t>(plan->info.cols) * N > INT32_MAX ||
                plan->info.dev_pairs * plan->rs * 2048 > INT32_MAX))
        return BCSR_ERR_UNSUPPORTED;
    ls.epl = tile_n / 16;
    ls.n_tiles = (N + tile_n - 1) / tile_n;
    const bool xcd_ok = ls.n_tiles == 1 || ls.n_tiles == 2 || ls.n_tiles == 4 || ls.n_tiles % 8 == 0;
    int map = cfg ? cfg->xcd_map : -1;
    if (map < 0) {
        // Column tiles pinned to XCDs keep each XCD's slice of B in its L2;
        // rows pinned fetch each block-row's A into one L2 for all tiles.
        // Measured on MI355X at 16384^2: rows win for 64-row device blocks
        // (N = 256: 1.15-1.34x, N = 512 at 20 %: 1.11x) and for dense
        // matrices with at most two tiles of <= 128 columns (20 %, N = 256:
        // 1.04-1.11x); tiles win elsewhere (16 x 16, 5 %, N = 512: 52 vs 84
        // us; two 256-column tiles at 20 %: 143 vs 165 us).  Tile counts that
        // 8 does not divide take the chunked runs (N = 384: 41.5 vs 59.1 us
        // in the linear order).
        // Products with few work items (above) do best with rows pinned in
        // the 16-bit modes and with the chunked runs in FP8 (the timing
        // database's most frequent winners there).
        if (few_items)
            map = fp8 ? 3 : 2;
        else if (plan->rs == 4 || (ls.n_tiles <= 2 && tile_n <= 128 && density >= 0.15))
            map = 2;
        else
            map = xcd_ok ? 1 : 3;
    }
    if (map == 1 && !xcd_ok) return BCSR_ERR_INVALID_ARG;
    if (map == 4 && ls.n_tiles % 2) return BCSR_ERR_INVALID_ARG;
    if (map < 0 || map > 4) return BCSR_ERR_INVALID_ARG;
    ls.map_mode = map;
    const int wpb = cfg && cfg->waves_per_block > 0 ? cfg->waves_per_block : 1;
    if (wpb != 1 && wpb != 2 && wpb != 4) return BCSR_ERR_INVALID_ARG;
    if (static_cast<int64_t>(ls.n_tiles) * plan->dev_rows + 8 > INT32_MAX) return BCSR_ERR_UNSUPPORTED;  // work items
    ls.wpb = wpb;
    const int rows = plan->dev_rows;
    if (map == 0) {
        ls.grid = (ls.n_tiles * rows + wpb - 1) / wpb;
    } else if (map == 2) {
        const int per_xcd = (rows + 7) / 8 * ls.n_tiles;  // items of one XCD
        ls.grid = 8 * ((per_xcd + wpb - 1) / wpb);
    } else if (map == 3) {
        const int per_xcd = static_cast<int>((static_cast<int64_t>(ls.n_tiles) * rows + 7) / 8);
        ls.grid = 8 * ((per_xcd + wpb - 1) / wpb);
    } else if (map == 4) {
        const int per_xcd = (rows + 3) / 4 * (ls.n_tiles / 2);
        ls.grid = 8 * ((per_xcd + wpb - 1) / wpb);
    } else if (ls.n_tiles >= 8) {
        const int per_xcd = ls.n_tiles / 8 * rows;  // items of one XCD
        ls.grid = 8 * ((per_xcd + wpb - 1) / wpb);
    } else {
        const int q = 8 / ls.n_tiles;
        const int per_xcd = (rows + q - 1) / q;
        ls.grid = 8 * ((per_xcd + wpb - 1) / wpb);
    }
    return BCSR_OK;
}

int ensure_panel(bcsr_plan *plan, hipStream_t stream) {
    if (!plan->panel_mu) return BCSR_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(*plan->panel_mu);
    if (plan->d_panel) return BCSR_OK;
    // an allocation and a synchronous upload cannot be recorded into a graph
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    int rc = hip_status(hipStreamIsCapturing(stream, &cap));
    if (rc != BCSR_OK) return rc;
    if (cap != hipStreamCaptureStatusNone) return BCSR_ERR_UNSUPPORTED;
    std::vector<int32_t> panel_ptr, panel_rec;
    try {
        build_panel_schedule(plan->h_grp_ptr, plan->h_grp_col, plan->h_nblk, plan->dev_rows, panel_ptr, panel_rec);
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
    const size_t off_rec = align_up(sizeof(int32_t) * panel_ptr.size(), 256);
    const size_t bytes = off_rec + sizeof(int32_t) * panel_rec.size();
    void *d = nullptr;
    rc = hip_status(hipMalloc(&d, bytes));
    if (rc != BCSR_OK) return rc;
    char *base = static_cast<char *>(d);
    rc = hip_status(hipMemcpy(base, panel_ptr.data(), sizeof(int32_t) * panel_ptr.size(), hipMemcpyHostToDevice));
    if (rc == BCSR_OK)
        rc = hip_status(
            hipMemcpy(base + off_rec, panel_rec.data(), sizeof(int32_t) * panel_rec.size(), hipMemcpyHostToDevice));
    if (rc != BCSR_OK) {
        (void)hipFree(d);
        return rc;
    }
    plan->d_panel_ptr = reinterpret_cast<int32_t *>(base);
    plan->d_panel_rec = reinterpret_cast<int32_t *>(base + off_rec);
    plan->panel_chunks = panel_ptr.back();
    plan->info.device_bytes += static_cast<int64_t>(bytes);
    plan->d_panel = d;  // set last: a plan whose d_panel is set is complete
    return BCSR_OK;
}

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
        case BCSR_ERR_SCRATCH_LIMIT: return "product exceeds the no-allocation scratch";
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
    if (plan->d_panel) {
        const int rc2 = hip_status(hipFree(plan->d_panel));
        if (rc == BCSR_OK) rc = rc2;
    }
    delete plan->panel_mu;
    if (plan->d_ws) plan->ws_retired.push_back(plan->d_ws);
    for (void *ws : plan->ws_retired) {
        int rc2 = hip_status(hipFree(ws));
        if (rc == BCSR_OK) rc = rc2;
    }
    delete plan->ws_mu;
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

int bcsr_spmm_b8(const bcsr_plan *plan, const void *B8, const void *scales, float *C, int32_t N,
                 const bcsr_launch_cfg *cfg, void *stream) {
    int rc = check_operands(plan, B8, C, N);
    if (rc != BCSR_OK) return rc;
    if (plan->info.precision != BCSR_PREC_FP8_E4M3) return BCSR_ERR_UNSUPPORTED;
    if (N == 0 || plan->info.rows == 0) return BCSR_OK;
    if (!scales) return BCSR_ERR_INVALID_ARG;
    if (reinterpret_cast<uintptr_t>(scales) & 15u) return BCSR_ERR_SHAPE;  // read 16 bytes at a time
    LaunchShape ls;
    rc = resolve_launch(plan, N, cfg, ls);
    if (rc != BCSR_OK) return rc;
    return launch_fp8_product(plan, ls, static_cast<const uint8_t *>(B8), static_cast<const uint8_t *>(scales), C, N,
                              static_cast<hipStream_t>(stream));
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

int bcsr_spmm_time(const bcsr_plan *plan, const void *B, float *C, int32_t N, const bcsr_launch_cfg *cfg,
                   int32_t warmup, int32_t reps, void *stream, float *ms) {
    if (!ms || reps <= 0 || warmup < 0) return BCSR_ERR_INVALID_ARG;
    int rc = bcsr_spmm_repeat(plan, B, C, N, cfg, warmup, stream);
    if (rc != BCSR_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = hip_status(hipEventCreate(&e0));
    if (rc == BCSR_OK) rc = hip_status(hipEventCreate(&e1));
    if (rc == BCSR_OK) rc = hip_status(hipEventRecord(e0, s));
    if (rc == BCSR_OK) rc = bcsr_spmm_repeat(plan, B, C, N, cfg, reps, stream);
    if (rc == BCSR_OK) rc = hip_status(hipEventRecord(e1, s));
    if (rc == BCSR_OK) rc = hip_status(hipEventSynchronize(e1));
    float t = 0.0f;
    if (rc == BCSR_OK) rc = hip_status(hipEventElapsedTime(&t, e0, e1));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == BCSR_OK) *ms = t / static_cast<float>(reps);
    return rc;
}

int bcsr_spmm_kernel_name(const bcsr_plan *plan, int32_t N, const bcsr_launch_cfg *cfg, char *buf,
                          size_t buflen) {
    if (!plan || !buf || buflen == 0) return BCSR_ERR_INVALID_ARG;
    LaunchShape ls;
    int rc = resolve_launch(plan, N, cfg, ls);
    if (rc != BCSR_OK) return rc;
    if (plan->info.precision == BCSR_PREC_FP8_E4M3)
        fp8_kernel_name(plan, ls, buf, buflen);
    else if (ls.variant == 3)
        panel_kernel_name(plan, buf, buflen);
    else
        pair_kernel_name(plan, ls, buf, buflen);
    return BCSR_OK;
}

int bcsr_copy_to_device(const int32_t *hdr, const int32_t *idx, const void *data_f16, int32_t block_rows,
                        int32_t block_cols, int32_t bh, int32_t bw, int **gpu_hdr, int **gpu_idx,
                        void **gpu_data) {
    if (!gpu_hdr || !gpu_idx || !gpu_data) return BCSR_ERR_INVALID_ARG;
    bcsr_plan *plan = nullptr;
    int rc = bcsr_plan_create(hdr, idx, data_f16, block_rows, block_cols, bh, bw, &plan);
    if (rc != BCSR_OK) return rc;
    try {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        registry()[plan->d_hdr] = plan;
    } catch (const std::bad_alloc &) {
        (void)bcsr_plan_destroy(plan);
        return BCSR_ERR_OUT_OF_MEMORY;
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
