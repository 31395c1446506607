// bcsr_internal.h -- types shared by the plan builder and the kernel files.
//
// Device layout of a plan
//   A bh x bw BCSR matrix is viewed as dev_bh x 16 device blocks (dev_bh = 64,
//   32 or 16, the largest that divides bh; bw is cut into 16-column pieces).
//   A device block-row's blocks (ascending block-column) are taken G at a time
//   (G = 2 "pairs" for the 16-bit modes: one K = 32 step of the 16x16x32 MFMA;
//   G = 8 "octets" for FP8: one K = 128 step of the scaled 16x16x128 MFMA);
//   short groups are padded with zero blocks that repeat the last column.
//     grp_ptr[dev_rows + 1]  prefix count of groups per device block-row
//     grp_col[G * groups]    16-column block index of every group slot
//     vals                   the A operands, 16 bytes per lane and operand:
//                            pairs: [group][row-subtile s][lane l][8 x 16-bit]
//                              = rows 16 s + l % 16, columns 8 (l / 16) .. + 7
//                              of the pair's 16 x 32 slice (hi+lo mode stores
//                              [group][s][hi, lo][lane]);
//                            octets: [group][s][lane l][32 x fp8]: bytes
//                              16 h .. 16 h + 15 of lane l = row 16 s + l % 16
//                              of the block in slot 4 h + l / 16;
//     ascale                 FP8 only: per (group, s) the E8M0 scales of the
//                            four slot pairs (2 u, 2 u + 1), byte u;
//     row_order[dev_rows]    device block-rows by descending group count.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "bcsr_mi355x.h"
#include "panel_schedule.h"

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x8 __attribute__((ext_vector_type(8)));

struct bcsr_plan {
    bcsr_plan_info info{};
    int device = 0;
    int rs = 1;        // dev_bh / 16
    int dev_rows = 0;  // M / dev_bh
    int group = 2;     // blocks per group: 2 (pairs) or 8 (FP8 octets)
    void *dev_base = nullptr;
    int32_t *d_hdr = nullptr;  // reference prefix array (API view)
    int32_t *d_idx = nullptr;  // reference block-column indices (API view)
    int32_t *d_grp_ptr = nullptr;
    int32_t *d_grp_col = nullptr;
    int32_t *d_row_order = nullptr;
    int32_t *d_mix_order = nullptr;  // row_order taken 32 longest, 32 shortest, next 32 longest, ...
    void *d_vals = nullptr;
    uint32_t *d_ascale = nullptr;
    // FP8: workspace holding the quantised B of the current call.  A
    // workspace outgrown by a larger N is kept (ws_retired) until the plan is
    // destroyed: a graph that captured an earlier call still points at it,
    // and freeing it would synchronise the device.
    std::mutex *ws_mu = nullptr;
    void *d_ws = nullptr;
    size_t ws_bytes = 0;
    std::vector<void *> ws_retired;
    // host copies of grp_ptr and row_order (work-item construction)
    std::vector<int32_t> h_grp_ptr, h_row_order;
    // panel schedule of launch variant 3 (16-row device blocks, fp16 / bf16
    // only; n_panels = 0 otherwise): chunk ranges per panel of 16 device
    // block-rows and one record of kPanelRecInts ints per chunk (layout in
    // kernels_panel.hip).  Built by the first variant-3 launch (ensure_panel),
    // from the host copies of the pair columns and block counts kept for it;
    // until then d_panel is null.
    int n_panels = 0;
    std::mutex *panel_mu = nullptr;
    std::vector<int32_t> h_grp_col, h_nblk;
    void *d_panel = nullptr;
    int32_t *d_panel_ptr = nullptr;
    int32_t *d_panel_rec = nullptr;
    int64_t panel_chunks = 0;
};


struct PanelArgs {
    const int32_t *chunk_ptr;  // panels + 1
    const int32_t *records;    // kPanelRecInts per chunk (+ one spare)
    const void *vals;          // the plan's pair operands
    const _Float16 *B;
    float *C;
    int N, n_tiles, panels, dev_rows, map_mode;
    int b_bytes, a_bytes, rec_bytes;
};

// Launch geometry resolved from a bcsr_launch_cfg.
struct LaunchShape {
    int variant = 0;
    int wpb = 1;  // waves (work items) per workgroup
    int epl = 8;  // output columns per lane; tile width = 16 * epl
    int map_mode = 0;
    int n_tiles = 1;
    int grid = 0;
};

struct SpmmArgs {
    const int32_t *grp_ptr;
    const int32_t *grp_col;
    const u32x4 *vals;
    const uint32_t *ascale;
    const int32_t *row_order;
    const _Float16 *B;       // fp16 B (pair kernels)
    const uint8_t *B8;       // quantised B, k-minor panels (FP8 kernel)
    const uint8_t *b8scale;  // E8M0 per (16-row slab, 256-column chunk), chunk-major
    float *C;
    int N;
    int n_tiles;
    int dev_rows;
    int map_mode;
    int b_bytes;  // bytes of B (or B8) when below 2^31: buffer-load bound
    int b_chunks;  // 256-column chunks of B8 scales
    int b_slabs;   // 16-row slabs of B8 (K / 16)
    int wpb;       // waves per workgroup
    int a_bytes;   // bytes of the FP8 A operands (buffer-load bound)
};

int hip_status(hipError_t e);
// The plan's device is current, N % 16 == 0, B and C present and 16-byte
// aligned (BCSR_OK also for the empty products N == 0 or M == 0).
int check_operands(const bcsr_plan *plan, const void *B, const float *C, int N);
// The FP8 mode's one-pass B quantiser (B8 and scales laid out as for
// bcsr_quantize_b_e4m3); K > 0.
int enqueue_quant(const void *B, int K, int N, uint8_t *B8, uint8_t *scales, hipStream_t stream);
// Bytes of the B scales of a K x N FP8 product: one per (16-row slab,
// 256-column chunk), each chunk's row padded to 16 bytes.
size_t b8_scale_bytes(int K, int N);
int resolve_launch(const bcsr_plan *plan, int N, const bcsr_launch_cfg *cfg, LaunchShape &ls);
int launch_pairs(const bcsr_plan *plan, const LaunchShape &ls, const void *B, float *C, int N,
                 hipStream_t stream);
int launch_fp8(bcsr_plan *plan, const LaunchShape &ls, const void *B, float *C, int N, hipStream_t stream);
// The FP8 product on an already quantised B (bcsr_quantize_b_e4m3's layout).
int launch_fp8_product(const bcsr_plan *plan, const LaunchShape &ls, const uint8_t *B8, const uint8_t *bscale,
                       float *C, int N, hipStream_t stream);
bool pair_shape_supported(int precision, int rs, int epl, int variant);
// FP8 product shapes (row sets x 16-column blocks per tile) with a kernel:
// all but 64-row blocks at 256 columns, whose 256 accumulator registers leave
// no room for the operands (the compiler spilled it, and now fails on it)
constexpr bool fp8_shape_supported(int rs, int epl) { return !(rs == 4 && epl == 16); }
// The slab-streaming panel kernel (variant 3).
bool panel_supported(const bcsr_plan *plan);
int panel_grid(const bcsr_plan *plan, int n_tiles, int map_mode);
int launch_panel(const bcsr_plan *plan, const LaunchShape &ls, const void *B, float *C, int N, hipStream_t stream);
// Builds the plan's panel schedule if it has none yet (a synchronous upload;
// BCSR_ERR_UNSUPPORTED while `stream` is being captured into a graph).
int ensure_panel(bcsr_plan *plan, hipStream_t stream);
void panel_kernel_name(const bcsr_plan *plan, char *buf, size_t len);
void pair_kernel_name(const bcsr_plan *plan, const LaunchShape &ls, char *buf, size_t len);
void fp8_kernel_name(const bcsr_plan *plan, const LaunchShape &ls, char *buf, size_t len);

// Work item of the calling wave: (column tile, slot in row_order).  Wave w of
// workgroup b is item (b >> 3) * wpb + w of XCD b & 7 (workgroups are
// dispatched round-robin over the 8 XCDs).  map_mode 0 ignores the XCD and
// numbers items b * wpb + w, tile fastest; map_mode 1 pins every column tile
// to XCDs: n_tiles % 8 == 0 gives XCD x the tiles x, x + 8, ...; n_tiles in
// {1, 2, 4} lets 8 / n_tiles XCDs share each tile.  map_mode 2 pins rows
// instead: XCD x takes the slots x, x + 8, ... with all their column tiles,
// so a block-row's A is fetched into one XCD's L2 for every tile.  map_mode 3
// is for tile counts that 8 does not divide: the items in tile-major order
// (rows in the plan's mixed order, so that any run of them, and any CU's
// share of it, carries about the average work) are cut into 8 equal runs,
// one per XCD, so each XCD serves at most two column tiles.  map_mode 4 (even
// tile counts) is a 2 x 4 grid of XCDs: XCD x takes the column tiles of half
// x & 1 for every fourth slot, starting at x >> 1, the tiles of one row
// next to each other, so they share that row's A in the XCD's L2.
__device__ __forceinline__ bool map_work(const SpmmArgs &a, int &tile, int &slot) {
    // threadIdx.x >> 6 is the same in every lane of a wave; readfirstlane
    // tells the compiler so, which keeps everything derived from it (row,
    // group range, column indices) in scalar registers and scalar loads
    const int bid = blockIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (a.map_mode == 0) {
        const int item = bid * a.wpb + w;
        tile = item % a.n_tiles;
        slot = item / a.n_tiles;
    } else if (a.map_mode == 2) {
        const int x = bid & 7, i = (bid >> 3) * a.wpb + w;
        tile = i % a.n_tiles;
        slot = (i / a.n_tiles) * 8 + x;
    } else if (a.map_mode == 3) {
        const int x = bid & 7, i = (bid >> 3) * a.wpb + w;
        const int total = a.n_tiles * a.dev_rows, per = (total + 7) / 8;  // total < 2^31 (checked at launch)
        if (i >= per) return false;
        const int64_t g = static_cast<int64_t>(x) * per + i;
        if (g >= total) return false;
        tile = static_cast<int>(g / a.dev_rows);
        slot = static_cast<int>(g % a.dev_rows);
    } else if (a.map_mode == 4) {
        const int x = bid & 7, i = (bid >> 3) * a.wpb + w;
        const int tpg = a.n_tiles >> 1;
        tile = (x & 1) * tpg + i % tpg;
        slot = (i / tpg) * 4 + (x >> 1);
    } else {
        const int x = bid & 7, i = (bid >> 3) * a.wpb + w;
        if (a.n_tiles >= 8) {
            const int per = a.n_tiles >> 3;
            tile = x + 8 * (i % per);
            slot = i / per;
        } else {
            const int q = 8 / a.n_tiles;
            tile = x % a.n_tiles;
            slot = i * q + x / a.n_tiles;
        }
    }
    return slot < a.dev_rows;
}

// Offsets at or above this are out of range for every buffer (num_records is
// at most INT32_MAX): a load there returns zeros without touching memory.
// Only voffset (plus the instruction offset) is range-checked, so kernels keep
// soffset strides for loads already known to be in bounds.
constexpr unsigned kOutOfRange = 0x80000000u;

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

// p[i] read through the constant address space, so that a wave-uniform
// address always becomes a scalar load.  The plan's row pointers, block
// columns and A scales are written before any launch and never by a kernel;
// read as generic pointers, the compiler proves that only while the kernel
// holds no store it cannot place, and otherwise turns these reads into vector
// loads that wait for every load in flight.
template <typename T>
__device__ __forceinline__ T const_load(const T *p, size_t i) {
    return reinterpret_cast<const __attribute__((address_space(4))) T *>(reinterpret_cast<uintptr_t>(p))[i];
}

// p as a global-memory pointer.  A pointer that a kernel reads from memory
// (the batched kernel's per-problem arguments) is generic to the compiler,
// and its loads become flat loads, which count on both the vector-memory and
// the scalar-memory counters: waiting for one then waits for every load in
// flight.  Global loads count on the vector-memory counter only.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *global_ptr(const T *p) {
    return reinterpret_cast<const __attribute__((address_space(1))) T *>(reinterpret_cast<uintptr_t>(p));
}
// (stores: a flat store also counts on both counters, so the next scalar
// load's wait waits for the stores in flight)
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *global_ptr(T *p) {
    return reinterpret_cast<__attribute__((address_space(1))) T *>(reinterpret_cast<uintptr_t>(p));
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, int bytes) {
    // gfx9 buffer descriptor word 3; num_records = bytes (reads past it return 0)
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), static_cast<short>(0), bytes, 0x00020000);
}
