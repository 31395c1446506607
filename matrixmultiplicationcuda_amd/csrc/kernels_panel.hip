// kernels_panel.hip -- the slab-streaming "panel" kernel for 16-row device
// blocks in fp16 / bf16 (no reference counterpart: the reference's kernels
// read B per block from global memory, src/cuda_kernels.cu:301-613).
//
// One 1024-thread workgroup owns a panel of 16 consecutive device block-rows
// (wave w = block-row 16 p + w) and one 128-column tile of C.  The columns
// that any of the 16 rows uses (the panel's union, in ascending order) are
// streamed through LDS in chunks of at most 8 slabs (16 rows x 128 columns
// of B, 4 KB each) together with the A blocks that fall in those slabs (at
// most 32 blocks of 512 bytes).  Every slab is loaded once per panel and
// read from LDS by each row that needs it, so B moves once per panel and
// tile instead of once per block.  Three stages of 49 KB are in flight; the
// loads go straight to LDS (buffer_load ... lds) and the only waits are one
// vmcnt + barrier per chunk.  Each stage also carries the record of the chunk
// two ahead, so the loop reads its metadata from LDS: the loads of chunk
// c + 2 are addressed from the record that arrived with chunk c, and each
// wave keeps its own row's fields of the next two chunks in registers.
// The kernel runs only when a launch configuration asks for variant 3: at
// the shapes measured on MI355X it is slower than the pair kernels, because
// its per-chunk load latency, barrier and LDS read burst add up instead of
// overlapping.
//
// A chunk record (kPanelRecInts = 96 ints, built on the host with the plan):
//   [0, 8)        block-columns of the chunk's slabs (-1: none)
//   [8]           number of A blocks in the chunk
//   [16, 48)      byte offsets of those blocks in the plan's vals (512-byte
//                 halves of the pair operands: rows 0..15, columns
//                 8 (l / 16) .. + 7 for l = 0..31), ordered by row, then column
//   [64 + 2 w]    row w: index of its first block in the chunk | count << 8
//   [65 + 2 w]    row w: slab slot (0..7) of its i-th block at bits 3 i
// A row's blocks in a chunk are taken two at a time into one 16x16x32 MFMA
// step (lanes 0..31 of the A operand = first block, 32..63 = second); an odd
// last block is paired with a zero record kept in LDS.
#include <algorithm>
#include <cstdio>

#include "pair_device.h"

namespace {

using namespace pairdev;

constexpr int kPanelRows = 16;
constexpr int kSlabBytes = 16 * 256;                      // 16 rows x 128 fp16 columns
constexpr int kStageB = kMaxSlabsPerChunk * kSlabBytes;   // 32 KB
constexpr int kStageRec = kStageB + kMaxBlocksPerChunk * 512;  // after 16 KB of A blocks
constexpr int kStageBytes = kStageRec + 1024;                 // + one 1 KB load of a record
constexpr int kStages = 3;
constexpr int kDummyOff = kStages * kStageBytes;  // target of loads that carry nothing
constexpr int kZeroOff = kDummyOff + 1024;
constexpr int kLdsBytes = kZeroOff + 512;
constexpr int kLoadsPerChunk = 4;  // per thread: two B pieces, one A pair, one record piece
static_assert(kPanelRecInts * 4 <= 1024, "a record fits one 1 KB load");

typedef __attribute__((address_space(3))) unsigned char lds_u8;

// 16 bytes per lane from a buffer straight into LDS at lds_ptr + 16 * lane.
// The asm has no memory clobber, so the compiler neither waits for it nor
// reorders LDS reads around it on its account; the kernel waits explicitly
// (vmcnt + barrier) before it reads a stage.  No VGPR is written.
__device__ __forceinline__ void lds_dma16(__amdgpu_buffer_rsrc_t rsrc, unsigned voff, const unsigned char *lds_ptr) {
    const unsigned m0v = __builtin_amdgcn_readfirstlane(
        static_cast<unsigned>(reinterpret_cast<uintptr_t>((const lds_u8 *)lds_ptr)));
    asm volatile("s_mov_b32 m0, %0\n\tbuffer_load_dwordx4 %1, %2, 0 offen lds" ::"s"(m0v), "v"(voff), "s"(rsrc));
}

// s_waitcnt vmcnt(n) leaving expcnt and lgkmcnt alone (gfx9 encoding)
#define PANEL_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n)&15) | (((n) >> 4) << 14) | (7 << 4) | (15 << 8))

// Chunk records and panel ranges are never written while a kernel runs:
// reading them through the constant address space makes them scalar loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *cmem(const T *p) {
    return (const __attribute__((address_space(4))) T *)p;
}

// Workgroup -> (panel, tile): every column tile pinned to a fixed set of
// XCDs as in the pair kernels (workgroups go round-robin over the 8 XCDs).
__device__ __forceinline__ bool map_panel(const PanelArgs &a, int &tile, int &panel) {
    const int bid = blockIdx.x;
    if (a.map_mode == 0) {
        tile = bid % a.n_tiles;
        panel = bid / a.n_tiles;
    } else {
        const int x = bid & 7, i = bid >> 3;
        if (a.n_tiles >= 8) {
            const int per = a.n_tiles >> 3;
            tile = x + 8 * (i % per);
            panel = i / per;
        } else {
            const int q = 8 / a.n_tiles;
            tile = x % a.n_tiles;
            panel = i * q + x / a.n_tiles;
        }
    }
    return panel < a.panels;
}

template <int PREC>
__global__ __launch_bounds__(1024) void bcsr_spmm_panel(PanelArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    int tile_in, panel_in;
    if (!map_panel(a, tile_in, panel_in)) return;  // whole workgroup: no barrier reached
    const int tile = __builtin_amdgcn_readfirstlane(tile_in);
    const int panel = __builtin_amdgcn_readfirstlane(panel_in);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int g = lane >> 4, j = lane & 15;
    const int c0 = cmem(a.chunk_ptr)[panel];
    const int nch = cmem(a.chunk_ptr)[panel + 1] - c0;
    const int br = panel * kPanelRows + w;
    const unsigned row_bytes = static_cast<unsigned>(a.N) * 2u;
    const __amdgpu_buffer_rsrc_t brs = make_rsrc(a.B, a.b_bytes);
    const __amdgpu_buffer_rsrc_t ars = make_rsrc(a.vals, a.a_bytes);

    if (threadIdx.x < 32) *reinterpret_cast<u32x4 *>(smem + kZeroOff + 16 * threadIdx.x) = u32x4{0u, 0u, 0u, 0u};

    // B piece q = 16 m + w (m = 0, 1) of a chunk: slab q / 4, rows
    // 4 (q % 4) .. + 3; lane -> row + lane / 16, columns 8 (lane % 16) .. + 7
    const unsigned b_lane = static_cast<unsigned>(lane >> 4) * row_bytes + static_cast<unsigned>(tile) * 256u +
                            static_cast<unsigned>(lane & 15) * 16u;
    const __amdgpu_buffer_rsrc_t rrs = make_rsrc(a.records, a.rec_bytes);
    // Loads of chunk c into stage c % 3, addressed by the record fields
    // (slab columns of pieces w and 16 + w, block count, sources of A blocks
    // 2 w and 2 w + 1), plus the record of chunk c + 2 (wave 0; the other
    // waves' fourth load carries nothing).  Past the panel's last chunk every
    // load is out of range: zeros, no memory traffic.
    auto issue = [&](int c, int col_a, int col_b, int nblk, int src0, int src1) {
        unsigned char *stage = smem + (c % kStages) * kStageBytes;
        const bool live = c < nch;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int q = 16 * m + w;
            const int col = live ? (m ? col_b : col_a) : -1;
            const unsigned off = col >= 0 ? static_cast<unsigned>(col * 16 + 4 * (q & 3)) * row_bytes + b_lane
                                          : kOutOfRange;
            lds_dma16(brs, off, stage + q * 1024);
        }
        const int r = 2 * w + (lane >> 5);
        const int src = (lane >> 5) ? src1 : src0;
        const unsigned aoff = live && r < nblk ? static_cast<unsigned>(src) + 16u * (lane & 31) : kOutOfRange;
        lds_dma16(ars, aoff, stage + kStageB + w * 1024);
        const bool rec_live = w == 0 && c + 2 < nch && lane < kPanelRecInts / 4;
        const unsigned roff =
            rec_live ? static_cast<unsigned>(c0 + c + 2) * (kPanelRecInts * 4u) + 16u * lane : kOutOfRange;
        lds_dma16(rrs, roff, w == 0 ? stage + kStageRec : smem + kDummyOff);
    };

    Acc<1, 8, PREC> acc;
#pragma unroll
    for (int e = 0; e < 8; ++e) acc.v[0][0][e] = f32x4{0.f, 0.f, 0.f, 0.f};

    // chunks 0 and 1: fields from scalar loads (a record index clamped to a
    // valid record, the spare one when the panel has no chunks)
    int info0, slots0, info1, slots1;
    {
        const __attribute__((address_space(4))) int *r0 = cmem(a.records) + static_cast<size_t>(c0) * kPanelRecInts;
        const __attribute__((address_space(4))) int *r1 = r0 + (nch > 1 ? kPanelRecInts : 0);
        info0 = r0[64 + 2 * w];
        slots0 = r0[65 + 2 * w];
        info1 = r1[64 + 2 * w];
        slots1 = r1[65 + 2 * w];
        issue(0, r0[w >> 2], r0[4 + (w >> 2)], r0[8], r0[16 + 2 * w], r0[17 + 2 * w]);
        issue(1, r1[w >> 2], r1[4 + (w >> 2)], r1[8], r1[16 + 2 * w], r1[17 + 2 * w]);
    }
    for (int c = 0; c < nch; ++c) {
        PANEL_WAIT_VMCNT(kLoadsPerChunk);  // chunk c's loads done; chunk c + 1's may be in flight
        __syncthreads();                   // ... in every wave, and the stage of chunk c - 1 is free
        const unsigned char *stage = smem + (c % kStages) * kStageBytes;
        // the record of chunk c + 2 arrived with chunk c
        const int *rec2 = reinterpret_cast<const int *>(stage + kStageRec);
        const int2 src2 = *reinterpret_cast<const int2 *>(rec2 + 16 + 2 * w);
        const int2 row2 = *reinterpret_cast<const int2 *>(rec2 + 64 + 2 * w);
        issue(c + 2, rec2[w >> 2], rec2[4 + (w >> 2)], rec2[8], src2.x, src2.y);
        const int info = info0, slots = slots0;
        info0 = info1;
        slots0 = slots1;
        info1 = row2.x;
        slots1 = row2.y;
        const int first = info & 0xff, nb = (info >> 8) & 0xff;
        for (int b = 0; b < nb; b += 2) {
            const int r0 = first + b;
            const bool two = b + 1 < nb;
            const int s0 = (slots >> (3 * b)) & 7;
            const int s1 = two ? (slots >> (3 * b + 3)) & 7 : s0;
            const unsigned char *a_ptr =
                (lane >> 5) ? (two ? stage + kStageB + (r0 + 1) * 512 : smem + kZeroOff) : stage + kStageB + r0 * 512;
            AOps<1, PREC> av;
            av.v[0][0] = *reinterpret_cast<const u32x4 *>(a_ptr + 16 * (lane & 31));
            const unsigned char *b_ptr = stage + ((g >> 1) ? s1 : s0) * kSlabBytes + (g & 1) * 8 * 256 + 16 * j;
            BRow<8> raw[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(b_ptr + i * 256);
                raw[i].d[0] = v.x;
                raw[i].d[1] = v.y;
                raw[i].d[2] = v.z;
                raw[i].d[3] = v.w;
            }
            consume<1, 8, PREC>(acc, raw, av);
        }
    }
    PANEL_WAIT_VMCNT(0);  // the trailing (out-of-range) loads land before the workgroup's LDS is released
    const int n0 = tile * 128 + 8 * j;
    if (br >= a.dev_rows || n0 >= a.N) return;
    store_tile<1, 8, PREC>(acc, a.C, static_cast<size_t>(a.N), br, g, n0);
}

}  // namespace

bool panel_supported(const bcsr_plan *plan) {
    return plan->n_panels > 0 && (plan->info.precision == BCSR_PREC_F16 || plan->info.precision == BCSR_PREC_BF16);
}

void panel_kernel_name(const bcsr_plan *plan, char *buf, size_t len) {
    std::snprintf(buf, len, "bcsr_spmm_panel<%s>", plan->info.precision == BCSR_PREC_BF16 ? "bf16" : "f16");
}

int panel_grid(const bcsr_plan *plan, int n_tiles, int map_mode) {
    const int panels = plan->n_panels;
    if (map_mode == 0) return n_tiles * panels;
    if (n_tiles >= 8) return 8 * (n_tiles / 8 * panels);
    const int q = 8 / n_tiles;
    return 8 * ((panels + q - 1) / q);
}

int launch_panel(const bcsr_plan *plan, const LaunchShape &ls, const void *B, float *C, int N, hipStream_t stream) {
    if (!panel_supported(plan) || ls.epl != 8) return BCSR_ERR_UNSUPPORTED;
    if (static_cast<int64_t>(plan->info.cols) * N * 2 > INT32_MAX) return BCSR_ERR_UNSUPPORTED;
    PanelArgs a{};
    a.chunk_ptr = plan->d_panel_ptr;
    a.records = plan->d_panel_rec;
    a.vals = plan->d_vals;
    a.B = static_cast<const _Float16 *>(B);
    a.C = C;
    a.N = N;
    a.n_tiles = ls.n_tiles;
    a.panels = plan->n_panels;
    a.dev_rows = plan->dev_rows;
    a.map_mode = ls.map_mode;
    a.b_bytes = static_cast<int>(static_cast<int64_t>(plan->info.cols) * N * 2);
    a.a_bytes = static_cast<int>(plan->info.dev_pairs * 1024);
    a.rec_bytes = static_cast<int>((plan->panel_chunks + 1) * kPanelRecInts * 4);
    auto k = plan->info.precision == BCSR_PREC_BF16 ? bcsr_spmm_panel<PREC_BF16> : bcsr_spmm_panel<PREC_F16>;
    hipLaunchKernelGGL(k, dim3(ls.grid), dim3(1024), kLdsBytes, stream, a);
    return hip_status(hipGetLastError());
}
