// fp8_device.h -- device code of the FP8 E4M3 storage mode on the scaled MFMA
// v_mfma_scale_f32_16x16x128_f8f6f4 (no reference counterpart: the
// reference's kernels are fp16 only).
//
// The scaled 16x16x128 MFMA reads, per lane, 32 bytes of A (row l % 16) and
// of B (column l % 16).  K index of byte b in lane group g (lanes 16 g ..
// 16 g + 15) is 64 (b >> 4) + 16 g + (b & 15), and the MFMA applies one E8M0
// scale per 32 consecutive K (scale group t = K 32 t .. 32 t + 31 = bytes
// 16 (t >> 1) .. + 15 of lane groups 2 (t & 1) and 2 (t & 1) + 1), taken from
// lane 16 t + row (A) / column (B) of the scale operands.
//
// Layout that follows: an octet is 8 consecutive 16 x 16 device blocks of a
// device block-row; block slot s sits in lane group s & 3, byte half s >> 2,
// so slots (2 t, 2 t + 1) form scale group t and share one A scale (the
// scale of a 16 x 32 tile).  B is quantised per call in one pass, with one
// E8M0 scale per (16-row slab, 256-column chunk), stored chunk-major with
// rows padded to 16 bytes: bscale[chunk * scale_stride(slabs) + slab].
// Where every slab of a chunk has the same scale (checked per row at its
// start), any two slabs can share a scale group and an octet is one MFMA per
// 16 columns; otherwise each octet takes two, one over its even slots (A's
// odd lane groups zero) with the even slabs' scales and one over its odd
// slots with theirs, in the same pipelined loop.  B is stored k-minor,
// B8[slab][n][16]: the 16 rows of slab c for column n are 16 consecutive
// bytes, exactly one lane's half operand.  All 64 lanes carry data: one octet
// step moves 8 blocks through one MFMA per 16 columns.
#pragma once

#include "bcsr_internal.h"

namespace fp8dev {

constexpr int kChunk = 256;  // columns per B scale

// Bytes from one chunk's slab scales to the next (16-byte aligned rows).
__host__ __device__ constexpr int scale_stride(int slabs) { return (slabs + 15) & ~15; }

// Smallest power of two 2^k (k in [-127, 127]) with absmax / 2^k <= 448, as
// the E8M0 byte k + 127; 127 for an all-zero group.
__device__ __forceinline__ unsigned e8m0_for_absmax(float amax) {
    if (!(amax > 0.0f)) return 127u;
    int e;
    (void)frexpf(amax, &e);  // amax = m 2^e, m in [0.5, 1)
    int k = (ldexpf(448.0f, e - 9) >= amax) ? e - 9 : e - 8;
    k = k < -127 ? -127 : (k > 127 ? 127 : k);
    return static_cast<unsigned>(k + 127);
}

// Quantiser geometry: one 128-thread workgroup per (pair of 16-row slabs
// 2 x, 2 x + 1, 256-column chunk h).  Thread t takes row quad m = t & 3 and
// the 8 columns 256 h + 8 (t >> 2) .. + 7 of both slabs: eight 16-byte loads
// issued together, each instruction reading 1 KB (4 rows x 256 contiguous
// bytes).  N % 16 == 0, so a thread's 8 columns are all in or all out.
struct SlabPart {
    u32x4 v[2][4];  // slab 2 x + u, row 4 m + k: columns n .. n + 7
    int n, m;
    bool ok, ok1;   // columns exist; second slab exists
};

__device__ __forceinline__ void load_slab_part(const _Float16 *__restrict__ B, int N, int slabs, int x, int h,
                                               SlabPart &sp) {
    const int t = threadIdx.x;
    sp.m = t & 3;
    sp.n = h * kChunk + 8 * (t >> 2);
    sp.ok = sp.n < N;
    sp.ok1 = 2 * x + 1 < slabs;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = sp.ok && (u == 0 || sp.ok1);
            sp.v[u][k] = ok ? *global_ptr(reinterpret_cast<const u32x4 *>(
                                  B + static_cast<size_t>((2 * x + u) * 16 + 4 * sp.m + k) * N + sp.n))
                            : u32x4{0u, 0u, 0u, 0u};
        }
}

__device__ __forceinline__ float half_at(const u32x4 &v, int e) {
    const unsigned d = v[e >> 1];
    const auto bits = static_cast<unsigned short>((e & 1) ? d >> 16 : d & 0xffffu);
    return static_cast<float>(__builtin_bit_cast(_Float16, bits));
}

// Maximum over a 64-lane wave, then over the two waves of the workgroup
// (every thread gets it).
__device__ __forceinline__ float block_max128(float v, float *wmax) {
#pragma unroll
    for (int b = 32; b >= 1; b >>= 1) v = fmaxf(v, __shfl_xor(v, b, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    v = fmaxf(wmax[0], wmax[1]);
    __syncthreads();  // wmax is reused by the next call
    return v;
}

// One pass: slabs 2 x, 2 x + 1 of chunk h, each with its own scale (the
// absmax of its 16 rows x 256 columns) in bscale[h * scale_stride + slab].  Each
// thread turns its 4 rows x 8 columns into 8 dwords (4 rows of one column
// each) and puts them in shared memory in the k-minor record order (column
// record p(c) = c ^ ((c >> 3) & 7) spreads the writes over the banks); thread
// t then writes the 16-byte records of columns t and t + 128, so the
// workgroup's global stores are contiguous.
__device__ __forceinline__ void quant_slab(const _Float16 *__restrict__ B, uint8_t *__restrict__ B8,
                                           uint8_t *__restrict__ scale, int N, int slabs, int x, int h) {
    __shared__ float wmax[2];
    __shared__ unsigned rec[2][kChunk * 4];
    SlabPart sp;
    load_slab_part(B, N, slabs, x, h, sp);
    float inv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        float m = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(half_at(sp.v[u][k], e)));
        m = block_max128(m, wmax);
        const unsigned se = e8m0_for_absmax(m);
        if (threadIdx.x == 0 && (u == 0 || sp.ok1))
            global_ptr(scale)[static_cast<size_t>(h) * scale_stride(slabs) + 2 * x + u] = static_cast<uint8_t>(se);
        inv[u] = ldexpf(1.0f, 127 - static_cast<int>(se));
    }
    const int g8 = threadIdx.x >> 2;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int q = __builtin_amdgcn_cvt_pk_fp8_f32(half_at(sp.v[u][0], e) * inv[u], half_at(sp.v[u][1], e) * inv[u],
                                                    0, false);
            q = __builtin_amdgcn_cvt_pk_fp8_f32(half_at(sp.v[u][2], e) * inv[u], half_at(sp.v[u][3], e) * inv[u], q,
                                                true);
            const int col = 8 * g8 + e;
            const int p = col ^ ((col >> 3) & 7);
            rec[u][4 * p + sp.m] = static_cast<unsigned>(q);
        }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        if (u == 1 && !sp.ok1) break;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int col = threadIdx.x + 128 * half;
            const int p = col ^ ((col >> 3) & 7);
            const int n = h * kChunk + col;
            if (n < N)
                *global_ptr(reinterpret_cast<u32x4 *>(B8 + (static_cast<size_t>(2 * x + u) * N + n) * 16)) =
                    u32x4{rec[u][4 * p], rec[u][4 * p + 1], rec[u][4 * p + 2], rec[u][4 * p + 3]};
        }
    }
    __syncthreads();  // rec is reused by the next call
}

// The scale shared by every slab of chunk h, or -1 if the slabs' scales
// differ (wave-uniform), in two steps: scale_row_load issues the read of
// the chunk's first 1024 slab scales (16 per lane, from the chunk's 16-byte
// aligned row; the buffer bound is the padded row, since buffer loads check
// bounds per dword), scale_row_uniform compares them (and reads the rest of
// a longer row).  Lane 0 always holds slab 0; the chunk is uniform if no
// lane saw a scale other than slab 0's (one ballot, no cross-lane shuffles).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t scale_row_rsrc(const SpmmArgs &a, int h) {
    const int stride = scale_stride(a.b_slabs);
    return make_rsrc(a.b8scale + static_cast<size_t>(h) * stride, stride);
}

__device__ __forceinline__ u32x4 scale_row_load(const SpmmArgs &a, int h) {
    return __builtin_amdgcn_raw_buffer_load_b128(scale_row_rsrc(a, h), 16 * (threadIdx.x & 63), 0, 0);
}

__device__ __forceinline__ int scale_row_uniform(const SpmmArgs &a, int h, u32x4 v) {
    const int slabs = a.b_slabs, lane = threadIdx.x & 63;
    const int s0 = __builtin_amdgcn_readfirstlane(static_cast<int>(v.x & 0xffu));
    bool differs = false;
    for (int base = 0;;) {
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const int sc = static_cast<int>((v[b >> 2] >> (8 * (b & 3))) & 0xffu);
            differs |= base + 16 * lane + b < slabs && sc != s0;
        }
        base += 1024;
        if (base >= slabs) break;
        v = __builtin_amdgcn_raw_buffer_load_b128(scale_row_rsrc(a, h), base + 16 * lane, 0, 0);
    }
    return __builtin_amdgcn_ballot_w64(differs) ? -1 : s0;
}

// Work item (device block-row br, tile of 16 * EPL columns).  MFMA e covers
// columns tile + 16 e + j for lane j of each lane group.  The B loads of the
// row's octets are issued half an octet ahead of the MFMAs that use them
// (two half-octet buffers), A one octet ahead, so loads stay in flight across
// octet boundaries.  A load past the row's end aims at an out-of-range
// offset (zeros, no traffic); with one row set and tiles of 64 columns or
// more the loop leaves before issuing any (see below).
//
// kSplitA chooses how a chunk whose slab scales differ feeds A to its two
// MFMAs: split by lane group as A loads (one row set; 153 VGPRs at 128
// columns, and 3 % faster than masking at 256), or masked at the MFMA (more
// row sets, and the batched kernel, where two more A buffers push its widest
// shapes past 256 VGPRs into spills).
template <int RS, int EPL, bool kSplitA = (RS == 1)>
__device__ __forceinline__ void fp8_work(const SpmmArgs &a, int br_in, int tile_in) {
    // wave-uniform by construction (scalar index chain)
    const int br = __builtin_amdgcn_readfirstlane(br_in);
    const int tile = __builtin_amdgcn_readfirstlane(tile_in);
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int n_base = tile * (16 * EPL);
    const unsigned N = static_cast<unsigned>(a.N);
    const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(a.B8, a.b_bytes);
    const __amdgpu_buffer_rsrc_t rsrc_a = make_rsrc(a.vals, a.a_bytes);
    const unsigned lane_off = static_cast<unsigned>(n_base + j) * 16u;

    f32x4 acc[RS][EPL];
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
        for (int e = 0; e < EPL; ++e) acc[s][e] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int q0 = const_load(a.grp_ptr, br), q1 = const_load(a.grp_ptr, br + 1);
    if (q0 < q1) {
        // B's scales exist only where B has rows (a row with octets implies
        // K > 0; with K = 0 no quantised B is made); sb, the chunk's one
        // scale, is read below, after the first loads
        int sb = 0;
        // the chunk's slab scales differ (set after the first loads); lanes of
        // the even (odd) lane groups hold the octet's even (odd) slots
        bool mixed = false;
        // For such a chunk, A is either split by lane group as it loads
        // (kSplitA: two loads, zeros from loads aimed out of range) or masked
        // at the MFMA (keep_first).
        const unsigned keep_even = (g & 1) ? 0u : ~0u;
        unsigned keep_first = ~0u;
        // the chunk's slab scales, read before the pipeline's first loads so
        // that waiting for them does not wait for those
        const u32x4 srow = scale_row_load(a, n_base / kChunk);

        // slab offsets of this lane's two halves (slots g and 4 + g) in octet q;
        // the eight columns come in as two scalar 16-byte loads (clamped to the
        // last octet) and are picked per lane group by selects, not branches
        auto bases = [&](int q, unsigned &lo, unsigned &hi) {
            const int qc = q < q1 ? q : q1 - 1;
            const i32x4 c03 = const_load(reinterpret_cast<const i32x4 *>(a.grp_col), 2 * qc);
            const i32x4 c47 = const_load(reinterpret_cast<const i32x4 *>(a.grp_col), 2 * qc + 1);
            int cl = c03.x, ch = c47.x;
            cl = g == 1 ? c03.y : cl;
            ch = g == 1 ? c47.y : ch;
            cl = g == 2 ? c03.z : cl;
            ch = g == 2 ? c47.z : ch;
            cl = g == 3 ? c03.w : cl;
            ch = g == 3 ? c47.w : ch;
            lo = q < q1 ? static_cast<unsigned>(cl) * (16u * N) + lane_off : kOutOfRange;
            hi = q < q1 ? static_cast<unsigned>(ch) * (16u * N) + lane_off : kOutOfRange;
        };
        // both halves land in one 8-register operand (no copies before the
        // MFMA).  Only voffset is range-checked, so column offsets go there and
        // columns past N (ragged last tile) read as zeros; no branches, so the
        // loads stay where the schedule puts them.
        auto load_b = [&](unsigned lo, unsigned hi, int e, u32x8 &b) {
            const bool ok = n_base + 16 * e + j < a.N;
            const unsigned ol = ok && lo != kOutOfRange ? lo + 256u * e : kOutOfRange;
            const unsigned oh = ok && hi != kOutOfRange ? hi + 256u * e : kOutOfRange;
            b.lo = __builtin_amdgcn_raw_buffer_load_b128(rsrc, static_cast<int>(ol), 0, 0);
            b.hi = __builtin_amdgcn_raw_buffer_load_b128(rsrc, static_cast<int>(oh), 0, 0);
        };
        // A of octet q for the first MFMA of each 16 columns (for a chunk
        // whose slab scales differ only the even lane groups load, the others
        // read zeros), and its A scales
        auto load_a = [&](int q, u32x8 (&av)[RS], int (&sa)[RS]) {
#pragma unroll
            for (int s = 0; s < RS; ++s) {
                const unsigned aoff =
                    q < q1 ? static_cast<unsigned>(((static_cast<size_t>(q) * RS + s) * 64 + lane) * 32) : kOutOfRange;
                const unsigned a1 = kSplitA && mixed && (g & 1) ? kOutOfRange : aoff;
                av[s].lo = __builtin_amdgcn_raw_buffer_load_b128(rsrc_a, static_cast<int>(a1), 0, 0);
                av[s].hi = __builtin_amdgcn_raw_buffer_load_b128(rsrc_a, static_cast<int>(a1), 16, 0);
                const int qc = q < q1 ? q : q1 - 1;
                const unsigned sc = const_load(a.ascale, static_cast<size_t>(qc) * RS + s);
                sa[s] = static_cast<int>((sc >> (8 * g)) & 0xffu);
            }
        };
        // for a chunk whose slab scales differ: A of octet q in the odd lane
        // groups only, for the second MFMA
        auto load_a_odd = [&](int q, u32x8 (&av)[RS]) {
#pragma unroll
            for (int s = 0; s < RS; ++s) {
                const unsigned aoff =
                    q < q1 ? static_cast<unsigned>(((static_cast<size_t>(q) * RS + s) * 64 + lane) * 32) : kOutOfRange;
                const unsigned a2 = (g & 1) ? aoff : kOutOfRange;
                av[s].lo = __builtin_amdgcn_raw_buffer_load_b128(rsrc_a, static_cast<int>(a2), 0, 0);
                av[s].hi = __builtin_amdgcn_raw_buffer_load_b128(rsrc_a, static_cast<int>(a2), 16, 0);
            }
        };
        // For a chunk whose slab scales differ: the B scales that lane group
        // g provides for scale group g of octet q, those of the slabs of its
        // slots 2 g (sbe) and 2 g + 1 (sbo), read from the chunk's scale row
        // with the octet's A (an octet past the row's end reads the last one's)
        const __amdgpu_buffer_rsrc_t rsrc_s = scale_row_rsrc(a, n_base / kChunk);
        auto load_bscale = [&](int q, int &sbe, int &sbo) {
            const int qc = q < q1 ? q : q1 - 1;
            const i32x4 c03 = const_load(reinterpret_cast<const i32x4 *>(a.grp_col), 2 * qc);
            const i32x4 c47 = const_load(reinterpret_cast<const i32x4 *>(a.grp_col), 2 * qc + 1);
            const int ce = g < 2 ? (g == 0 ? c03.x : c03.z) : (g == 2 ? c47.x : c47.z);
            const int co = g < 2 ? (g == 0 ? c03.y : c03.w) : (g == 2 ? c47.y : c47.w);
            sbe = __builtin_amdgcn_raw_buffer_load_b8(rsrc_s, ce, 0, 0);
            sbo = __builtin_amdgcn_raw_buffer_load_b8(rsrc_s, co, 0, 0);
        };

        // B for half an octet (EPL / 2 column blocks) per step; steps
        // alternate between two buffers: the loads of step u + 1 are issued
        // before the MFMAs of step u (the schedule barriers keep them there).
        // Two octets per loop iteration give every buffer a fixed role, so no
        // loaded register is copied (a copy would wait for its load).
        constexpr int HB = EPL / 2 > 0 ? EPL / 2 : 1;
        u32x8 bufX[HB], bufY[HB];
        u32x8 avA[RS], avB[RS], aoA[RS], aoB[RS];
        int saA[RS], saB[RS];
        // B scales of the first and second MFMA, per octet (A's buffers)
        int sbeA = 0, sboA = 0, sbeB = 0, sboB = 0;
        unsigned lo_c, hi_c, lo_n, hi_n;
        bases(q0, lo_c, hi_c);
        bases(q0 + 1, lo_n, hi_n);
        auto issue_half = [&](unsigned lo, unsigned hi, int half, u32x8 (&buf)[HB]) {
#pragma unroll
            for (int k = 0; k < HB; ++k) load_b(lo, hi, half * HB + k, buf[k]);
        };
        // The first MFMA of each 16 columns runs always: for a uniform chunk
        // over all slots with the chunk's scale, else over the even slots (av
        // holds zeros in the odd lane groups) with the even slabs' scales; for
        // a chunk whose slab scales differ a second one adds the odd slots (ao)
        // with theirs, so every slab keeps its own scale.  Zeros come from
        // loads aimed out of range, so no VALU work splits A.  mixed is
        // wave-uniform.
        auto compute_half = [&](int half, const u32x8 (&buf)[HB], const u32x8 (&av)[RS], const u32x8 (&ao)[RS],
                                const int (&sa)[RS], int sbe, int sbo) {
#pragma unroll
            for (int k = 0; k < HB; ++k) {
                const int e = half * HB + k;
                if (e >= EPL) continue;
                const i32x8 bv = __builtin_bit_cast(i32x8, buf[k]);
#pragma unroll
                for (int s = 0; s < RS; ++s)
                    acc[s][e] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        __builtin_bit_cast(i32x8, kSplitA ? av[s] : av[s] & keep_first), bv, acc[s][e], 0, 0, 0,
                        sa[s], 0, sbe);
            }
            // The empty statements use each result where it is made.  Without
            // them the compiler sinks these MFMAs into the blocks that follow
            // the next loads, past the schedule barriers: the buffer just read
            // then lives on beside the next one, and an accumulator whose old
            // value is still live gets a second register set (the MFMA form
            // whose result may not overlap its input) -- over 210 VGPRs and 2
            // waves per SIMD for one row set at 128 columns, against 153 and 3.
#pragma unroll
            for (int k = 0; k < HB; ++k) {
                const int e = half * HB + k;
                if (e >= EPL) continue;
#pragma unroll
                for (int s = 0; s < RS; ++s) asm("" : "+v"(acc[s][e]));
            }
            if (mixed) {
#pragma unroll
                for (int k = 0; k < HB; ++k) {
                    const int e = half * HB + k;
                    if (e >= EPL) continue;
                    const i32x8 bv = __builtin_bit_cast(i32x8, buf[k]);
#pragma unroll
                    for (int s = 0; s < RS; ++s)
                        acc[s][e] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                            __builtin_bit_cast(i32x8, kSplitA ? ao[s] : av[s] & ~keep_first), bv, acc[s][e], 0,
                            0, 0, sa[s], 0, sbo);
                }
#pragma unroll
                for (int k = 0; k < HB; ++k) {
                    const int e = half * HB + k;
                    if (e >= EPL) continue;
#pragma unroll
                    for (int s = 0; s < RS; ++s) asm("" : "+v"(acc[s][e]));
                }
            }
        };
        auto advance = [&](int q) {
            lo_c = lo_n;
            hi_c = hi_n;
            bases(q, lo_n, hi_n);
        };
        // The octet loop.  For a chunk whose slab scales differ, each octet's
        // A load also brings the octet's B scales.
        auto octets = [&]() {
            auto next_a = [&](int q, u32x8 (&av)[RS], u32x8 (&ao)[RS], int (&sa)[RS], int &sbe, int &sbo) {
                load_a(q, av, sa);
                if (mixed) {
                    if constexpr (kSplitA) load_a_odd(q, ao);
                    load_bscale(q, sbe, sbo);
                }
            };
            auto half_x = [&](int half, const u32x8 (&buf)[HB], bool second) {
                if (second)
                    compute_half(half, buf, avB, aoB, saB, sbeB, sboB);
                else
                    compute_half(half, buf, avA, aoA, saA, sbeA, sboA);
            };
            if constexpr (RS == 1 && EPL >= 4) {
                // The loop runs while the row has two more octets after the
                // current one, so every load it issues is inside the row and it
                // has one exit; the last one or two octets follow it.  Nothing is
                // loaded past the row's end (such loads would read zeros but still
                // occupy the load path).
                int q = q0;
                for (; q + 2 < q1; q += 2) {
                    // octet q (A in avA)
                    issue_half(lo_c, hi_c, 1, bufY);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(0, bufX, false);
                    __builtin_amdgcn_sched_barrier(0);
                    next_a(q + 1, avB, aoB, saB, sbeB, sboB);
                    issue_half(lo_n, hi_n, 0, bufX);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(1, bufY, false);
                    __builtin_amdgcn_sched_barrier(0);
                    advance(q + 2);
                    // octet q + 1 (A in avB)
                    issue_half(lo_c, hi_c, 1, bufY);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(0, bufX, true);
                    __builtin_amdgcn_sched_barrier(0);
                    next_a(q + 2, avA, aoA, saA, sbeA, sboA);
                    issue_half(lo_n, hi_n, 0, bufX);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(1, bufY, true);
                    __builtin_amdgcn_sched_barrier(0);
                    advance(q + 3);
                }
                // octet q (A in avA, first half of B in bufX), then octet q + 1
                // if the row has it
                issue_half(lo_c, hi_c, 1, bufY);
                __builtin_amdgcn_sched_barrier(0);
                half_x(0, bufX, false);
                __builtin_amdgcn_sched_barrier(0);
                if (q + 1 < q1) {
                    next_a(q + 1, avB, aoB, saB, sbeB, sboB);
                    issue_half(lo_n, hi_n, 0, bufX);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(1, bufY, false);
                    __builtin_amdgcn_sched_barrier(0);
                    issue_half(lo_n, hi_n, 1, bufY);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(0, bufX, true);
                    half_x(1, bufY, true);
                } else {
                    half_x(1, bufY, false);
                }
            } else {
                // (with more row sets or narrow tiles the early exits below cost
                // registers, and so occupancy: the loads past the row's end stay)
                for (int q = q0; q < q1; q += 2) {
                    // octet q (A in avA)
                    issue_half(lo_c, hi_c, 1, bufY);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(0, bufX, false);
                    __builtin_amdgcn_sched_barrier(0);
                    next_a(q + 1, avB, aoB, saB, sbeB, sboB);
                    issue_half(lo_n, hi_n, 0, bufX);
                    __builtin_amdgcn_sched_barrier(0);
                    half_x(1, bufY, false);
                    __builtin_amdgcn_sched_barrier(0);
                    advance(q + 2);
                    // octet q + 1 (A in avB); past the row's end its loads read zeros
                    issue_half(lo_c, hi_c, 1, bufY);
                    __builtin_amdgcn_sched_barrier(0);
                    if (q + 1 < q1) half_x(0, bufX, true);
                    __builtin_amdgcn_sched_barrier(0);
                    next_a(q + 2, avA, aoA, saA, sbeA, sboA);
                    issue_half(lo_n, hi_n, 0, bufX);
                    __builtin_amdgcn_sched_barrier(0);
                    if (q + 1 < q1) half_x(1, bufY, true);
                    __builtin_amdgcn_sched_barrier(0);
                    advance(q + 3);
                }
            }
        };
        load_a(q0, avA, saA);
        issue_half(lo_c, hi_c, 0, bufX);
        // checked after the pipeline's first loads are issued
        sb = scale_row_uniform(a, n_base / kChunk, srow);
        mixed = sb < 0;
        if (mixed) {
            keep_first = keep_even;
            // the first octet's A came in whole: split it by lane group
            if constexpr (kSplitA) {
#pragma unroll
                for (int s = 0; s < RS; ++s) {
                    aoA[s] = avA[s] & ~keep_even;
                    avA[s] = avA[s] & keep_even;
                }
            }
            load_bscale(q0, sbeA, sboA);
        } else {
            sbeA = sbeB = sb;
        }
        octets();
    }

#pragma unroll
    for (int e = 0; e < EPL; ++e) {
        const int n = n_base + 16 * e + j;
        if (n >= a.N) continue;
#pragma unroll
        for (int s = 0; s < RS; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const size_t row = static_cast<size_t>(br) * (16 * RS) + 16 * s + 4 * g + r;
                global_ptr(a.C)[row * static_cast<size_t>(a.N) + n] = acc[s][e][r];
            }
    }
}

}  // namespace fp8dev
