// pair_device.h -- device code of the pair-layout kernels (16-bit MFMA),
// shared by the per-matrix kernels (kernels_pair.hip) and the batched
// persistent kernel (kernels_batch.hip).
//
// One 64-lane wave owns (device block-row, column tile of 16 * EPL).  Lane
// l = 16 g + j loads, for every block pair, 8 consecutive rows of B
// (rows 8 (g % 2) .. + 7 of the slab of block g / 2) at the EPL columns
// tile + EPL j .. + EPL - 1, and transposes that 8 x EPL patch in registers
// into EPL MFMA B operands (8 consecutive k at one column each) with one
// v_perm_b32 per dword.  MFMA number e of the pair computes output columns
// tile + EPL j + e, j = 0..15, so no LDS is needed and each lane ends with EPL
// consecutive fp32 outputs per row: C is written with 16-byte stores.
#pragma once

#include "bcsr_internal.h"

namespace pairdev {

enum { PREC_F16 = 0, PREC_BF16 = 1, PREC_HILO = 3 };

// EPL consecutive halves of one B row.
template <int EPL>
struct BRow {
    unsigned int d[EPL / 2];
};

// The 8 x EPL patch through a buffer resource: row r at soffset r * row_bytes.
template <int EPL>
__device__ __forceinline__ void load_patch(__amdgpu_buffer_rsrc_t rsrc, unsigned off, unsigned row_bytes,
                                           BRow<EPL> (&raw)[8]) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int so = static_cast<int>(r * row_bytes);
        if constexpr (EPL == 2) {
            raw[r].d[0] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, static_cast<int>(off), so, 0);
        } else if constexpr (EPL == 4) {
            u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, static_cast<int>(off), so, 0);
            raw[r].d[0] = v.x;
            raw[r].d[1] = v.y;
        } else {
#pragma unroll
            for (int h = 0; h < EPL / 8; ++h) {
                u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, static_cast<int>(off + 16 * h), so, 0);
                raw[r].d[4 * h + 0] = v.x;
                raw[r].d[4 * h + 1] = v.y;
                raw[r].d[4 * h + 2] = v.z;
                raw[r].d[4 * h + 3] = v.w;
            }
        }
    }
}

// The same patch through 64-bit global addresses (any B size).
template <int EPL>
__device__ __forceinline__ void load_patch_global(const _Float16 *p, size_t N, BRow<EPL> (&raw)[8]) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const _Float16 *q = p + r * N;
        if constexpr (EPL == 2) {
            raw[r].d[0] = *reinterpret_cast<const unsigned int *>(q);
        } else if constexpr (EPL == 4) {
            u32x2 v = *reinterpret_cast<const u32x2 *>(q);
            raw[r].d[0] = v.x;
            raw[r].d[1] = v.y;
        } else {
#pragma unroll
            for (int h = 0; h < EPL / 8; ++h) {
                u32x4 v = reinterpret_cast<const u32x4 *>(q)[h];
                raw[r].d[4 * h + 0] = v.x;
                raw[r].d[4 * h + 1] = v.y;
                raw[r].d[4 * h + 2] = v.z;
                raw[r].d[4 * h + 3] = v.w;
            }
        }
    }
}

// Column e of the patch as an fp16 MFMA B operand.
template <int EPL>
__device__ __forceinline__ f16x8 patch_column_f16(const BRow<EPL> (&raw)[8], int e) {
    const unsigned sel = (e & 1) ? 0x07060302u : 0x05040100u;
    const int w = e >> 1;
    u32x4 v;
    v.x = __builtin_amdgcn_perm(raw[1].d[w], raw[0].d[w], sel);
    v.y = __builtin_amdgcn_perm(raw[3].d[w], raw[2].d[w], sel);
    v.z = __builtin_amdgcn_perm(raw[5].d[w], raw[4].d[w], sel);
    v.w = __builtin_amdgcn_perm(raw[7].d[w], raw[6].d[w], sel);
    return __builtin_bit_cast(f16x8, v);
}

// Column e of the patch rounded to bf16 (round to nearest even) for the bf16 MFMA.
template <int EPL>
__device__ __forceinline__ bf16x8 patch_column_bf16(const BRow<EPL> (&raw)[8], int e) {
    const int w = e >> 1, sh = (e & 1) * 16;
    bf16x8 v;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const unsigned short bits = static_cast<unsigned short>(raw[r].d[w] >> sh);
        v[r] = static_cast<__bf16>(static_cast<float>(__builtin_bit_cast(_Float16, bits)));
    }
    return v;
}

// PREC_F16 / PREC_BF16: one operand per row-subtile; PREC_HILO: hi and lo.
template <int RS, int PREC>
struct AOps {
    static constexpr int H = PREC == PREC_HILO ? 2 : 1;
    u32x4 v[RS][H];
};

template <int RS, int EPL, int PREC>
struct Acc {
    static constexpr int H = PREC == PREC_HILO ? 2 : 1;
    f32x4 v[H][RS][EPL];
};

template <int RS, int EPL, int PREC>
__device__ __forceinline__ void consume(Acc<RS, EPL, PREC> &acc, const BRow<EPL> (&raw)[8],
                                        const AOps<RS, PREC> &av) {
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
        if constexpr (PREC == PREC_BF16) {
            const bf16x8 bv = patch_column_bf16<EPL>(raw, e);
#pragma unroll
            for (int s = 0; s < RS; ++s)
                acc.v[0][s][e] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av.v[s][0]), bv,
                                                                         acc.v[0][s][e], 0, 0, 0);
        } else {
            const f16x8 bv = patch_column_f16<EPL>(raw, e);
#pragma unroll
            for (int h = 0; h < AOps<RS, PREC>::H; ++h)
#pragma unroll
                for (int s = 0; s < RS; ++s)
                    acc.v[h][s][e] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, av.v[s][h]), bv,
                                                                            acc.v[h][s][e], 0, 0, 0);
        }
    }
}

// Writes the (16 RS) x (16 EPL) tile: lane (g, j) holds rows 16 s + 4 g + r,
// columns n0 .. n0 + EPL - 1.  hi+lo combines acc_hi + 2^-11 acc_lo.
template <int RS, int EPL, int PREC>
__device__ __forceinline__ void store_tile(const Acc<RS, EPL, PREC> &acc, float *C, size_t N, int br, int g, int n0) {
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const size_t row = static_cast<size_t>(br) * (16 * RS) + 16 * s + 4 * g + r;
            auto *cp = global_ptr(C + row * N + n0);
            float o[EPL];
#pragma unroll
            for (int e = 0; e < EPL; ++e) {
                if constexpr (PREC == PREC_HILO)
                    o[e] = acc.v[0][s][e][r] + acc.v[1][s][e][r] * 0x1p-11f;
                else
                    o[e] = acc.v[0][s][e][r];
            }
            if constexpr (EPL == 2) {
                *reinterpret_cast<__attribute__((address_space(1))) f32x2 *>(cp) = f32x2{o[0], o[1]};
            } else {
#pragma unroll
                for (int h = 0; h < EPL / 4; ++h)
                    reinterpret_cast<__attribute__((address_space(1))) f32x4 *>(cp)[h] =
                        f32x4{o[4 * h], o[4 * h + 1], o[4 * h + 2], o[4 * h + 3]};
            }
        }
}

// One work item of the buffer-load kernels (variants 0 and 1): output tile
// `tile` of device block-row br.  B is read through a buffer
// resource with the row stride in soffset, so a pair costs two VALU ops of
// addressing; lanes whose columns fall past N read column 0 (their outputs
// are never stored), so loads need no predication.  The two columns of the
// next pair are fetched by a scalar load one step ahead (scalar loads count
// in lgkmcnt, so waiting for them never waits for B).  PF double-buffers: the loads of pair p + 1 are issued before the
// MFMAs of pair p, unconditionally inside the loop, so the wait counts are the
// same on every path through it and each wait covers exactly one pair's
// loads; a load past the row's end aims at an offset beyond the buffer (which
// returns zeros without touching memory).  Up to 128-column tiles the last one
// or two pairs leave the loop and issue nothing past the row's end.
template <int RS, int EPL, bool PF, int PREC>
__device__ __forceinline__ void pair_work(const SpmmArgs &a, int br_in, int tile_in) {
    // wave-uniform by construction; scalar registers keep the index chain
    // (group range, column indices, B offsets) off the vector memory path
    const int br = __builtin_amdgcn_readfirstlane(br_in);
    const int tile = __builtin_amdgcn_readfirstlane(tile_in);
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int n0 = tile * (16 * EPL) + EPL * j;
    const bool nvalid = n0 < a.N;
    const unsigned row_bytes = static_cast<unsigned>(a.N) * 2u;
    const unsigned slab_bytes = 16u * row_bytes;
    const unsigned lane_off = static_cast<unsigned>((g & 1) * 8) * row_bytes + (nvalid ? n0 : 0) * 2u;
    const int sel_hi = g >> 1;
    const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(a.B, a.b_bytes);
    constexpr int H = AOps<RS, PREC>::H;

    Acc<RS, EPL, PREC> acc;
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int s = 0; s < RS; ++s)
#pragma unroll
            for (int e = 0; e < EPL; ++e) acc.v[h][s][e] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int p0 = const_load(a.grp_ptr, br), p1 = const_load(a.grp_ptr, br + 1);
    if (p0 < p1) {
        // global (not generic) loads for A, scalar ones for the columns, also
        // where the pointers come from memory (the batched kernel)
        const auto *vals = global_ptr(a.vals + lane);
        const i32x2 *cols = reinterpret_cast<const i32x2 *>(a.grp_col);
        // Loads of pair p with columns cc (p >= p1: an out-of-range B offset
        // and the last pair's A).
        auto issue = [&](int p, i32x2 cc, BRow<EPL>(&raw)[8], AOps<RS, PREC> &av) {
            const int pc = p < p1 ? p : p1 - 1;
            const unsigned off =
                p < p1 ? static_cast<unsigned>(sel_hi ? cc.y : cc.x) * slab_bytes + lane_off : kOutOfRange;
            load_patch<EPL>(rsrc, off, row_bytes, raw);
#pragma unroll
            for (int s = 0; s < RS; ++s)
#pragma unroll
                for (int h = 0; h < H; ++h) av.v[s][h] = vals[((static_cast<size_t>(pc) * RS + s) * H + h) * 64];
        };
        auto col_of = [&](int p) { return const_load(cols, p < p1 ? p : p1 - 1); };
        if constexpr (!PF) {
            i32x2 cn = col_of(p0);
            for (int p = p0; p < p1; ++p) {
                BRow<EPL> raw[8];
                AOps<RS, PREC> av;
                issue(p, cn, raw, av);
                cn = col_of(p + 1);
                consume<RS, EPL, PREC>(acc, raw, av);
            }
        } else {
            BRow<EPL> rawA[8], rawB[8];
            AOps<RS, PREC> avA, avB;
            i32x2 cn = col_of(p0);
            issue(p0, cn, rawA, avA);
            cn = col_of(p0 + 1);
            // sched_barrier keeps the compiler from sinking a pair's loads
            // below the previous pair's MFMAs, which would undo the pipeline
            if constexpr (EPL <= 8) {
                // the loop runs while both pairs it loads exist; the last one
                // or two pairs follow without loads past the row's end (those
                // read zeros but still occupy the load path: 1.5 % of the
                // time at 16384^2, 5 %, N = 512)
                int p = p0;
                for (; p + 2 < p1; p += 2) {
                    issue(p + 1, cn, rawB, avB);
                    cn = col_of(p + 2);
                    __builtin_amdgcn_sched_barrier(0);
                    consume<RS, EPL, PREC>(acc, rawA, avA);
                    __builtin_amdgcn_sched_barrier(0);
                    issue(p + 2, cn, rawA, avA);
                    cn = col_of(p + 3);
                    __builtin_amdgcn_sched_barrier(0);
                    if (p + 1 < p1) consume<RS, EPL, PREC>(acc, rawB, avB);
                    __builtin_amdgcn_sched_barrier(0);
                }
                const bool two = p + 1 < p1;
                if (two) issue(p + 1, cn, rawB, avB);
                __builtin_amdgcn_sched_barrier(0);
                consume<RS, EPL, PREC>(acc, rawA, avA);
                __builtin_amdgcn_sched_barrier(0);
                if (two) consume<RS, EPL, PREC>(acc, rawB, avB);
            } else {
                // 256-column tiles keep the single loop (the split measured
                // 1 % slower there)
                for (int p = p0; p < p1; p += 2) {
                    issue(p + 1, cn, rawB, avB);
                    cn = col_of(p + 2);
                    __builtin_amdgcn_sched_barrier(0);
                    consume<RS, EPL, PREC>(acc, rawA, avA);
                    __builtin_amdgcn_sched_barrier(0);
                    issue(p + 2, cn, rawA, avA);
                    cn = col_of(p + 3);
                    __builtin_amdgcn_sched_barrier(0);
                    if (p + 1 < p1) consume<RS, EPL, PREC>(acc, rawB, avB);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
    }
    if (!nvalid) return;
    store_tile<RS, EPL, PREC>(acc, a.C, static_cast<size_t>(a.N), br, g, n0);
}

}  // namespace pairdev
