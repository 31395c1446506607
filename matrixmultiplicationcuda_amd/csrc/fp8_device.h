// fp8_device.h -- device code of the FP8 E4M3 storage mode on the scaled MFMA
// v_mfma_scale_f32_16x16x128_f8f6f4 (no reference counterpart: the
// reference's kernels are fp16 only).
//
// Every call first quantises B into a workspace (quant_b_fp8): one E8M0
// scale per 16-row slab and 256-column chunk, and the values stored k-minor,
// B8[slab][n][16] (the 16 rows of a slab for column n in 16 consecutive
// bytes), so that 16 bytes are exactly one lane's half of an MFMA B operand
// and no transpose is left for the product kernel.
//
// The product kernel walks quads of A blocks.  The scaled 16x16x128 MFMA
// takes one E8M0 scale per row (A) or column (B) for each of four 32-wide K
// groups; K group t is read from lane group t's scale operand and covers bytes
// 16 (t >> 1) .. + 15 of lane groups 2 (t & 1) and 2 (t & 1) + 1.  Quad slot t
// therefore puts its block in those bytes of lane group 2 (t & 1) and zeros in
// lane group 2 (t & 1) + 1, so the per-block A scale and the per-slab B scale
// are applied by the MFMA itself at no VALU cost.  Half of every operand is
// padding: lane groups 1 and 3 read through an out-of-range buffer offset,
// which returns zeros without memory traffic.  The MFMA pipe is not the bound
// (fp8 runs at twice the fp16 rate).
#pragma once

#include "bcsr_internal.h"

namespace fp8dev {

constexpr int kChunk = 256;  // columns per B scale

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

// One workgroup of 256 threads quantises 16-row slab c, 256-column chunk h:
// thread t takes column n = 256 h + t.
__device__ __forceinline__ void quant_slab(const _Float16 *__restrict__ B, uint8_t *__restrict__ B8,
                                           uint8_t *__restrict__ scale, int N, int chunks, int c, int h) {
    __shared__ float wmax[4];
    const int t = threadIdx.x;
    const int n = h * kChunk + t;
    const bool valid = n < N;
    float v[16];
    float amax = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        v[r] = valid ? static_cast<float>(B[static_cast<size_t>(c * 16 + r) * N + n]) : 0.0f;
        amax = fmaxf(amax, fabsf(v[r]));
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) amax = fmaxf(amax, __shfl_xor(amax, m, 64));
    if ((t & 63) == 0) wmax[t >> 6] = amax;
    __syncthreads();
    amax = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    const unsigned se = e8m0_for_absmax(amax);
    if (t == 0) scale[static_cast<size_t>(c) * chunks + h] = static_cast<uint8_t>(se);
    if (!valid) return;
    const float inv = ldexpf(1.0f, 127 - static_cast<int>(se));
    u32x4 q;
    unsigned w[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        int p = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * d] * inv, v[4 * d + 1] * inv, 0, false);
        p = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * d + 2] * inv, v[4 * d + 3] * inv, p, true);
        w[d] = static_cast<unsigned>(p);
    }
    q = u32x4{w[0], w[1], w[2], w[3]};
    *reinterpret_cast<u32x4 *>(B8 + (static_cast<size_t>(c) * N + n) * 16) = q;
}

// Work item (device block-row br, tile of 16 * EPL columns); MFMA e covers
// columns tile + 16 e + j for lane j of each lane group.
template <int RS, int EPL>
__device__ __forceinline__ void fp8_work(const SpmmArgs &a, int br_in, int tile_in) {
    // wave-uniform by construction (scalar index chain)
    const int br = __builtin_amdgcn_readfirstlane(br_in);
    const int tile = __builtin_amdgcn_readfirstlane(tile_in);
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int n_base = tile * (16 * EPL);
    const int chunk = n_base / kChunk;
    const unsigned N = static_cast<unsigned>(a.N);
    const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(a.B8, a.b_bytes);
    const unsigned lane_off = static_cast<unsigned>(n_base + j) * 16u;

    f32x4 acc[RS][EPL];
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
        for (int e = 0; e < EPL; ++e) acc[s][e] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int q0 = a.grp_ptr[br], q1 = a.grp_ptr[br + 1];
    const bool data_lane = (g & 1) == 0;       // lane groups 0 and 2 carry blocks
    const bool full_tile = n_base + 16 * EPL <= a.N;
    const int lane32 = (g >> 1) * 16 + j;      // index among the stored lanes
    const __amdgpu_buffer_rsrc_t rsrc_a = make_rsrc(a.vals, a.a_bytes);
    for (int q = q0; q < q1; ++q) {
        const int4 cq = reinterpret_cast<const int4 *>(a.grp_col)[q];
        const int c_scale = g == 0 ? cq.x : (g == 1 ? cq.y : (g == 2 ? cq.z : cq.w));
        const int c_lo = (g >> 1) ? cq.y : cq.x;  // slots 2 (g >> 1) .. for halves 0 / 1
        const int c_hi = (g >> 1) ? cq.w : cq.z;
        const int sa = static_cast<int>((a.ascale[q] >> (8 * g)) & 0xffu);
        const int sb = a.b8scale[static_cast<size_t>(c_scale) * a.b_chunks + chunk];
        const unsigned base_lo = static_cast<unsigned>(c_lo) * (16u * N) + lane_off;
        const unsigned base_hi = static_cast<unsigned>(c_hi) * (16u * N) + lane_off;
        u32x4 blo[EPL], bhi[EPL];
        if (full_tile) {
            // every column of the tile exists: step columns through soffset
            const unsigned off_lo = data_lane ? base_lo : kOutOfRange;
            const unsigned off_hi = data_lane ? base_hi : kOutOfRange;
#pragma unroll
            for (int e = 0; e < EPL; ++e) {
                blo[e] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, static_cast<int>(off_lo), 256 * e, 0);
                bhi[e] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, static_cast<int>(off_hi), 256 * e, 0);
            }
        } else {
            // ragged last tile: the range check covers voffset only, so every
            // column offset goes there and missing columns read as zeros
#pragma unroll
            for (int e = 0; e < EPL; ++e) {
                const bool ok = data_lane && n_base + 16 * e + j < a.N;
                const unsigned ol = ok ? base_lo + 256u * e : kOutOfRange;
                const unsigned oh = ok ? base_hi + 256u * e : kOutOfRange;
                blo[e] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, static_cast<int>(ol), 0, 0);
                bhi[e] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, static_cast<int>(oh), 0, 0);
            }
        }
        i32x8 av[RS];
#pragma unroll
        for (int s = 0; s < RS; ++s) {
            const unsigned aoff = data_lane ? static_cast<unsigned>(((static_cast<size_t>(q) * RS + s) * 32 + lane32) * 32)
                                            : kOutOfRange;
            const u32x4 x0 = __builtin_amdgcn_raw_buffer_load_b128(rsrc_a, static_cast<int>(aoff), 0, 0);
            const u32x4 x1 = __builtin_amdgcn_raw_buffer_load_b128(rsrc_a, static_cast<int>(aoff), 16, 0);
            av[s] = __builtin_bit_cast(i32x8, (u32x8{x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w}));
        }
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            const i32x8 bv = __builtin_bit_cast(
                i32x8, (u32x8{blo[e].x, blo[e].y, blo[e].z, blo[e].w, bhi[e].x, bhi[e].y, bhi[e].z, bhi[e].w}));
#pragma unroll
            for (int s = 0; s < RS; ++s)
                acc[s][e] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av[s], bv, acc[s][e], 0, 0, 0, sa, 0, sb);
        }
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
                a.C[row * static_cast<size_t>(a.N) + n] = acc[s][e][r];
            }
    }
}

}  // namespace fp8dev
