// CPU emulation of the panel kernel (matrixmultiplicationcuda_amd/csrc/
// kernels_panel.hip) on the chunk schedule that panel_schedule.cpp builds.
//
// For random BCSR matrices with 16-row blocks it builds the plan's pair
// groups, runs build_panel_schedule, and then replays the kernel literally:
// each chunk's loads into a staged copy of LDS (B pieces q = 16 m + w, lane
// -> row 4 (q % 4) + lane / 16, octet lane % 16; A records 2 w + lane / 32),
// every index checked against its buffer, and each wave's 16x16x32 steps
// with the MFMA operand layout (A lane l: row l % 16, k 8 (l / 16) .. + 7; B
// lane (g, j): rows 8 (g % 2) .. + 7 of slab g / 2, output columns
// 8 j .. 8 j + 7).  B columns past N are NaN in the staged copy, so a leak
// into a stored output shows.  The result must equal a direct BCSR x B
// product (both in double), and the schedule must consume every block once.
//
// usage: panel_emulator  (prints "ok <cases>" or the first failure)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "panel_schedule.h"

namespace {

int fail(const char *what, int a = 0, int b = 0) {
    std::printf("FAIL %s (%d, %d)\n", what, a, b);
    std::exit(1);
}

struct Case {
    int block_rows, block_cols, bw, N;
    double density;
    const char *pattern;  // "random", "band", "empty_rows"
    unsigned seed;
};

void run_case(const Case &cs) {
    std::mt19937 rng(cs.seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0), U01(0.0, 1.0);
    const int pieces = cs.bw / 16, bh = 16;
    const int M = cs.block_rows * bh, K = cs.block_cols * cs.bw, N = cs.N;
    // BCSR (hdr, idx, values as double)
    std::vector<int32_t> hdr(cs.block_rows + 1, 0), idx;
    std::vector<double> data;
    for (int R = 0; R < cs.block_rows; ++R) {
        for (int c = 0; c < cs.block_cols; ++c) {
            bool keep;
            if (!std::strcmp(cs.pattern, "band"))
                keep = std::abs(c - R * cs.block_cols / std::max(cs.block_rows, 1)) <= 3;
            else if (!std::strcmp(cs.pattern, "empty_rows"))
                keep = (R % 5 != 2) && U01(rng) < cs.density;
            else
                keep = U01(rng) < cs.density;
            if (!keep) continue;
            idx.push_back(c);
            for (int e = 0; e < bh * cs.bw; ++e) data.push_back(U(rng));
        }
        hdr[R + 1] = static_cast<int32_t>(idx.size());
    }
    std::vector<double> B(static_cast<size_t>(K) * N);
    for (auto &v : B) v = U(rng);

    // the plan's pair groups over 16 x 16 device blocks
    const int dev_rows = cs.block_rows;
    std::vector<int32_t> nblk(dev_rows), grp_ptr(dev_rows + 1, 0), grp_col;
    for (int r = 0; r < dev_rows; ++r) {
        nblk[r] = (hdr[r + 1] - hdr[r]) * pieces;
        grp_ptr[r + 1] = grp_ptr[r] + (nblk[r] + 1) / 2;
    }
    grp_col.resize(2 * static_cast<size_t>(grp_ptr[dev_rows]));
    // value of device block (r, jj) at (row i, column k)
    auto dev_val = [&](int r, int jj, int i, int k) {
        const int64_t blk = hdr[r] + jj / pieces;
        const int q = jj % pieces;
        return data[static_cast<size_t>(blk) * bh * cs.bw + static_cast<size_t>(i) * cs.bw + 16 * q + k];
    };
    auto dev_col = [&](int r, int jj) { return idx[hdr[r] + jj / pieces] * pieces + jj % pieces; };
    for (int r = 0; r < dev_rows; ++r)
        for (int p = grp_ptr[r]; p < grp_ptr[r + 1]; ++p)
            for (int s = 0; s < 2; ++s) {
                int jj = 2 * (p - grp_ptr[r]) + s;
                if (jj >= nblk[r]) jj = nblk[r] - 1;
                grp_col[2 * static_cast<size_t>(p) + s] = dev_col(r, jj);
            }
    // byte offset of a 512-byte operand half -> (row, block); the kernel's
    // A operand half of lane l (< 32) is row l % 16, columns 8 (l / 16) .. + 7
    const int64_t a_bytes = static_cast<int64_t>(grp_ptr[dev_rows]) * 1024;
    std::vector<int32_t> owner_row(grp_ptr[dev_rows]);
    for (int r = 0; r < dev_rows; ++r)
        for (int p = grp_ptr[r]; p < grp_ptr[r + 1]; ++p) owner_row[p] = r;

    std::vector<int32_t> panel_ptr, recs;
    build_panel_schedule(grp_ptr, grp_col, nblk, dev_rows, panel_ptr, recs);
    const int panels = (dev_rows + 15) / 16;
    if (static_cast<int>(panel_ptr.size()) != panels + 1) fail("panel_ptr size");
    if (recs.size() != static_cast<size_t>(panel_ptr[panels] + 1) * kPanelRecInts) fail("records size");

    const int n_tiles = (N + 127) / 128;
    const double nan = std::nan("");
    std::vector<double> C(static_cast<size_t>(M) * N, nan);
    std::vector<int> used(static_cast<size_t>(grp_ptr[dev_rows]) * 2, 0);
    for (int p = 0; p < panels; ++p) {
        const int c0 = panel_ptr[p], nch = panel_ptr[p + 1] - c0;
        if (nch < 0) fail("chunk range", p);
        for (int t = 0; t < n_tiles; ++t) {
            std::vector<double> acc(16 * 16 * 128, 0.0);  // [w][i][column in tile]
            // three stages: B slabs [8][16][128], A records [32][16][16] and
            // the carried chunk record (kPanelRecInts ints)
            std::vector<double> bst(3 * 8 * 16 * 128, -1e300), ast(3 * 32 * 16 * 16, -1e300);
            std::vector<int32_t> rst(3 * kPanelRecInts, 0x7f7f7f7f);
            auto record = [&](int c) { return recs.data() + static_cast<size_t>(c0 + c) * kPanelRecInts; };
            // issue(c): wave w's loads of chunk c, addressed by fields f[w] =
            // (col_a, col_b, nblk, src0, src1), and the record of chunk c + 2
            auto issue = [&](int c, const int (*f)[5]) {
                const int st = c % 3;
                const bool live = c < nch;
                for (int w = 0; w < 16; ++w) {
                    for (int m = 0; m < 2; ++m) {
                        const int q = 16 * m + w, col = live ? f[w][m] : -1;
                        for (int lane = 0; lane < 64; ++lane) {
                            const int lds = q * 1024 + 16 * lane;  // stage byte
                            if (lds + 16 > 8 * 4096) fail("B stage overflow", q, lane);
                            const int slab = lds / 4096, row = (lds % 4096) / 256, oct = (lds % 256) / 16;
                            if (slab != q >> 2 || row != 4 * (q & 3) + (lane >> 4) || oct != (lane & 15))
                                fail("B piece mapping", q, lane);
                            for (int e = 0; e < 8; ++e) {
                                double v = 0.0;  // out-of-range load: zeros
                                if (col >= 0) {
                                    const int64_t grow = static_cast<int64_t>(col) * 16 + 4 * (q & 3) + (lane >> 4);
                                    const int64_t gcol = static_cast<int64_t>(t) * 128 + 8 * (lane & 15) + e;
                                    const int64_t byte = (grow * N + gcol) * 2;
                                    if (grow >= K) fail("B row out of range", col, q);
                                    if (byte < static_cast<int64_t>(K) * N * 2)
                                        v = gcol < N ? B[static_cast<size_t>(grow) * N + gcol] : nan;
                                }
                                bst[((st * 8 + slab) * 16 + row) * 128 + 8 * oct + e] = v;
                            }
                        }
                    }
                    const int nblk_c = f[w][2];
                    if (live && (nblk_c < 0 || nblk_c > kMaxBlocksPerChunk)) fail("block count", nblk_c);
                    for (int lane = 0; lane < 64; ++lane) {
                        const int r = 2 * w + (lane >> 5);
                        const int src = (lane >> 5) ? f[w][4] : f[w][3];
                        const int lds = 32768 + w * 1024 + 16 * lane;
                        const int rr = (lds - 32768) / 512, part = ((lds - 32768) % 512) / 16;
                        if (rr != r || part != (lane & 31)) fail("A record mapping", w, lane);
                        const int i = (lane & 31) % 16, k0 = 8 * ((lane & 31) / 16);
                        double *dst = &ast[((st * 32 + r) * 16 + i) * 16 + k0];
                        if (!(live && r < nblk_c)) {
                            for (int e = 0; e < 8; ++e) dst[e] = 0.0;
                            continue;
                        }
                        const int64_t off = static_cast<int64_t>(src) + 16 * (lane & 31);
                        if (src < 0 || src % 512 || off + 16 > a_bytes) fail("A offset", src, lane);
                        const int pair = src / 1024, half = (src % 1024) / 512;
                        const int orow = owner_row[pair];
                        const int jj = 2 * (pair - grp_ptr[orow]) + half;
                        if (jj >= nblk[orow]) fail("A record is padding", pair, half);
                        for (int e = 0; e < 8; ++e) dst[e] = dev_val(orow, jj, i, k0 + e);
                        if ((lane & 31) == 0) used[2 * static_cast<size_t>(pair) + half] += 1;
                    }
                }
                // wave 0 lanes < kPanelRecInts / 4: the record of chunk c + 2
                for (int k = 0; k < kPanelRecInts; ++k)
                    rst[st * kPanelRecInts + k] = c + 2 < nch ? record(c + 2)[k] : 0;
            };
            auto fields_of = [&](const int32_t *rec, int (*f)[5]) {
                for (int w = 0; w < 16; ++w) {
                    f[w][0] = rec[w >> 2];
                    f[w][1] = rec[4 + (w >> 2)];
                    f[w][2] = rec[8];
                    f[w][3] = rec[16 + 2 * w];
                    f[w][4] = rec[17 + 2 * w];
                }
            };
            // prologue: chunks 0 and 1 from the records in memory (clamped)
            const int32_t *r0p = recs.data() + static_cast<size_t>(c0) * kPanelRecInts;
            const int32_t *r1p = r0p + (nch > 1 ? kPanelRecInts : 0);
            int f[16][5];
            int info0[16], slots0[16], info1[16], slots1[16];
            for (int w = 0; w < 16; ++w) {
                info0[w] = r0p[64 + 2 * w];
                slots0[w] = r0p[65 + 2 * w];
                info1[w] = r1p[64 + 2 * w];
                slots1[w] = r1p[65 + 2 * w];
            }
            fields_of(r0p, f);
            issue(0, f);
            fields_of(r1p, f);
            issue(1, f);
            for (int c = 0; c < nch; ++c) {
                const int st = c % 3;
                const int32_t *rec2 = &rst[st * kPanelRecInts];  // record of chunk c + 2
                const int32_t *recc = record(c);                  // reference copy, for checks only
                fields_of(rec2, f);
                int info[16], slots[16];
                for (int w = 0; w < 16; ++w) {
                    info[w] = info0[w];
                    slots[w] = slots0[w];
                    info0[w] = info1[w];
                    slots0[w] = slots1[w];
                    info1[w] = rec2[64 + 2 * w];
                    slots1[w] = rec2[65 + 2 * w];
                    if (info[w] != recc[64 + 2 * w] || slots[w] != recc[65 + 2 * w]) fail("carried row fields", c, w);
                }
                issue(c + 2, f);  // into stage (c + 2) % 3, before the compute as in the kernel
                // compute reads stage st
                for (int w = 0; w < 16; ++w) {
                    const int first = info[w] & 0xff, nb = (info[w] >> 8) & 0xff;
                    const int br = 16 * p + w;
                    if (nb && br >= dev_rows) fail("blocks for a missing row", br);
                    if (first + nb > recc[8]) fail("row range", first, nb);
                    for (int b = 0; b < nb; b += 2) {
                        const int r0 = first + b;
                        const bool two = b + 1 < nb;
                        const int s0 = (slots[w] >> (3 * b)) & 7;
                        const int s1 = two ? (slots[w] >> (3 * b + 3)) & 7 : s0;
                        for (int h = 0; h < (two ? 2 : 1); ++h) {
                            const int src = recc[16 + r0 + h];
                            const int pair = src / 1024, half = (src % 1024) / 512;
                            const int orow = owner_row[pair];
                            const int jj = 2 * (pair - grp_ptr[orow]) + half;
                            const int sl = h ? s1 : s0;
                            if (orow != br) fail("block of another row", orow, br);
                            if (recc[sl] != dev_col(orow, jj)) fail("slot column", recc[sl], dev_col(orow, jj));
                        }
                        double Aop[16][32];
                        for (int lane = 0; lane < 64; ++lane) {
                            const int i = lane % 16, k0 = 8 * (lane / 16), l32 = lane & 31;
                            for (int e = 0; e < 8; ++e) {
                                double v = 0.0;  // zero record
                                if (lane < 32)
                                    v = ast[((st * 32 + r0) * 16 + l32 % 16) * 16 + 8 * (l32 / 16) + e];
                                else if (two)
                                    v = ast[((st * 32 + r0 + 1) * 16 + l32 % 16) * 16 + 8 * (l32 / 16) + e];
                                Aop[i][k0 + e] = v;
                            }
                        }
                        for (int i = 0; i < 16; ++i)
                            for (int jcol = 0; jcol < 16; ++jcol)
                                for (int e = 0; e < 8; ++e) {
                                    double sum = 0.0;
                                    for (int g = 0; g < 4; ++g) {
                                        const int slab = (g >> 1) ? s1 : s0;
                                        for (int ii = 0; ii < 8; ++ii)
                                            sum += Aop[i][8 * g + ii] *
                                                   bst[((st * 8 + slab) * 16 + 8 * (g & 1) + ii) * 128 + 8 * jcol + e];
                                    }
                                    acc[(w * 16 + i) * 128 + 8 * jcol + e] += sum;
                                }
                    }
                }
            }
            // store: columns < N of rows < dev_rows
            for (int w = 0; w < 16; ++w) {
                const int br = 16 * p + w;
                if (br >= dev_rows) continue;
                for (int i = 0; i < 16; ++i)
                    for (int cc = 0; cc < 128; ++cc) {
                        const int n = t * 128 + cc;
                        if (n >= N) continue;
                        C[static_cast<size_t>(br * 16 + i) * N + n] = acc[(w * 16 + i) * 128 + cc];
                    }
            }
        }
    }
    // every real block consumed once per tile, padding never
    for (int r = 0; r < dev_rows; ++r)
        for (int jj = 0; jj < 2 * (grp_ptr[r + 1] - grp_ptr[r]); ++jj) {
            const size_t u = 2 * static_cast<size_t>(grp_ptr[r]) + jj;
            const int want = jj < nblk[r] ? n_tiles : 0;
            if (used[u] != want) fail("block use count", r, jj);
        }
    // direct product
    for (int R = 0; R < cs.block_rows; ++R)
        for (int i = 0; i < bh; ++i)
            for (int n = 0; n < N; ++n) {
                double ref = 0.0;
                for (int k = hdr[R]; k < hdr[R + 1]; ++k)
                    for (int cc = 0; cc < cs.bw; ++cc)
                        ref += data[static_cast<size_t>(k) * bh * cs.bw + i * cs.bw + cc] *
                               B[static_cast<size_t>(idx[k] * cs.bw + cc) * N + n];
                const double got = C[static_cast<size_t>(R * bh + i) * N + n];
                if (!(std::fabs(got - ref) <= 1e-9 * (1.0 + std::fabs(ref)))) fail("product", R * bh + i, n);
            }
}

}  // namespace

int main() {
    const Case cases[] = {
        {40, 96, 16, 256, 0.05, "random", 1},   // panels of 16 + a partial one, ragged union
        {33, 64, 16, 200 - 200 % 16, 0.25, "random", 2},  // dense-ish, N not a multiple of 128
        {32, 48, 32, 128, 0.20, "random", 3},  // 16 x 32 blocks (two device blocks each)
        {48, 64, 16, 384, 0.0, "band", 4},     // band: few slabs per panel
        {20, 40, 16, 144, 0.3, "empty_rows", 5},
        {16, 8, 16, 128, 1.0, "random", 6},    // every row uses every slab: chunks cut by A capacity
        {5, 30, 16, 64, 0.5, "random", 7},     // one partial panel, one ragged tile
        {17, 1, 16, 16, 1.0, "random", 8},     // single block-column
        {18, 50, 16, 128, 0.0, "random", 9},   // no blocks at all
    };
    int n = 0;
    for (const Case &c : cases) {
        run_case(c);
        ++n;
    }
    std::printf("ok %d\n", n);
    return 0;
}
