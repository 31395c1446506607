// panel_schedule.cpp -- host-side chunk schedule of the slab-streaming panel
// kernel (kernels_panel.hip); plain C++ so that it can also be built into a
// CPU emulation of the kernel.
#include "panel_schedule.h"

#include <algorithm>
#include <utility>

// Chunk schedule of the panel kernel (16-row device blocks, pair operands of
// 1024 bytes): for each panel of 16 device block-rows, the union of their
// block-columns in ascending order is cut greedily into chunks of at most
// kMaxSlabsPerChunk columns and kMaxBlocksPerChunk blocks; each chunk gets a
// record of kPanelRecInts ints (layout in kernels_panel.hip).  Block jj of device row r is
// slot jj % 2 of pair grp_ptr[r] + jj / 2, so its 512-byte operand half sits
// at byte (grp_ptr[r] + jj / 2) * 1024 + (jj % 2) * 512 of vals.
void build_panel_schedule(const std::vector<int32_t> &grp_ptr, const std::vector<int32_t> &grp_col,
                          const std::vector<int32_t> &nblk, int dev_rows, std::vector<int32_t> &panel_ptr,
                          std::vector<int32_t> &recs) {
    const int panels = (dev_rows + 15) / 16;
    panel_ptr.assign(panels + 1, 0);
    recs.clear();
    std::vector<int32_t> cols;
    std::vector<std::pair<int32_t, int32_t>> uc;  // (column, rows using it)
    for (int p = 0; p < panels; ++p) {
        const int r0 = 16 * p, nr = std::min(16, dev_rows - r0);
        cols.clear();
        for (int w = 0; w < nr; ++w) {
            const int64_t base = 2 * static_cast<int64_t>(grp_ptr[r0 + w]);
            for (int jj = 0; jj < nblk[r0 + w]; ++jj) cols.push_back(grp_col[base + jj]);
        }
        std::sort(cols.begin(), cols.end());
        uc.clear();
        for (int32_t c : cols) {
            if (uc.empty() || uc.back().first != c)
                uc.push_back({c, 1});
            else
                ++uc.back().second;
        }
        int cur[16] = {0};
        size_t u = 0;
        int nchunks = 0;
        while (u < uc.size()) {
            const size_t rec0 = recs.size();
            recs.resize(rec0 + kPanelRecInts, 0);
            int32_t *rec = recs.data() + rec0;
            for (int i = 0; i < kMaxSlabsPerChunk; ++i) rec[i] = -1;
            int ns = 0, na = 0;
            while (u < uc.size() && ns < kMaxSlabsPerChunk && na + uc[u].second <= kMaxBlocksPerChunk) {
                rec[ns++] = uc[u].first;
                na += uc[u].second;
                ++u;
            }
            const int32_t last = rec[ns - 1];
            int k = 0;
            for (int w = 0; w < nr; ++w) {
                const int r = r0 + w;
                const int64_t base = 2 * static_cast<int64_t>(grp_ptr[r]);
                const int first = k;
                int slots = 0;
                while (cur[w] < nblk[r] && grp_col[base + cur[w]] <= last) {
                    const int jj = cur[w]++;
                    const int32_t c = grp_col[base + jj];
                    const int slot = static_cast<int>(std::find(rec, rec + ns, c) - rec);
                    rec[16 + k] = (grp_ptr[r] + jj / 2) * 1024 + (jj % 2) * 512;
                    slots |= slot << (3 * (k - first));
                    ++k;
                }
                rec[64 + 2 * w] = first | (k - first) << 8;
                rec[65 + 2 * w] = slots;
            }
            rec[8] = k;
            ++nchunks;
        }
        panel_ptr[p + 1] = panel_ptr[p] + nchunks;
    }
    recs.resize(recs.size() + kPanelRecInts, 0);  // spare record for panels without chunks
}
