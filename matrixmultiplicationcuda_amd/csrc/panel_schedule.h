// panel_schedule.h -- chunk schedule of the slab-streaming panel kernel
// (16-row device blocks, fp16 / bf16 pair operands).  Plain C++.
#pragma once

#include <cstdint>
#include <vector>

// Chunk limits: slabs of B (16 rows x 128 columns) and 512-byte A blocks
// staged per chunk, and the ints of one chunk record.
constexpr int kMaxSlabsPerChunk = 8;
constexpr int kMaxBlocksPerChunk = 32;
constexpr int kPanelRecInts = 96;

// grp_ptr / grp_col: the plan's pair groups (2 columns per group, padding
// repeats the last column); nblk[r]: real device blocks of device row r.
// Fills panel_ptr (panels + 1 chunk offsets) and recs (kPanelRecInts per
// chunk plus one spare record).
void build_panel_schedule(const std::vector<int32_t> &grp_ptr, const std::vector<int32_t> &grp_col,
                          const std::vector<int32_t> &nblk, int dev_rows, std::vector<int32_t> &panel_ptr,
                          std::vector<int32_t> &recs);
