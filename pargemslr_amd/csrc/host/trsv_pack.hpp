// Level-set packing of the block-diagonal ILU factors of one GeMSLR level for
// the device triangular solves.
//
// The reference builds level sets for its OpenMP solve in IluClass::BuildLevels
// (ParGeMSLR/SRC/preconditioners/ilu.cpp:832-1057) and leaves the device solve
// to cuSPARSE csrsv2 with USE_LEVEL analysis (ilu_cuda.cu:225-459).  Here every
// diagonal block gets its own level sets, for L (forward) and U (backward).
// Two schedules are valid for a sweep: as soon as possible (level(i) = 1 +
// max level of the rows i depends on) and as late as possible (level(j) =
// min level of the rows depending on j, minus 1); both have the same number
// of levels.  The one whose values stay alive over fewer sweep positions is
// used (see window below).  A level set with more than `max_rows` rows is cut
// into consecutive steps (rows of one level are independent, so any split is
// valid).  Each step is stored as a small row-major ELL slab of
// (column, value) pairs:
//   entry (row t of the step, slot j) at ent_off + t * width + j,
// padded with (column 0, value 0), plus one (row, 1/pivot) pair per row.
// Offsets are relative to the block's first entry / first row pair so that a
// device kernel can address a block with 32-bit offsets.
//
// Sweep window.  Number the row pairs of a block in storage order (all L rows,
// then all U rows); a row's position is its pair index.  The window of a
// sweep is the largest (end position of a step) - (position of a row that
// step reads).  When it is at most W (a power of two), a solve can keep x in
// a ring of W slots, slot = position & (W - 1): no value is overwritten before
// its last reader has run.  With `max_window` > 0 and every sweep's window
// within it, the ELL columns hold those slots instead of row ids and
// `window` is the ring size; otherwise columns are block-local row ids and
// `window` is 0.  Ring-packed slabs have a fixed width, so a lane that reads
// slot j of its row never reaches into the next row, and padded slots weigh
// zero.
#pragma once

#include <cstdint>
#include <vector>

#include "gemslr_host.hpp"

namespace pgs {

template <typename T>
struct TrsvPack {
  int nblocks = 0;
  int nloc = 0;
  int max_block_rows = 0;
  int max_steps = 0;                 // deepest block, L + U steps
  int max_width = 0;                 // widest step
  int window_needed = 0;             // largest sweep window over all blocks
  int window = 0;                    // ring size when columns are ring slots, else 0
  std::vector<int> blk_row0;         // nblocks+1, level-local first row of each block
  std::vector<int> blk_step;         // nblocks+1, step ranges into desc
  std::vector<int> blk_stepU;        // nblocks, first U step of each block
  std::vector<int64_t> blk_ent0;     // nblocks+1, first entry of each block
  std::vector<int64_t> blk_ri0;      // nblocks+1, first row pair of each block
  std::vector<int> desc;             // 4 per step: row_off, ent_off (block relative), nrows, width
  std::vector<int> col;              // ELL columns (block-local row ids or ring slots)
  std::vector<T> val;                // ELL values
  std::vector<int> rows;             // per step row: block-local row id
  std::vector<T> dinv;               // per step row: 1/pivot for U steps, 0 for L steps
  int64_t nnz_real = 0;              // entries before padding
};

// max_window: largest ring (a power of two, in rows) the caller can hold; 0
// keeps row ids.  ring_width: with ring slots, every step's slab is exactly
// this wide (rows padded with (0, 0)), and rows longer than it keep row ids.
template <typename T>
TrsvPack<T> trsv_pack(const HostLevel<T>& L, int lev_start, int nloc, int max_rows, int max_window = 0,
                      int ring_width = 0);

// Half steps of a ring packing whose steps hold at most 32 rows, for a solve
// that takes 16 rows per pass: a step of more than 16 rows becomes two
// consecutive half steps (rows 0-15, then the rest) over the same slabs and
// row pairs.  Rows of a step are independent and a half step reads and
// writes a subset of its step's positions, so the ring window still holds.
// desc_h / step_h / stepU_h have the layout of desc / blk_step / blk_stepU.
void half_steps(const std::vector<int>& desc, const std::vector<int>& blk_step, const std::vector<int>& blk_stepU,
                std::vector<int>& desc_h, std::vector<int>& step_h, std::vector<int>& stepU_h);

}  // namespace pgs
