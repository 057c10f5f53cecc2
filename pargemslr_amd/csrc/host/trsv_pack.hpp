// Level-set packing of the block-diagonal ILU factors of one GeMSLR level for
// the device triangular solves.
//
// The reference builds level sets for its OpenMP solve in IluClass::BuildLevels
// (ParGeMSLR/SRC/preconditioners/ilu.cpp:832-1057) and leaves the device solve
// to cuSPARSE csrsv2 with USE_LEVEL analysis (ilu_cuda.cu:225-459).  Here every
// diagonal block gets its own level sets (L: level(i) = 1 + max level(j) over
// L(i,j) != 0; U: the same over U(i,j) != 0 walking from the last row).  A
// level set with more than `max_rows` rows is cut into consecutive steps (rows
// of one level are independent, so any split is valid).  Each step is stored
// as a small column-major ELL slab of (column, value) pairs:
//   entry (row t of the step, slot j) at ent_off + j * nrows + t,
// padded with (column 0, value 0), plus one (row, 1/pivot) pair per row.
// Offsets are relative to the block's first entry / first row pair so that a
// device kernel can address a block with 32-bit offsets.
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
  std::vector<int> blk_row0;         // nblocks+1, level-local first row of each block
  std::vector<int> blk_step;         // nblocks+1, step ranges into desc
  std::vector<int> blk_stepU;        // nblocks, first U step of each block
  std::vector<int64_t> blk_ent0;     // nblocks+1, first entry of each block
  std::vector<int64_t> blk_ri0;      // nblocks+1, first row pair of each block
  std::vector<int> desc;             // 4 per step: row_off, ent_off (block relative), nrows, width
  std::vector<int> col;              // ELL columns (block-local row ids)
  std::vector<T> val;                // ELL values
  std::vector<int> rows;             // per step row: block-local row id
  std::vector<T> dinv;               // per step row: 1/pivot for U steps, 0 for L steps
  int64_t nnz_real = 0;              // entries before padding
};

template <typename T>
TrsvPack<T> trsv_pack(const HostLevel<T>& L, int lev_start, int nloc, int max_rows);

}  // namespace pgs
