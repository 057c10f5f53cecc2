// Level-set packing of the block-diagonal ILU factors of one GeMSLR level for
// the device triangular solves.
//
// The reference builds level sets for its OpenMP solve in IluClass::BuildLevels
// (ParGeMSLR/SRC/preconditioners/ilu.cpp:832-1057) and leaves the device solve
// to cuSPARSE csrsv2 with USE_LEVEL analysis (ilu_cuda.cu:225-459).  Here every
// diagonal block gets its own level sets (L: level(i) = 1 + max level(j) over
// L(i,j) != 0; U: the same over U(i,j) != 0 walking from the last row), and each
// level set is stored as a small column-major ELL slab:
//   entry (row t of the set, slot j) at ent_off + j * nrows + t,
// padded with (col 0, value 0).  The device kernel processes one block per
// workgroup and one level set per step.
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
  int max_sets = 0;                  // deepest block (L or U)
  std::vector<int> blk_row0;         // nblocks+1, level-local first row of each block
  std::vector<int> blk_setL, blk_setU;  // nblocks+1, ranges into desc
  std::vector<int> desc;             // 4 per set: row_off, ent_off, nrows, width
  std::vector<int> rows;             // block-local row ids, set by set
  std::vector<int> col;              // block-local column ids (ELL)
  std::vector<T> val;                // ELL values
  std::vector<T> dinv;               // aligned with rows; Dinv of the row for U sets, 0 for L sets
  int64_t nnz_real = 0;              // entries before padding
};

template <typename T>
TrsvPack<T> trsv_pack(const HostLevel<T>& L, int lev_start, int nloc);

}  // namespace pgs
