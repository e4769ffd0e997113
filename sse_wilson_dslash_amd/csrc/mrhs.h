// Work list of the multi-RHS / persistent Dslash launch (passed by value as a
// kernel argument, so a captured graph replays it without device tables).
#pragma once

#include <hip/hip_runtime.h>

#include "dslash_common.h"
#include "dslash_mi355x.h"

namespace wd {

constexpr int MRHS_MAX_LATTICES = 8;
constexpr int MRHS_T_BLOCK = 4;  // t-slices per thread in t-marching tiles

struct MrhsLattice {
  Geom g;
  const v4u* gauge;  // bundle of the target checkerboard
  const v4u* psi;    // source field 0; field r at psi + r * fstride
  v4u* res;          // target field 0
  uint64_t fstride;  // 16-byte vectors between consecutive fields
  int nrhs;
  int tcb;
  int isign;
  uint32_t tile0;    // first tile of this lattice in the flattened tile range
  uint32_t nsb;      // site blocks (t_block = 1) or spatial-site blocks (t_block > 1)
  uint32_t nrb;      // rhs blocks
  uint32_t ntc;      // t-chunks (t_block > 1)
  uint32_t S;        // target sites per t-slice (Lx/2 * Ly * Lz)
  FastDiv dlxh, dly; // for the spatial coordinates of a t-slice site
};

struct MrhsList {
  MrhsLattice lat[MRHS_MAX_LATTICES];
  int n;
  uint32_t total_tiles;
};

// t_block = 1: a tile is (site block, rhs block).  t_block > 1: a tile is
// (t-chunk of t_block slices, block of spatial sites, rhs block) and each
// thread walks the chunk's slices.
int mrhs_build_list(const dslash_mrhs_problem* probs, int nprob, int prec, int R, int block, int t_block,
                    MrhsList* list);
int mrhs_launch(const MrhsList& list, int prec, int G, int R, int block, int t_block, uint32_t grid, hipStream_t s);

}  // namespace wd
