// Work list of the multi-RHS / persistent Dslash launch (passed by value as a
// kernel argument, so a captured graph replays it without device tables).
#pragma once

#include <hip/hip_runtime.h>

#include "dslash_common.h"
#include "dslash_mi355x.h"

namespace wd {

constexpr int MRHS_MAX_LATTICES = 8;
constexpr int MRHS_T_BLOCK = 4;      // t-slices per thread in t-marching tiles
constexpr int MRHS_T_BLOCK_LDS = 8;  // t-slices per block in LDS-staged t-marching tiles
constexpr int MRHS_LDS_THREADS = 128;  // threads (= spatial sites) per LDS-staged block
constexpr int MRHS_LDS_GROUP = 8;    // RHS blocks of one spatial block run together
constexpr int MRHS_T_BOX = 16;       // t_block value of LDS-staged box tiles
constexpr int MRHS_BOX_SLICES = 8;   // t-slices per box tile
constexpr int MRHS_BOX_WIDE = 768;   // most threads of a wide box tile (FP32, one field per tile)

// Tile shapes: t_block 1 (one site per thread, any rhs_block), MRHS_T_BLOCK
// (t-marching, rhs_block <= 2), MRHS_T_BLOCK_LDS (t-marching with an LDS slab,
// F32, rhs_block <= 2, MRHS_LDS_THREADS threads), MRHS_T_BOX (LDS-staged box
// tiles, F32 with rhs_block <= 2 or F64 with rhs_block 1, >= 128 threads; the
// single-field path also takes F64S).  Blocks hold 64..256 threads, except wide
// box tiles: F32, rhs_block x threads <= MRHS_BOX_WIDE (one block per CU, 144 KiB
// of LDS at 768 field-sites).
inline bool mrhs_shape_ok(int prec, int R, int block, int t_block) {
  if (!(R == 1 || R == 2 || R == 4) || block < 64 || block % 64) return false;
  if (block > 256)
    return t_block == MRHS_T_BOX && prec == F32 && R <= 2 && R * block <= MRHS_BOX_WIDE;
  if (t_block == 1) return true;
  if (t_block == MRHS_T_BLOCK) return R <= 2;
  if (t_block == MRHS_T_BLOCK_LDS) return prec == F32 && R <= 2 && block == MRHS_LDS_THREADS;
  if (t_block == MRHS_T_BOX) return (prec == F32 ? R <= 2 : (prec == F64 || prec == F64S) && R == 1) && block >= 128;
  return false;
}

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
  uint32_t rg;       // RHS blocks per group (LDS-staged tiles; nrb otherwise)
  uint32_t S;        // target sites per t-slice (Lx/2 * Ly * Lz)
  FastDiv dlxh, dly; // for the spatial coordinates of a t-slice site
  uint32_t ty, tz;   // box tiles: a tile is Lx/2 x ty x tz sites of a slice
  uint32_t nty;      // box tiles along y (ntz = nsb / nty along z)
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
// LDS-staged tiles also need every problem to have nrhs % R == 0 and one isign.
inline bool mrhs_lds_compatible(const dslash_mrhs_problem* probs, int nprob, int R) {
  for (int i = 0; i < nprob; ++i)
    if (probs[i].nrhs % R != 0 || probs[i].isign != probs[0].isign) return false;
  return nprob >= 1;
}
// Box tiles need a whole x row of the target checkerboard in one block.
inline bool mrhs_box_compatible(const dslash_mrhs_problem* probs, int nprob, int R, int block) {
  if (!mrhs_lds_compatible(probs, nprob, R)) return false;
  for (int i = 0; i < nprob; ++i)
    if (probs[i].latt[0] / 2 > block) return false;
  return true;
}
int mrhs_launch(const MrhsList& list, int prec, int G, int R, int block, int t_block, uint32_t grid, hipStream_t s);

// Epilogue operands of the single-field box launch: with xp set, res = xp + a D psi
// (xp may alias res); with norm_partial also set, tile t writes its (|res|^2, 0)
// pair to norm_partial[2 t], norm_partial[2 t + 1].
struct BoxEpi {
  const v4u* xp;
  double a;
  double* norm_partial;
};
// res = D psi (or an epilogue form) for one FP32, FP64 or sloppy-FP64 4D field
// pair with box tiles and non-temporal links (the single-RHS path's march 16, and
// march 32 with the halo below);
// needs box_tiles_single(latt, block) > 0 and block >= 128.
// halo: FP32 and sloppy-FP64 tiles whose y / z hops that leave the box read an LDS
// halo of the neighbouring sites, loaded one slice ahead (needs
// box_halo_single(latt, block) > 0; ignored for FP64, which has no halo form).
int box_apply_single(void* res_cb, const void* psi_cb, const void* gauge_cb, const int latt[4], int prec, int G,
                     int isign, int tcb, int block, const BoxEpi& epi, hipStream_t s, bool halo = false);
// Tiles (= blocks and |res|^2 partial sums) of that launch; 0 when Lx/2 > block,
// when a box would hold fewer than 64 sites (the lanes would mostly idle) or when
// the tiles would outnumber ceil(Vh / 64), the one-site kernel's most blocks.
uint32_t box_tiles_single(const int latt[4], int block);
// Halo entries of those tiles (one per thread at most, so 0 also when they would
// exceed block, or when the box spans y and z or box tiles do not apply).
uint32_t box_halo_single(const int latt[4], int block);

}  // namespace wd
