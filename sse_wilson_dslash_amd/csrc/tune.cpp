// Launch-shape autotuner and the batched (multi-RHS / multi-lattice) entry point.
//
// The tuner times every candidate shape with HIP events on the caller's
// stream, on the caller's own fields, and keeps the fastest per kernel family:
//   single-RHS Dslash:  (nd, format, link reals) -> threads per block, XCD remap
//   batched Dslash:     (format, link reals, persistent) -> RHS block, threads
//                        per block, persistent grid size, t-slices per thread
// Choices can be saved to and reloaded from a small text file, so a run tunes
// once per machine and later processes start tuned.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include "dslash_common.h"
#include "dslash_internal.h"
#include "dslash_mi355x.h"
#include "mrhs.h"
#include "status.h"

namespace wd {

struct MrhsCfg {
  int R;
  int block;
  int grid;  // persistent grid blocks (0 for the one-tile-per-block launch)
  int tb;    // tile shape: 1, MRHS_T_BLOCK (t-marching), MRHS_T_BLOCK_LDS (t-marching with an LDS
             // slab) or MRHS_T_BOX (the same with box tiles)
};

// [format F32/F64][G==12][persistent]; defaults measured at 32 RHS on 24^3 x 48
static MrhsCfg g_mrhs[2][2][2] = {
    {{{2, 192, 0, MRHS_T_BOX}, {1, MRHS_LDS_THREADS, 1024, MRHS_T_BLOCK_LDS}},
     {{2, 192, 0, MRHS_T_BOX}, {1, MRHS_LDS_THREADS, 1024, MRHS_T_BLOCK_LDS}}},
    {{{2, 128, 0, 1}, {2, 128, 2048, 1}}, {{2, 128, 0, 1}, {2, 128, 2048, 1}}}};

static MrhsCfg& mrhs_entry(int prec, int G, int persistent) {
  return g_mrhs[prec == F64][G == 12][persistent ? 1 : 0];
}

static int mrhs_run(const dslash_mrhs_problem* probs, int nprob, int prec, int G, int R, int block, int grid,
                    int tb, hipStream_t s) {
  MrhsList list;
  if (int e = mrhs_build_list(probs, nprob, prec, R, block, tb, &list)) return e;
  uint32_t g = grid > 0 ? (uint32_t)grid : list.total_tiles;
  return mrhs_launch(list, prec, G, R, block, tb, g, s);
}

template <typename F>
static int time_it(hipStream_t s, int reps, F launch, float* ms) {
  hipEvent_t e0, e1;
  WD_HIP(hipEventCreate(&e0));
  WD_HIP(hipEventCreate(&e1));
  int err = launch();  // warm-up (also validates the shape)
  if (!err) {
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < reps && !err; ++i) err = launch();
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return err;
}

}  // namespace wd

using namespace wd;

extern "C" {

int dslash_apply_mrhs(const dslash_mrhs_problem* probs, int nprob, int prec, int G, int rhs_block, int block_threads,
                      int grid_blocks, void* stream) {
  clear_error();
  WD_REQUIRE(prec == F32 || prec == F64, "multi-RHS Dslash is built for F32 and F64");
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(grid_blocks >= 0, "grid_blocks must be >= 0");
  const MrhsCfg& c = mrhs_entry(prec, G, grid_blocks > 0);
  int R = rhs_block ? rhs_block : c.R;
  int block = block_threads ? block_threads : c.block;
  // the tabled tile shape, or the closest one this launch supports
  int tb = c.tb;
  if (tb == MRHS_T_BOX && (!mrhs_shape_ok(prec, R, block, tb) || !probs || nprob < 1 ||
                           !mrhs_box_compatible(probs, nprob, R, block)))
    tb = MRHS_T_BLOCK_LDS;
  if (tb == MRHS_T_BLOCK_LDS && (!mrhs_shape_ok(prec, R, block, tb) || !probs || nprob < 1 ||
                                 !mrhs_lds_compatible(probs, nprob, R)))
    tb = MRHS_T_BLOCK;
  if (!mrhs_shape_ok(prec, R, block, tb)) tb = 1;
  if (tb != MRHS_T_BOX && block > 256) block = 256;  // a wide box shape that this launch cannot run
  return mrhs_run(probs, nprob, prec, G, R, block, grid_blocks, tb, (hipStream_t)stream);
}

int dslash_get_mrhs_launch(int prec, int G, int persistent, int* rhs_block, int* block_threads, int* grid_blocks,
                           int* t_block) {
  clear_error();
  WD_REQUIRE(prec == F32 || prec == F64, "multi-RHS Dslash is built for F32 and F64");
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  const MrhsCfg& c = mrhs_entry(prec, G, persistent);
  if (rhs_block) *rhs_block = c.R;
  if (block_threads) *block_threads = c.block;
  if (grid_blocks) *grid_blocks = c.grid;
  if (t_block) *t_block = c.tb;
  return 0;
}

int dslash_set_mrhs_launch(int prec, int G, int persistent, int rhs_block, int block_threads, int grid_blocks,
                           int t_block) {
  clear_error();
  WD_REQUIRE(prec == F32 || prec == F64, "multi-RHS Dslash is built for F32 and F64");
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(rhs_block == 1 || rhs_block == 2 || rhs_block == 4, "rhs_block must be 1, 2 or 4");
  WD_REQUIRE(block_threads >= 64 && block_threads <= MRHS_BOX_WIDE && block_threads % 64 == 0,
             "block_threads must be a multiple of 64 in 64..256 (wide box tiles: up to %d)", MRHS_BOX_WIDE);
  WD_REQUIRE(mrhs_shape_ok(prec, rhs_block, block_threads, t_block),
             "t_block must be 1, %d (rhs_block 1 or 2), %d (F32, rhs_block 1 or 2, %d threads) or %d (F32, "
             "rhs_block 1 or 2, >= 128 threads; rhs_block 1 above 256 threads, up to %d)", MRHS_T_BLOCK,
             MRHS_T_BLOCK_LDS, MRHS_LDS_THREADS, MRHS_T_BOX, MRHS_BOX_WIDE);
  WD_REQUIRE(persistent ? grid_blocks > 0 : grid_blocks == 0, "grid_blocks must be > 0 exactly when persistent");
  mrhs_entry(prec, G, persistent) = MrhsCfg{rhs_block, block_threads, grid_blocks, t_block};
  return 0;
}

int dslash_tune_apply(void* d_res, const void* d_psi, const void* d_gauge, const int latt[4], int nd, int prec, int G,
                      void* stream, int* block_threads, int* xcd_remap, float* best_ms) {
  clear_error();
  if (int e = dslash_check_lattice(latt, nd)) return e;
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(prec >= F32 && prec < NUM_PREC, "unknown precision");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t vh = (uint32_t)((uint64_t)latt[0] * latt[1] * latt[2] * latt[3] / 2);
  const int reps = vh > (1u << 18) ? 24 : 200;
  const LaunchCfg saved = launch_cfg(nd, prec, G);
  // Candidates are timed in three interleaved rounds and scored by their best
  // round, so a transient slowdown of the device does not pick a shape.
  const int blocks[4] = {64, 128, 192, 256};
  const int marches[5] = {1, 4, 8, MARCH_BOX, MARCH_BOX_HALO};
  // box tiles (16): 4D FP32, FP64, sloppy FP64; with the LDS halo (32): 4D FP32, sloppy FP64
  const int nm = nd == 4 && (prec == F32 || prec == F64S) ? 5 : (nd == 4 && prec == F64) ? 4 : 3;
  float tmin[4][2][5];
  for (auto& a : tmin)
    for (auto& b : a)
      for (float& t : b) t = 1e30f;
  for (int round = 0; round < 3; ++round)
    for (int bi = 0; bi < 4; ++bi)
      for (int remap : {0, 1})
        for (int mi = 0; mi < nm; ++mi) {
          // box tiles ignore the remap flag (their tile order is XCD-contiguous by
          // construction); tuning them with remap 1 only keeps the flag right for
          // the one-site kernels that share this entry (the domain-decomposition
          // interior, lattices below the box thread floor)
          if (marches[mi] >= MARCH_BOX && (blocks[bi] < 128 || remap == 0)) continue;
          set_launch_cfg(nd, prec, G, LaunchCfg{blocks[bi], remap, marches[mi]});
          float t = 0;
          int k = 0;
          // the same alternation of isign and source checkerboard as a solver sees
          int e = time_it(s, reps, [&]() {
            ++k;
            return dslash_apply_full(d_res, d_psi, d_gauge, latt, nd, prec, G, (k & 2) ? -1 : 1, k & 1, stream);
          }, &t);
          if (e) {
            set_launch_cfg(nd, prec, G, saved);
            return e;
          }
          if (t < tmin[bi][remap][mi]) tmin[bi][remap][mi] = t;
        }
  LaunchCfg best = saved;
  float best_t = 1e30f;
  for (int bi = 0; bi < 4; ++bi)
    for (int remap : {0, 1})
      for (int mi = 0; mi < nm; ++mi)
        if (tmin[bi][remap][mi] < best_t) {
          best_t = tmin[bi][remap][mi];
          best = LaunchCfg{blocks[bi], remap, marches[mi]};
        }
  set_launch_cfg(nd, prec, G, best);
  if (block_threads) *block_threads = best.block;
  if (xcd_remap) *xcd_remap = best.remap;
  if (best_ms) *best_ms = best_t;
  return 0;
}

int dslash_tune_mrhs(const dslash_mrhs_problem* probs, int nprob, int prec, int G, int persistent, void* stream,
                     int* rhs_block, int* block_threads, int* grid_blocks, float* best_ms) {
  clear_error();
  WD_REQUIRE(prec == F32 || prec == F64, "multi-RHS Dslash is built for F32 and F64");
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(probs && nprob >= 1 && nprob <= MRHS_MAX_LATTICES, "1..%d problems per launch", MRHS_MAX_LATTICES);
  hipStream_t s = (hipStream_t)stream;
  MrhsCfg best{0, 0, 0, 1};
  float best_t = 1e30f;
  const int grids_p[] = {512, 1024, 2048, 4096};
  const int grids_np[] = {0};
  const int* grids = persistent ? grids_p : grids_np;
  const int ng = persistent ? 4 : 1;
  for (int tb : {1, MRHS_T_BLOCK, MRHS_T_BLOCK_LDS, MRHS_T_BOX})
    for (int R : {1, 2, 4}) {
      if (tb == MRHS_T_BLOCK_LDS && !mrhs_lds_compatible(probs, nprob, R)) continue;
      for (int block : {64, 128, 192, 256, 384, 512, 768})
        for (int gi = 0; gi < ng; ++gi) {
          if (!mrhs_shape_ok(prec, R, block, tb)) continue;
          if (tb == MRHS_T_BOX && !mrhs_box_compatible(probs, nprob, R, block)) continue;
          if (tb != MRHS_T_BOX && block == 192) continue;
          float t = 0;
          int grid = grids[gi];
          int e = time_it(s, 10, [&]() { return mrhs_run(probs, nprob, prec, G, R, block, grid, tb, s); }, &t);
          if (e) return e;
          if (t < best_t) {
            best_t = t;
            best = MrhsCfg{R, block, grid, tb};
          }
        }
    }
  mrhs_entry(prec, G, persistent) = best;
  if (rhs_block) *rhs_block = best.R;
  if (block_threads) *block_threads = best.block;
  if (grid_blocks) *grid_blocks = best.grid;
  if (best_ms) *best_ms = best_t;
  return 0;
}

int dslash_tune_save(const char* path) {
  clear_error();
  WD_REQUIRE(path, "path is NULL");
  FILE* f = fopen(path, "w");
  WD_REQUIRE(f, "cannot write %s", path);
  fprintf(f, "# dslash_mi355x launch shapes\n");
  for (int nd : {3, 4})
    for (int p = 0; p < NUM_PREC; ++p)
      for (int G : {18, 12}) {
        LaunchCfg c = launch_cfg(nd, p, G);
        fprintf(f, "apply %d %d %d %d %d %d\n", nd, p, G, c.block, c.remap, c.march);
      }
  for (int p : {F32, F64})
    for (int G : {18, 12})
      for (int per : {0, 1}) {
        MrhsCfg c = mrhs_entry(p, G, per);
        fprintf(f, "mrhs %d %d %d %d %d %d %d\n", p, G, per, c.R, c.block, c.grid, c.tb);
      }
  fclose(f);
  return 0;
}

int dslash_tune_load(const char* path) {
  clear_error();
  WD_REQUIRE(path, "path is NULL");
  FILE* f = fopen(path, "r");
  WD_REQUIRE(f, "cannot read %s", path);
  char line[256];
  int bad = 0;
  while (fgets(line, sizeof(line), f)) {
    int a[7];
    if (line[0] == '#') continue;
    if (int n = sscanf(line, "apply %d %d %d %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]); n >= 5) {
      const int march = n == 6 ? a[5] : 1;  // files written before marching carry 5 fields
      if ((a[0] == 3 || a[0] == 4) && a[1] >= 0 && a[1] < NUM_PREC && (a[2] == 18 || a[2] == 12) && a[3] >= 64 &&
          a[3] <= 256 && a[3] % 64 == 0 && march_ok(march))
        set_launch_cfg(a[0], a[1], a[2], LaunchCfg{a[3], a[4] ? 1 : 0, march});
      else
        ++bad;
    } else if (int n = sscanf(line, "mrhs %d %d %d %d %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &a[6]);
               n >= 6) {
      const int tb = n == 7 ? a[6] : 1;  // files written before t-marching tiles carry 6 fields
      if ((a[0] == F32 || a[0] == F64) && (a[1] == 18 || a[1] == 12) && a[5] >= 0 &&
          mrhs_shape_ok(a[0], a[3], a[4], tb))  // the RHS block, thread count and tile shape together
        mrhs_entry(a[0], a[1], a[2]) = MrhsCfg{a[3], a[4], a[2] ? (a[5] > 0 ? a[5] : 2048) : 0, tb};
      else
        ++bad;
    } else if (line[0] != '\n') {
      ++bad;
    }
  }
  fclose(f);
  WD_REQUIRE(bad == 0, "%d malformed line(s) in %s", bad, path);
  return 0;
}

}  // extern "C"
