/* A C caller of the drop-in API, shaped like the reference's own full-lattice test
 * (tests/testDslashFull.cc:24-117 and tests/testDslash3D.cc): it registers
 * site-order callbacks with init_sse_su3dslash (or _3d), applies
 * sse_su3dslash_wilson (or _3d) for isign +1 / -1 and source checkerboard 0 / 1
 * to host arrays, frees the library and writes gauge, source and results as raw
 * SSEREAL arrays for the test harness to compare with the oracle.
 *
 *   dropin_check nd Lx Ly Lz Lt outdir
 *
 * Site order: the canonical cb2 order (4D, lib/shift_tables_scalar.c:99-123) or
 * cb3d order (3D, lib/shift_tables_scalar_3d.c:97-137).  Gauge and source are
 * uniform in [-1, 1) from a 64-bit LCG (the operator does not need unitary
 * links); every result array starts as 7.0 so that a write to the source
 * checkerboard shows up.  A fifth call applies the operator in place (res == psi),
 * which the legacy contract allows because the two checkerboards are disjoint. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sse_dslash.h"
#include "sse_dslash_3d.h"

static int L[4], ND, VH;

static int parity(const int c[]) {
  int s = c[0] + c[1] + c[2] + (ND == 4 ? c[3] : 0);
  return s & 1;
}

static int linear_of(const int c[]) {
  const int xh = c[0] / 2;
  if (ND == 4) return parity(c) * VH + xh + (L[0] / 2) * (c[1] + L[1] * (c[2] + L[2] * c[3]));
  return parity(c) * VH + c[3] + L[3] * (xh + (L[0] / 2) * (c[1] + L[1] * c[2]));
}

static void coords_of(int coord[], int node, int linear) {
  (void)node;
  const int cb = linear / VH;
  int r = linear - cb * VH, xh;
  if (ND == 4) {
    xh = r % (L[0] / 2);
    r /= L[0] / 2;
    coord[1] = r % L[1];
    r /= L[1];
    coord[2] = r % L[2];
    coord[3] = r / L[2];
  } else {
    coord[3] = r % L[3];
    r /= L[3];
    xh = r % (L[0] / 2);
    r /= L[0] / 2;
    coord[1] = r % L[1];
    coord[2] = r / L[1];
  }
  coord[0] = 2 * xh;
  if (parity(coord) != cb) coord[0] += 1;
}

static int node_of(const int coord[]) {
  (void)coord;
  return 0;
}

static unsigned long long lcg = 0x2545F4914F6CDD1DULL;
static SSEREAL uniform(void) {
  lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
  return (SSEREAL)((double)(lcg >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

static int dump(const char* dir, const char* name, const SSEREAL* a, size_t n) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE* f = fopen(path, "wb");
  if (!f) return 1;
  const size_t w = fwrite(a, sizeof(SSEREAL), n, f);
  fclose(f);
  return w != n;
}

int main(int argc, char** argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: %s nd Lx Ly Lz Lt outdir\n", argv[0]);
    return 2;
  }
  ND = atoi(argv[1]);
  for (int mu = 0; mu < 4; ++mu) L[mu] = atoi(argv[2 + mu]);
  const char* out = argv[6];
  if (ND != 3 && ND != 4) return 2;
  const int V = L[0] * L[1] * L[2] * L[3];
  VH = V / 2;
  const size_t nu = (size_t)V * 4 * 18, ns = (size_t)V * 24;
  SSEREAL* u = (SSEREAL*)aligned_alloc(64, nu * sizeof(SSEREAL));
  SSEREAL* psi = (SSEREAL*)aligned_alloc(64, ns * sizeof(SSEREAL));
  SSEREAL* res = (SSEREAL*)aligned_alloc(64, ns * sizeof(SSEREAL));
  if (!u || !psi || !res) return 3;
  for (size_t i = 0; i < nu; ++i) u[i] = uniform();
  for (size_t i = 0; i < ns; ++i) psi[i] = uniform();
  if (dump(out, "u.bin", u, nu) || dump(out, "psi.bin", psi, ns)) return 4;

  if (ND == 4)
    init_sse_su3dslash(L, coords_of, linear_of, node_of);
  else
    init_sse_su3dslash_3d(L, coords_of, linear_of, node_of);
  for (int isign = 1; isign >= -1; isign -= 2)
    for (int cb = 0; cb < 2; ++cb) {
      for (size_t i = 0; i < ns; ++i) res[i] = (SSEREAL)7.0;
      if (ND == 4)
        sse_su3dslash_wilson(u, psi, res, isign, cb);
      else
        sse_su3dslash_wilson_3d(u, psi, res, isign, cb);
      char name[64];
      snprintf(name, sizeof(name), "res_%d_%d.bin", isign > 0 ? 1 : -1, cb);
      if (dump(out, name, res, ns)) return 4;
    }
  /* in place: psi doubles as res; the source checkerboard (cb 0) stays psi */
  if (ND == 4)
    sse_su3dslash_wilson(u, psi, psi, 1, 0);
  else
    sse_su3dslash_wilson_3d(u, psi, psi, 1, 0);
  if (dump(out, "inplace_1_0.bin", psi, ns)) return 4;
  if (ND == 4)
    free_sse_su3dslash();
  else
    free_sse_su3dslash_3d();
  printf("SSEREAL %d\n", (int)sizeof(SSEREAL));
  free(u);
  free(psi);
  free(res);
  return 0;
}
