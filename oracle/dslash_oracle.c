/*
 * CPU restatement of the reference Wilson Dslash -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: the checker that tests/ and __graft_entry__.smoke()
 * compare the HIP library and the CPU build against.  It is never linked into,
 * loaded by or called from the product libraries.
 *
 * It restates, in plain C with OpenMP over target sites, what the reference's
 * scalar path computes (JeffersonLab/sse_wilson_dslash):
 *   - lib/sse_su3dslash_64bit_scalar.c:150-170 / 32bit:159-180
 *       (sse_su3dslash_wilson: cb is the SOURCE checkerboard, sites of 1-cb are
 *        overwritten, isign selects D or D^dagger, other isign = no-op)
 *   - lib/sse_su3dslash_64bit_scalar.c:175-464 / 32bit:185-385
 *       (D_psi_fun_plus/minus: site loop, forward term U(x) P psi(x+mu),
 *        backward term U(x-mu)^dagger P psi(x-mu), gauge indexed [site][mu])
 *   - lib/site_dslash_64bit_scalar.c:20-4926
 *       (per-direction projection, SU(3) multiply with the TRANSPOSED packed
 *        link: forward r_i = sum_j u[j][i] h_j, backward r_i = sum_j conj(u[i][j]) h_j,
 *        reconstruction)
 *   - lib/shift_tables_scalar.c:99-175 (cb2 order, periodic offsets) and
 *     lib/shift_tables_scalar_3d.c:97-200 (cb3d order, t fastest)
 *   - lib/sse_su3dslash_64bit_scalar_3d.c:144-304 (3D: mu = 0,1,2 only)
 * The spin algebra is encoded as a data table (which spinor components feed the
 * two-spinor and with which unit factor, and how rows 2,3 are rebuilt) rather
 * than as code paths, so it shares no structure with the HIP kernel.
 *
 * Arithmetic is done in the precision of the instantiation (float for the
 * 32-bit path, double for the 64-bit path), as the reference does.
 *
 * Pinning (see tests/test_oracle.py): against a dense gamma-matrix evaluation
 * of the operator definition in the DeGrand-Rossi basis (numpy), the closed
 * form U=1, constant psi (D psi = 8 psi in 4D, 6 psi in 3D), the adjoint
 * identity <chi, D psi> = <D^dagger chi, psi>, gamma_5 hermiticity, gauge
 * covariance, and a brute-force neighbour/layout round trip.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- spin table -------------------------------------------------------
 * For the projector (1 + s*gamma_mu), s = +-1 (DeGrand-Rossi basis):
 *   h_a = psi_a + s * pc[mu][a] * psi_{ps[mu][a]}            (a = 0,1)
 *   out_a = r_a                                              (a = 0,1)
 *   out_{2+a} = s * rc[mu][a] * r_{rs[mu][a]}                (a = 0,1)
 * Unit factors are encoded as (re, im) pairs: 1, -1, i, -i.
 * Source: lib/site_dslash_64bit_scalar.c (1-gamma_0 :34-166, 1+gamma_0
 * :306-634, 1-gamma_1 :638-922, 1-gamma_2 :1228-, 1+gamma_3 :2171-2474).
 */
static const int ps[4][2] = {{3, 2}, {3, 2}, {2, 3}, {2, 3}};
static const double pc[4][2][2] = {
    {{0, 1}, {0, 1}},   /* mu=0: +i psi3, +i psi2 */
    {{-1, 0}, {1, 0}},  /* mu=1: -psi3, +psi2 */
    {{0, 1}, {0, -1}},  /* mu=2: +i psi2, -i psi3 */
    {{1, 0}, {1, 0}}};  /* mu=3: +psi2, +psi3 */
static const int rs[4][2] = {{1, 0}, {1, 0}, {0, 1}, {0, 1}};
static const double rc[4][2][2] = {
    {{0, -1}, {0, -1}}, /* mu=0: -i r1, -i r0 */
    {{1, 0}, {-1, 0}},  /* mu=1: +r1, -r0 */
    {{0, -1}, {0, 1}},  /* mu=2: -i r0, +i r1 */
    {{1, 0}, {1, 0}}};  /* mu=3: +r0, +r1 */

/* ---- geometry ---------------------------------------------------------- */
typedef struct {
  int L[4];
  long vh;
  int nd;
} geom_t;

static void set_geom(geom_t* g, const int L[4], int nd) {
  for (int i = 0; i < 4; ++i) g->L[i] = L[i];
  g->vh = (long)L[0] * L[1] * L[2] * L[3] / 2;
  g->nd = nd;
}

/* my full index (cb*Vh + idx) -> coordinates; lib/shift_tables_scalar.c:127-160
 * (4D cb2) and lib/shift_tables_scalar_3d.c:140-176 (3D cb3d). */
static void coords_of(const geom_t* g, long full, int c[4]) {
  int cb = (int)(full / g->vh);
  long i = full % g->vh;
  int lxh = g->L[0] / 2, xh;
  if (g->nd == 4) {
    xh = (int)(i % lxh); i /= lxh;
    c[1] = (int)(i % g->L[1]); i /= g->L[1];
    c[2] = (int)(i % g->L[2]); i /= g->L[2];
    c[3] = (int)i;
    c[0] = 2 * xh + ((cb + c[1] + c[2] + c[3]) & 1);
  } else {
    c[3] = (int)(i % g->L[3]); i /= g->L[3];
    xh = (int)(i % lxh); i /= lxh;
    c[1] = (int)(i % g->L[1]); i /= g->L[1];
    c[2] = (int)i;
    c[0] = 2 * xh + ((cb + c[1] + c[2]) & 1);
  }
}

/* coordinates -> my full index; lib/shift_tables_scalar.c:99-123 and
 * lib/shift_tables_scalar_3d.c:95-137. */
static long index_of(const geom_t* g, const int c[4]) {
  int lxh = g->L[0] / 2;
  if (g->nd == 4) {
    int cb = (c[0] + c[1] + c[2] + c[3]) & 1;
    return (long)cb * g->vh + (c[0] / 2) + (long)lxh * (c[1] + (long)g->L[1] * (c[2] + (long)g->L[2] * c[3]));
  } else {
    int cb = (c[0] + c[1] + c[2]) & 1;
    return (long)cb * g->vh + c[3] + (long)g->L[3] * ((c[0] / 2) + (long)lxh * (c[1] + (long)g->L[1] * c[2]));
  }
}

/* periodic offset, lib/shift_tables_scalar.c:165-175 */
static long neighbour(const geom_t* g, const int c[4], int mu, int dir) {
  int t[4] = {c[0], c[1], c[2], c[3]};
  t[mu] = (t[mu] + dir + 2 * g->L[mu]) % g->L[mu];
  return index_of(g, t);
}

/* ---- the operator, instantiated for float and double ------------------- */
#define DEFINE_ORACLE(REAL, SFX)                                                                          \
  static void term_##SFX(REAL out[4][3][2], const REAL* sp, const REAL* u, int mu, int s, int dagger) {   \
    REAL h[2][3][2], r[2][3][2];                                                                           \
    for (int a = 0; a < 2; ++a) {                                                                          \
      const REAL cr = (REAL)(s * pc[mu][a][0]), ciq = (REAL)(s * pc[mu][a][1]);                           \
      const REAL* q = sp + ps[mu][a] * 6;                                                                  \
      for (int k = 0; k < 3; ++k) {                                                                        \
        h[a][k][0] = sp[a * 6 + 2 * k] + (cr * q[2 * k] - ciq * q[2 * k + 1]);                           \
        h[a][k][1] = sp[a * 6 + 2 * k + 1] + (cr * q[2 * k + 1] + ciq * q[2 * k]);                       \
      }                                                                                                    \
    }                                                                                                      \
    for (int a = 0; a < 2; ++a)                                                                            \
      for (int i = 0; i < 3; ++i) {                                                                        \
        REAL re = 0, im = 0;                                                                               \
        for (int j = 0; j < 3; ++j) {                                                                      \
          /* U(i,j): forward = u[j][i]; backward uses U^dagger(i,j) = conj(u[i][j]) */                    \
          REAL ur, ui;                                                                                     \
          if (!dagger) {                                                                                   \
            ur = u[(j * 3 + i) * 2]; ui = u[(j * 3 + i) * 2 + 1];                                         \
          } else {                                                                                         \
            ur = u[(i * 3 + j) * 2]; ui = -u[(i * 3 + j) * 2 + 1];                                        \
          }                                                                                                \
          re += ur * h[a][j][0] - ui * h[a][j][1];                                                        \
          im += ur * h[a][j][1] + ui * h[a][j][0];                                                        \
        }                                                                                                  \
        r[a][i][0] = re;                                                                                   \
        r[a][i][1] = im;                                                                                   \
      }                                                                                                    \
    for (int a = 0; a < 2; ++a) {                                                                          \
      const REAL cr = (REAL)(s * rc[mu][a][0]), ciq = (REAL)(s * rc[mu][a][1]);                           \
      const int b = rs[mu][a];                                                                             \
      for (int k = 0; k < 3; ++k) {                                                                        \
        out[a][k][0] += r[a][k][0];                                                                        \
        out[a][k][1] += r[a][k][1];                                                                        \
        out[2 + a][k][0] += cr * r[b][k][0] - ciq * r[b][k][1];                                           \
        out[2 + a][k][1] += cr * r[b][k][1] + ciq * r[b][k][0];                                           \
      }                                                                                                    \
    }                                                                                                      \
  }                                                                                                        \
                                                                                                           \
  static void apply_##SFX(const REAL* u, const REAL* psi, REAL* res, const int L[4], int nd, int isign,   \
                          int cb, const int* site_table) {                                                 \
    if (isign != 1 && isign != -1) return; /* no-op, lib/sse_su3dslash_32bit_scalar.c:162-179 */          \
    geom_t g;                                                                                              \
    set_geom(&g, L, nd);                                                                                   \
    const int tcb = 1 - cb;                                                                                \
    long n;                                                                                                \
    _Pragma("omp parallel for schedule(static)") for (n = 0; n < g.vh; ++n) {                            \
      long me = (long)tcb * g.vh + n;                                                                      \
      int c[4];                                                                                            \
      coords_of(&g, me, c);                                                                                \
      long here = site_table ? site_table[me] : me;                                                        \
      REAL out[4][3][2];                                                                                   \
      memset(out, 0, sizeof(out));                                                                         \
      for (int mu = 0; mu < nd; ++mu) {                                                                    \
        long f = neighbour(&g, c, mu, +1), b = neighbour(&g, c, mu, -1);                                   \
        if (site_table) { f = site_table[f]; b = site_table[b]; }                                          \
        /* forward: U_mu(x) (1 - isign gamma_mu) psi(x+mu) */                                             \
        term_##SFX(out, psi + f * 24, u + (here * 4 + mu) * 18, mu, -isign, 0);                            \
        /* backward: U_mu(x-mu)^dagger (1 + isign gamma_mu) psi(x-mu) */                                  \
        term_##SFX(out, psi + b * 24, u + (b * 4 + mu) * 18, mu, +isign, 1);                               \
      }                                                                                                    \
      memcpy(res + here * 24, out, sizeof(out));                                                           \
    }                                                                                                      \
  }

DEFINE_ORACLE(double, f64)
DEFINE_ORACLE(float, f32)

/* ---- exported entry points (ctypes) ------------------------------------ */
void oracle_dslash_f64(const double* u, const double* psi, double* res, const int L[4], int isign, int cb,
                       const int* site_table) {
  apply_f64(u, psi, res, L, 4, isign, cb, site_table);
}
void oracle_dslash_f32(const float* u, const float* psi, float* res, const int L[4], int isign, int cb,
                       const int* site_table) {
  apply_f32(u, psi, res, L, 4, isign, cb, site_table);
}
void oracle_dslash_3d_f64(const double* u, const double* psi, double* res, const int L[4], int isign, int cb,
                          const int* site_table) {
  apply_f64(u, psi, res, L, 3, isign, cb, site_table);
}
void oracle_dslash_3d_f32(const float* u, const float* psi, float* res, const int L[4], int isign, int cb,
                          const int* site_table) {
  apply_f32(u, psi, res, L, 3, isign, cb, site_table);
}

/* my full index <-> coordinates, exported so tests can check the layout */
void oracle_coords(const int L[4], int nd, long full, int c[4]) {
  geom_t g;
  set_geom(&g, L, nd);
  coords_of(&g, full, c);
}
long oracle_index(const int L[4], int nd, const int c[4]) {
  geom_t g;
  set_geom(&g, L, nd);
  return index_of(&g, c);
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
