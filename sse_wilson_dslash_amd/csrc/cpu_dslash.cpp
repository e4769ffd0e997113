// Host (CPU) Wilson Dslash behind the reference's drop-in C API, for nodes
// without a GPU: the same symbols as libsse_dslash_f32.so / _f64.so
// (include/sse_dslash.h, include/sse_dslash_3d.h), built with g++ into
// libsse_dslash_cpu_f32.so (-DSSE_PRECISION=32) and libsse_dslash_cpu_f64.so
// (-DSSE_PRECISION=64).  It replaces the reference's scalar SSE path
// (lib/sse_su3dslash_32bit_scalar.c:159-180 with the site loop of :185-385 and
// the OpenMP dispatch of lib/dispatch_scalar.c:37-67) and is selected
// explicitly by linking it; nothing falls back to it.
//
// Arithmetic layout: a "pair" vector holds one colour of two spin components,
// (re, im, re, im) -- __m128 for float, __m256d for double (AVX2 + FMA).  For
// every direction:
//   h = (1 + S gamma_mu) psi restricted to spins 0,1:  h = psi_up + c (.) psi_q,
//       where psi_q is the partner pair ((3,2) for mu = 0,1; (2,3) for mu = 2,3)
//       and c a pair of complex units fixed by (mu, S);
//   r = W h for the link W = U_mu(x) (forward) or U_mu(x - mu)^dagger (backward),
//       r_i = sum_j W_ij h_j as broadcast real/imaginary parts times h_j and
//       its re/im-swapped copy, combined with one addsub;
//   spins 0,1 accumulate r, spins 2,3 accumulate conj(c) (.) r (halves swapped
//   for mu = 0,1), which is rows 2,3 of (1 + S gamma_mu) psi times W.
// S = -isign for the forward hop (x + mu) and +isign for the backward hop, the
// operator of lib/sse_su3dslash_32bit_scalar.c:39-64 in the DeGrand-Rossi basis.
//
// Neighbours: for the canonical cb2 / cb3d orders (what QDP++ supplies) they are
// computed per row of sites (4D: an x row at fixed y, z, t; 3D: a t run at fixed
// x, y, z); for any other caller order a table of caller indices per target
// site is built at init from the callbacks, as lib/shift_tables_scalar.c:191-305
// builds its shift table.  The result is written with non-temporal stores when
// res is aligned for them (the target spinors are not read again in the call).
#include <immintrin.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "dslash_mi355x.h"
#include "sse_dslash.h"
#include "sse_dslash_3d.h"

namespace {

typedef SSEREAL Real;

// ---- pair-vector primitives -------------------------------------------------
inline __m128 pair(const float* a, const float* b) {
  return _mm_loadh_pi(_mm_castpd_ps(_mm_load_sd((const double*)a)), (const __m64*)b);
}
inline __m256d pair(const double* a, const double* b) {
  return _mm256_insertf128_pd(_mm256_castpd128_pd256(_mm_loadu_pd(a)), _mm_loadu_pd(b), 1);
}
inline __m128 bcast(const float* p) { return _mm_broadcast_ss(p); }
inline __m256d bcast(const double* p) { return _mm256_broadcast_sd(p); }
inline __m128 zero(float) { return _mm_setzero_ps(); }
inline __m256d zero(double) { return _mm256_setzero_pd(); }
inline __m128 vset(float, float a, float b, float c, float d) { return _mm_setr_ps(a, b, c, d); }
inline __m256d vset(double, double a, double b, double c, double d) { return _mm256_setr_pd(a, b, c, d); }
inline __m128 add(__m128 a, __m128 b) { return _mm_add_ps(a, b); }
inline __m256d add(__m256d a, __m256d b) { return _mm256_add_pd(a, b); }
inline __m128 sub(__m128 a, __m128 b) { return _mm_sub_ps(a, b); }
inline __m256d sub(__m256d a, __m256d b) { return _mm256_sub_pd(a, b); }
inline __m128 mul(__m128 a, __m128 b) { return _mm_mul_ps(a, b); }
inline __m256d mul(__m256d a, __m256d b) { return _mm256_mul_pd(a, b); }
inline __m128 fma(__m128 a, __m128 b, __m128 c) { return _mm_fmadd_ps(a, b, c); }
inline __m256d fma(__m256d a, __m256d b, __m256d c) { return _mm256_fmadd_pd(a, b, c); }
// (a0 - b0, a1 + b1, a2 - b2, a3 + b3)
inline __m128 addsub(__m128 a, __m128 b) { return _mm_addsub_ps(a, b); }
inline __m256d addsub(__m256d a, __m256d b) { return _mm256_addsub_pd(a, b); }
// swap re and im of both complex numbers
inline __m128 swap_ri(__m128 a) { return _mm_shuffle_ps(a, a, _MM_SHUFFLE(2, 3, 0, 1)); }
inline __m256d swap_ri(__m256d a) { return _mm256_permute_pd(a, 0x5); }
// swap the two complex numbers
inline __m128 swap_halves(__m128 a) { return _mm_shuffle_ps(a, a, _MM_SHUFFLE(1, 0, 3, 2)); }
inline __m256d swap_halves(__m256d a) { return _mm256_permute2f128_pd(a, a, 0x01); }
// both swaps at once
inline __m128 swap_both(__m128 a) { return _mm_shuffle_ps(a, a, _MM_SHUFFLE(0, 1, 2, 3)); }
inline __m256d swap_both(__m256d a) { return _mm256_permute4x64_pd(a, _MM_SHUFFLE(0, 1, 2, 3)); }

template <class T> struct VT;
template <> struct VT<float> { typedef __m128 v; };
template <> struct VT<double> { typedef __m256d v; };

// ---- one hop ----------------------------------------------------------------
// Projector data for h = (1 + S gamma_mu) psi (spins 0,1), DeGrand-Rossi basis:
//   mu 0: h = up + S i (psi3, psi2)         rows 2,3 = -S i (r1, r0)
//   mu 1: h = up + S (-psi3, psi2)          rows 2,3 = ( S r1, -S r0)
//   mu 2: h = up + S i (psi2, -psi3)        rows 2,3 = (-S i r0, S i r1)
//   mu 3: h = up + S (psi2, psi3)           rows 2,3 = ( S r0,  S r1)
// (the entries of lib/site_dslash_64bit_scalar.c:34-166, 638-926, 1228-1549,
// 2171-2474 for S = -1, and their S = +1 partners).
template <class T, int MU, int S>
inline typename VT<T>::v project(typename VT<T>::v up, typename VT<T>::v q) {
  const T s = (T)S;
  if constexpr (MU == 0) return fma(swap_ri(q), vset(T(), -s, s, -s, s), up);
  if constexpr (MU == 1) return fma(q, vset(T(), -s, -s, s, s), up);
  if constexpr (MU == 2) return fma(swap_ri(q), vset(T(), -s, s, s, -s), up);
  if constexpr (S > 0) return add(up, q);
  return sub(up, q);
}

template <class T, int MU, int S>
inline typename VT<T>::v lower(typename VT<T>::v r, typename VT<T>::v acc) {
  const T s = (T)S;
  if constexpr (MU == 0) return fma(swap_both(r), vset(T(), s, -s, s, -s), acc);
  if constexpr (MU == 1) return fma(swap_halves(r), vset(T(), s, s, -s, -s), acc);
  if constexpr (MU == 2) return fma(swap_ri(r), vset(T(), s, -s, -s, s), acc);
  if constexpr (S > 0) return add(acc, r);
  return sub(acc, r);
}

// r = W h; the packed link stores W's transpose, l[(3a + b) * 2] = U(b, a)
// (lib/qdp_packer_nopad.cc:7-19).  Forward: W = U, W_ij = l[j][i].
// Backward: W = U^dagger, W_ij = conj(U_ji) = conj(l[i][j]).
template <class T, bool DAG>
inline void su3_mul(const T* __restrict l, const typename VT<T>::v h[3], typename VT<T>::v r[3]) {
  typedef typename VT<T>::v V;
  const V sh[3] = {swap_ri(h[0]), swap_ri(h[1]), swap_ri(h[2])};
#pragma GCC unroll 3
  for (int i = 0; i < 3; ++i) {
    const int e0 = DAG ? 6 * i : 2 * i, e1 = DAG ? 6 * i + 2 : 6 + 2 * i, e2 = DAG ? 6 * i + 4 : 12 + 2 * i;
    V t = mul(bcast(l + e0), h[0]);
    V w = mul(bcast(l + e0 + 1), sh[0]);
    t = fma(bcast(l + e1), h[1], t);
    w = fma(bcast(l + e1 + 1), sh[1], w);
    t = fma(bcast(l + e2), h[2], t);
    w = fma(bcast(l + e2 + 1), sh[2], w);
    // forward: (a + ib) h -> (a hr - b hi, a hi + b hr); conjugate: (a hr + b hi, a hi - b hr)
    r[i] = DAG ? fma(w, vset(T(), 1, -1, 1, -1), t) : addsub(t, w);
  }
}

template <class T, int MU, int S, bool DAG>
inline void hop(const T* __restrict src, const T* __restrict link, typename VT<T>::v up_acc[3],
                typename VT<T>::v lo_acc[3]) {
  typedef typename VT<T>::v V;
  constexpr int q0 = MU < 2 ? 3 : 2, q1 = MU < 2 ? 2 : 3;
  V h[3], r[3];
#pragma GCC unroll 3
  for (int j = 0; j < 3; ++j)
    h[j] = project<T, MU, S>(pair(src + 2 * j, src + 6 + 2 * j), pair(src + 6 * q0 + 2 * j, src + 6 * q1 + 2 * j));
  su3_mul<T, DAG>(link, h, r);
#pragma GCC unroll 3
  for (int j = 0; j < 3; ++j) {
    up_acc[j] = add(up_acc[j], r[j]);
    lo_acc[j] = lower<T, MU, S>(r[j], lo_acc[j]);
  }
}

// ---- store one site's 24 reals ----------------------------------------------
// up[j] = (spin0 c_j, spin1 c_j), lo[j] = (spin2 c_j, spin3 c_j); memory order is
// [spin][colour][re, im].
template <bool NT>
inline void store_site(float* o, const __m128 up[3], const __m128 lo[3]) {
  const __m128* g[2] = {up, lo};
  for (int k = 0; k < 2; ++k) {
    const __m128 a = _mm_movelh_ps(g[k][0], g[k][1]);                           // s c0, s c1
    const __m128 b = _mm_shuffle_ps(g[k][2], g[k][0], _MM_SHUFFLE(3, 2, 1, 0));  // s c2, s' c0
    const __m128 c = _mm_movehl_ps(g[k][2], g[k][1]);                           // s' c1, s' c2
    if constexpr (NT) {
      _mm_stream_ps(o + 12 * k, a);
      _mm_stream_ps(o + 12 * k + 4, b);
      _mm_stream_ps(o + 12 * k + 8, c);
    } else {
      _mm_storeu_ps(o + 12 * k, a);
      _mm_storeu_ps(o + 12 * k + 4, b);
      _mm_storeu_ps(o + 12 * k + 8, c);
    }
  }
}
template <bool NT>
inline void store_site(double* o, const __m256d up[3], const __m256d lo[3]) {
  const __m256d* g[2] = {up, lo};
  for (int k = 0; k < 2; ++k) {
    const __m256d a = _mm256_permute2f128_pd(g[k][0], g[k][1], 0x20);  // s c0, s c1
    const __m256d b = _mm256_permute2f128_pd(g[k][2], g[k][0], 0x30);  // s c2, s' c0
    const __m256d c = _mm256_permute2f128_pd(g[k][1], g[k][2], 0x31);  // s' c1, s' c2
    if constexpr (NT) {
      _mm256_stream_pd(o + 12 * k, a);
      _mm256_stream_pd(o + 12 * k + 4, b);
      _mm256_stream_pd(o + 12 * k + 8, c);
    } else {
      _mm256_storeu_pd(o + 12 * k, a);
      _mm256_storeu_pd(o + 12 * k + 4, b);
      _mm256_storeu_pd(o + 12 * k + 8, c);
    }
  }
}

// One target site: f[mu] / b[mu] are the caller indices of x + mu and x - mu,
// self the caller index of x.
template <class T, int ND, int IS, bool NT>
inline void site(const T* __restrict psi, const T* __restrict u, T* __restrict res, size_t self, const size_t* f,
                 const size_t* b) {
  typedef typename VT<T>::v V;
  V up[3] = {zero(T()), zero(T()), zero(T())}, lo[3] = {zero(T()), zero(T()), zero(T())};
  const T* ux = u + self * 72;
  hop<T, 0, -IS, false>(psi + f[0] * 24, ux, up, lo);
  hop<T, 0, IS, true>(psi + b[0] * 24, u + b[0] * 72, up, lo);
  hop<T, 1, -IS, false>(psi + f[1] * 24, ux + 18, up, lo);
  hop<T, 1, IS, true>(psi + b[1] * 24, u + b[1] * 72 + 18, up, lo);
  hop<T, 2, -IS, false>(psi + f[2] * 24, ux + 36, up, lo);
  hop<T, 2, IS, true>(psi + b[2] * 24, u + b[2] * 72 + 36, up, lo);
  if constexpr (ND == 4) {
    hop<T, 3, -IS, false>(psi + f[3] * 24, ux + 54, up, lo);
    hop<T, 3, IS, true>(psi + b[3] * 24, u + b[3] * 72 + 54, up, lo);
  }
  store_site<NT>(res + self * 24, up, lo);
}

// ---- library state ------------------------------------------------------------
struct CpuDslash {
  const int nd;
  const char* name;
  int refcount = 0;
  int last_error = 0;
  int latt[4] = {0, 0, 0, 0};
  size_t V = 0, Vh = 0;
  bool canonical = true;
  // non-canonical orders: per target site (both checkerboards, my canonical
  // order) the caller index of the site, then of its nd forward and nd
  // backward neighbours
  std::vector<int32_t> table;
  const Real* resident_u = nullptr;
  // Work item shape (set at init): blk_z z-planes x blk_t t-slices (4D) or
  // x blk_t3 t values (3D).  Measured at 32^3 x 64 on 8 cores with 2 MB of L2
  // each: 2 planes x 8 slices (FP32) / x 4 slices (FP64) in 4D, 2 planes x the
  // whole t extent in 3D (shorter t chunks break the contiguous t runs).
  int blk_z = 2, blk_t = 8, blk_t3 = 0;

  CpuDslash(int nd_, const char* name_) : nd(nd_), name(name_) {}

  size_t my_index(const int c[4]) const {
    const size_t lxh = (size_t)latt[0] / 2;
    const int cb = nd == 4 ? ((c[0] + c[1] + c[2] + c[3]) & 1) : ((c[0] + c[1] + c[2]) & 1);
    size_t idx;
    if (nd == 4)
      idx = (size_t)(c[0] / 2) + lxh * ((size_t)c[1] + (size_t)latt[1] * ((size_t)c[2] + (size_t)latt[2] * (size_t)c[3]));
    else
      idx = (size_t)c[3] + (size_t)latt[3] * ((size_t)(c[0] / 2) + lxh * ((size_t)c[1] + (size_t)latt[1] * (size_t)c[2]));
    return (size_t)cb * Vh + idx;
  }
  void coords_of(size_t full, int c[4]) const {
    const size_t cb = full / Vh;
    size_t i = full % Vh;
    const size_t lxh = (size_t)latt[0] / 2;
    int xh;
    if (nd == 4) {
      xh = (int)(i % lxh); i /= lxh;
      c[1] = (int)(i % latt[1]); i /= latt[1];
      c[2] = (int)(i % latt[2]); i /= latt[2];
      c[3] = (int)i;
      c[0] = 2 * xh + (int)((cb + c[1] + c[2] + c[3]) & 1);
    } else {
      c[3] = (int)(i % latt[3]); i /= latt[3];
      xh = (int)(i % lxh); i /= lxh;
      c[1] = (int)(i % latt[1]); i /= latt[1];
      c[2] = (int)i;
      c[0] = 2 * xh + (int)((cb + c[1] + c[2]) & 1);
    }
  }

  void init(const int* ls, void (*getSiteCoords)(int[], int, int), int (*getLinearSiteIndex)(const int[])) {
    if (refcount > 0) {  // lib/sse_su3dslash_32bit_scalar.c:95-99
      ++refcount;
      return;
    }
    for (int mu = 0; mu < nd; ++mu)
      if (ls[mu] % 2 != 0) {  // lib/sse_su3dslash_32bit_scalar.c:103-110
        if (nd == 4)
          fprintf(stderr,
                  "This is a Dslash with checkerboarding in 4 dimensions. Each GLOBAL dimension must be even,  "
                  "Your lattice is not like this: latt_size[%d]=%d\n",
                  mu, ls[mu]);
        else
          fprintf(stderr,
                  "This is a Dslash with checkerboarding in 3 dimensions. GLOBAL dimension 0,1,2 (corresponding to "
                  "x,y,z) must be even,  Your lattice is not like this: latt_size[%d]=%d\n",
                  mu, ls[mu]);
        exit(1);
      }
    for (int mu = 0; mu < 4; ++mu) {
      if (ls[mu] <= 0) {
        fprintf(stderr, "%s: lattice extent %d in direction %d is not positive\n", name, ls[mu], mu);
        exit(1);
      }
      latt[mu] = ls[mu];
    }
    V = (size_t)latt[0] * latt[1] * latt[2] * latt[3];
    Vh = V / 2;
    if (V > (size_t)INT32_MAX) {
      fprintf(stderr, "%s: %zu sites exceed the int site index of the API\n", name, V);
      exit(1);
    }
    // caller index of every site, visited in my canonical order
    std::vector<int32_t> caller(V);
    std::vector<char> seen(V, 0);
    canonical = true;
    int c[4];
    for (size_t i = 0; i < V; ++i) {
      coords_of(i, c);
      const int q = getLinearSiteIndex(c);
      if (q < 0 || (size_t)q >= V || seen[q]) {
        fprintf(stderr, "%s: getLinearSiteIndex is not a permutation of 0..V-1 (site %zu -> %d)\n", name, i, q);
        exit(1);
      }
      seen[q] = 1;
      caller[i] = q;
      if ((size_t)q != i) canonical = false;
      if (getSiteCoords) {
        int cc[4] = {0, 0, 0, 0};
        getSiteCoords(cc, 0, q);
        if (my_index(cc) != i) {
          fprintf(stderr, "%s: getSiteCoords(%d) is inconsistent with getLinearSiteIndex\n", name, q);
          exit(1);
        }
      }
    }
    table.clear();
    if (!canonical) {
      const int w = 1 + 2 * nd;
      table.resize(V * (size_t)w);
      for (size_t i = 0; i < V; ++i) {
        coords_of(i, c);
        int32_t* e = &table[i * (size_t)w];
        e[0] = caller[i];
        for (int mu = 0; mu < nd; ++mu) {
          int cf[4] = {c[0], c[1], c[2], c[3]}, cbk[4] = {c[0], c[1], c[2], c[3]};
          cf[mu] = (cf[mu] + 1) % latt[mu];
          cbk[mu] = (cbk[mu] + latt[mu] - 1) % latt[mu];
          e[1 + mu] = caller[my_index(cf)];
          e[1 + nd + mu] = caller[my_index(cbk)];
        }
      }
    }
    resident_u = nullptr;
    blk_t = sizeof(Real) == 4 ? 8 : 4;
    blk_t3 = latt[3];
    // enough items for every thread: fall back to single planes if needed
    const long threads = omp_get_max_threads();
    const long items = nd == 4 ? (long)((latt[2] + 1) / 2) * ((latt[3] + blk_t - 1) / blk_t) : (long)(latt[2] + 1) / 2;
    blk_z = items >= threads ? 2 : 1;
    // SSE_DSLASH_CPU_BLOCK="bz,bt,bt3" overrides the shape (tuning on other hosts)
    if (const char* e = getenv("SSE_DSLASH_CPU_BLOCK")) sscanf(e, "%d,%d,%d", &blk_z, &blk_t, &blk_t3);
    refcount = 1;
    last_error = 0;
  }

  void release() {
    if (refcount == 0) return;  // lib/sse_su3dslash_32bit_scalar.c:127-129
    if (--refcount > 0) return;
    table.clear();
    table.shrink_to_fit();
    resident_u = nullptr;
  }

  int fail(int code, const char* what) {
    last_error = code;
    fprintf(stderr, "%s: %s\n", name, what);
    return code;
  }

  template <int ND, int IS, bool NT>
  void run_canonical(const Real* psi, const Real* u, Real* res, int cb) const {
    const int tcb = 1 - cb;
    const size_t lxh = (size_t)latt[0] / 2;
    const size_t src0 = (size_t)cb * Vh, tgt0 = (size_t)tcb * Vh;
    if constexpr (ND == 4) {
      const int Ly = latt[1], Lz = latt[2], Lt = latt[3];
      // Work items are blocks of bz z-planes x tc t-slices, walked t-major:
      // the z +- 1 neighbours of a row are in the plane just done or next, and
      // its t - 1 neighbours were touched bz planes earlier, close enough to
      // still sit in the core's L2.
      const int bz = blk_z < 1 ? 1 : (blk_z > Lz ? Lz : blk_z);
      const int tc = blk_t < 1 ? 1 : (blk_t > Lt ? Lt : blk_t);
      const int nzb = (Lz + bz - 1) / bz, ntc = (Lt + tc - 1) / tc;
#pragma omp parallel for schedule(static)
      for (long item = 0; item < (long)nzb * ntc; ++item) {
        const int zb = (int)(item % nzb), tcn = (int)(item / nzb);
        const int z0 = zb * bz, z1 = z0 + bz < Lz ? z0 + bz : Lz;
        const int t0 = tcn * tc, t1 = t0 + tc < Lt ? t0 + tc : Lt;
        for (int t = t0; t < t1; ++t)
          for (int z = z0; z < z1; ++z)
            for (int y = 0; y < Ly; ++y) {
              const size_t rw = (size_t)y + (size_t)Ly * ((size_t)z + (size_t)Lz * t);
              const int p = (tcb + y + z + t) & 1;  // parity of x at this row's target sites
              const int yp = y + 1 == Ly ? 0 : y + 1, ym = y == 0 ? Ly - 1 : y - 1;
              const int zp = z + 1 == Lz ? 0 : z + 1, zm = z == 0 ? Lz - 1 : z - 1;
              const int tp = t + 1 == Lt ? 0 : t + 1, tm = t == 0 ? Lt - 1 : t - 1;
              auto row = [&](int yy, int zz, int tt) {
                return src0 + lxh * ((size_t)yy + (size_t)Ly * ((size_t)zz + (size_t)Lz * tt));
              };
              const size_t r0 = row(y, z, t), ryp = row(yp, z, t), rym = row(ym, z, t), rzp = row(y, zp, t),
                           rzm = row(y, zm, t), rtp = row(y, z, tp), rtm = row(y, z, tm);
              const size_t self0 = tgt0 + lxh * rw;
              for (size_t xh = 0; xh < lxh; ++xh) {
                const size_t xf = p ? (xh + 1 == lxh ? 0 : xh + 1) : xh;
                const size_t xb = p ? xh : (xh == 0 ? lxh - 1 : xh - 1);
                const size_t f[4] = {r0 + xf, ryp + xh, rzp + xh, rtp + xh};
                const size_t b[4] = {r0 + xb, rym + xh, rzm + xh, rtm + xh};
                site<Real, 4, IS, NT>(psi, u, res, self0 + xh, f, b);
              }
            }
      }
    } else {
      const int Lt = latt[3], Ly = latt[1], Lz = latt[2];
      // Work items are blocks of bz z-planes x a chunk of tc3 t values: the
      // order is (z, y, x, t) inside an item, so a site's z +- 1 neighbours were
      // touched Lx/2 * Ly * tc3 sites earlier and stay in L2, where whole
      // t runs would put them a full z-plane apart.
      const int bz = blk_z < 1 ? 1 : (blk_z > Lz ? Lz : blk_z);
      const int tc = blk_t3 < 1 ? 1 : (blk_t3 > Lt ? Lt : blk_t3);
      const int nzb = (Lz + bz - 1) / bz, ntc = (Lt + tc - 1) / tc;
#pragma omp parallel for schedule(static)
      for (long item = 0; item < (long)nzb * ntc; ++item) {
        const int zb = (int)(item / ntc), tcn = (int)(item % ntc);
        const int z0 = zb * bz, z1 = z0 + bz < Lz ? z0 + bz : Lz;
        const int t0 = tcn * tc, t1 = t0 + tc < Lt ? t0 + tc : Lt;
        for (int z = z0; z < z1; ++z)
          for (int y = 0; y < Ly; ++y)
            for (size_t xh = 0; xh < lxh; ++xh) {
              const int p = (tcb + y + z) & 1;
              const size_t xf = p ? (xh + 1 == lxh ? 0 : xh + 1) : xh;
              const size_t xb = p ? xh : (xh == 0 ? lxh - 1 : xh - 1);
              const int yp = y + 1 == Ly ? 0 : y + 1, ym = y == 0 ? Ly - 1 : y - 1;
              const int zp = z + 1 == Lz ? 0 : z + 1, zm = z == 0 ? Lz - 1 : z - 1;
              auto run = [&](size_t xx, int yy, int zz) {
                return src0 + (size_t)Lt * (xx + lxh * ((size_t)yy + (size_t)Ly * zz));
              };
              const size_t fx = run(xf, y, z), bx = run(xb, y, z), fy = run(xh, yp, z), by = run(xh, ym, z),
                           fz = run(xh, y, zp), bz_ = run(xh, y, zm);
              const size_t self0 = tgt0 + (size_t)Lt * (xh + lxh * ((size_t)y + (size_t)Ly * z));
              for (int t = t0; t < t1; ++t) {
                const size_t f[3] = {fx + t, fy + t, fz + t};
                const size_t b[3] = {bx + t, by + t, bz_ + t};
                site<Real, 3, IS, NT>(psi, u, res, self0 + t, f, b);
              }
            }
      }
    }
  }

  template <int ND, int IS, bool NT>
  void run_table(const Real* psi, const Real* u, Real* res, int cb) const {
    const size_t tgt0 = (size_t)(1 - cb) * Vh;
    const int w = 1 + 2 * ND;
    const int32_t* tab = table.data();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)Vh; ++i) {
      const int32_t* e = tab + (tgt0 + (size_t)i) * (size_t)w;
      size_t f[4], b[4];
      for (int mu = 0; mu < ND; ++mu) {
        f[mu] = (size_t)e[1 + mu];
        b[mu] = (size_t)e[1 + ND + mu];
      }
      site<Real, ND, IS, NT>(psi, u, res, (size_t)e[0], f, b);
    }
  }

  template <int ND, int IS>
  void dispatch(const Real* psi, const Real* u, Real* res, int cb) const {
    // non-temporal stores need 16-byte (float) / 32-byte (double) aligned sites
    const bool nt = canonical && ((uintptr_t)res % (sizeof(Real) == 4 ? 16 : 32)) == 0;
    if (!canonical)
      run_table<ND, IS, false>(psi, u, res, cb);
    else if (nt)
      run_canonical<ND, IS, true>(psi, u, res, cb);
    else
      run_canonical<ND, IS, false>(psi, u, res, cb);
    _mm_sfence();
  }

  int apply(const Real* u, const Real* psi, Real* res, int isign, int cb) {
    if (refcount == 0) return fail(DSLASH_ERR_NOT_INIT, "called before init");
    if (cb != 0 && cb != 1) return fail(DSLASH_ERR_INVALID, "cb must be 0 or 1");
    if (!u || !psi || !res) return fail(DSLASH_ERR_INVALID, "NULL array");
    if (nd == 4) {
      if (isign > 0) dispatch<4, 1>(psi, u, res, cb);
      else dispatch<4, -1>(psi, u, res, cb);
    } else {
      if (isign > 0) dispatch<3, 1>(psi, u, res, cb);
      else dispatch<3, -1>(psi, u, res, cb);
    }
    last_error = 0;
    return 0;
  }

  int wilson(Real* u, Real* psi, Real* res, int isign, int cb) {
    if (isign != 1 && isign != -1) return 0;  // lib/sse_su3dslash_32bit_scalar.c:162-179 does nothing
    return apply(u, psi, res, isign, cb);
  }
};

CpuDslash g4(4, "sse_su3dslash (cpu)");
CpuDslash g3(3, "sse_su3dslash_3d (cpu)");

int unsupported(CpuDslash& g, const char* what) { return g.fail(DSLASH_ERR_UNSUPPORTED, what); }

}  // namespace

extern "C" {

void init_sse_su3dslash(const int* latt_size, void (*getSiteCoords)(int coord[], int node, int linearsite),
                        int (*getLinearSiteIndex)(const int coord[]), int (*nodeNumber)(const int coord[])) {
  (void)nodeNumber;
  g4.init(latt_size, getSiteCoords, getLinearSiteIndex);
}
void free_sse_su3dslash(void) { g4.release(); }
void sse_su3dslash_wilson(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  (void)g4.wilson(u, psi, res, isign, cb);
}
void sse_su3dslash_prepost_receives(void) {}
int sse_su3dslash_last_error(void) { return g4.last_error; }
// Extension of the CPU build only: the OpenMP threads a call will use.
int sse_su3dslash_cpu_threads(void) { return omp_get_max_threads(); }

void init_sse_su3dslash_3d(const int* latt_size, void (*getSiteCoords)(int coord[], int node, int linearsite),
                           int (*getLinearSiteIndex)(const int coord[]), int (*nodeNumber)(const int coord[])) {
  (void)nodeNumber;
  g3.init(latt_size, getSiteCoords, getLinearSiteIndex);
}
void free_sse_su3dslash_3d(void) { g3.release(); }
void sse_su3dslash_wilson_3d(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  (void)g3.wilson(u, psi, res, isign, cb);
}
int sse_su3dslash_3d_last_error(void) { return g3.last_error; }

// The GPU libraries' extensions, so that a program links against either build.
// "Resident" here keeps the caller's host pointer (the links are read in place
// on every call); the device-resident and sloppy entry points have no CPU form.
int sse_su3dslash_gauge_upload(SSEREAL* u, int gauge_reals) {
  if (g4.refcount == 0) return g4.fail(DSLASH_ERR_NOT_INIT, "called before init");
  if (gauge_reals != 18 && gauge_reals != 12) return g4.fail(DSLASH_ERR_INVALID, "gauge_reals must be 18 or 12");
  g4.resident_u = u;
  g4.last_error = 0;
  return 0;
}
int sse_su3dslash_gauge_release(void) {
  g4.resident_u = nullptr;
  return 0;
}
int sse_su3dslash_wilson_resident(SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  if (!g4.resident_u) return g4.fail(DSLASH_ERR_NOT_INIT, "no resident gauge: call sse_su3dslash_gauge_upload first");
  if (isign != 1 && isign != -1) return g4.fail(DSLASH_ERR_INVALID, "isign must be +1 or -1");
  return g4.apply(g4.resident_u, psi, res, isign, cb);
}
int sse_su3dslash_wilson_device(void*, const void*, int, int, void*) {
  return unsupported(g4, "device-resident fields need the GPU library");
}
int sse_su3dslash_set_sloppy(int on) {
  if (!on) return g4.last_error = 0;
  return unsupported(g4, "sloppy mode is a GPU-library option");
}
int sse_su3dslash_set_gauge_reuse(int mode) {
  if (mode < 0 || mode > 2) return g4.fail(DSLASH_ERR_INVALID, "gauge reuse mode must be 0, 1 or 2");
  return g4.last_error = 0;  // the links are read in place on every call
}

int sse_su3dslash_3d_gauge_upload(SSEREAL* u, int gauge_reals) {
  if (g3.refcount == 0) return g3.fail(DSLASH_ERR_NOT_INIT, "called before init");
  if (gauge_reals != 18 && gauge_reals != 12) return g3.fail(DSLASH_ERR_INVALID, "gauge_reals must be 18 or 12");
  g3.resident_u = u;
  g3.last_error = 0;
  return 0;
}
int sse_su3dslash_3d_gauge_release(void) {
  g3.resident_u = nullptr;
  return 0;
}
int sse_su3dslash_wilson_3d_resident(SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  if (!g3.resident_u) return g3.fail(DSLASH_ERR_NOT_INIT, "no resident gauge: call sse_su3dslash_3d_gauge_upload first");
  if (isign != 1 && isign != -1) return g3.fail(DSLASH_ERR_INVALID, "isign must be +1 or -1");
  return g3.apply(g3.resident_u, psi, res, isign, cb);
}
int sse_su3dslash_wilson_3d_device(void*, const void*, int, int, void*) {
  return unsupported(g3, "device-resident fields need the GPU library");
}
int sse_su3dslash_3d_set_sloppy(int on) {
  if (!on) return g3.last_error = 0;
  return unsupported(g3, "sloppy mode is a GPU-library option");
}
int sse_su3dslash_3d_set_gauge_reuse(int mode) {
  if (mode < 0 || mode > 2) return g3.fail(DSLASH_ERR_INVALID, "gauge reuse mode must be 0, 1 or 2");
  return g3.last_error = 0;
}

#include "gauge_packers.h"

}  // extern "C"
