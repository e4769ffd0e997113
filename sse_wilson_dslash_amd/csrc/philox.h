// Counter-based random fields for synthetic inputs.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011): a 4x32-bit counter is mixed with a 2x32-bit key by ten
// rounds of two 32x32->64 multiplies; the key is bumped by the Weyl constants
// between rounds.  Being counter-based, the value for (seed, field, site, draw)
// is the same on host and device and independent of the thread schedule, so a
// field can be generated in HBM or on the host and the two agree.
//
// Gaussian numbers use the Box-Muller transform in double precision.  The
// SU(3) link generator follows the Gram-Schmidt reunitarisation of
// tests/reunit.cc:184-241: normalise column 0, orthogonalise column 1 against
// it and normalise, column 2 = conj(col0 x col1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

namespace wd {

struct U4 {
  uint32_t v[4];
};

__host__ __device__ inline void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__host__ __device__ inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, ctr.v[0], hi0, lo0);
    philox_mulhilo(M1, ctr.v[2], hi1, lo1);
    U4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// Four N(0,1) numbers for (seed, stream, site, draw).
__host__ __device__ inline void gauss4(uint64_t seed, uint32_t stream, uint32_t site, uint32_t draw, double out[4]) {
  U4 c;
  c.v[0] = site;
  c.v[1] = draw;
  c.v[2] = stream;
  c.v[3] = 0x5157u;
  U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double two_pi = 6.283185307179586476925286766559;
  const double inv32 = 2.3283064365386962890625e-10;  // 2^-32
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    double u1 = ((double)r.v[2 * p] + 0.5) * inv32;  // (0,1)
    double u2 = ((double)r.v[2 * p + 1] + 0.5) * inv32;
    double rad = sqrt(-2.0 * log(u1));
    out[2 * p] = rad * cos(two_pi * u2);
    out[2 * p + 1] = rad * sin(two_pi * u2);
  }
}

// Streams: 1 + mu for gauge direction mu, 16 + k for spinor field k.
enum : uint32_t { RNG_GAUGE = 1, RNG_SPINOR = 16 };

// An SU(3) matrix for lexicographic site `lex`, direction mu.  Output U[i][j]
// as (re, im) pairs, row-major.
__host__ __device__ inline void random_su3(uint64_t seed, uint32_t lex, int mu, double U[3][3][2]) {
  double g[20];
#pragma unroll
  for (int d = 0; d < 5; ++d) gauss4(seed, RNG_GAUGE + (uint32_t)mu, lex, (uint32_t)d, &g[4 * d]);
  double a[3][3][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j][0] = g[2 * (3 * i + j)];
      a[i][j][1] = g[2 * (3 * i + j) + 1];
    }
  // normalise column 0
  double t1 = 0;
  for (int c = 0; c < 3; ++c) t1 += a[c][0][0] * a[c][0][0] + a[c][0][1] * a[c][0][1];
  double t4 = 1.0 / sqrt(t1);
  for (int c = 0; c < 3; ++c) {
    a[c][0][0] *= t4;
    a[c][0][1] *= t4;
  }
  // t2 = col0^dagger col1 ; col1 -= t2 col0
  double t2r = 0, t2i = 0;
  for (int c = 0; c < 3; ++c) {
    t2r += a[c][0][0] * a[c][1][0] + a[c][0][1] * a[c][1][1];
    t2i += a[c][0][0] * a[c][1][1] - a[c][0][1] * a[c][1][0];
  }
  for (int c = 0; c < 3; ++c) {
    double pr = t2r * a[c][0][0] - t2i * a[c][0][1];
    double pi = t2r * a[c][0][1] + t2i * a[c][0][0];
    a[c][1][0] -= pr;
    a[c][1][1] -= pi;
  }
  double t3 = 0;
  for (int c = 0; c < 3; ++c) t3 += a[c][1][0] * a[c][1][0] + a[c][1][1] * a[c][1][1];
  t4 = 1.0 / sqrt(t3);
  for (int c = 0; c < 3; ++c) {
    a[c][1][0] *= t4;
    a[c][1][1] *= t4;
  }
  // column 2 = conj(col0 x col1)
  const int i1[3] = {1, 2, 0}, i2[3] = {2, 0, 1};
  for (int k = 0; k < 3; ++k) {
    int p = i1[k], q = i2[k];
    // u_p v_q - u_q v_p
    double re = (a[p][0][0] * a[q][1][0] - a[p][0][1] * a[q][1][1]) - (a[q][0][0] * a[p][1][0] - a[q][0][1] * a[p][1][1]);
    double im = (a[p][0][0] * a[q][1][1] + a[p][0][1] * a[q][1][0]) - (a[q][0][0] * a[p][1][1] + a[q][0][1] * a[p][1][0]);
    a[k][2][0] = re;
    a[k][2][1] = -im;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      U[i][j][0] = a[i][j][0];
      U[i][j][1] = a[i][j][1];
    }
}

}  // namespace wd
