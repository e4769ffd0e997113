// Shared device-side building blocks of the MI355X Wilson Dslash:
// storage formats, complex arithmetic, lattice geometry (checkerboarded site
// order and periodic neighbours), the DeGrand-Rossi spin projectors and the
// per-site link bundle layout.  Everything here is header-only so the same
// definitions serve the hot kernel, the pack/unpack kernels, the solvers and
// the force kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace wd {

// ---------------------------------------------------------------------------
// Storage formats.  A field is stored per checkerboard as [nvec][Vh] 16-byte
// vectors ("planes"), so lane i of a wavefront reads site i of a plane and a
// wavefront load is one contiguous 1 KiB segment.
// ---------------------------------------------------------------------------
// F64S ("sloppy FP64") keeps FP64 spinors but FP32 links and FP32 arithmetic:
// the on-device analogue of the reference's SSEDSLASH_SLOPPY build, whose FP64
// Dslash keeps its half-spinor temporaries in float (include/types64.h:26-31).
enum Prec : int { F32 = 0, F64 = 1, F16 = 2, BF16 = 3, F64S = 4 };
constexpr int NUM_PREC = 5;

// Epilogues of the single-field operator: EPI_PLAIN res = D psi; EPI_XPAY
// res = x + a D psi; EPI_XPAY_NORM also writes one (|res|^2, 0) partial sum per
// block (one-site kernel) or per tile (box tiles), so the solvers get
// <p, Mhat^dag Mhat p> = |Mhat p|^2 without another pass.
enum { EPI_PLAIN = 0, EPI_XPAY = 1, EPI_XPAY_NORM = 2 };

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

template <int P> struct Fmt;
template <> struct Fmt<F32> {
  typedef float real;     // stored scalar
  typedef float compute;  // arithmetic type
  static constexpr int bytes = 4;
  static constexpr bool half = false;
};
template <> struct Fmt<F64> {
  typedef double real;
  typedef double compute;
  static constexpr int bytes = 8;
  static constexpr bool half = false;
};
template <> struct Fmt<F16> {
  typedef uint16_t real;
  typedef float compute;
  static constexpr int bytes = 2;
  static constexpr bool half = true;
};
template <> struct Fmt<BF16> {
  typedef uint16_t real;
  typedef float compute;
  static constexpr int bytes = 2;
  static constexpr bool half = true;
};
template <> struct Fmt<F64S> {
  typedef double real;
  typedef float compute;
  static constexpr int bytes = 8;
  static constexpr bool half = false;
};

// Bytes of one stored spinor real.
__host__ __device__ constexpr int prec_bytes(int p) {
  return (p == F64 || p == F64S) ? 8 : (p == F32 ? 4 : 2);
}
// Format of the link bundle that goes with a spinor format, and of spinor
// storage (F64S spinors are stored exactly as F64 ones).
__host__ __device__ constexpr int link_prec(int p) { return p == F64S ? F32 : p; }
__host__ __device__ constexpr int spinor_store_prec(int p) { return p == F64S ? F64 : p; }
template <int P> struct LinkP {
  static constexpr int value = link_prec(P);
};
__host__ __device__ constexpr bool prec_is_half(int p) { return p == F16 || p == BF16; }

// 16-bit float conversions.  FP16 uses the compiler's _Float16 (IEEE binary16,
// round to nearest even); BF16 keeps the top 16 bits of the binary32 pattern
// with round to nearest even (NaN kept quiet).
__host__ __device__ inline float h16_to_f(uint16_t b) {
  return (float)__builtin_bit_cast(_Float16, b);
}
__host__ __device__ inline uint16_t f_to_h16(float f) {
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}
__host__ __device__ inline float bf16_to_f(uint16_t b) {
  return __builtin_bit_cast(float, ((uint32_t)b) << 16);
}
__host__ __device__ inline uint16_t f_to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <int P> __host__ __device__ inline typename Fmt<P>::compute to_compute(typename Fmt<P>::real v);
template <> __host__ __device__ inline float to_compute<F32>(float v) { return v; }
template <> __host__ __device__ inline double to_compute<F64>(double v) { return v; }
template <> __host__ __device__ inline float to_compute<F16>(uint16_t v) { return h16_to_f(v); }
template <> __host__ __device__ inline float to_compute<BF16>(uint16_t v) { return bf16_to_f(v); }
template <> __host__ __device__ inline float to_compute<F64S>(double v) { return (float)v; }

template <int P> __host__ __device__ inline typename Fmt<P>::real from_compute(typename Fmt<P>::compute v);
template <> __host__ __device__ inline float from_compute<F32>(float v) { return v; }
template <> __host__ __device__ inline double from_compute<F64>(double v) { return v; }
template <> __host__ __device__ inline uint16_t from_compute<F16>(float v) { return f_to_h16(v); }
template <> __host__ __device__ inline uint16_t from_compute<BF16>(float v) { return f_to_bf16(v); }
template <> __host__ __device__ inline double from_compute<F64S>(float v) { return (double)v; }

// ---------------------------------------------------------------------------
// Complex arithmetic on plain structs (no std::complex in device code).
// ---------------------------------------------------------------------------
template <typename T> struct C {
  T re, im;
};
template <typename T> __host__ __device__ inline C<T> cadd(C<T> a, C<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __host__ __device__ inline C<T> csub(C<T> a, C<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __host__ __device__ inline C<T> cmul(C<T> a, C<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ inline C<T> cconj(C<T> a) { return {a.re, -a.im}; }
// i*a and -i*a
template <typename T> __host__ __device__ inline C<T> ci(C<T> a) { return {-a.im, a.re}; }
template <typename T> __host__ __device__ inline C<T> cmi(C<T> a) { return {a.im, -a.re}; }
template <typename T> __host__ __device__ inline C<T> cscale(C<T> a, T s) { return {a.re * s, a.im * s}; }
// acc += a*b
template <typename T> __host__ __device__ inline void cmac(C<T>& acc, C<T> a, C<T> b) {
  acc.re += a.re * b.re - a.im * b.im;
  acc.im += a.re * b.im + a.im * b.re;
}

// ---------------------------------------------------------------------------
// Division by an invariant 31-bit divisor (Granlund & Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, round-up variant):
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1, q = (mulhi(n,m) + n) >> s,
// exact for 0 <= n, d < 2^31.
// ---------------------------------------------------------------------------
struct FastDiv {
  uint32_t d, m, s;
};
inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  f.s = s;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
  return f;
}
__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
  uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
  return (hi + n) >> f.s;
}

// ---------------------------------------------------------------------------
// Lattice geometry.  Sites of one checkerboard are numbered exactly as the
// reference's canonical orders:
//  * 4D "cb2": idx = xh + Lxh*(y + Ly*(z + Lz*t)), cb = (x+y+z+t)&1, x = 2*xh + b
//    (lib/shift_tables_scalar.c:99-123)
//  * 3D "cb3d": idx = t + Lt*(xh + Lxh*(y + Ly*z)), cb3 = (x+y+z)&1, T fastest
//    (lib/shift_tables_scalar_3d.c:97-137)
// The full-lattice index of a site is cb*Vh + idx.
// ---------------------------------------------------------------------------
struct Geom {
  int L[4];      // global extents x,y,z,t
  int lxh;       // L[0]/2
  uint32_t vh;   // sites per checkerboard
  FastDiv d0, d1, d2;  // the three fastest extents of the site order
  int nd;        // 4 or 3
};

inline Geom make_geom(const int latt[4], int nd) {
  Geom g;
  for (int i = 0; i < 4; ++i) g.L[i] = latt[i];
  g.lxh = latt[0] / 2;
  g.vh = (uint32_t)((uint64_t)latt[0] * latt[1] * latt[2] * latt[3] / 2);
  g.nd = nd;
  if (nd == 4) {
    g.d0 = make_fastdiv(g.lxh);
    g.d1 = make_fastdiv(latt[1]);
    g.d2 = make_fastdiv(latt[2]);
  } else {
    g.d0 = make_fastdiv(latt[3]);
    g.d1 = make_fastdiv(g.lxh);
    g.d2 = make_fastdiv(latt[1]);
  }
  return g;
}

struct Coord {
  int x, y, z, t;
};

// Coordinates of site idx on checkerboard cb.
template <int ND> __host__ __device__ inline Coord site_coords(const Geom& g, uint32_t idx, int cb) {
  Coord c;
  if (ND == 4) {
    uint32_t q0 = fdiv(idx, g.d0);
    int xh = (int)(idx - q0 * (uint32_t)g.lxh);
    uint32_t q1 = fdiv(q0, g.d1);
    c.y = (int)(q0 - q1 * (uint32_t)g.L[1]);
    uint32_t q2 = fdiv(q1, g.d2);
    c.z = (int)(q1 - q2 * (uint32_t)g.L[2]);
    c.t = (int)q2;
    c.x = 2 * xh + ((cb + c.y + c.z + c.t) & 1);
  } else {
    uint32_t q0 = fdiv(idx, g.d0);
    c.t = (int)(idx - q0 * (uint32_t)g.L[3]);
    uint32_t q1 = fdiv(q0, g.d1);
    int xh = (int)(q0 - q1 * (uint32_t)g.lxh);
    uint32_t q2 = fdiv(q1, g.d2);
    c.y = (int)(q1 - q2 * (uint32_t)g.L[1]);
    c.z = (int)q2;
    c.x = 2 * xh + ((cb + c.y + c.z) & 1);
  }
  return c;
}

// Checkerboard index of a site from its coordinates (coordinates in range).
template <int ND> __host__ __device__ inline uint32_t site_index(const Geom& g, Coord c) {
  if (ND == 4)
    return (uint32_t)(c.x >> 1) + (uint32_t)g.lxh * ((uint32_t)c.y + (uint32_t)g.L[1] * ((uint32_t)c.z + (uint32_t)g.L[2] * (uint32_t)c.t));
  else
    return (uint32_t)c.t + (uint32_t)g.L[3] * ((uint32_t)(c.x >> 1) + (uint32_t)g.lxh * ((uint32_t)c.y + (uint32_t)g.L[1] * (uint32_t)c.z));
}

template <int ND> __host__ __device__ inline int site_parity(Coord c) {
  return ND == 4 ? ((c.x + c.y + c.z + c.t) & 1) : ((c.x + c.y + c.z) & 1);
}

// Coordinates shifted by +-1 in direction mu with periodic wrap.
__host__ __device__ inline Coord shift_coord(const Geom& g, Coord c, int mu, int dir) {
  int* p = mu == 0 ? &c.x : mu == 1 ? &c.y : mu == 2 ? &c.z : &c.t;
  int L = g.L[mu];
  int v = *p + dir;
  *p = v < 0 ? v + L : (v >= L ? v - L : v);
  return c;
}

// Neighbour indices of one target site, all on the opposite checkerboard.
// nb[2*mu] = x + mu^, nb[2*mu+1] = x - mu^.
template <int ND> __host__ __device__ inline void neighbours(const Geom& g, Coord c, uint32_t nb[2 * ND]) {
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) {
    nb[2 * mu] = site_index<ND>(g, shift_coord(g, c, mu, +1));
    nb[2 * mu + 1] = site_index<ND>(g, shift_coord(g, c, mu, -1));
  }
}

// ---------------------------------------------------------------------------
// Spin structure (DeGrand-Rossi basis, gamma_5 = diag(1,1,-1,-1)).  For the
// projector (1 + s*gamma_mu), s = +-1, the two-spinor h = (h0, h1) is built
// from the upper components and the lower two rows of (1+s*gamma_mu)psi are
// functions of h, so the full spinor is rebuilt from r = W h.  This is the table
// of lib/site_dslash_64bit_scalar.c:34-166 (1-gamma_0), :638-926 (1-gamma_1),
// :1228-1549 (1-gamma_2), :2171-2474 (1+gamma_3), and of
// tests/dslashm_w.cc:37-121 (spinProjectDir*/spinReconstructDir*).
// ---------------------------------------------------------------------------
template <typename T> struct Spinor {
  C<T> s[4][3];
};
template <typename T> struct Half {
  C<T> h[2][3];
};

template <int MU, int S, typename T> __host__ __device__ inline Half<T> spin_project(const Spinor<T>& p) {
  Half<T> h;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (MU == 0) {  // h0 = p0 + s i p3, h1 = p1 + s i p2
      h.h[0][c] = S > 0 ? cadd(p.s[0][c], ci(p.s[3][c])) : csub(p.s[0][c], ci(p.s[3][c]));
      h.h[1][c] = S > 0 ? cadd(p.s[1][c], ci(p.s[2][c])) : csub(p.s[1][c], ci(p.s[2][c]));
    } else if (MU == 1) {  // h0 = p0 - s p3, h1 = p1 + s p2
      h.h[0][c] = S > 0 ? csub(p.s[0][c], p.s[3][c]) : cadd(p.s[0][c], p.s[3][c]);
      h.h[1][c] = S > 0 ? cadd(p.s[1][c], p.s[2][c]) : csub(p.s[1][c], p.s[2][c]);
    } else if (MU == 2) {  // h0 = p0 + s i p2, h1 = p1 - s i p3
      h.h[0][c] = S > 0 ? cadd(p.s[0][c], ci(p.s[2][c])) : csub(p.s[0][c], ci(p.s[2][c]));
      h.h[1][c] = S > 0 ? csub(p.s[1][c], ci(p.s[3][c])) : cadd(p.s[1][c], ci(p.s[3][c]));
    } else {  // h0 = p0 + s p2, h1 = p1 + s p3
      h.h[0][c] = S > 0 ? cadd(p.s[0][c], p.s[2][c]) : csub(p.s[0][c], p.s[2][c]);
      h.h[1][c] = S > 0 ? cadd(p.s[1][c], p.s[3][c]) : csub(p.s[1][c], p.s[3][c]);
    }
  }
  return h;
}

// acc += reconstruct_{mu,s}(r): rows 0,1 get r0,r1; rows 2,3 per projector.
template <int MU, int S, typename T> __host__ __device__ inline void spin_recon_add(Spinor<T>& a, const Half<T>& r) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    a.s[0][c] = cadd(a.s[0][c], r.h[0][c]);
    a.s[1][c] = cadd(a.s[1][c], r.h[1][c]);
    if (MU == 0) {  // row2 = -s i r1, row3 = -s i r0
      a.s[2][c] = S > 0 ? cadd(a.s[2][c], cmi(r.h[1][c])) : cadd(a.s[2][c], ci(r.h[1][c]));
      a.s[3][c] = S > 0 ? cadd(a.s[3][c], cmi(r.h[0][c])) : cadd(a.s[3][c], ci(r.h[0][c]));
    } else if (MU == 1) {  // row2 = s r1, row3 = -s r0
      a.s[2][c] = S > 0 ? cadd(a.s[2][c], r.h[1][c]) : csub(a.s[2][c], r.h[1][c]);
      a.s[3][c] = S > 0 ? csub(a.s[3][c], r.h[0][c]) : cadd(a.s[3][c], r.h[0][c]);
    } else if (MU == 2) {  // row2 = -s i r0, row3 = s i r1
      a.s[2][c] = S > 0 ? cadd(a.s[2][c], cmi(r.h[0][c])) : cadd(a.s[2][c], ci(r.h[0][c]));
      a.s[3][c] = S > 0 ? cadd(a.s[3][c], ci(r.h[1][c])) : cadd(a.s[3][c], cmi(r.h[1][c]));
    } else {  // row2 = s r0, row3 = s r1
      a.s[2][c] = S > 0 ? cadd(a.s[2][c], r.h[0][c]) : csub(a.s[2][c], r.h[0][c]);
      a.s[3][c] = S > 0 ? cadd(a.s[3][c], r.h[1][c]) : csub(a.s[3][c], r.h[1][c]);
    }
  }
}

// r = W h for both spin components (W row-major: W[i][j]).
template <typename T> __host__ __device__ inline Half<T> su3_mul(const C<T> W[3][3], const Half<T>& h) {
  Half<T> r;
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      C<T> a = cmul(W[i][0], h.h[sp][0]);
      cmac(a, W[i][1], h.h[sp][1]);
      cmac(a, W[i][2], h.h[sp][2]);
      r.h[sp][i] = a;
    }
  return r;
}

// Third row of an SU(3) matrix from the first two: row2 = conj(row0 x row1).
template <typename T> __host__ __device__ inline void su3_reconstruct_row2(C<T> W[3][3]) {
  W[2][0] = cconj(csub(cmul(W[0][1], W[1][2]), cmul(W[0][2], W[1][1])));
  W[2][1] = cconj(csub(cmul(W[0][2], W[1][0]), cmul(W[0][0], W[1][2])));
  W[2][2] = cconj(csub(cmul(W[0][0], W[1][1]), cmul(W[0][1], W[1][0])));
}

// ---------------------------------------------------------------------------
// Layout of the per-target-site link bundle.  For target site x and direction
// mu the bundle holds the forward link W_f = U_mu(x) and the backward link
// already daggered, W_b = U_mu(x - mu^)^dagger, so every one of the 2*ND terms
// is r = W * h.  The two links of one direction form a "pair block" of
// PairVec<P,G> 16-byte vectors: link f occupies reals [0,G), link b [G,2G); a
// link keeps G = 18 reals (full 3x3) or G = 12 (rows 0 and 1, row 2 rebuilt).
// ---------------------------------------------------------------------------
template <int P, int G> struct PairVec {
  static constexpr int value = (2 * G * Fmt<LinkP<P>::value>::bytes + 15) / 16;
};
__host__ __device__ constexpr int pair_vec(int prec, int G) { return (2 * G * prec_bytes(link_prec(prec)) + 15) / 16; }
__host__ __device__ constexpr int spinor_vec(int prec) { return 24 * prec_bytes(prec) / 16; }

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of one checkerboard of a spinor field: planes, then (16-bit formats)
// one float norm per site.
inline size_t spinor_cb_bytes(uint32_t vh, int prec) {
  size_t b = (size_t)spinor_vec(prec) * 16 * vh;
  if (prec_is_half(prec)) b += (size_t)4 * vh;
  return align256(b);
}
// Bytes of the link bundles of one target checkerboard.
inline size_t gauge_cb_bytes(uint32_t vh, int nd, int prec, int G) {
  return align256((size_t)nd * pair_vec(prec, G) * 16 * vh);
}

// ---------------------------------------------------------------------------
// Plane access helpers.
// ---------------------------------------------------------------------------
template <typename T> __device__ inline T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T> __device__ inline void st_nt(T v, T* p) { __builtin_nontemporal_store(v, p); }

// Load one site's spinor (24 reals) from planes with stride vh vectors.
template <int P> __device__ inline Spinor<typename Fmt<P>::compute> load_spinor(const v4u* __restrict__ base, const float* __restrict__ norm, uint32_t idx, uint32_t vh) {
  typedef typename Fmt<P>::compute T;
  typedef typename Fmt<P>::real R;
  constexpr int NV = 24 * Fmt<P>::bytes / 16;
  v4u buf[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) buf[v] = base[(size_t)v * vh + idx];
  const R* r = reinterpret_cast<const R*>(buf);
  Spinor<T> s;
  T scale = (T)1;
  if constexpr (Fmt<P>::half) scale = norm[idx];
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s.s[sp][c].re = to_compute<P>(r[(sp * 3 + c) * 2 + 0]) * scale;
      s.s[sp][c].im = to_compute<P>(r[(sp * 3 + c) * 2 + 1]) * scale;
    }
  return s;
}

// Store one site's spinor; 16-bit formats store x/max|x| and the max as norm.
template <int P, bool NT> __device__ inline void store_spinor(v4u* __restrict__ base, float* __restrict__ norm, uint32_t idx, uint32_t vh, const Spinor<typename Fmt<P>::compute>& s) {
  typedef typename Fmt<P>::compute T;
  typedef typename Fmt<P>::real R;
  constexpr int NV = 24 * Fmt<P>::bytes / 16;
  v4u buf[NV];
  R* r = reinterpret_cast<R*>(buf);
  T inv = (T)1;
  if constexpr (Fmt<P>::half) {
    T m = (T)0;
#pragma unroll
    for (int sp = 0; sp < 4; ++sp)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        m = fmaxf(m, fabsf(s.s[sp][c].re));
        m = fmaxf(m, fabsf(s.s[sp][c].im));
      }
    norm[idx] = (float)m;
    inv = m > (T)0 ? (T)1 / m : (T)0;
  }
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[(sp * 3 + c) * 2 + 0] = from_compute<P>(s.s[sp][c].re * inv);
      r[(sp * 3 + c) * 2 + 1] = from_compute<P>(s.s[sp][c].im * inv);
    }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (NT)
      st_nt(buf[v], &base[(size_t)v * vh + idx]);
    else
      base[(size_t)v * vh + idx] = buf[v];
  }
}

// Load the two links of direction mu for one target site (links are stored in
// the format link_prec(P) and converted to P's arithmetic type).  NT: the
// links are read once per launch (non-temporal loads, the single-RHS kernel);
// kernels that reuse a link from L2 (batched right-hand sides) load normally.
template <int P, int G, bool NT = true> __device__ inline void load_pair(const v4u* __restrict__ pair_base, uint32_t idx, uint32_t vh, C<typename Fmt<P>::compute> Wf[3][3], C<typename Fmt<P>::compute> Wb[3][3]) {
  constexpr int PL = LinkP<P>::value;
  typedef typename Fmt<PL>::real R;
  constexpr int NV = PairVec<P, G>::value;
  v4u buf[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) buf[v] = NT ? ld_nt(&pair_base[(size_t)v * vh + idx]) : pair_base[(size_t)v * vh + idx];
  const R* r = reinterpret_cast<const R*>(buf);
  constexpr int ROWS = G / 6;
#pragma unroll
  for (int i = 0; i < ROWS; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Wf[i][j].re = to_compute<PL>(r[2 * (3 * i + j)]);
      Wf[i][j].im = to_compute<PL>(r[2 * (3 * i + j) + 1]);
      Wb[i][j].re = to_compute<PL>(r[G + 2 * (3 * i + j)]);
      Wb[i][j].im = to_compute<PL>(r[G + 2 * (3 * i + j) + 1]);
    }
  if (G == 12) {
    su3_reconstruct_row2(Wf);
    su3_reconstruct_row2(Wb);
  }
}

// XCD-aware block order: hardware hands consecutive workgroup ids to the 8
// XCDs round-robin; remap so each XCD walks one contiguous range of the site
// order and neighbouring blocks share that XCD's L2.
__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t nb) {
  const uint32_t NX = 8;
  uint32_t per = nb / NX, rem = nb % NX;
  uint32_t x = bid % NX, k = bid / NX;
  return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

}  // namespace wd
