// The fused Wilson Dslash kernel for gfx950 and the layout kernels around it.
//
// One thread owns one target site.  For each direction it streams the site's
// link pair (U_mu(x), U_mu(x-mu)^dagger) from its own bundle planes with
// non-temporal loads (every link is read exactly once per call), gathers the
// two neighbour spinors from the source checkerboard planes (shifted but
// contiguous across a wavefront, so they coalesce and the 8-fold reuse of each
// source spinor is served from L2/MALL), spin-projects, multiplies by the link
// and accumulates the reconstructed spinor in registers.  The result is
// written once with non-temporal stores.  No MFMA: 3x3 complex mat-vec work at
// 1.7 flop/byte is bound by HBM bandwidth.
//
// The reference's per-direction SSE routines
// (lib/site_dslash_{32,64}bit_scalar.c, 16 functions per precision), its shift
// tables (lib/shift_tables_scalar.c:191-305) and its thread dispatch
// (lib/dispatch_scalar.c:37-67) collapse into k_dslash below.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <thread>
#include <vector>

#include "dslash_common.h"
#include "dslash_mi355x.h"
#include "philox.h"
#include "status.h"
#include "dslash_internal.h"
#include "mrhs.h"

namespace wd {

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
static thread_local int t_err = 0;
static thread_local char t_msg[512] = "";

int set_error(int code, const char* fmt, ...) {
  t_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
  va_end(ap);
  return code;
}
void clear_error() {
  t_err = 0;
  t_msg[0] = 0;
}

// ---------------------------------------------------------------------------
// launch configuration
// ---------------------------------------------------------------------------
// Threads per block, XCD remapping and marching length per (4D/3D, format,
// link reals); set by dslash_set_launch / dslash_set_launch_march or by the
// autotuner.  Defaults are the shapes measured fastest at 32^3 x 64 on MI355X
// (XCD-contiguous order is worth ~25%; marching 6-16% for FP64, sloppy FP64 and
// the 3D operator, nothing for 16-bit spinors; for 4D FP32, LDS-staged box tiles
// (march 16) are 6-7% faster than one site per thread at 32^3 x 64 and within
// 1% at 48^3 x 96, where marching is 3-5% slower, and box tiles with the LDS halo
// (march 32) another 5% (18-real links) and 9% (12-real) faster at 32^3 x 64;
// for 4D FP64 box tiles are 3-4% and for sloppy FP64 6-10% faster than a march
// of 8, and sloppy FP64 with the halo another 4.6% (18-real) / 1.4% (12-real)).
// Index: [nd == 4][F32, F64, F16, BF16, F64S][G == 12].
static LaunchCfg g_cfg[2][NUM_PREC][2] = {
    {{{256, 1, 8}, {256, 1, 8}},
     {{64, 1, 4}, {256, 1, 8}},
     {{64, 1, 1}, {64, 1, 8}},
     {{64, 1, 1}, {64, 1, 1}},
     {{256, 1, 8}, {256, 1, 8}}},
    {{{256, 1, 32}, {256, 1, 32}},
     {{256, 1, 16}, {256, 1, 16}},
     {{256, 1, 1}, {256, 1, 1}},
     {{256, 1, 1}, {256, 1, 1}},
     {{256, 1, 32}, {256, 1, 32}}}};

static int prec_slot(int prec) { return prec >= 0 && prec < NUM_PREC ? prec : 0; }
LaunchCfg launch_cfg(int nd, int prec, int G) { return g_cfg[nd == 4][prec_slot(prec)][G == 12]; }
void set_launch_cfg(int nd, int prec, int G, LaunchCfg c) { g_cfg[nd == 4][prec_slot(prec)][G == 12] = c; }

static int check_lattice(const int latt[4], int nd) {
  WD_REQUIRE(latt != nullptr, "latt is NULL");
  WD_REQUIRE(nd == 4 || nd == 3, "nd must be 4 or 3 (got %d)", nd);
  uint64_t v = 1;
  for (int mu = 0; mu < 4; ++mu) {
    WD_REQUIRE(latt[mu] > 0, "latt[%d]=%d must be positive", mu, latt[mu]);
    if (mu < nd) WD_REQUIRE(latt[mu] % 2 == 0, "latt[%d]=%d must be even for the %dD checkerboard", mu, latt[mu], nd);
    v *= (uint64_t)latt[mu];
  }
  WD_REQUIRE(v / 2 < (1ull << 31) / 32, "lattice volume %llu too large for 32-bit plane indexing", (unsigned long long)v);
  return 0;
}

static int check_prec(int prec) {
  WD_REQUIRE(prec == F32 || prec == F64 || prec == F16 || prec == BF16 || prec == F64S, "unknown precision %d", prec);
  return 0;
}

// ---------------------------------------------------------------------------
// Dslash kernel
// ---------------------------------------------------------------------------
struct ApplyArgs {
  const v4u* psi;
  const float* psi_norm;
  v4u* res;
  float* res_norm;
  const v4u* gauge;
  const v4u* xp;  // optional target-cb field of the epilogue res = xp + a * D psi
  const float* xp_norm;
  double a;
  double* norm_partial;  // optional per-block |res|^2 (re, im=0 pairs)
  Geom g;
  int tcb;
  int remap;
  int block;
  int march;        // slices per thread of k_dslash_march (1 = k_dslash)
  uint32_t slice;   // sites per slice of the slowest coordinate (t in 4D, z in 3D)
  FastDiv dslice;
  uint32_t site0 = 0;       // k_dslash: target sites [site0, site_end) of the checkerboard
  uint32_t site_end = 0;
};

template <int ND, int P, int G, int IS, int MU>
__device__ __forceinline__ void hop_dir(const ApplyArgs& a, const Coord& c, uint32_t idx,
                                        Spinor<typename Fmt<P>::compute>& acc) {
  typedef typename Fmt<P>::compute T;
  constexpr int PV = PairVec<P, G>::value;
  const uint32_t vh = a.g.vh;
  C<T> Wf[3][3], Wb[3][3];
  load_pair<P, G>(a.gauge + (size_t)MU * PV * vh, idx, vh, Wf, Wb);
  {  // U_mu(x) (1 - isign gamma_mu) psi(x + mu)
    uint32_t n = site_index<ND>(a.g, shift_coord(a.g, c, MU, +1));
    Spinor<T> p = load_spinor<P>(a.psi, a.psi_norm, n, vh);
    spin_recon_add<MU, -IS>(acc, su3_mul(Wf, spin_project<MU, -IS>(p)));
  }
  {  // U_mu(x - mu)^dagger (1 + isign gamma_mu) psi(x - mu)
    uint32_t n = site_index<ND>(a.g, shift_coord(a.g, c, MU, -1));
    Spinor<T> p = load_spinor<P>(a.psi, a.psi_norm, n, vh);
    spin_recon_add<MU, IS>(acc, su3_mul(Wb, spin_project<MU, IS>(p)));
  }
}

// Epilogues (EPI_*): the |res|^2 partial sum of EPI_XPAY_NORM goes to
// norm_partial[blockIdx.x].

// Minimum waves per SIMD the register allocator must leave room for: the
// 16-bit formats have small loads that the scheduler would otherwise hoist
// into >200 VGPRs (2 waves/SIMD); 4 waves keeps them at <= 128 without spills.
template <int P> struct MinWaves {
  static constexpr int value = Fmt<P>::half ? 4 : 1;
};

template <int ND, int P, int G, int IS, int EPI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves<P>::value, 8)))
k_dslash(ApplyArgs a) {
  typedef typename Fmt<P>::compute T;
  uint32_t b = a.remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  uint32_t idx = a.site0 + b * blockDim.x + threadIdx.x;
  const bool live = idx < a.site_end;
  if constexpr (EPI != EPI_XPAY_NORM) {
    if (!live) return;
  }
  double s2 = 0.0;
  if (live) {
    Coord c = site_coords<ND>(a.g, idx, a.tcb);
    Spinor<T> acc;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int k = 0; k < 3; ++k) acc.s[s][k] = {(T)0, (T)0};
    hop_dir<ND, P, G, IS, 0>(a, c, idx, acc);
    if constexpr (Fmt<P>::half) __builtin_amdgcn_sched_barrier(0);
    hop_dir<ND, P, G, IS, 1>(a, c, idx, acc);
    if constexpr (Fmt<P>::half) __builtin_amdgcn_sched_barrier(0);
    hop_dir<ND, P, G, IS, 2>(a, c, idx, acc);
    if constexpr (ND == 4) {
      if constexpr (Fmt<P>::half) __builtin_amdgcn_sched_barrier(0);
      hop_dir<ND, P, G, IS, 3>(a, c, idx, acc);
    }
    if constexpr (EPI != EPI_PLAIN) {
      Spinor<T> x = load_spinor<P>(a.xp, a.xp_norm, idx, a.g.vh);
      const T s = (T)a.a;
      T n2 = (T)0;
#pragma unroll
      for (int sp = 0; sp < 4; ++sp)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          acc.s[sp][k].re = x.s[sp][k].re + s * acc.s[sp][k].re;
          acc.s[sp][k].im = x.s[sp][k].im + s * acc.s[sp][k].im;
          n2 += acc.s[sp][k].re * acc.s[sp][k].re + acc.s[sp][k].im * acc.s[sp][k].im;
        }
      s2 = (double)n2;
      store_spinor<P, true>(a.res, a.res_norm, idx, a.g.vh, acc);
    } else {
      store_spinor<P, true>(a.res, a.res_norm, idx, a.g.vh, acc);
    }
  }
  if constexpr (EPI == EPI_XPAY_NORM) {
    __shared__ double lds[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s2 += __shfl_down(s2, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += lds[w];
      a.norm_partial[2 * blockIdx.x] = t;
      a.norm_partial[2 * blockIdx.x + 1] = 0.0;
    }
  }
}

// Fewest threads a marching launch may have: 512 blocks of 256, two per CU.
constexpr uint64_t MARCH_MIN_THREADS = 131072;

// Marching form of res = D psi.  The site order runs slowest in t (4D) or z
// (3D), so the neighbours of site s + S*m in that direction are s + S*(m +- 1)
// with S = Vh / L: a thread owns one site s of a slice and walks TB consecutive
// slices m.  The source spinor it gathers for the forward hop of slice m is the
// backward neighbour of slice m + 2, so its projection (1 + isign gamma) psi
// stays in registers for two slices instead of being gathered again from S
// sites away, where at large S the second read misses L2.
template <int ND, int P, int G, int IS, int TB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Fmt<P>::half ? 2 : 1, 8)))
k_dslash_march(ApplyArgs a) {
  typedef typename Fmt<P>::compute T;
  constexpr int MD = ND - 1;  // marching direction: t (4D), z (3D)
  constexpr int PV = PairVec<P, G>::value;
  const uint32_t b = a.remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t p = b * blockDim.x + threadIdx.x;
  const uint32_t S = a.slice;
  const uint32_t chunk = fdiv(p, a.dslice);
  const uint32_t s = p - chunk * S;
  const int L = a.g.L[MD];
  const int m0 = (int)chunk * TB;
  if (m0 >= L) return;
  const Coord c0 = site_coords<ND>(a.g, s, 0);  // s < S: coordinate MD is 0
  const uint32_t vh = a.g.vh;
  Half<T> keep[2];
#pragma unroll
  for (int k = 0; k < TB; ++k) {
    const int m = m0 + k;
    if (m >= L) break;
    Coord c = c0;
    if constexpr (ND == 4)
      c.t = m;
    else
      c.z = m;
    c.x = (c0.x & ~1) | ((a.tcb + c.y + c.z + (ND == 4 ? c.t : 0)) & 1);
    const uint32_t idx = s + S * (uint32_t)m;
    Spinor<T> acc;
#pragma unroll
    for (int sp = 0; sp < 4; ++sp)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) acc.s[sp][cc] = {(T)0, (T)0};
    hop_dir<ND, P, G, IS, 0>(a, c, idx, acc);
    if constexpr (Fmt<P>::half) __builtin_amdgcn_sched_barrier(0);
    hop_dir<ND, P, G, IS, 1>(a, c, idx, acc);
    if constexpr (ND == 4) {
      if constexpr (Fmt<P>::half) __builtin_amdgcn_sched_barrier(0);
      hop_dir<ND, P, G, IS, 2>(a, c, idx, acc);
    }
    if constexpr (Fmt<P>::half) __builtin_amdgcn_sched_barrier(0);
    {
      C<T> Wf[3][3], Wb[3][3];
      load_pair<P, G>(a.gauge + (size_t)MD * PV * vh, idx, vh, Wf, Wb);
      const int mf = m + 1 == L ? 0 : m + 1;
      const Spinor<T> pf = load_spinor<P>(a.psi, a.psi_norm, s + S * (uint32_t)mf, vh);
      spin_recon_add<MD, -IS>(acc, su3_mul(Wf, spin_project<MD, -IS>(pf)));
      Half<T> hb;
      if (k < 2) {
        const int mb = m == 0 ? L - 1 : m - 1;
        hb = spin_project<MD, IS>(load_spinor<P>(a.psi, a.psi_norm, s + S * (uint32_t)mb, vh));
      } else {
        hb = keep[k & 1];
      }
      spin_recon_add<MD, IS>(acc, su3_mul(Wb, hb));
      if (k + 2 < TB) keep[k & 1] = spin_project<MD, IS>(pf);
    }
    store_spinor<P, true>(a.res, a.res_norm, idx, vh, acc);
    // keeps the next slice's loads from being hoisted into this one's registers
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int ND, int P, int G, int IS>
static void launch_dslash_t(const ApplyArgs& a, hipStream_t s) {
  uint32_t blocks = (a.site_end - a.site0 + a.block - 1) / a.block;
  if (!a.xp && a.march > 1) {
    const uint32_t chunks = (uint32_t)(a.g.L[ND - 1] + a.march - 1) / (uint32_t)a.march;
    const uint32_t mblocks = (a.slice * chunks + a.block - 1) / a.block;
    if (a.march == 4)
      hipLaunchKernelGGL((k_dslash_march<ND, P, G, IS, 4>), dim3(mblocks), dim3(a.block), 0, s, a);
    else
      hipLaunchKernelGGL((k_dslash_march<ND, P, G, IS, 8>), dim3(mblocks), dim3(a.block), 0, s, a);
    return;
  }
  if (a.xp) {
    if constexpr (ND == 4) {
      if (a.norm_partial)
        hipLaunchKernelGGL((k_dslash<ND, P, G, IS, EPI_XPAY_NORM>), dim3(blocks), dim3(a.block), 0, s, a);
      else
        hipLaunchKernelGGL((k_dslash<ND, P, G, IS, EPI_XPAY>), dim3(blocks), dim3(a.block), 0, s, a);
    }
  } else {
    hipLaunchKernelGGL((k_dslash<ND, P, G, IS, EPI_PLAIN>), dim3(blocks), dim3(a.block), 0, s, a);
  }
}

template <int ND, int P, int G>
static void launch_dslash_g(const ApplyArgs& a, int isign, hipStream_t s) {
  if (isign > 0)
    launch_dslash_t<ND, P, G, 1>(a, s);
  else
    launch_dslash_t<ND, P, G, -1>(a, s);
}

template <int ND, int P>
static void launch_dslash_p(const ApplyArgs& a, int G, int isign, hipStream_t s) {
  if (G == 18)
    launch_dslash_g<ND, P, 18>(a, isign, s);
  else
    launch_dslash_g<ND, P, 12>(a, isign, s);
}

template <int ND>
static void launch_dslash_nd(const ApplyArgs& a, int prec, int G, int isign, hipStream_t s) {
  switch (prec) {
    case F32: launch_dslash_p<ND, F32>(a, G, isign, s); break;
    case F64: launch_dslash_p<ND, F64>(a, G, isign, s); break;
    case F16: launch_dslash_p<ND, F16>(a, G, isign, s); break;
    case F64S: launch_dslash_p<ND, F64S>(a, G, isign, s); break;
    default: launch_dslash_p<ND, BF16>(a, G, isign, s); break;
  }
}

// Marching divides the thread count by the march length; on lattices with few
// sites per slice or few slices that would leave CUs idle, so a shorter march
// (down to one site per thread) is used when the launch would fall below
// MARCH_MIN_THREADS threads.
// Box tiles (16) need the 4D FP32, FP64 or sloppy-FP64 operator, a whole x row per block and
// the same thread floor; otherwise they fall back to a march of 8.  Box tiles with
// the LDS halo (32) need FP32 or sloppy FP64 (FP64 would spill: its tile already
// takes 256 VGPRs) and a halo of at most one entry per thread; otherwise they fall
// back to 16.
static int effective_march(const int latt[4], int nd, int prec, int block, int march) {
  const uint64_t L = (uint64_t)latt[nd - 1];
  const uint64_t slice = (uint64_t)latt[0] * latt[1] * latt[2] * latt[3] / 2 / L;
  int m = march_ok(march) ? march : 1;
  if (m == MARCH_BOX_HALO)
    m = (prec == F32 || prec == F64S) && box_halo_single(latt, block) > 0 ? MARCH_BOX_HALO : MARCH_BOX;
  if (m == MARCH_BOX || m == MARCH_BOX_HALO) {
    if (nd == 4 && (prec == F32 || prec == F64 || prec == F64S) && block >= 128 &&
        (uint64_t)box_tiles_single(latt, block) * (uint64_t)block >= MARCH_MIN_THREADS)
      return m;
    m = 8;
  }
  while (m > 1 && slice * ((L + m - 1) / m) < MARCH_MIN_THREADS) m = m == 8 ? 4 : 1;
  return m;
}

int apply_blocks(const int latt[4], int nd, int prec, int G) {
  const uint32_t vh = (uint32_t)((uint64_t)latt[0] * latt[1] * latt[2] * latt[3] / 2);
  const LaunchCfg lc = launch_cfg(nd, prec, G);
  const int m = effective_march(latt, nd, prec, lc.block, lc.march);
  if (m == MARCH_BOX || m == MARCH_BOX_HALO) return (int)box_tiles_single(latt, lc.block);
  return (int)((vh + lc.block - 1) / lc.block);
}

int apply_cb_ex(void* res_cb, const void* psi_cb, const void* gauge_cb, const int latt[4], int nd, int prec, int G,
                int isign, int tcb, const void* xp_cb, double xa, hipStream_t s, double* norm_partial) {
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12 (got %d)", G);
  WD_REQUIRE(isign == 1 || isign == -1, "isign must be +1 or -1 (got %d)", isign);
  WD_REQUIRE(tcb == 0 || tcb == 1, "checkerboard must be 0 or 1 (got %d)", tcb);
  WD_REQUIRE(res_cb && psi_cb && gauge_cb, "NULL field pointer");
  WD_REQUIRE(res_cb != psi_cb, "res and psi must be different checkerboard buffers");
  WD_REQUIRE(!xp_cb || nd == 4, "the fused x + a*D psi epilogue is built for the 4D operator only");
  WD_REQUIRE(xp_cb != psi_cb, "x and psi must be different checkerboard buffers");
  ApplyArgs a;
  a.g = make_geom(latt, nd);
  const size_t planes = (size_t)spinor_vec(prec) * 16 * a.g.vh;
  a.psi = (const v4u*)psi_cb;
  a.psi_norm = (const float*)((const char*)psi_cb + planes);
  a.res = (v4u*)res_cb;
  a.res_norm = (float*)((char*)res_cb + planes);
  a.gauge = (const v4u*)gauge_cb;
  a.xp = (const v4u*)xp_cb;
  a.xp_norm = xp_cb ? (const float*)((const char*)xp_cb + planes) : nullptr;
  a.a = xa;
  WD_REQUIRE(!norm_partial || xp_cb, "the |res|^2 epilogue needs the x + a*D psi form");
  a.norm_partial = norm_partial;
  a.tcb = tcb;
  const LaunchCfg lc = launch_cfg(nd, prec, G);
  a.remap = lc.remap;
  a.block = lc.block;
  a.slice = a.g.vh / (uint32_t)latt[nd - 1];
  a.dslice = make_fastdiv(a.slice);
  a.site0 = 0;
  a.site_end = a.g.vh;
  a.march = effective_march(latt, nd, prec, lc.block, lc.march);
  if (a.march == MARCH_BOX || a.march == MARCH_BOX_HALO)  // LDS-staged box tiles, every epilogue form
    return box_apply_single(res_cb, psi_cb, gauge_cb, latt, prec, G, isign, tcb, lc.block,
                            BoxEpi{a.xp, xa, norm_partial}, s, a.march == MARCH_BOX_HALO);
  if (nd == 4)
    launch_dslash_nd<4>(a, prec, G, isign, s);
  else
    launch_dslash_nd<3>(a, prec, G, isign, s);
  WD_LAUNCHED();
  return 0;
}

// res = D psi on the target sites [site0, site1) of the checkerboard only, one
// site per thread (the other target sites are not written).
static int apply_cb_sites(void* res_cb, const void* psi_cb, const void* gauge_cb, const int latt[4], int nd, int prec,
                          int G, int isign, int tcb, uint64_t site0, uint64_t site1, hipStream_t s) {
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12 (got %d)", G);
  WD_REQUIRE(isign == 1 || isign == -1, "isign must be +1 or -1 (got %d)", isign);
  WD_REQUIRE(tcb == 0 || tcb == 1, "checkerboard must be 0 or 1 (got %d)", tcb);
  WD_REQUIRE(res_cb && psi_cb && gauge_cb, "NULL field pointer");
  WD_REQUIRE(res_cb != psi_cb, "res and psi must be different checkerboard buffers");
  ApplyArgs a;
  a.g = make_geom(latt, nd);
  WD_REQUIRE(site0 <= site1 && site1 <= a.g.vh, "site range [%llu, %llu) outside the %u sites of a checkerboard",
             (unsigned long long)site0, (unsigned long long)site1, a.g.vh);
  if (site0 == site1) return 0;
  const size_t planes = (size_t)spinor_vec(prec) * 16 * a.g.vh;
  a.psi = (const v4u*)psi_cb;
  a.psi_norm = (const float*)((const char*)psi_cb + planes);
  a.res = (v4u*)res_cb;
  a.res_norm = (float*)((char*)res_cb + planes);
  a.gauge = (const v4u*)gauge_cb;
  a.xp = nullptr;
  a.xp_norm = nullptr;
  a.a = 0.0;
  a.norm_partial = nullptr;
  a.tcb = tcb;
  const LaunchCfg lc = launch_cfg(nd, prec, G);
  a.remap = lc.remap;
  a.block = lc.block;
  a.slice = a.g.vh / (uint32_t)latt[nd - 1];
  a.dslice = make_fastdiv(a.slice);
  a.site0 = (uint32_t)site0;
  a.site_end = (uint32_t)site1;
  a.march = 1;
  if (nd == 4)
    launch_dslash_nd<4>(a, prec, G, isign, s);
  else
    launch_dslash_nd<3>(a, prec, G, isign, s);
  WD_LAUNCHED();
  return 0;
}

static int apply_cb(void* res_cb, const void* psi_cb, const void* gauge_cb, const int latt[4], int nd,
                    int prec, int G, int isign, int tcb, hipStream_t s) {
  return apply_cb_ex(res_cb, psi_cb, gauge_cb, latt, nd, prec, G, isign, tcb, nullptr, 0.0, s, nullptr);
}

// ---------------------------------------------------------------------------
// layout kernels
// ---------------------------------------------------------------------------
template <int P> __device__ inline typename Fmt<P>::real cvt_out(double x) {
  return from_compute<P>((typename Fmt<P>::compute)x);
}

template <int ND, int PIN, int POUT, int G>
__global__ void k_pack_gauge(v4u* __restrict__ out, const typename Fmt<PIN>::real* __restrict__ u,
                             const int* __restrict__ table, Geom g, int tcb) {
  typedef typename Fmt<POUT>::real RO;
  constexpr int PV = PairVec<POUT, G>::value;
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.vh) return;
  Coord c = site_coords<ND>(g, idx, tcb);
  uint32_t mx = (uint32_t)tcb * g.vh + idx;
  size_t cx = table ? (size_t)table[mx] : mx;
  for (int mu = 0; mu < ND; ++mu) {
    uint32_t mb = (uint32_t)(1 - tcb) * g.vh + site_index<ND>(g, shift_coord(g, c, mu, -1));
    size_t cb = table ? (size_t)table[mb] : mb;
    const typename Fmt<PIN>::real* uf = u + (cx * 4 + mu) * 18;
    const typename Fmt<PIN>::real* ub = u + (cb * 4 + mu) * 18;
    v4u buf[PV];
#pragma unroll
    for (int v = 0; v < PV; ++v) buf[v] = v4u{0u, 0u, 0u, 0u};
    RO* r = reinterpret_cast<RO*>(buf);
#pragma unroll
    for (int i = 0; i < G / 6; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        // packed links are stored transposed: stored[j][i] = U(i,j)
        r[2 * (3 * i + j)] = cvt_out<POUT>((double)uf[(j * 3 + i) * 2]);
        r[2 * (3 * i + j) + 1] = cvt_out<POUT>((double)uf[(j * 3 + i) * 2 + 1]);
        // (U^dagger)(i,j) = conj(U(j,i)) = conj(stored[i][j])
        r[G + 2 * (3 * i + j)] = cvt_out<POUT>((double)ub[(i * 3 + j) * 2]);
        r[G + 2 * (3 * i + j) + 1] = cvt_out<POUT>(-(double)ub[(i * 3 + j) * 2 + 1]);
      }
#pragma unroll
    for (int v = 0; v < PV; ++v) out[((size_t)mu * PV + v) * g.vh + idx] = buf[v];
  }
}

template <int PIN, int POUT>
__global__ void k_pack_spinor(v4u* __restrict__ out, float* __restrict__ out_norm,
                              const typename Fmt<PIN>::real* __restrict__ aos, const int* __restrict__ table,
                              uint32_t vh, int cb, uint32_t i0, uint32_t i1) {
  typedef typename Fmt<POUT>::compute T;
  uint32_t idx = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= i1) return;
  uint32_t mx = (uint32_t)cb * vh + idx;
  size_t cx = table ? (size_t)table[mx] : mx;
  const typename Fmt<PIN>::real* p = aos + cx * 24;
  Spinor<T> s;
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s.s[sp][c].re = (T)p[(sp * 3 + c) * 2];
      s.s[sp][c].im = (T)p[(sp * 3 + c) * 2 + 1];
    }
  store_spinor<POUT, false>(out, out_norm, idx, vh, s);
}

template <int PCB, int POUT>
__global__ void k_unpack_spinor(typename Fmt<POUT>::real* __restrict__ aos, const v4u* __restrict__ in,
                                const float* __restrict__ in_norm, const int* __restrict__ table, uint32_t vh,
                                int cb, int compact, uint32_t i0, uint32_t i1) {
  uint32_t idx = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= i1) return;
  Spinor<typename Fmt<PCB>::compute> s = load_spinor<PCB>(in, in_norm, idx, vh);
  uint32_t mx = (uint32_t)cb * vh + idx;
  size_t dst = compact ? idx : (table ? (size_t)table[mx] : mx);
  typename Fmt<POUT>::real* p = aos + dst * 24;
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      p[(sp * 3 + c) * 2] = (typename Fmt<POUT>::real)s.s[sp][c].re;
      p[(sp * 3 + c) * 2 + 1] = (typename Fmt<POUT>::real)s.s[sp][c].im;
    }
}

template <int PIN, int POUT>
__global__ void k_convert_spinor(v4u* __restrict__ out, float* __restrict__ out_norm, const v4u* __restrict__ in,
                                 const float* __restrict__ in_norm, uint32_t vh) {
  typedef typename Fmt<POUT>::compute TO;
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= vh) return;
  Spinor<typename Fmt<PIN>::compute> s = load_spinor<PIN>(in, in_norm, idx, vh);
  Spinor<TO> o;
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < 3; ++c) o.s[sp][c] = {(TO)s.s[sp][c].re, (TO)s.s[sp][c].im};
  store_spinor<POUT, false>(out, out_norm, idx, vh, o);
}

// ---------------------------------------------------------------------------
// synthetic fields
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t lex_index(const Geom& g, Coord c) {
  return (uint32_t)c.x + (uint32_t)g.L[0] * ((uint32_t)c.y + (uint32_t)g.L[1] * ((uint32_t)c.z + (uint32_t)g.L[2] * (uint32_t)c.t));
}

template <int ND> __host__ __device__ inline Coord full_coords(const Geom& g, uint32_t i) {
  int cb = i >= g.vh ? 1 : 0;
  return site_coords<ND>(g, i - (uint32_t)cb * g.vh, cb);
}

template <typename R, int ND>
__host__ __device__ inline void gen_gauge_site(R* u, const Geom& g, uint64_t seed, uint32_t i) {
  uint32_t lex = lex_index(g, full_coords<ND>(g, i));
  for (int mu = 0; mu < 4; ++mu) {
    double U[3][3][2];
    random_su3(seed, lex, mu, U);
    R* d = u + ((size_t)i * 4 + mu) * 18;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        d[(b * 3 + a) * 2] = (R)U[a][b][0];  // stored transposed
        d[(b * 3 + a) * 2 + 1] = (R)U[a][b][1];
      }
  }
}

template <typename R, int ND>
__host__ __device__ inline void gen_spinor_site(R* psi, const Geom& g, uint64_t seed, int field, uint32_t i) {
  uint32_t lex = lex_index(g, full_coords<ND>(g, i));
  R* d = psi + (size_t)i * 24;
  for (int k = 0; k < 6; ++k) {
    double z[4];
    gauss4(seed, RNG_SPINOR + (uint32_t)field, lex, (uint32_t)k, z);
    for (int q = 0; q < 4; ++q) d[4 * k + q] = (R)z[q];
  }
}

template <typename R, int ND>
__global__ void k_gen_gauge(R* u, Geom g, uint64_t seed) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * g.vh) return;
  gen_gauge_site<R, ND>(u, g, seed, i);
}

template <typename R, int ND>
__global__ void k_gen_spinor(R* psi, Geom g, uint64_t seed, int field) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * g.vh) return;
  gen_spinor_site<R, ND>(psi, g, seed, field, i);
}

template <typename F>
static void host_parallel(uint32_t n, F f) {
  unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([=]() {
      uint32_t lo = (uint32_t)((uint64_t)n * t / nt), hi = (uint32_t)((uint64_t)n * (t + 1) / nt);
      for (uint32_t i = lo; i < hi; ++i) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace wd

using namespace wd;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int dslash_last_error(void) { return t_err; }
const char* dslash_last_error_message(void) { return t_msg; }
int dslash_version(void) { return 10000; }

int dslash_check_lattice(const int latt[4], int nd) {
  clear_error();
  return check_lattice(latt, nd);
}

size_t dslash_spinor_cb_bytes(const int latt[4], int nd, int prec) {
  if (check_lattice(latt, nd) || check_prec(prec)) return 0;
  return spinor_cb_bytes(make_geom(latt, nd).vh, prec);
}

size_t dslash_gauge_cb_bytes(const int latt[4], int nd, int prec, int G) {
  if (check_lattice(latt, nd) || check_prec(prec)) return 0;
  if (G != 18 && G != 12) return 0;
  return gauge_cb_bytes(make_geom(latt, nd).vh, nd, prec, G);
}

int dslash_set_launch(int nd, int prec, int gauge_reals, int block_threads, int xcd_remap) {
  clear_error();
  WD_REQUIRE(block_threads >= 64 && block_threads <= 256 && block_threads % 64 == 0,
             "block_threads must be 64, 128, 192 or 256 (got %d)", block_threads);
  if (nd == 0) {
    for (int n : {3, 4})
      for (int p = 0; p < NUM_PREC; ++p)
        for (int G : {18, 12}) {
          LaunchCfg c = launch_cfg(n, p, G);
          c.block = block_threads;
          c.remap = xcd_remap ? 1 : 0;
          set_launch_cfg(n, p, G, c);
        }
    return 0;
  }
  WD_REQUIRE(nd == 3 || nd == 4, "nd must be 0 (all), 3 or 4");
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(gauge_reals == 18 || gauge_reals == 12, "gauge_reals must be 18 or 12");
  LaunchCfg c = launch_cfg(nd, prec, gauge_reals);
  c.block = block_threads;
  c.remap = xcd_remap ? 1 : 0;
  set_launch_cfg(nd, prec, gauge_reals, c);
  return 0;
}

int dslash_set_launch_march(int nd, int prec, int gauge_reals, int march) {
  clear_error();
  WD_REQUIRE(march_ok(march), "march must be 1, 4, 8, 16 or 32 (got %d)", march);
  WD_REQUIRE(nd == 3 || nd == 4, "nd must be 3 or 4");
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(gauge_reals == 18 || gauge_reals == 12, "gauge_reals must be 18 or 12");
  LaunchCfg c = launch_cfg(nd, prec, gauge_reals);
  c.march = march;
  set_launch_cfg(nd, prec, gauge_reals, c);
  return 0;
}

int dslash_launch_march_for(const int latt[4], int nd, int prec, int gauge_reals, int* march) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(gauge_reals == 18 || gauge_reals == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(march, "march is NULL");
  const LaunchCfg c = launch_cfg(nd, prec, gauge_reals);
  *march = effective_march(latt, nd, prec, c.block, c.march);
  return 0;
}

int dslash_get_launch_march(int nd, int prec, int gauge_reals, int* march) {
  clear_error();
  WD_REQUIRE(nd == 3 || nd == 4, "nd must be 3 or 4");
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(gauge_reals == 18 || gauge_reals == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(march, "march is NULL");
  *march = launch_cfg(nd, prec, gauge_reals).march;
  return 0;
}

int dslash_get_launch(int nd, int prec, int gauge_reals, int* block_threads, int* xcd_remap) {
  clear_error();
  WD_REQUIRE(nd == 3 || nd == 4, "nd must be 3 or 4");
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(gauge_reals == 18 || gauge_reals == 12, "gauge_reals must be 18 or 12");
  LaunchCfg c = launch_cfg(nd, prec, gauge_reals);
  if (block_threads) *block_threads = c.block;
  if (xcd_remap) *xcd_remap = c.remap;
  return 0;
}

int dslash_apply(void* d_res_cb, const void* d_psi_cb, const void* d_gauge_cb, const int latt[4], int nd,
                 int prec, int gauge_reals, int isign, int target_cb, void* stream) {
  clear_error();
  return apply_cb(d_res_cb, d_psi_cb, d_gauge_cb, latt, nd, prec, gauge_reals, isign, target_cb,
                  (hipStream_t)stream);
}

int dslash_apply_xpay(void* d_res_cb, const void* d_psi_cb, const void* d_gauge_cb, const void* d_x_cb, double a,
                      const int latt[4], int prec, int gauge_reals, int isign, int target_cb, void* stream) {
  clear_error();
  WD_REQUIRE(d_x_cb, "x is NULL");
  return apply_cb_ex(d_res_cb, d_psi_cb, d_gauge_cb, latt, 4, prec, gauge_reals, isign, target_cb, d_x_cb, a,
                     (hipStream_t)stream);
}

int dslash_apply_full(void* d_res, const void* d_psi, const void* d_gauge, const int latt[4], int nd, int prec,
                      int gauge_reals, int isign, int cb, void* stream) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(cb == 0 || cb == 1, "cb must be 0 or 1 (got %d)", cb);
  WD_REQUIRE(gauge_reals == 18 || gauge_reals == 12, "gauge_reals must be 18 or 12");
  Geom g = make_geom(latt, nd);
  size_t sb = spinor_cb_bytes(g.vh, prec), gb = gauge_cb_bytes(g.vh, nd, prec, gauge_reals);
  int tcb = 1 - cb;
  return apply_cb((char*)d_res + tcb * sb, (const char*)d_psi + cb * sb, (const char*)d_gauge + tcb * gb, latt, nd,
                  prec, gauge_reals, isign, tcb, (hipStream_t)stream);
}

int dslash_pack_gauge(void* d_bundle, int target_cb, const void* d_u_packed, int prec_in, const int latt[4], int nd,
                      int prec_out, int G, const int* d_site_table, void* stream) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec_out)) return e;
  prec_out = link_prec(prec_out);
  WD_REQUIRE(prec_in == F32 || prec_in == F64, "packed gauge must be F32 or F64");
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(target_cb == 0 || target_cb == 1, "target_cb must be 0 or 1");
  WD_REQUIRE(d_bundle && d_u_packed, "NULL pointer");
  Geom g = make_geom(latt, nd);
  dim3 grid((g.vh + 255) / 256), blk(256);
  hipStream_t s = (hipStream_t)stream;
#define PACK_G(ND, PI, PO, GG)                                                                          \
  hipLaunchKernelGGL((k_pack_gauge<ND, PI, PO, GG>), grid, blk, 0, s, (v4u*)d_bundle,                  \
                     (const typename Fmt<PI>::real*)d_u_packed, d_site_table, g, target_cb)
#define PACK_G_PO(ND, PI)                                   \
  switch (prec_out * 2 + (G == 12)) {                       \
    case F32 * 2: PACK_G(ND, PI, F32, 18); break;           \
    case F32 * 2 + 1: PACK_G(ND, PI, F32, 12); break;       \
    case F64 * 2: PACK_G(ND, PI, F64, 18); break;           \
    case F64 * 2 + 1: PACK_G(ND, PI, F64, 12); break;       \
    case F16 * 2: PACK_G(ND, PI, F16, 18); break;           \
    case F16 * 2 + 1: PACK_G(ND, PI, F16, 12); break;       \
    case BF16 * 2: PACK_G(ND, PI, BF16, 18); break;         \
    default: PACK_G(ND, PI, BF16, 12); break;               \
  }
  if (nd == 4) {
    if (prec_in == F32) { PACK_G_PO(4, F32) } else { PACK_G_PO(4, F64) }
  } else {
    if (prec_in == F32) { PACK_G_PO(3, F32) } else { PACK_G_PO(3, F64) }
  }
#undef PACK_G_PO
#undef PACK_G
  WD_LAUNCHED();
  return 0;
}

int dslash_pack_spinor_sites(void* d_cb, int cb, const void* d_aos, int prec_in, const int latt[4], int nd,
                             int prec_out, const int* d_site_table, size_t site_begin, size_t site_end,
                             void* stream) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec_out)) return e;
  prec_out = spinor_store_prec(prec_out);
  WD_REQUIRE(prec_in == F32 || prec_in == F64, "legacy spinor must be F32 or F64");
  WD_REQUIRE(cb == 0 || cb == 1, "cb must be 0 or 1");
  WD_REQUIRE(d_cb && d_aos, "NULL pointer");
  Geom g = make_geom(latt, nd);
  WD_REQUIRE(site_begin <= site_end && site_end <= g.vh, "site range [%zu, %zu) outside the %u sites of a checkerboard",
             site_begin, site_end, g.vh);
  if (site_begin == site_end) return 0;
  const uint32_t i0 = (uint32_t)site_begin, i1 = (uint32_t)site_end;
  size_t planes = (size_t)spinor_vec(prec_out) * 16 * g.vh;
  v4u* out = (v4u*)d_cb;
  float* nrm = (float*)((char*)d_cb + planes);
  dim3 grid((i1 - i0 + 255) / 256), blk(256);
  hipStream_t s = (hipStream_t)stream;
#define PACK_S(PI, PO)                                                                                        \
  hipLaunchKernelGGL((k_pack_spinor<PI, PO>), grid, blk, 0, s, out, nrm, (const typename Fmt<PI>::real*)d_aos, \
                     d_site_table, g.vh, cb, i0, i1)
#define PACK_S_PO(PI)                          \
  switch (prec_out) {                          \
    case F32: PACK_S(PI, F32); break;          \
    case F64: PACK_S(PI, F64); break;          \
    case F16: PACK_S(PI, F16); break;          \
    default: PACK_S(PI, BF16); break;          \
  }
  if (prec_in == F32) { PACK_S_PO(F32) } else { PACK_S_PO(F64) }
#undef PACK_S_PO
#undef PACK_S
  WD_LAUNCHED();
  return 0;
}

int dslash_pack_spinor(void* d_cb, int cb, const void* d_aos, int prec_in, const int latt[4], int nd, int prec_out,
                       const int* d_site_table, void* stream) {
  if (int e = check_lattice(latt, nd)) return e;
  return dslash_pack_spinor_sites(d_cb, cb, d_aos, prec_in, latt, nd, prec_out, d_site_table, 0,
                                  make_geom(latt, nd).vh, stream);
}

int dslash_unpack_spinor_sites(void* d_aos, int cb, const void* d_cb, int prec_cb, const int latt[4], int nd,
                               int prec_out, const int* d_site_table, int compact, size_t site_begin,
                               size_t site_end, void* stream) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec_cb)) return e;
  prec_cb = spinor_store_prec(prec_cb);
  WD_REQUIRE(prec_out == F32 || prec_out == F64, "legacy spinor must be F32 or F64");
  WD_REQUIRE(cb == 0 || cb == 1, "cb must be 0 or 1");
  WD_REQUIRE(d_cb && d_aos, "NULL pointer");
  Geom g = make_geom(latt, nd);
  WD_REQUIRE(site_begin <= site_end && site_end <= g.vh, "site range [%zu, %zu) outside the %u sites of a checkerboard",
             site_begin, site_end, g.vh);
  if (site_begin == site_end) return 0;
  const uint32_t i0 = (uint32_t)site_begin, i1 = (uint32_t)site_end;
  size_t planes = (size_t)spinor_vec(prec_cb) * 16 * g.vh;
  const v4u* in = (const v4u*)d_cb;
  const float* nrm = (const float*)((const char*)d_cb + planes);
  dim3 grid((i1 - i0 + 255) / 256), blk(256);
  hipStream_t s = (hipStream_t)stream;
#define UNPACK_S(PC, PO)                                                                                      \
  hipLaunchKernelGGL((k_unpack_spinor<PC, PO>), grid, blk, 0, s, (typename Fmt<PO>::real*)d_aos, in, nrm, \
                     d_site_table, g.vh, cb, compact, i0, i1)
#define UNPACK_S_PC(PO)                         \
  switch (prec_cb) {                            \
    case F32: UNPACK_S(F32, PO); break;         \
    case F64: UNPACK_S(F64, PO); break;         \
    case F16: UNPACK_S(F16, PO); break;         \
    default: UNPACK_S(BF16, PO); break;         \
  }
  if (prec_out == F32) { UNPACK_S_PC(F32) } else { UNPACK_S_PC(F64) }
#undef UNPACK_S_PC
#undef UNPACK_S
  WD_LAUNCHED();
  return 0;
}

int dslash_unpack_spinor(void* d_aos, int cb, const void* d_cb, int prec_cb, const int latt[4], int nd,
                         int prec_out, const int* d_site_table, int compact, void* stream) {
  if (int e = check_lattice(latt, nd)) return e;
  return dslash_unpack_spinor_sites(d_aos, cb, d_cb, prec_cb, latt, nd, prec_out, d_site_table, compact, 0,
                                    make_geom(latt, nd).vh, stream);
}

int dslash_apply_sites(void* d_res_cb, const void* d_psi_cb, const void* d_gauge_cb, const int latt[4], int nd,
                       int prec, int gauge_reals, int isign, int target_cb, size_t site_begin, size_t site_end,
                       void* stream) {
  clear_error();
  return apply_cb_sites(d_res_cb, d_psi_cb, d_gauge_cb, latt, nd, prec, gauge_reals, isign, target_cb, site_begin,
                        site_end, (hipStream_t)stream);
}

int dslash_convert_spinor(void* d_out, int prec_out, const void* d_in, int prec_in, const int latt[4], int nd,
                          void* stream) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec_out)) return e;
  if (int e = check_prec(prec_in)) return e;
  prec_out = spinor_store_prec(prec_out);
  prec_in = spinor_store_prec(prec_in);
  WD_REQUIRE(d_out && d_in && d_out != d_in, "bad pointers");
  Geom g = make_geom(latt, nd);
  const v4u* in = (const v4u*)d_in;
  const float* in_n = (const float*)((const char*)d_in + (size_t)spinor_vec(prec_in) * 16 * g.vh);
  v4u* out = (v4u*)d_out;
  float* out_n = (float*)((char*)d_out + (size_t)spinor_vec(prec_out) * 16 * g.vh);
  dim3 grid((g.vh + 255) / 256), blk(256);
  hipStream_t s = (hipStream_t)stream;
#define CONV(PI, PO) hipLaunchKernelGGL((k_convert_spinor<PI, PO>), grid, blk, 0, s, out, out_n, in, in_n, g.vh)
#define CONV_PO(PI)                        \
  switch (prec_out) {                      \
    case F32: CONV(PI, F32); break;        \
    case F64: CONV(PI, F64); break;        \
    case F16: CONV(PI, F16); break;        \
    default: CONV(PI, BF16); break;        \
  }
  switch (prec_in) {
    case F32: CONV_PO(F32) break;
    case F64: CONV_PO(F64) break;
    case F16: CONV_PO(F16) break;
    default: CONV_PO(BF16) break;
  }
#undef CONV_PO
#undef CONV
  WD_LAUNCHED();
  return 0;
}

int dslash_gen_gauge_host(void* u, int prec, const int latt[4], int nd, uint64_t seed) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  WD_REQUIRE(prec == F32 || prec == F64, "packed gauge must be F32 or F64");
  WD_REQUIRE(u, "NULL pointer");
  Geom g = make_geom(latt, nd);
  if (prec == F64) {
    if (nd == 4) host_parallel(2 * g.vh, [=](uint32_t i) { gen_gauge_site<double, 4>((double*)u, g, seed, i); });
    else host_parallel(2 * g.vh, [=](uint32_t i) { gen_gauge_site<double, 3>((double*)u, g, seed, i); });
  } else {
    if (nd == 4) host_parallel(2 * g.vh, [=](uint32_t i) { gen_gauge_site<float, 4>((float*)u, g, seed, i); });
    else host_parallel(2 * g.vh, [=](uint32_t i) { gen_gauge_site<float, 3>((float*)u, g, seed, i); });
  }
  return 0;
}

int dslash_gen_spinor_host(void* psi, int prec, const int latt[4], int nd, uint64_t seed, int field) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  WD_REQUIRE(prec == F32 || prec == F64, "legacy spinor must be F32 or F64");
  WD_REQUIRE(psi, "NULL pointer");
  Geom g = make_geom(latt, nd);
  if (prec == F64) {
    if (nd == 4) host_parallel(2 * g.vh, [=](uint32_t i) { gen_spinor_site<double, 4>((double*)psi, g, seed, field, i); });
    else host_parallel(2 * g.vh, [=](uint32_t i) { gen_spinor_site<double, 3>((double*)psi, g, seed, field, i); });
  } else {
    if (nd == 4) host_parallel(2 * g.vh, [=](uint32_t i) { gen_spinor_site<float, 4>((float*)psi, g, seed, field, i); });
    else host_parallel(2 * g.vh, [=](uint32_t i) { gen_spinor_site<float, 3>((float*)psi, g, seed, field, i); });
  }
  return 0;
}

int dslash_gen_gauge_device(void* d_u, int prec, const int latt[4], int nd, uint64_t seed, void* stream) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  WD_REQUIRE(prec == F32 || prec == F64, "packed gauge must be F32 or F64");
  WD_REQUIRE(d_u, "NULL pointer");
  Geom g = make_geom(latt, nd);
  dim3 grid((2 * g.vh + 127) / 128), blk(128);
  hipStream_t s = (hipStream_t)stream;
  if (prec == F64) {
    if (nd == 4) hipLaunchKernelGGL((k_gen_gauge<double, 4>), grid, blk, 0, s, (double*)d_u, g, seed);
    else hipLaunchKernelGGL((k_gen_gauge<double, 3>), grid, blk, 0, s, (double*)d_u, g, seed);
  } else {
    if (nd == 4) hipLaunchKernelGGL((k_gen_gauge<float, 4>), grid, blk, 0, s, (float*)d_u, g, seed);
    else hipLaunchKernelGGL((k_gen_gauge<float, 3>), grid, blk, 0, s, (float*)d_u, g, seed);
  }
  WD_LAUNCHED();
  return 0;
}

int dslash_gen_spinor_device(void* d_psi, int prec, const int latt[4], int nd, uint64_t seed, int field,
                             void* stream) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  WD_REQUIRE(prec == F32 || prec == F64, "legacy spinor must be F32 or F64");
  WD_REQUIRE(d_psi, "NULL pointer");
  Geom g = make_geom(latt, nd);
  dim3 grid((2 * g.vh + 127) / 128), blk(128);
  hipStream_t s = (hipStream_t)stream;
  if (prec == F64) {
    if (nd == 4) hipLaunchKernelGGL((k_gen_spinor<double, 4>), grid, blk, 0, s, (double*)d_psi, g, seed, field);
    else hipLaunchKernelGGL((k_gen_spinor<double, 3>), grid, blk, 0, s, (double*)d_psi, g, seed, field);
  } else {
    if (nd == 4) hipLaunchKernelGGL((k_gen_spinor<float, 4>), grid, blk, 0, s, (float*)d_psi, g, seed, field);
    else hipLaunchKernelGGL((k_gen_spinor<float, 3>), grid, blk, 0, s, (float*)d_psi, g, seed, field);
  }
  WD_LAUNCHED();
  return 0;
}

void dslash_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  U4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  U4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

static Geom g_layout;
static int g_layout_nd = 0;

int dslash_layout_set(const int latt[4], int nd) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  g_layout = make_geom(latt, nd);
  g_layout_nd = nd;
  return 0;
}

int dslash_layout_linear_index(const int coord[]) {
  Coord c{coord[0], coord[1], coord[2], coord[3]};
  if (g_layout_nd == 4) return (int)(site_parity<4>(c) * g_layout.vh + site_index<4>(g_layout, c));
  return (int)(site_parity<3>(c) * g_layout.vh + site_index<3>(g_layout, c));
}

void dslash_layout_site_coords(int coord[], int node, int linearsite) {
  (void)node;
  uint32_t i = (uint32_t)linearsite;
  Coord c = g_layout_nd == 4 ? full_coords<4>(g_layout, i) : full_coords<3>(g_layout, i);
  coord[0] = c.x;
  coord[1] = c.y;
  coord[2] = c.z;
  coord[3] = c.t;
}

int dslash_layout_node_number(const int coord[]) {
  (void)coord;
  return 0;
}

int dslash_time_apply(void* d_res_full, const void* d_psi_full, const void* d_gauge_full, const int latt[4], int nd,
                      int prec, int G, int n, int alternate, void* stream, float* ms) {
  clear_error();
  if (int e = check_lattice(latt, nd)) return e;
  if (int e = check_prec(prec)) return e;
  WD_REQUIRE(n > 0 && ms, "n must be positive and ms non-NULL");
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  hipStream_t s = (hipStream_t)stream;
  Geom g = make_geom(latt, nd);
  size_t sb = spinor_cb_bytes(g.vh, prec), gb = gauge_cb_bytes(g.vh, nd, prec, G);
  hipEvent_t e0, e1;
  WD_HIP(hipEventCreate(&e0));
  WD_HIP(hipEventCreate(&e1));
  WD_HIP(hipEventRecord(e0, s));
  for (int k = 0; k < n; ++k) {
    int isign = (alternate && (k & 2)) ? -1 : 1;
    int cb = alternate ? (k & 1) : 0;
    int tcb = 1 - cb;
    int e = apply_cb((char*)d_res_full + tcb * sb, (const char*)d_psi_full + cb * sb, (const char*)d_gauge_full + tcb * gb,
                     latt, nd, prec, G, isign, tcb, s);
    if (e) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return e;
    }
  }
  WD_HIP(hipEventRecord(e1, s));
  WD_HIP(hipEventSynchronize(e1));
  float t = 0;
  WD_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / n;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

}  // extern "C"
