// Domain-decomposition stages of the Wilson hopping term on one GPU.
//
// The reference's parscalar path (lib/sse_su3dslash_32bit_parscalar.c:108-820,
// driven at :1185-1392) splits D into
//   decomp      project the source spinor for the forward hops (1 -+ gamma_mu)
//   decomp_hvv  project for the backward hops and multiply by U_mu^dagger at
//               the source site
//   (exchange the boundary half-spinors with the +-mu neighbour nodes)
//   mvv_recons  U_mu(x) * half-spinor, reconstruct and accumulate
//   recons      reconstruct and accumulate (the link was applied by the sender)
// Here the same split works on faces only: the bulk of the operator is the
// fused interior kernel, which skips exactly the hops that cross the boundary
// of a partitioned direction, and the faces carry those hops:
//   dslash_halo_pack      boundary source sites -> send buffer (decomp, decomp_hvv)
//   dslash_halo_interior  every hop that stays inside (fused, as dslash_apply)
//   dslash_halo_exterior  recv buffer -> boundary target sites (mvv_recons, recons)
// With one GPU the exchange is the identity (recv = send, the periodic wrap of
// a ring of one); with several, recv[mu][fwd] comes from the +mu neighbour's
// send[mu][fwd] and recv[mu][bwd] from the -mu neighbour's send[mu][bwd].
//
// Face numbering: the sites of one checkerboard on the slice x_mu = c are
// numbered by their other coordinates, in increasing dimension order with the
// first one (always an even spatial extent) halved.  A target site and its
// neighbour across the boundary share these coordinates, so the sender's and
// the receiver's numbering agree without a table.
#include <hip/hip_runtime.h>

#include "dslash_common.h"
#include "dslash_internal.h"
#include "dslash_mi355x.h"
#include "status.h"

namespace wd {

struct FaceGeom {
  int d[3];      // the other dimensions, d[0] halved
  int n[3];      // extents along them
  uint32_t nf;   // sites per face (one checkerboard of the slice)
};

static FaceGeom make_face(const Geom& g, int mu) {
  FaceGeom f;
  int k = 0;
  for (int d = 0; d < 4; ++d)
    if (d != mu) f.d[k++] = d;
  f.n[0] = g.L[f.d[0]] / 2;
  f.n[1] = g.L[f.d[1]];
  f.n[2] = g.L[f.d[2]];
  f.nf = (uint32_t)f.n[0] * f.n[1] * f.n[2];
  return f;
}

template <int ND>
__device__ inline Coord face_site(const FaceGeom& f, int mu, uint32_t i, int slice, int parity) {
  int v[4];
  const int k0 = (int)(i % (uint32_t)f.n[0]);
  const uint32_t q = i / (uint32_t)f.n[0];
  v[f.d[1]] = (int)(q % (uint32_t)f.n[1]);
  v[f.d[2]] = (int)(q / (uint32_t)f.n[1]);
  v[mu] = slice;
  v[f.d[0]] = 0;
  const int partial = ND == 4 ? v[0] + v[1] + v[2] + v[3] : v[0] + v[1] + v[2];
  v[f.d[0]] = 2 * k0 + ((parity - partial) & 1);
  return Coord{v[0], v[1], v[2], v[3]};
}

__device__ inline int coord_of(const Coord& c, int mu) {
  return mu == 0 ? c.x : mu == 1 ? c.y : mu == 2 ? c.z : c.t;
}

// Half-spinor (12 reals) in SoA planes of one face: [hv][nf] 16-byte vectors.
template <int P> struct HalfVec {
  static constexpr int value = 12 * Fmt<P>::bytes / 16;
};

template <int P>
__device__ inline void store_half(v4u* base, uint32_t i, uint32_t nf, const Half<typename Fmt<P>::compute>& h) {
  typedef typename Fmt<P>::real R;
  constexpr int NV = HalfVec<P>::value;
  v4u buf[NV];
  R* r = reinterpret_cast<R*>(buf);
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[(s * 3 + c) * 2] = from_compute<P>(h.h[s][c].re);
      r[(s * 3 + c) * 2 + 1] = from_compute<P>(h.h[s][c].im);
    }
#pragma unroll
  for (int v = 0; v < NV; ++v) base[(size_t)v * nf + i] = buf[v];
}

template <int P>
__device__ inline Half<typename Fmt<P>::compute> load_half(const v4u* base, uint32_t i, uint32_t nf) {
  typedef typename Fmt<P>::real R;
  constexpr int NV = HalfVec<P>::value;
  v4u buf[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) buf[v] = base[(size_t)v * nf + i];
  const R* r = reinterpret_cast<const R*>(buf);
  Half<typename Fmt<P>::compute> h;
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      h.h[s][c].re = to_compute<P>(r[(s * 3 + c) * 2]);
      h.h[s][c].im = to_compute<P>(r[(s * 3 + c) * 2 + 1]);
    }
  return h;
}

// r = W^dagger h
template <typename T> __device__ inline Half<T> su3_adj_mul(const C<T> W[3][3], const Half<T>& h) {
  Half<T> r;
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      C<T> a = cmul(cconj(W[0][i]), h.h[sp][0]);
      cmac(a, cconj(W[1][i]), h.h[sp][1]);
      cmac(a, cconj(W[2][i]), h.h[sp][2]);
      r.h[sp][i] = a;
    }
  return r;
}

struct HaloArgs {
  const v4u* psi;      // source checkerboard
  v4u* res;            // target checkerboard
  const v4u* gauge_t;  // link bundle of the target checkerboard
  const v4u* gauge_s;  // link bundle of the source checkerboard (links U_mu(y) at source sites)
  v4u* face;           // this direction's faces: [fwd planes][bwd planes]
  Geom g;
  FaceGeom f;
  int mu;
  int tcb;
  int mask;
  int remap;
};

// ---- decomp / decomp_hvv -----------------------------------------------
template <int ND, int P, int G, int IS, int MU>
__device__ inline void pack_one(const HaloArgs& a, uint32_t i) {
  typedef typename Fmt<P>::compute T;
  constexpr int PV = PairVec<P, G>::value;
  constexpr int HV = HalfVec<P>::value;
  const uint32_t nf = a.f.nf;
  const bool fwd = i < nf;
  const uint32_t k = fwd ? i : i - nf;
  // forward face: source sites on x_mu = 0 (needed by the -mu side's forward
  // hops); backward face: source sites on x_mu = L-1, already multiplied by
  // U_mu(y)^dagger (needed by the +mu side's backward hops).
  const Coord c = face_site<ND>(a.f, MU, k, fwd ? 0 : a.g.L[MU] - 1, 1 - a.tcb);
  const uint32_t idx = site_index<ND>(a.g, c);
  const Spinor<T> p = load_spinor<P>(a.psi, nullptr, idx, a.g.vh);
  Half<T> h;
  if (fwd) {
    h = spin_project<MU, -IS>(p);
  } else {
    C<T> Wf[3][3], Wb[3][3];
    load_pair<P, G>(a.gauge_s + (size_t)MU * PV * a.g.vh, idx, a.g.vh, Wf, Wb);
    h = su3_adj_mul(Wf, spin_project<MU, IS>(p));
  }
  store_half<P>(a.face + (fwd ? 0 : (size_t)HV * nf), k, nf, h);
}

template <int ND, int P, int G, int IS>
__global__ void __launch_bounds__(256) k_halo_pack(HaloArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * a.f.nf) return;
  switch (a.mu) {
    case 0: pack_one<ND, P, G, IS, 0>(a, i); break;
    case 1: pack_one<ND, P, G, IS, 1>(a, i); break;
    case 2: pack_one<ND, P, G, IS, 2>(a, i); break;
    default: if constexpr (ND == 4) pack_one<ND, P, G, IS, 3>(a, i); break;
  }
}

// ---- mvv_recons / recons -------------------------------------------------
template <int ND, int P, int G, int IS, int MU>
__device__ inline void exterior_one(const HaloArgs& a, uint32_t i) {
  typedef typename Fmt<P>::compute T;
  constexpr int PV = PairVec<P, G>::value;
  constexpr int HV = HalfVec<P>::value;
  const uint32_t nf = a.f.nf;
  const bool fwd = i < nf;
  const uint32_t k = fwd ? i : i - nf;
  // forward hops leave through x_mu = L-1, backward hops through x_mu = 0
  const Coord c = face_site<ND>(a.f, MU, k, fwd ? a.g.L[MU] - 1 : 0, a.tcb);
  const uint32_t idx = site_index<ND>(a.g, c);
  const Half<T> h = load_half<P>(a.face + (fwd ? 0 : (size_t)HV * nf), k, nf);
  Spinor<T> acc = load_spinor<P>(a.res, nullptr, idx, a.g.vh);
  if (fwd) {
    C<T> Wf[3][3], Wb[3][3];
    load_pair<P, G>(a.gauge_t + (size_t)MU * PV * a.g.vh, idx, a.g.vh, Wf, Wb);
    spin_recon_add<MU, -IS>(acc, su3_mul(Wf, h));
  } else {
    spin_recon_add<MU, IS>(acc, h);
  }
  store_spinor<P, false>(a.res, nullptr, idx, a.g.vh, acc);
}

template <int ND, int P, int G, int IS>
__global__ void __launch_bounds__(256) k_halo_exterior(HaloArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * a.f.nf) return;
  switch (a.mu) {
    case 0: exterior_one<ND, P, G, IS, 0>(a, i); break;
    case 1: exterior_one<ND, P, G, IS, 1>(a, i); break;
    case 2: exterior_one<ND, P, G, IS, 2>(a, i); break;
    default: if constexpr (ND == 4) exterior_one<ND, P, G, IS, 3>(a, i); break;
  }
}

// ---- interior: every hop that does not cross a partitioned boundary -----
template <int ND, int P, int G, int IS, int MU>
__device__ __forceinline__ void hop_inside(const HaloArgs& a, const Coord& c, uint32_t idx,
                                           Spinor<typename Fmt<P>::compute>& acc) {
  typedef typename Fmt<P>::compute T;
  constexpr int PV = PairVec<P, G>::value;
  const uint32_t vh = a.g.vh;
  const bool cut = (a.mask >> MU) & 1;
  const int cm = coord_of(c, MU);
  const bool do_f = !(cut && cm == a.g.L[MU] - 1);
  const bool do_b = !(cut && cm == 0);
  C<T> Wf[3][3], Wb[3][3];
  load_pair<P, G>(a.gauge_t + (size_t)MU * PV * vh, idx, vh, Wf, Wb);
  if (do_f) {
    const uint32_t n = site_index<ND>(a.g, shift_coord(a.g, c, MU, +1));
    const Spinor<T> p = load_spinor<P>(a.psi, nullptr, n, vh);
    spin_recon_add<MU, -IS>(acc, su3_mul(Wf, spin_project<MU, -IS>(p)));
  }
  if (do_b) {
    const uint32_t n = site_index<ND>(a.g, shift_coord(a.g, c, MU, -1));
    const Spinor<T> p = load_spinor<P>(a.psi, nullptr, n, vh);
    spin_recon_add<MU, IS>(acc, su3_mul(Wb, spin_project<MU, IS>(p)));
  }
}

template <int ND, int P, int G, int IS>
__global__ void __launch_bounds__(256) k_halo_interior(HaloArgs a) {
  typedef typename Fmt<P>::compute T;
  const uint32_t b = a.remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t idx = b * blockDim.x + threadIdx.x;
  if (idx >= a.g.vh) return;
  const Coord c = site_coords<ND>(a.g, idx, a.tcb);
  Spinor<T> acc;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int k = 0; k < 3; ++k) acc.s[s][k] = {(T)0, (T)0};
  hop_inside<ND, P, G, IS, 0>(a, c, idx, acc);
  hop_inside<ND, P, G, IS, 1>(a, c, idx, acc);
  hop_inside<ND, P, G, IS, 2>(a, c, idx, acc);
  if constexpr (ND == 4) hop_inside<ND, P, G, IS, 3>(a, c, idx, acc);
  store_spinor<P, true>(a.res, nullptr, idx, a.g.vh, acc);
}

// ---- host side -------------------------------------------------------------
enum HaloStage { H_PACK, H_INTERIOR, H_EXTERIOR };

template <int ND, int P, int G, int IS>
static void launch_stage(HaloStage st, const HaloArgs& a, uint32_t blocks, int block, hipStream_t s) {
  if (st == H_PACK)
    hipLaunchKernelGGL((k_halo_pack<ND, P, G, IS>), dim3(blocks), dim3(block), 0, s, a);
  else if (st == H_EXTERIOR)
    hipLaunchKernelGGL((k_halo_exterior<ND, P, G, IS>), dim3(blocks), dim3(block), 0, s, a);
  else
    hipLaunchKernelGGL((k_halo_interior<ND, P, G, IS>), dim3(blocks), dim3(block), 0, s, a);
}

template <int ND>
static void dispatch_stage(HaloStage st, const HaloArgs& a, uint32_t blocks, int block, int prec, int G, int isign,
                           hipStream_t s) {
#define WD_STAGE(P, GG)                                             \
  if (isign > 0) launch_stage<ND, P, GG, 1>(st, a, blocks, block, s); \
  else launch_stage<ND, P, GG, -1>(st, a, blocks, block, s);
  if (prec == F32) {
    if (G == 18) { WD_STAGE(F32, 18) } else { WD_STAGE(F32, 12) }
  } else {
    if (G == 18) { WD_STAGE(F64, 18) } else { WD_STAGE(F64, 12) }
  }
#undef WD_STAGE
}

static int check_halo(const int latt[4], int nd, int prec, int G, int isign, int cb, int mask) {
  if (int e = dslash_check_lattice(latt, nd)) return e;
  WD_REQUIRE(prec == F32 || prec == F64, "the domain-decomposition stages are built for F32 and F64");
  WD_REQUIRE(G == 18 || G == 12, "gauge_reals must be 18 or 12");
  WD_REQUIRE(isign == 1 || isign == -1, "isign must be +1 or -1 (got %d)", isign);
  WD_REQUIRE(cb == 0 || cb == 1, "cb must be 0 or 1 (got %d)", cb);
  WD_REQUIRE(mask >= 0 && mask < (1 << nd), "dir_mask 0x%x names a direction the %dD operator does not hop in", mask, nd);
  return 0;
}

// Byte layout of the face buffer: per partitioned mu (increasing), the forward
// face then the backward face, each [hv][nf] vectors, each direction's pair
// 256-byte aligned.
static void face_layout(const Geom& g, int prec, int mask, size_t off[8], size_t face_b[4], size_t* total) {
  const size_t hb = (size_t)12 * prec_bytes(prec);
  size_t o = 0;
  for (int mu = 0; mu < 4; ++mu) {
    off[2 * mu] = off[2 * mu + 1] = 0;
    face_b[mu] = 0;
    if (!((mask >> mu) & 1)) continue;
    const size_t fb = hb * make_face(g, mu).nf;
    face_b[mu] = fb;
    off[2 * mu] = o;
    off[2 * mu + 1] = o + fb;
    o += align256(2 * fb);
  }
  *total = o;
}

}  // namespace wd

using namespace wd;

extern "C" {

int dslash_face_layout(const int latt[4], int nd, int prec, int dir_mask, size_t offsets[8], size_t face_bytes[4],
                       size_t* total_bytes) {
  clear_error();
  if (int e = check_halo(latt, nd, prec, 18, 1, 0, dir_mask)) return e;
  size_t off[8], fb[4], tot;
  face_layout(make_geom(latt, nd), prec, dir_mask, off, fb, &tot);
  if (offsets)
    for (int i = 0; i < 8; ++i) offsets[i] = off[i];
  if (face_bytes)
    for (int i = 0; i < 4; ++i) face_bytes[i] = fb[i];
  if (total_bytes) *total_bytes = tot;
  return 0;
}

static int halo_run(HaloStage st, void* d_res, const void* d_psi, void* d_face, const void* d_gauge, const int latt[4],
                    int nd, int prec, int G, int isign, int cb, int mask, void* stream) {
  clear_error();
  if (int e = check_halo(latt, nd, prec, G, isign, cb, mask)) return e;
  WD_REQUIRE(d_gauge, "gauge is NULL");
  HaloArgs a{};
  a.g = make_geom(latt, nd);
  const size_t sb = spinor_cb_bytes(a.g.vh, prec), gb = gauge_cb_bytes(a.g.vh, nd, prec, G);
  const int tcb = 1 - cb;
  a.tcb = tcb;
  a.mask = mask;
  a.psi = d_psi ? (const v4u*)((const char*)d_psi + (size_t)cb * sb) : nullptr;
  a.res = d_res ? (v4u*)((char*)d_res + (size_t)tcb * sb) : nullptr;
  a.gauge_t = (const v4u*)((const char*)d_gauge + (size_t)tcb * gb);
  a.gauge_s = (const v4u*)((const char*)d_gauge + (size_t)cb * gb);
  hipStream_t s = (hipStream_t)stream;
  if (st == H_INTERIOR) {
    WD_REQUIRE(d_res && d_psi, "NULL field pointer");
    WD_REQUIRE(d_res != d_psi, "res and psi must be different fields");
    const LaunchCfg lc = launch_cfg(nd, prec, G);
    a.remap = lc.remap;
    const uint32_t blocks = (a.g.vh + lc.block - 1) / lc.block;
    if (nd == 4) dispatch_stage<4>(st, a, blocks, lc.block, prec, G, isign, s);
    else dispatch_stage<3>(st, a, blocks, lc.block, prec, G, isign, s);
    WD_LAUNCHED();
    return 0;
  }
  WD_REQUIRE(d_face, "face buffer is NULL");
  WD_REQUIRE(st == H_PACK ? d_psi != nullptr : d_res != nullptr, "NULL field pointer");
  size_t off[8], fb[4], tot;
  face_layout(a.g, prec, mask, off, fb, &tot);
  // Directions one after another: a corner target site gets its exterior
  // terms in increasing mu, from launches ordered on the stream.
  for (int mu = 0; mu < nd; ++mu) {
    if (!((mask >> mu) & 1)) continue;
    a.mu = mu;
    a.f = make_face(a.g, mu);
    a.face = (v4u*)((char*)d_face + off[2 * mu]);
    const uint32_t blocks = (2 * a.f.nf + 255) / 256;
    if (nd == 4) dispatch_stage<4>(st, a, blocks, 256, prec, G, isign, s);
    else dispatch_stage<3>(st, a, blocks, 256, prec, G, isign, s);
    WD_LAUNCHED();
  }
  return 0;
}

int dslash_halo_pack(void* d_send, const void* d_psi, const void* d_gauge, const int latt[4], int nd, int prec,
                     int gauge_reals, int isign, int cb, int dir_mask, void* stream) {
  return halo_run(H_PACK, nullptr, d_psi, d_send, d_gauge, latt, nd, prec, gauge_reals, isign, cb, dir_mask, stream);
}

int dslash_halo_interior(void* d_res, const void* d_psi, const void* d_gauge, const int latt[4], int nd, int prec,
                         int gauge_reals, int isign, int cb, int dir_mask, void* stream) {
  return halo_run(H_INTERIOR, d_res, d_psi, nullptr, d_gauge, latt, nd, prec, gauge_reals, isign, cb, dir_mask,
                  stream);
}

int dslash_halo_exterior(void* d_res, const void* d_recv, const void* d_gauge, const int latt[4], int nd, int prec,
                         int gauge_reals, int isign, int cb, int dir_mask, void* stream) {
  return halo_run(H_EXTERIOR, d_res, nullptr, (void*)d_recv, d_gauge, latt, nd, prec, gauge_reals, isign, cb,
                  dir_mask, stream);
}

}  // extern "C"
