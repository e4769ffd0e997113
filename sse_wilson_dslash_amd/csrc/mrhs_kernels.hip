// Multi-right-hand-side Wilson Dslash and the persistent launch over a list of
// lattices of uneven volumes.
//
// A tile is (lattice, block of BS target sites, block of R right-hand sides).
// A thread owns one site and R right-hand sides: it loads each link pair once
// and applies it to R neighbour spinors, so link bytes are paid once per R
// fields from HBM (and the tiles of the other RHS blocks of the same sites run
// next to each other on the same XCD, where they find the links in L2: links
// are therefore loaded with the normal cache policy here, not non-temporal).
// Tiles are numbered lattice by lattice; each XCD walks one contiguous range
// of tiles, so the work list is balanced over the 8 XCDs whatever the mix of
// volumes.  With a persistent grid each block strides over its XCD's range
// until the range is exhausted (every block exits after a bounded loop).
//
// t-marching tiles (t_block TB > 1): a tile is (t-chunk, block of spatial
// sites, RHS block) and a thread walks TB consecutive t-slices of one spatial
// site.  The source spinor psi(t+1) loaded for the forward t hop of slice t is
// the backward t neighbour of slice t+2 (same spatial site, same x parity), so
// its projected half-spinor is kept in registers for that hop: with 32 fields
// in flight the t neighbours (Lx/2*Ly*Lz sites apart) no longer fit L2 between
// their two uses, and this halves the t-direction reloads.
//
// LDS-staged t-marching tiles (t_block 8, FP32): a 128-thread block owns 128
// consecutive spatial sites of one (t-chunk, RHS block) and walks 8 slices
// together.  At each step every thread loads the spinor psi(t+1) at its own
// spatial site (its forward t neighbour) and stores it into a double-buffered
// LDS slab; one step later the slab is slice t+1 of the source, and the x and y
// hops of the whole block read their neighbours from it (neighbours outside the
// block, and all z hops, load from global memory; a generic pointer selects LDS
// or global per lane without a divergent branch).  The slot psi(t+1) goes into
// held psi(t-1), the backward t neighbour, which the thread reads first, so the
// t hops need no registers across steps.  Tiles are ordered (t-chunk, group of
// 8 RHS blocks, spatial block, RHS block in group) so the blocks that exchange
// z neighbours run at the same time.
#include <hip/hip_runtime.h>

#include <string.h>

#include "dslash_common.h"
#include "dslash_mi355x.h"
#include "mrhs.h"
#include "status.h"

namespace wd {

template <int P, int G, int IS, int R, int MU>
__device__ __forceinline__ void mrhs_dir(const MrhsLattice& L, const Coord& c, uint32_t idx, int r0, int nr,
                                         Spinor<typename Fmt<P>::compute> (&acc)[R]) {
  typedef typename Fmt<P>::compute T;
  constexpr int PV = PairVec<P, G>::value;
  const uint32_t vh = L.g.vh;
  C<T> Wf[3][3], Wb[3][3];
  load_pair<P, G, false>(L.gauge + (size_t)MU * PV * vh, idx, vh, Wf, Wb);
  const uint32_t nf = site_index<4>(L.g, shift_coord(L.g, c, MU, +1));
  const uint32_t nb = site_index<4>(L.g, shift_coord(L.g, c, MU, -1));
  const size_t planes = (size_t)spinor_vec(P) * vh;  // in 16-byte vectors
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r < nr) {
      const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride;
      const float* fn = (const float*)(f + planes);
      Spinor<T> p = load_spinor<P>(f, fn, nf, vh);
      spin_recon_add<MU, -IS>(acc[r], su3_mul(Wf, spin_project<MU, -IS>(p)));
      Spinor<T> q = load_spinor<P>(f, fn, nb, vh);
      spin_recon_add<MU, IS>(acc[r], su3_mul(Wb, spin_project<MU, IS>(q)));
    }
  }
}

template <int P, int G, int IS, int R>
__device__ __forceinline__ void mrhs_tile(const MrhsLattice& L, uint32_t idx, int r0) {
  typedef typename Fmt<P>::compute T;
  const int nr = min(R, L.nrhs - r0);
  Coord c = site_coords<4>(L.g, idx, L.tcb);
  Spinor<T> acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[r].s[s][k] = {(T)0, (T)0};
  mrhs_dir<P, G, IS, R, 0>(L, c, idx, r0, nr, acc);
  mrhs_dir<P, G, IS, R, 1>(L, c, idx, r0, nr, acc);
  mrhs_dir<P, G, IS, R, 2>(L, c, idx, r0, nr, acc);
  mrhs_dir<P, G, IS, R, 3>(L, c, idx, r0, nr, acc);
  const size_t planes = (size_t)spinor_vec(P) * L.g.vh;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r < nr) {
      v4u* o = L.res + (size_t)(r0 + r) * L.fstride;
      store_spinor<P, true>(o, (float*)(o + planes), idx, L.g.vh, acc[r]);
    }
}

// One thread: spatial site s of t-slices t0 .. t0+TB-1 (clipped to T), R fields.
template <int P, int G, int IS, int R, int TB>
__device__ __forceinline__ void mrhs_march(const MrhsLattice& L, uint32_t s, int t0, int r0) {
  typedef typename Fmt<P>::compute T;
  constexpr int PV = PairVec<P, G>::value;
  const int nr = min(R, L.nrhs - r0);
  const uint32_t vh = L.g.vh, S = L.S;
  const uint32_t q0 = fdiv(s, L.dlxh);
  const int xh = (int)(s - q0 * (uint32_t)L.g.lxh);
  const uint32_t q1 = fdiv(q0, L.dly);
  const int y = (int)(q0 - q1 * (uint32_t)L.g.L[1]), z = (int)q1;
  const int LT = L.g.L[3];
  const size_t planes = (size_t)spinor_vec(P) * vh;
  Half<T> fifo[2][R];
#pragma unroll
  for (int k = 0; k < TB; ++k) {
    const int t = t0 + k;
    if (t >= LT) break;
    const Coord c{2 * xh + ((L.tcb + y + z + t) & 1), y, z, t};
    const uint32_t idx = s + S * (uint32_t)t;
    Spinor<T> acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int sp = 0; sp < 4; ++sp)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) acc[r].s[sp][cc] = {(T)0, (T)0};
    mrhs_dir<P, G, IS, R, 0>(L, c, idx, r0, nr, acc);
    mrhs_dir<P, G, IS, R, 1>(L, c, idx, r0, nr, acc);
    mrhs_dir<P, G, IS, R, 2>(L, c, idx, r0, nr, acc);
    {
      C<T> Wf[3][3], Wb[3][3];
      load_pair<P, G, false>(L.gauge + (size_t)3 * PV * vh, idx, vh, Wf, Wb);
      const uint32_t nf = s + S * (uint32_t)(t + 1 == LT ? 0 : t + 1);
      const uint32_t nb = s + S * (uint32_t)(t == 0 ? LT - 1 : t - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r < nr) {
          const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride;
          const float* fn = (const float*)(f + planes);
          const Spinor<T> pf = load_spinor<P>(f, fn, nf, vh);
          spin_recon_add<3, -IS>(acc[r], su3_mul(Wf, spin_project<3, -IS>(pf)));
          Half<T> hb;
          if (k < 2) hb = spin_project<3, IS>(load_spinor<P>(f, fn, nb, vh));
          else hb = fifo[k & 1][r];
          spin_recon_add<3, IS>(acc[r], su3_mul(Wb, hb));
          if (k + 2 < TB) fifo[k & 1][r] = spin_project<3, IS>(pf);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r < nr) {
        v4u* o = L.res + (size_t)(r0 + r) * L.fstride;
        store_spinor<P, true>(o, (float*)(o + planes), idx, vh, acc[r]);
      }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// FP32 spinor from its six 16-byte vectors.
__device__ inline Spinor<float> spinor_of(const v4u (&buf)[6]) {
  const float* fr = reinterpret_cast<const float*>(buf);
  Spinor<float> q;
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) q.s[sp][cc] = {fr[(sp * 3 + cc) * 2], fr[(sp * 3 + cc) * 2 + 1]};
  return q;
}

// One spinor through a generic pointer: an LDS slab (stride = block) or global
// planes (stride = vh), FP32.
__device__ inline Spinor<float> spinor_at(const v4u* p, size_t stride) {
  v4u buf[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) buf[v] = p[(size_t)v * stride];
  return spinor_of(buf);
}

template <int G, int IS, int R, int MU>
__device__ __forceinline__ void lds_hop(const MrhsLattice& L, v4u (*slab)[6][MRHS_LDS_THREADS], const Coord& c,
                                        uint32_t idx, uint32_t s0, uint32_t tS, int r0,
                                        Spinor<float> (&acc)[R]) {
  constexpr int BS = MRHS_LDS_THREADS;
  constexpr int PV = PairVec<F32, G>::value;
  const uint32_t vh = L.g.vh;
  C<float> Wf[3][3], Wb[3][3];
  load_pair<F32, G, false>(L.gauge + (size_t)MU * PV * vh, idx, vh, Wf, Wb);
  // spatial indices of the two neighbours within slice t
  const uint32_t nf = site_index<4>(L.g, shift_coord(L.g, c, MU, +1)) - tS;
  const uint32_t nb = site_index<4>(L.g, shift_coord(L.g, c, MU, -1)) - tS;
  const bool inf = MU < 2 && nf - s0 < (uint32_t)BS, inb = MU < 2 && nb - s0 < (uint32_t)BS;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride + tS;
    const v4u* pf = inf ? &slab[r][0][nf - s0] : f + nf;
    const v4u* pb = inb ? &slab[r][0][nb - s0] : f + nb;
    const Spinor<float> p = spinor_at(pf, inf ? (size_t)BS : (size_t)vh);
    const Spinor<float> q = spinor_at(pb, inb ? (size_t)BS : (size_t)vh);
    spin_recon_add<MU, -IS>(acc[r], su3_mul(Wf, spin_project<MU, -IS>(p)));
    spin_recon_add<MU, IS>(acc[r], su3_mul(Wb, spin_project<MU, IS>(q)));
  }
}

// One LDS-staged tile: spatial sites [s0, s0 + BS) of slices t0 .. t0+7 (clipped
// to T), fields r0 .. r0+R-1 (LDS tiles need nrhs to be a multiple of R, so the
// RHS block is always full).  Every thread of the block executes the same
// barriers: the step loop bound and the tile are uniform over the block.
template <int G, int IS, int R>
__device__ __forceinline__ void mrhs_march_lds(const MrhsLattice& L, v4u (*slab)[R][6][MRHS_LDS_THREADS],
                                               uint32_t s0, int t0, int r0) {
  constexpr int BS = MRHS_LDS_THREADS;
  constexpr int TB = MRHS_T_BLOCK_LDS;
  constexpr int PV = PairVec<F32, G>::value;
  const uint32_t vh = L.g.vh, S = L.S;
  const int LT = L.g.L[3];
  const uint32_t s = s0 + threadIdx.x;
  const bool live = s < S;
  const uint32_t sc = live ? s : S - 1;
  const uint32_t q0 = fdiv(sc, L.dlxh);
  const int xh = (int)(sc - q0 * (uint32_t)L.g.lxh);
  const uint32_t q1 = fdiv(q0, L.dly);
  const int y = (int)(q0 - q1 * (uint32_t)L.g.L[1]), z = (int)q1;
  // preamble: slices t0 and t0 - 1 of the source at the own spatial site -> slabs 0 and 1
  if (live) {
    const uint32_t tm = (uint32_t)(t0 == 0 ? LT - 1 : t0 - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride + s;
#pragma unroll
      for (int v = 0; v < 6; ++v) {
        slab[0][r][v][threadIdx.x] = f[(size_t)v * vh + S * (uint32_t)t0];
        slab[1][r][v][threadIdx.x] = f[(size_t)v * vh + S * tm];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < TB; ++k) {
    const int t = t0 + k;
    if (t >= LT) break;
    const int cur = k & 1;
    if (live) {
      const Coord c{2 * xh + ((L.tcb + y + z + t) & 1), y, z, t};
      const uint32_t tS = S * (uint32_t)t;
      const uint32_t idx = s + tS;
      Spinor<float> acc[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int sp = 0; sp < 4; ++sp)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) acc[r].s[sp][cc] = {0.f, 0.f};
      {  // t hops.  The own slot of the other slab holds slice t - 1 (the
         // backward neighbour); it is read, then overwritten with slice t + 1
         // (the forward neighbour), which is the next step's slab.  No other
         // lane reads that slab during this step.
        C<float> Wf[3][3], Wb[3][3];
        load_pair<F32, G, false>(L.gauge + (size_t)3 * PV * vh, idx, vh, Wf, Wb);
        const uint32_t nf = s + S * (uint32_t)(t + 1 == LT ? 0 : t + 1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride;
          const Spinor<float> q = spinor_at(&slab[cur ^ 1][r][0][threadIdx.x], (size_t)BS);
          spin_recon_add<3, IS>(acc[r], su3_mul(Wb, spin_project<3, IS>(q)));
          v4u buf[6];
#pragma unroll
          for (int v = 0; v < 6; ++v) buf[v] = f[(size_t)v * vh + nf];
#pragma unroll
          for (int v = 0; v < 6; ++v) slab[cur ^ 1][r][v][threadIdx.x] = buf[v];
          spin_recon_add<3, -IS>(acc[r], su3_mul(Wf, spin_project<3, -IS>(spinor_of(buf))));
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      lds_hop<G, IS, R, 0>(L, slab[cur], c, idx, s0, tS, r0, acc);
      __builtin_amdgcn_sched_barrier(0);
      lds_hop<G, IS, R, 1>(L, slab[cur], c, idx, s0, tS, r0, acc);
      __builtin_amdgcn_sched_barrier(0);
      lds_hop<G, IS, R, 2>(L, slab[cur], c, idx, s0, tS, r0, acc);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < R; ++r) store_spinor<F32, true>(L.res + (size_t)(r0 + r) * L.fstride, nullptr, idx, vh, acc[r]);
    }
    __syncthreads();  // slab[cur ^ 1] complete; slab[cur] free for the next step's writes
  }
  __syncthreads();  // the next tile's preamble overwrites slab 0
}

template <int G, int R, int IS>
__global__ void __launch_bounds__(MRHS_LDS_THREADS) k_mrhs_lds(MrhsList list) {
  __shared__ v4u slab[2][R][6][MRHS_LDS_THREADS];
  const uint32_t NX = gridDim.x < 8 ? gridDim.x : 8;
  const uint32_t total = list.total_tiles;
  const uint32_t x = blockIdx.x % NX;
  const uint32_t per_x = gridDim.x / NX + (x < (gridDim.x % NX) ? 1u : 0u);
  const uint32_t j = blockIdx.x / NX;
  const uint32_t t_lo = (uint32_t)((uint64_t)total * x / NX), t_hi = (uint32_t)((uint64_t)total * (x + 1) / NX);
  for (uint32_t t = t_lo + j; t < t_hi; t += per_x) {
    int l = 0;
    while (l + 1 < list.n && t >= list.lat[l + 1].tile0) ++l;
    const MrhsLattice& L = list.lat[l];
    // local tile = ((tc * ngroups + group) * nsb + sb) * rg + rb_in_group
    const uint32_t local = t - L.tile0;
    const uint32_t q = local / L.rg, rbi = local - q * L.rg;
    const uint32_t q2 = q / L.nsb, sb = q - q2 * L.nsb;
    const uint32_t ng = L.nrb / L.rg, tc = q2 / ng, grp = q2 - tc * ng;
    const int r0 = (int)(grp * L.rg + rbi) * R;
    mrhs_march_lds<G, IS, R>(L, slab, sb * MRHS_LDS_THREADS, (int)tc * MRHS_T_BLOCK_LDS, r0);
  }
}

// ---------------------------------------------------------------------------
// LDS-staged box tiles.  A block owns a box of Lx/2 x ty x tz spatial sites (a
// whole x row, ty rows in y, tz planes in z) of one (t-chunk, RHS block) and
// walks 8 slices together like the LDS-staged tiles above.  Every x neighbour
// and every y / z neighbour inside the box comes from the LDS slab; only the
// y / z hops that leave the box load from global memory.  With the contiguous
// 128-site blocks all z hops went to global memory, and at 32 fields in flight
// L2 (4 MB per XCD) turns over faster than the neighbour blocks reach the same
// slice, so those loads were re-fetched from past L2 (3.7 GB per launch at
// 32 RHS on 24^3 x 48 against 1.2 GB compulsory reads).
// Slab layout (dynamic LDS): [2 buffers][R][6 (FP32) or 12 (FP64) vectors][blockDim.x lanes].
// ---------------------------------------------------------------------------
struct BoxLane {
  int i;           // x row position (= x / 2)
  int y, z;        // lattice coordinates
  int y0, z0;      // the box's first y and z (uniform over the block)
  bool live;       // a site of the lattice (partial boxes at the y / z ends)
};

// One FP32 / FP64 (or sloppy FP64, converted to FP32) spinor from its 16-byte
// vectors at p, p + stride, ...
template <int P> __device__ inline Spinor<typename Fmt<P>::compute> box_spinor(const v4u* p, size_t stride) {
  typedef typename Fmt<P>::real Rl;
  constexpr int NV = 24 * Fmt<P>::bytes / 16;
  v4u buf[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) buf[v] = p[(size_t)v * stride];
  const Rl* r = reinterpret_cast<const Rl*>(buf);
  Spinor<typename Fmt<P>::compute> q;
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
      q.s[sp][cc] = {to_compute<P>(r[(sp * 3 + cc) * 2]), to_compute<P>(r[(sp * 3 + cc) * 2 + 1])};
  return q;
}

// Halo of a one-field box (HALO tiles): the source sites just outside the
// box in y and z (faces y0 - 1 and y0 + ty for every z of the box, z0 - 1 and
// z0 + tz for every y; a face is absent when the box spans that direction),
// each already projected for the one hop that uses it: (1 + isign gamma_mu) for
// a backward face, (1 - isign gamma_mu) for a forward face.  Entries are
// numbered [y backward | y forward | z backward | z forward], each face x-row
// fastest; thread h of the block loads entry h.  Entries hold the 12 reals in
// the compute type (3 vectors: the FP32 and sloppy-FP64 tiles both compute in FP32).  LDS layout:
// [2 buffers][HV vectors][H] after the slab, buffer k & 1 holding slice t0 + k.
struct BoxHalo {
  v4u* buf;  // buffer 0; buffer 1 at buf + HV H
  int Hy;    // entries per y face (0: the box spans y)
  int Hz;    // entries per z face (0: the box spans z)
  int H;     // 2 Hy + 2 Hz
};

// A half spinor as 12 reals of T in 3 (float) or 6 (double) 16-byte vectors at
// p, p + stride, ...
template <typename T> constexpr int half_vecs() { return 12 * (int)sizeof(T) / 16; }
template <typename T> __device__ inline Half<T> half_from(const v4u* p, int stride) {
  constexpr int HV = half_vecs<T>();
  v4u buf[HV];
#pragma unroll
  for (int v = 0; v < HV; ++v) buf[v] = p[v * stride];
  const T* r = reinterpret_cast<const T*>(buf);
  Half<T> h;
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) h.h[sp][cc] = {r[(sp * 3 + cc) * 2], r[(sp * 3 + cc) * 2 + 1]};
  return h;
}

template <typename T> __device__ inline void half_to(v4u* p, int stride, const Half<T>& h) {
  constexpr int HV = half_vecs<T>();
  v4u buf[HV];
  T* r = reinterpret_cast<T*>(buf);
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      r[(sp * 3 + cc) * 2] = h.h[sp][cc].re;
      r[(sp * 3 + cc) * 2 + 1] = h.h[sp][cc].im;
    }
#pragma unroll
  for (int v = 0; v < HV; ++v) p[v * stride] = buf[v];
}

template <int P, int G, int IS, int R, int MU, bool NTL, bool HALO = false>
__device__ __forceinline__ void box_hop(const MrhsLattice& L, const v4u* slab, const BoxLane& b, const Coord& c,
                                        uint32_t idx, uint32_t tS, int r0,
                                        Spinor<typename Fmt<P>::compute> (&acc)[R], const BoxHalo& hal = {}) {
  typedef typename Fmt<P>::compute T;
  constexpr int PV = PairVec<P, G>::value;
  constexpr int NV = 24 * Fmt<P>::bytes / 16;
  const uint32_t vh = L.g.vh, nt = blockDim.x;
  const int lxh = L.g.lxh;
  C<T> Wf[3][3], Wb[3][3];
  load_pair<P, G, NTL>(L.gauge + (size_t)MU * PV * vh, idx, vh, Wf, Wb);
  const Coord cf = shift_coord(L.g, c, MU, +1), cb = shift_coord(L.g, c, MU, -1);
  // neighbour lanes inside the box (x: always; y / z: when the shifted
  // coordinate stays in the box's range, wrap included)
  int lf, lb;
  bool inf, inb;
  const int lane = (int)threadIdx.x;
  if constexpr (MU == 0) {
    lf = lane - b.i + (cf.x >> 1);
    lb = lane - b.i + (cb.x >> 1);
    inf = inb = true;
  } else if constexpr (MU == 1) {
    inf = (unsigned)(cf.y - b.y0) < L.ty;
    inb = (unsigned)(cb.y - b.y0) < L.ty;
    lf = lane + (cf.y - b.y) * lxh;
    lb = lane + (cb.y - b.y) * lxh;
  } else {
    inf = (unsigned)(cf.z - b.z0) < L.tz;
    inb = (unsigned)(cb.z - b.z0) < L.tz;
    lf = lane + (cf.z - b.z) * lxh * (int)L.ty;
    lb = lane + (cb.z - b.z) * lxh * (int)L.ty;
  }
  if constexpr (HALO && MU != 0) {
    static_assert(R == 1, "halo tiles take one field");
    // in-box neighbours from the slab, out-of-box ones from the halo (both LDS):
    // every lane reads both, at its own slot / entry 0 where unused, and keeps one
    const int face = MU == 1 ? b.z - b.z0 : b.y - b.y0;
    const int h0 = (MU == 1 ? 0 : 2 * hal.Hy) + face * lxh + b.i;
    const int hs = MU == 1 ? hal.Hy : hal.Hz;
    const Spinor<T> p = box_spinor<P>(slab + (inf ? lf : lane), nt);
    const Spinor<T> q = box_spinor<P>(slab + (inb ? lb : lane), nt);
    const Half<T> xf = half_from<T>(hal.buf + (inf ? 0 : h0 + hs), hal.H);
    const Half<T> xb = half_from<T>(hal.buf + (inb ? 0 : h0), hal.H);
    Half<T> hf = spin_project<MU, -IS>(p), hb = spin_project<MU, IS>(q);
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        hf.h[sp][cc] = inf ? hf.h[sp][cc] : xf.h[sp][cc];
        hb.h[sp][cc] = inb ? hb.h[sp][cc] : xb.h[sp][cc];
      }
    spin_recon_add<MU, -IS>(acc[0], su3_mul(Wf, hf));
    spin_recon_add<MU, IS>(acc[0], su3_mul(Wb, hb));
    return;
  }
  const uint32_t nf = site_index<4>(L.g, cf) - tS, nb = site_index<4>(L.g, cb) - tS;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride + tS;
    const v4u* sr = slab + (size_t)r * NV * nt;
    const v4u* pf = inf ? sr + lf : f + nf;
    const v4u* pb = inb ? sr + lb : f + nb;
    const Spinor<T> p = box_spinor<P>(pf, inf ? (size_t)nt : (size_t)vh);
    const Spinor<T> q = box_spinor<P>(pb, inb ? (size_t)nt : (size_t)vh);
    spin_recon_add<MU, -IS>(acc[r], su3_mul(Wf, spin_project<MU, -IS>(p)));
    spin_recon_add<MU, IS>(acc[r], su3_mul(Wb, spin_project<MU, IS>(q)));
  }
}

// One box tile: box (bty, btz) of slices t0 .. t0 + MRHS_BOX_SLICES - 1 (clipped to T), fields
// r0 .. r0+R-1.  Every thread executes the same barriers.  NTL: links loaded
// non-temporally (one field: every link is used once), else cached for the
// other RHS blocks of the same box.  EPI (one field only): EPI_XPAY stores
// x + a D psi, EPI_XPAY_NORM also adds each of the lane's |res|^2 to s2.
// One halo spinor of format P (its vectors in registers), projected for its face
// (mu = 1 or 2, backward or forward) in the compute type, to dst, dst + H, ...
template <int P, int IS>
__device__ inline void box_halo_store(const v4u* raw, v4u* dst, int H, int hmu, int hfwd) {
  typedef typename Fmt<P>::compute T;
  const Spinor<T> q = box_spinor<P>(raw, 1);
  Half<T> h;
  if (hmu == 1)
    h = hfwd ? spin_project<1, -IS>(q) : spin_project<1, IS>(q);
  else
    h = hfwd ? spin_project<2, -IS>(q) : spin_project<2, IS>(q);
  half_to(dst, H, h);
}

// HALO (one field): the y / z hops that leave the box read the projected
// halo (BoxHalo) from LDS; thread h loads halo entry h of slice t + 1 while the
// step for slice t runs, i.e. in the same step in which the neighbouring box
// loads those sites for its own slab, and writes it into the other halo buffer.
template <int P, int G, int IS, int R, bool NTL, int EPI, bool HALO = false>
__device__ __forceinline__ void mrhs_box_tile(const MrhsLattice& L, v4u* slab, uint32_t bty, uint32_t btz, int t0,
                                              int r0, const BoxEpi& e, double& s2) {
  static_assert(EPI == EPI_PLAIN || R == 1, "the epilogue forms take one field");
  static_assert(!HALO || R == 1, "halo tiles take one field");
  typedef typename Fmt<P>::compute T;
  constexpr int HV = half_vecs<T>();  // vectors per halo entry
  constexpr int TB = MRHS_BOX_SLICES;
  constexpr int PV = PairVec<P, G>::value;
  constexpr int NV = 24 * Fmt<P>::bytes / 16;
  const uint32_t vh = L.g.vh, S = L.S, nt = blockDim.x;
  const int LT = L.g.L[3], lxh = L.g.lxh;
  BoxLane b;
  {
    const int lane = (int)threadIdx.x, rest = lane / lxh;
    const int j = rest % (int)L.ty, k = rest / (int)L.ty;
    b.i = lane - rest * lxh;
    b.y0 = (int)(bty * L.ty);
    b.z0 = (int)(btz * L.tz);
    b.y = b.y0 + j;
    b.z = b.z0 + k;
    b.live = k < (int)L.tz && b.y < L.g.L[1] && b.z < L.g.L[2];
  }
  const uint32_t s = (uint32_t)b.i + (uint32_t)lxh * ((uint32_t)b.y + (uint32_t)L.g.L[1] * (uint32_t)b.z);
  const size_t bufsz = (size_t)R * NV * nt;
  // this thread's halo entry: face (y or z, backward or forward) and source site
  BoxHalo hal{};
  bool hlive = false;
  int hmu = 1, hfwd = 0;
  uint32_t hsite = 0;
  if constexpr (HALO) {
    const int Ly = L.g.L[1], Lz = L.g.L[2], ty = (int)L.ty, tz = (int)L.tz;
    hal.Hy = ty < Ly ? lxh * tz : 0;
    hal.Hz = tz < Lz ? lxh * ty : 0;
    hal.H = 2 * hal.Hy + 2 * hal.Hz;
    hal.buf = slab + 2 * bufsz;
    const int h = (int)threadIdx.x;
    hlive = h < hal.H;
    if (hlive) {
      const int y0 = b.y0, z0 = b.z0;
      int yy, zz, rem;
      if (h < 2 * hal.Hy) {
        hmu = 1;
        hfwd = h >= hal.Hy;
        rem = h - hfwd * hal.Hy;
        yy = hfwd ? (y0 + ty == Ly ? 0 : y0 + ty) : (y0 == 0 ? Ly - 1 : y0 - 1);
        zz = z0 + rem / lxh;
      } else {
        hmu = 2;
        hfwd = h - 2 * hal.Hy >= hal.Hz;
        rem = h - 2 * hal.Hy - hfwd * hal.Hz;
        zz = hfwd ? (z0 + tz == Lz ? 0 : z0 + tz) : (z0 == 0 ? Lz - 1 : z0 - 1);
        yy = y0 + rem / lxh;
      }
      hsite = (uint32_t)(rem % lxh) + (uint32_t)lxh * ((uint32_t)yy + (uint32_t)Ly * (uint32_t)zz);
    }
  }
  if constexpr (HALO) {  // preamble: the halo of slice t0 -> halo buffer 0
    if (hlive) {
      const v4u* f = L.psi + (size_t)r0 * L.fstride + hsite + S * (uint32_t)t0;
      v4u raw[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) raw[v] = f[(size_t)v * vh];
      box_halo_store<P, IS>(raw, hal.buf + threadIdx.x, hal.H, hmu, hfwd);
    }
  }
  if (b.live) {  // preamble: slice t0 at the own site -> buffer 0, slice t0 - 1 -> buffer 1
    const uint32_t tm = (uint32_t)(t0 == 0 ? LT - 1 : t0 - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride + s;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        slab[((size_t)r * NV + v) * nt + threadIdx.x] = f[(size_t)v * vh + S * (uint32_t)t0];
        slab[bufsz + ((size_t)r * NV + v) * nt + threadIdx.x] = f[(size_t)v * vh + S * tm];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < TB; ++k) {
    const int t = t0 + k;
    if (t >= LT) break;
    const int cur = k & 1;
    // the next slice's halo entry, requested before this slice's work
    const bool hnext = HALO && hlive && k + 1 < TB && t + 1 < LT;
    v4u hraw[HALO ? NV : 1];
    if constexpr (HALO) {
      if (hnext) {
        const v4u* f = L.psi + (size_t)r0 * L.fstride + hsite + S * (uint32_t)(t + 1);
#pragma unroll
        for (int v = 0; v < NV; ++v) hraw[v] = f[(size_t)v * vh];
      }
    }
    if (b.live) {
      const Coord c{2 * b.i + ((L.tcb + b.y + b.z + t) & 1), b.y, b.z, t};
      const uint32_t tS = S * (uint32_t)t;
      const uint32_t idx = s + tS;
      Spinor<T> acc[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int sp = 0; sp < 4; ++sp)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) acc[r].s[sp][cc] = {(T)0, (T)0};
      {  // t hops.  The own slot of the other buffer holds slice t - 1 (the
         // backward neighbour); it is read, then overwritten with slice t + 1
         // (the forward neighbour), which is the next step's slab.  No other
         // lane reads that buffer during this step.
        C<T> Wf[3][3], Wb[3][3];
        load_pair<P, G, NTL>(L.gauge + (size_t)3 * PV * vh, idx, vh, Wf, Wb);
        const uint32_t nf = s + S * (uint32_t)(t + 1 == LT ? 0 : t + 1);
        v4u* other = slab + (size_t)(cur ^ 1) * bufsz;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const v4u* f = L.psi + (size_t)(r0 + r) * L.fstride;
          const Spinor<T> q = box_spinor<P>(other + (size_t)r * NV * nt + threadIdx.x, nt);
          spin_recon_add<3, IS>(acc[r], su3_mul(Wb, spin_project<3, IS>(q)));
          v4u buf[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) buf[v] = f[(size_t)v * vh + nf];
#pragma unroll
          for (int v = 0; v < NV; ++v) other[((size_t)r * NV + v) * nt + threadIdx.x] = buf[v];
          spin_recon_add<3, -IS>(acc[r], su3_mul(Wf, spin_project<3, -IS>(box_spinor<P>(buf, 1))));
        }
      }
      const v4u* here = slab + (size_t)cur * bufsz;
      // Scheduling fences between the hops keep each hop's loads and math
      // together.  The one-field FP64 tile runs without them: the allocator then
      // fits the step in 256 VGPRs + 13 AGPRs without spilling (34 AGPRs and 2
      // spills fenced) and the kernel is 2.4% faster at 32^3 x 64; for FP32 and
      // sloppy FP64 the difference is within 0.5%.
      constexpr bool FENCE = !(NTL && P == F64);
      BoxHalo hc = hal;
      if constexpr (HALO) hc.buf = hal.buf + (size_t)cur * HV * hal.H;
      if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
      box_hop<P, G, IS, R, 0, NTL, HALO>(L, here, b, c, idx, tS, r0, acc, hc);
      if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
      box_hop<P, G, IS, R, 1, NTL, HALO>(L, here, b, c, idx, tS, r0, acc, hc);
      if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
      box_hop<P, G, IS, R, 2, NTL, HALO>(L, here, b, c, idx, tS, r0, acc, hc);
      if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
      if constexpr (EPI == EPI_PLAIN) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          store_spinor<P, true>(L.res + (size_t)(r0 + r) * L.fstride, nullptr, idx, vh, acc[r]);
      } else {  // read x before the store: x may alias res (same site, same lane)
        const Spinor<T> x = box_spinor<P>(e.xp + idx, vh);
        const T sa = (T)e.a;
        T n2 = (T)0;
#pragma unroll
        for (int sp = 0; sp < 4; ++sp)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            acc[0].s[sp][cc].re = x.s[sp][cc].re + sa * acc[0].s[sp][cc].re;
            acc[0].s[sp][cc].im = x.s[sp][cc].im + sa * acc[0].s[sp][cc].im;
            n2 += acc[0].s[sp][cc].re * acc[0].s[sp][cc].re + acc[0].s[sp][cc].im * acc[0].s[sp][cc].im;
          }
        if constexpr (EPI == EPI_XPAY_NORM) s2 += (double)n2;
        store_spinor<P, true>(L.res, nullptr, idx, vh, acc[0]);
      }
    }
    if constexpr (HALO) {
      if (hnext) box_halo_store<P, IS>(hraw, hal.buf + (size_t)(cur ^ 1) * HV * hal.H + threadIdx.x, hal.H, hmu, hfwd);
    }
    __syncthreads();  // buffer cur ^ 1 complete; buffer cur free for the next step's writes
  }
  __syncthreads();  // the next tile's preamble overwrites buffer 0
}

// Sum of v over the block (<= 256 threads) to *out on thread 0.  Every thread
// calls it; the caller's next barrier orders the LDS reuse of a following call.
__device__ inline void box_block_sum(double v, double* out) {
  __shared__ double part[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
    out[0] = t;
    out[1] = 0.0;
  }
}

// The box-tile grid: each block walks its XCD's contiguous share of the tiles.
template <int P, int G, int R, int IS, bool NTL, int EPI, bool HALO = false>
__device__ __forceinline__ void mrhs_box_grid(const MrhsList& list, const BoxEpi& e) {
  extern __shared__ v4u box_slab[];
  const uint32_t NX = gridDim.x < 8 ? gridDim.x : 8;
  const uint32_t total = list.total_tiles;
  const uint32_t x = blockIdx.x % NX;
  const uint32_t per_x = gridDim.x / NX + (x < (gridDim.x % NX) ? 1u : 0u);
  const uint32_t j = blockIdx.x / NX;
  const uint32_t t_lo = (uint32_t)((uint64_t)total * x / NX), t_hi = (uint32_t)((uint64_t)total * (x + 1) / NX);
  for (uint32_t t = t_lo + j; t < t_hi; t += per_x) {
    int l = 0;
    while (l + 1 < list.n && t >= list.lat[l + 1].tile0) ++l;
    const MrhsLattice& L = list.lat[l];
    // local tile = ((tc * ngroups + group) * nboxes + box) * rg + rb_in_group,
    // box = bty + nty * btz
    const uint32_t local = t - L.tile0;
    const uint32_t q = local / L.rg, rbi = local - q * L.rg;
    const uint32_t q2 = q / L.nsb, box = q - q2 * L.nsb;
    const uint32_t ng = L.nrb / L.rg, tc = q2 / ng, grp = q2 - tc * ng;
    const uint32_t btz = box / L.nty, bty = box - btz * L.nty;
    double s2 = 0.0;
    mrhs_box_tile<P, G, IS, R, NTL, EPI, HALO>(L, box_slab, bty, btz, (int)tc * MRHS_BOX_SLICES,
                                               (int)(grp * L.rg + rbi) * R, e, s2);
    if constexpr (EPI == EPI_XPAY_NORM) box_block_sum(s2, e.norm_partial + 2 * (size_t)t);
  }
}

template <int P, int G, int R, int IS, bool NTL, int EPI = EPI_PLAIN>
__global__ void __launch_bounds__(256) k_mrhs_box(MrhsList list, BoxEpi e) {
  mrhs_box_grid<P, G, R, IS, NTL, EPI>(list, e);
}

// One FP32 field with the LDS halo for the y / z hops that leave the box.
template <int G, int IS, int EPI>
__global__ void __launch_bounds__(256) k_box_halo(MrhsList list, BoxEpi e) {
  mrhs_box_grid<F32, G, 1, IS, true, EPI, true>(list, e);
}

// The same for one sloppy-FP64 field: FP64 slab, halo entries in the FP32 compute type.
template <int P, int G, int IS, int EPI>
__global__ void __launch_bounds__(256) k_box_halo64(MrhsList list, BoxEpi e) {
  mrhs_box_grid<P, G, 1, IS, true, EPI, true>(list, e);
}

// Wide FP32 tiles: R x blockDim.x <= MRHS_BOX_WIDE (one block per CU), cached links.
template <int G, int R, int IS>
__global__ void __launch_bounds__(MRHS_BOX_WIDE / R) k_mrhs_box_wide(MrhsList list) {
  mrhs_box_grid<F32, G, R, IS, false, EPI_PLAIN>(list, BoxEpi{});
}

template <int P, int G, int R, int TB>
__global__ void __launch_bounds__(256) k_mrhs(MrhsList list) {
  const uint32_t NX = gridDim.x < 8 ? gridDim.x : 8;
  const uint32_t total = list.total_tiles;
  // this block's XCD and its contiguous share of the tile range
  const uint32_t x = blockIdx.x % NX;
  const uint32_t per_x = gridDim.x / NX + (x < (gridDim.x % NX) ? 1u : 0u);
  const uint32_t j = blockIdx.x / NX;
  const uint32_t t_lo = (uint32_t)((uint64_t)total * x / NX), t_hi = (uint32_t)((uint64_t)total * (x + 1) / NX);
  for (uint32_t t = t_lo + j; t < t_hi; t += per_x) {
    int l = 0;
    while (l + 1 < list.n && t >= list.lat[l + 1].tile0) ++l;
    const MrhsLattice& L = list.lat[l];
    const uint32_t local = t - L.tile0;
    const uint32_t q = local / L.nrb, rb = local - q * L.nrb;
    if constexpr (TB == 1) {
      const uint32_t idx = q * blockDim.x + threadIdx.x;
      if (idx < L.g.vh) {
        if (L.isign > 0)
          mrhs_tile<P, G, 1, R>(L, idx, (int)rb * R);
        else
          mrhs_tile<P, G, -1, R>(L, idx, (int)rb * R);
      }
    } else {
      const uint32_t tc = q / L.nsb, sb = q - tc * L.nsb;
      const uint32_t s = sb * blockDim.x + threadIdx.x;
      if (s < L.S) {
        if (L.isign > 0)
          mrhs_march<P, G, 1, R, TB>(L, s, (int)tc * TB, (int)rb * R);
        else
          mrhs_march<P, G, -1, R, TB>(L, s, (int)tc * TB, (int)rb * R);
      }
    }
  }
}

template <int P, int G>
static void launch_mrhs_pg(const MrhsList& list, int R, int block, int TB, uint32_t grid, hipStream_t s) {
  if (TB == 1) {
    switch (R) {
      case 1: hipLaunchKernelGGL((k_mrhs<P, G, 1, 1>), dim3(grid), dim3(block), 0, s, list); break;
      case 2: hipLaunchKernelGGL((k_mrhs<P, G, 2, 1>), dim3(grid), dim3(block), 0, s, list); break;
      default: hipLaunchKernelGGL((k_mrhs<P, G, 4, 1>), dim3(grid), dim3(block), 0, s, list); break;
    }
  } else if (TB == MRHS_T_BLOCK) {
    if (R == 1) hipLaunchKernelGGL((k_mrhs<P, G, 1, MRHS_T_BLOCK>), dim3(grid), dim3(block), 0, s, list);
    else hipLaunchKernelGGL((k_mrhs<P, G, 2, MRHS_T_BLOCK>), dim3(grid), dim3(block), 0, s, list);
  } else if (TB == MRHS_T_BOX) {  // validated by mrhs_shape_ok / mrhs_box_compatible; one isign per list
    const bool plus = list.lat[0].isign > 0;
    const size_t lds = (size_t)2 * R * spinor_vec(P) * block * sizeof(v4u);
    if constexpr (P == F64) {  // FP64: one field per thread
      if (plus) hipLaunchKernelGGL((k_mrhs_box<P, G, 1, 1, false>), dim3(grid), dim3(block), lds, s, list, BoxEpi{});
      else hipLaunchKernelGGL((k_mrhs_box<P, G, 1, -1, false>), dim3(grid), dim3(block), lds, s, list, BoxEpi{});
    } else if (block > 256) {  // wide tiles: R x block <= MRHS_BOX_WIDE
      if (R == 1) {
        if (plus) hipLaunchKernelGGL((k_mrhs_box_wide<G, 1, 1>), dim3(grid), dim3(block), lds, s, list);
        else hipLaunchKernelGGL((k_mrhs_box_wide<G, 1, -1>), dim3(grid), dim3(block), lds, s, list);
      } else {
        if (plus) hipLaunchKernelGGL((k_mrhs_box_wide<G, 2, 1>), dim3(grid), dim3(block), lds, s, list);
        else hipLaunchKernelGGL((k_mrhs_box_wide<G, 2, -1>), dim3(grid), dim3(block), lds, s, list);
      }
    } else {
      if (R == 1) {
        if (plus) hipLaunchKernelGGL((k_mrhs_box<P, G, 1, 1, false>), dim3(grid), dim3(block), lds, s, list, BoxEpi{});
        else hipLaunchKernelGGL((k_mrhs_box<P, G, 1, -1, false>), dim3(grid), dim3(block), lds, s, list, BoxEpi{});
      } else {
        if (plus) hipLaunchKernelGGL((k_mrhs_box<P, G, 2, 1, false>), dim3(grid), dim3(block), lds, s, list, BoxEpi{});
        else hipLaunchKernelGGL((k_mrhs_box<P, G, 2, -1, false>), dim3(grid), dim3(block), lds, s, list, BoxEpi{});
      }
    }
  } else if constexpr (P == F32) {  // MRHS_T_BLOCK_LDS (validated by mrhs_shape_ok; one isign per list)
    const bool plus = list.lat[0].isign > 0;
    if (R == 1) {
      if (plus) hipLaunchKernelGGL((k_mrhs_lds<G, 1, 1>), dim3(grid), dim3(MRHS_LDS_THREADS), 0, s, list);
      else hipLaunchKernelGGL((k_mrhs_lds<G, 1, -1>), dim3(grid), dim3(MRHS_LDS_THREADS), 0, s, list);
    } else {
      if (plus) hipLaunchKernelGGL((k_mrhs_lds<G, 2, 1>), dim3(grid), dim3(MRHS_LDS_THREADS), 0, s, list);
      else hipLaunchKernelGGL((k_mrhs_lds<G, 2, -1>), dim3(grid), dim3(MRHS_LDS_THREADS), 0, s, list);
    }
  }
}

// Box extents (ty, tz) for a block of nt threads: divisors of Ly and Lz (no
// partial boxes) with lxh * ty * tz <= nt, chosen to minimise the global loads
// per site: 1/ty + 1/tz hops leave the box (0 for a direction the box spans),
// plus the site's other traffic (taken as 0.5 hop), over the fraction of the
// block's lanes that hold a site.  Partial boxes (extents that do not divide Ly
// or Lz, costed by the lanes launched per site) were measured 1.7-2% slower at
// 48^3 x 96, where they fill 240 instead of 216 of 256 lanes.
static void box_shape(int lxh, int ly, int lz, int nt, uint32_t* ty, uint32_t* tz) {
  double best = 1e30;
  *ty = *tz = 1;
  for (int a = 1; a <= ly; ++a) {
    if (ly % a) continue;
    for (int c = 1; c <= lz; ++c) {
      if (lz % c || lxh * a * c > nt) continue;
      const double out = (a < ly ? 1.0 / a : 0.0) + (c < lz ? 1.0 / c : 0.0);
      const double cost = (out + 0.5) * nt / (double)(lxh * a * c);
      if (cost < best - 1e-12) {
        best = cost;
        *ty = (uint32_t)a;
        *tz = (uint32_t)c;
      }
    }
  }
}

int mrhs_build_list(const dslash_mrhs_problem* probs, int nprob, int prec, int R, int block, int t_block,
                    MrhsList* list) {
  WD_REQUIRE(probs && nprob >= 1 && nprob <= MRHS_MAX_LATTICES, "1..%d problems per launch", MRHS_MAX_LATTICES);
  WD_REQUIRE(mrhs_shape_ok(prec, R, block, t_block),
             "unsupported batched launch shape (rhs_block %d, %d threads, t_block %d): t_block 1 takes rhs_block "
             "1/2/4; t_block %d takes rhs_block 1/2; t_block %d takes F32, rhs_block 1/2 and %d threads; "
             "t_block %d takes F32, rhs_block 1/2 and >= 128 threads",
             R, block, t_block, MRHS_T_BLOCK, MRHS_T_BLOCK_LDS, MRHS_LDS_THREADS, MRHS_T_BOX);
  WD_REQUIRE(block >= 64 && block % 64 == 0 && (block <= 256 || t_block == MRHS_T_BOX),
             "block_threads must be 64..256, multiple of 64 (wide box tiles: up to %d)", MRHS_BOX_WIDE);
  memset(list, 0, sizeof(*list));
  uint64_t tiles = 0;
  for (int i = 0; i < nprob; ++i) {
    const dslash_mrhs_problem& p = probs[i];
    WD_REQUIRE(dslash_check_lattice(p.latt, 4) == 0, "problem %d: %s", i, dslash_last_error_message());
    WD_REQUIRE(p.nrhs >= 1, "problem %d: nrhs must be >= 1", i);
    WD_REQUIRE(p.isign == 1 || p.isign == -1, "problem %d: isign must be +-1", i);
    WD_REQUIRE(p.target_cb == 0 || p.target_cb == 1, "problem %d: target_cb must be 0 or 1", i);
    WD_REQUIRE(p.d_gauge_cb && p.d_psi && p.d_res, "problem %d: NULL pointer", i);
    MrhsLattice& L = list->lat[i];
    L.g = make_geom(p.latt, 4);
    const size_t cbb = spinor_cb_bytes(L.g.vh, prec);
    WD_REQUIRE(p.field_stride >= cbb && p.field_stride % 16 == 0,
               "problem %d: field_stride must be a multiple of 16 and >= %zu", i, cbb);
    WD_REQUIRE((const char*)p.d_res + p.field_stride * p.nrhs <= (const char*)p.d_psi ||
                   (const char*)p.d_psi + p.field_stride * p.nrhs <= (const char*)p.d_res,
               "problem %d: psi and res fields overlap", i);
    L.gauge = (const v4u*)p.d_gauge_cb;
    L.psi = (const v4u*)p.d_psi;
    L.res = (v4u*)p.d_res;
    L.fstride = p.field_stride / 16;
    L.nrhs = p.nrhs;
    L.tcb = p.target_cb;
    L.isign = p.isign;
    L.nrb = (p.nrhs + R - 1) / R;
    L.S = (uint32_t)(L.g.lxh * p.latt[1] * p.latt[2]);
    L.dlxh = make_fastdiv(L.g.lxh);
    L.dly = make_fastdiv(p.latt[1]);
    WD_REQUIRE((t_block != MRHS_T_BLOCK_LDS && t_block != MRHS_T_BOX) || (p.nrhs % R == 0 && p.isign == probs[0].isign),
               "problem %d: LDS-staged tiles need nrhs (%d) to be a multiple of rhs_block (%d) and one isign "
               "for the whole launch", i, p.nrhs, R);
    L.rg = L.nrb;
    if (t_block == 1) {
      L.nsb = (L.g.vh + block - 1) / block;
      L.ntc = 1;
    } else if (t_block == MRHS_T_BOX) {
      WD_REQUIRE(L.g.lxh <= block, "problem %d: box tiles need Lx/2 (%d) <= block_threads (%d)", i, L.g.lxh, block);
      box_shape(L.g.lxh, p.latt[1], p.latt[2], block, &L.ty, &L.tz);
      L.nty = (p.latt[1] + L.ty - 1) / L.ty;
      L.nsb = L.nty * ((p.latt[2] + L.tz - 1) / L.tz);
      L.ntc = (p.latt[3] + MRHS_BOX_SLICES - 1) / MRHS_BOX_SLICES;
      // A wide tile fills a CU, so an XCD runs ~32 at once: all fields of a box
      // in a row then share one fetch of its links.  256-thread tiles run two or
      // more per CU and take the fields of a box in groups.
      if (block <= 256 && L.nrb > MRHS_LDS_GROUP && L.nrb % MRHS_LDS_GROUP == 0) L.rg = MRHS_LDS_GROUP;
    } else {
      L.nsb = (L.S + block - 1) / block;
      L.ntc = (p.latt[3] + t_block - 1) / t_block;
      if (t_block == MRHS_T_BLOCK_LDS && L.nrb > MRHS_LDS_GROUP && L.nrb % MRHS_LDS_GROUP == 0) L.rg = MRHS_LDS_GROUP;
    }
    L.tile0 = (uint32_t)tiles;
    tiles += (uint64_t)L.ntc * L.nsb * L.nrb;
  }
  WD_REQUIRE(tiles < (1ull << 31), "too many tiles");
  list->n = nprob;
  list->total_tiles = (uint32_t)tiles;
  return 0;
}

uint32_t box_tiles_single(const int latt[4], int block) {
  const int lxh = latt[0] / 2;
  if (lxh > block) return 0;
  uint32_t ty, tz;
  box_shape(lxh, latt[1], latt[2], block, &ty, &tz);
  if ((uint32_t)lxh * ty * tz < 64) return 0;
  const uint64_t nbox = (uint64_t)((latt[1] + ty - 1) / ty) * (uint64_t)((latt[2] + tz - 1) / tz);
  const uint64_t tiles = nbox * (uint64_t)((latt[3] + MRHS_BOX_SLICES - 1) / MRHS_BOX_SLICES);
  const uint64_t vh = (uint64_t)lxh * latt[1] * latt[2] * latt[3];
  return tiles <= (vh + 63) / 64 ? (uint32_t)tiles : 0;
}

// Halo entries of a box of ty x tz rows and planes (0 when the box spans y and z).
static int box_halo_entries(int lxh, int ly, int lz, int ty, int tz) {
  const int hy = ty < ly ? lxh * tz : 0;
  const int hz = tz < lz ? lxh * ty : 0;
  return 2 * hy + 2 * hz;
}

uint32_t box_halo_single(const int latt[4], int block) {
  if (box_tiles_single(latt, block) == 0) return 0;
  uint32_t ty, tz;
  box_shape(latt[0] / 2, latt[1], latt[2], block, &ty, &tz);
  const int H = box_halo_entries(latt[0] / 2, latt[1], latt[2], (int)ty, (int)tz);
  return H <= block ? (uint32_t)H : 0;
}

template <int P, int G, int IS>
static void launch_box_single_s(const MrhsList& list, int block, const BoxEpi& e, hipStream_t s, bool halo) {
  const size_t lds = (size_t)2 * spinor_vec(P) * block * sizeof(v4u);
  const dim3 grid(list.total_tiles), blk(block);
  if constexpr (P == F32 || P == F64S) {  // FP64 has no halo form: its tile already fills 256 VGPRs
    const MrhsLattice& L = list.lat[0];
    const int H = box_halo_entries(L.g.lxh, L.g.L[1], L.g.L[2], (int)L.ty, (int)L.tz);
    if (halo && H >= 1 && H <= block) {  // one halo entry per thread at most
      const size_t lh = lds + (size_t)2 * half_vecs<typename Fmt<P>::compute>() * H * sizeof(v4u);
      if constexpr (P == F32) {
        if (!e.xp)
          hipLaunchKernelGGL((k_box_halo<G, IS, EPI_PLAIN>), grid, blk, lh, s, list, e);
        else if (!e.norm_partial)
          hipLaunchKernelGGL((k_box_halo<G, IS, EPI_XPAY>), grid, blk, lh, s, list, e);
        else
          hipLaunchKernelGGL((k_box_halo<G, IS, EPI_XPAY_NORM>), grid, blk, lh, s, list, e);
      } else {
        if (!e.xp)
          hipLaunchKernelGGL((k_box_halo64<P, G, IS, EPI_PLAIN>), grid, blk, lh, s, list, e);
        else if (!e.norm_partial)
          hipLaunchKernelGGL((k_box_halo64<P, G, IS, EPI_XPAY>), grid, blk, lh, s, list, e);
        else
          hipLaunchKernelGGL((k_box_halo64<P, G, IS, EPI_XPAY_NORM>), grid, blk, lh, s, list, e);
      }
      return;
    }
  }
  if (!e.xp)
    hipLaunchKernelGGL((k_mrhs_box<P, G, 1, IS, true, EPI_PLAIN>), grid, blk, lds, s, list, e);
  else if (!e.norm_partial)
    hipLaunchKernelGGL((k_mrhs_box<P, G, 1, IS, true, EPI_XPAY>), grid, blk, lds, s, list, e);
  else
    hipLaunchKernelGGL((k_mrhs_box<P, G, 1, IS, true, EPI_XPAY_NORM>), grid, blk, lds, s, list, e);
}

template <int P, int G>
static void launch_box_single(const MrhsList& list, int isign, int block, const BoxEpi& e, hipStream_t s,
                              bool halo) {
  if (isign > 0)
    launch_box_single_s<P, G, 1>(list, block, e, s, halo);
  else
    launch_box_single_s<P, G, -1>(list, block, e, s, halo);
}

int box_apply_single(void* res_cb, const void* psi_cb, const void* gauge_cb, const int latt[4], int prec, int G,
                     int isign, int tcb, int block, const BoxEpi& epi, hipStream_t s, bool halo) {
  WD_REQUIRE(prec == F32 || prec == F64 || prec == F64S, "box tiles are built for F32, F64 and F64S");
  WD_REQUIRE(block >= 128 && box_tiles_single(latt, block) > 0,
             "box tiles need >= 128 threads, Lx/2 <= block_threads and >= 64 sites per box");
  WD_REQUIRE(!epi.norm_partial || epi.xp, "the |res|^2 epilogue needs the x + a*D psi form");
  dslash_mrhs_problem p;
  for (int mu = 0; mu < 4; ++mu) p.latt[mu] = latt[mu];
  p.d_gauge_cb = gauge_cb;
  p.d_psi = psi_cb;
  p.d_res = res_cb;
  p.field_stride = spinor_cb_bytes(make_geom(latt, 4).vh, prec);
  p.nrhs = 1;
  p.isign = isign;
  p.target_cb = tcb;
  MrhsList list;
  if (int e = mrhs_build_list(&p, 1, prec, 1, block, MRHS_T_BOX, &list)) return e;
  if (prec == F32) {
    if (G == 18) launch_box_single<F32, 18>(list, isign, block, epi, s, halo);
    else launch_box_single<F32, 12>(list, isign, block, epi, s, halo);
  } else if (prec == F64) {
    if (G == 18) launch_box_single<F64, 18>(list, isign, block, epi, s, halo);
    else launch_box_single<F64, 12>(list, isign, block, epi, s, halo);
  } else {
    if (G == 18) launch_box_single<F64S, 18>(list, isign, block, epi, s, halo);
    else launch_box_single<F64S, 12>(list, isign, block, epi, s, halo);
  }
  WD_LAUNCHED();
  return 0;
}

int mrhs_launch(const MrhsList& list, int prec, int G, int R, int block, int t_block, uint32_t grid, hipStream_t s) {
  WD_REQUIRE(grid >= 1, "empty grid");
  if (prec == F32) {
    if (G == 18) launch_mrhs_pg<F32, 18>(list, R, block, t_block, grid, s);
    else launch_mrhs_pg<F32, 12>(list, R, block, t_block, grid, s);
  } else if (prec == F64) {
    if (G == 18) launch_mrhs_pg<F64, 18>(list, R, block, t_block, grid, s);
    else launch_mrhs_pg<F64, 12>(list, R, block, t_block, grid, s);
  } else {
    return set_error(DSLASH_ERR_UNSUPPORTED, "multi-RHS Dslash is built for F32 and F64");
  }
  WD_LAUNCHED();
  return 0;
}

}  // namespace wd
