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
  } else {
    if (R == 1) hipLaunchKernelGGL((k_mrhs<P, G, 1, MRHS_T_BLOCK>), dim3(grid), dim3(block), 0, s, list);
    else hipLaunchKernelGGL((k_mrhs<P, G, 2, MRHS_T_BLOCK>), dim3(grid), dim3(block), 0, s, list);
  }
}

int mrhs_build_list(const dslash_mrhs_problem* probs, int nprob, int prec, int R, int block, int t_block,
                    MrhsList* list) {
  WD_REQUIRE(probs && nprob >= 1 && nprob <= MRHS_MAX_LATTICES, "1..%d problems per launch", MRHS_MAX_LATTICES);
  WD_REQUIRE(R == 1 || R == 2 || R == 4, "rhs_block must be 1, 2 or 4");
  WD_REQUIRE(t_block == 1 || t_block == MRHS_T_BLOCK, "t_block must be 1 or %d", MRHS_T_BLOCK);
  WD_REQUIRE(t_block == 1 || R <= 2, "t-marching tiles are built for rhs_block 1 and 2");
  WD_REQUIRE(block >= 64 && block <= 256 && block % 64 == 0, "block_threads must be 64..256, multiple of 64");
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
    if (t_block == 1) {
      L.nsb = (L.g.vh + block - 1) / block;
      L.ntc = 1;
    } else {
      L.nsb = (L.S + block - 1) / block;
      L.ntc = (p.latt[3] + t_block - 1) / t_block;
    }
    L.tile0 = (uint32_t)tiles;
    tiles += (uint64_t)L.ntc * L.nsb * L.nrb;
  }
  WD_REQUIRE(tiles < (1ull << 31), "too many tiles");
  list->n = nprob;
  list->total_tiles = (uint32_t)tiles;
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
