// Backward of the Wilson Dslash with respect to the gauge links.
//
// For S = Re <chi, D psi>, D mapping the source checkerboard s to the target
// t = 1 - s with
//   (D psi)(x) = sum_mu U_mu(x) P-_mu psi(x+mu) + U_mu(x-mu)^dagger P+_mu psi(x-mu),
//   P-_mu = 1 - isign gamma_mu,  P+_mu = 1 + isign gamma_mu,
// a change dU of one link gives dS = Re tr(dU G) with the spin-traced outer
// products
//   y on t:  G_mu(y) = sum_spin [P-_mu psi(y+mu)] chi(y)^dagger      (forward term)
//   y on s:  G_mu(y) = sum_spin chi(y+mu) [P+_mu psi(y)]^dagger      (backward term)
// The kernel writes, per link, either or both of
//   * the gradient with respect to the real entries of the packed (transposed)
//     link array: stored[a][b] = U(b,a), so d S / d stored[a][b] = conj(G[a][b]);
//   * the su(3) force F = TA((U G)^dagger), TA(M) = (M - M^dagger)/2 - tr(M - M^dagger)/6,
//     the unique traceless anti-Hermitian matrix with dS = Re tr(X^dagger F)
//     for dU = X U, X in su(3).
// One thread per (site, direction); FP64.  A block owns 64 sites and their four
// directions, so its output records are one contiguous 36 KiB range that it
// writes with coalesced vector stores (see k_force).
#include <hip/hip_runtime.h>

#include "dslash_common.h"
#include "dslash_mi355x.h"
#include "status.h"

namespace wd {

template <int MU, int S>
__device__ inline Spinor<double> full_project(const Spinor<double>& p) {
  // (1 + S gamma_mu) p as a full 4-spinor, rebuilt from the two-spinor
  Half<double> h = spin_project<MU, S>(p);
  Spinor<double> o;
#pragma unroll
  for (int sp = 0; sp < 4; ++sp)
#pragma unroll
    for (int c = 0; c < 3; ++c) o.s[sp][c] = {0.0, 0.0};
  spin_recon_add<MU, S>(o, h);
  return o;
}

template <int MU, int IS>
__device__ inline void force_link(const Geom& g, const v4u* chi_t, const float* chi_tn, const v4u* psi_s,
                                  const float* psi_sn, int tcb, int ycb, uint32_t idx, Coord c, C<double> G[3][3]) {
  const uint32_t vh = g.vh;
  const uint32_t n = site_index<4>(g, shift_coord(g, c, MU, +1));
  Spinor<double> a, b;  // G = sum_spin a_s b_s^dagger
  if (ycb == tcb) {     // forward term: a = P- psi(y+mu), b = chi(y)
    a = full_project<MU, -IS>(load_spinor<F64>(psi_s, psi_sn, n, vh));
    b = load_spinor<F64>(chi_t, chi_tn, idx, vh);
  } else {              // backward term: a = chi(y+mu), b = P+ psi(y)
    a = load_spinor<F64>(chi_t, chi_tn, n, vh);
    b = full_project<MU, IS>(load_spinor<F64>(psi_s, psi_sn, idx, vh));
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      C<double> acc{0.0, 0.0};
#pragma unroll
      for (int sp = 0; sp < 4; ++sp) cmac(acc, a.s[sp][i], cconj(b.s[sp][j]));
      G[i][j] = acc;
    }
}

// Block = 64 sites x 4 directions (wave w = direction w, lane = site), i.e. the
// 256 consecutive 144-byte records [4*site + mu] of the output arrays.  Link
// records are read and output records written cooperatively through a 36 KiB
// LDS tile with fully coalesced 16-byte accesses; each thread works on its own
// record in LDS.
constexpr int FORCE_SITES = 64;
constexpr int FORCE_REC = 4 * FORCE_SITES;  // records per block
typedef double d2 __attribute__((ext_vector_type(2)));

// Cooperative copy of the block's records between global memory and the tile
// (9 16-byte vectors per record; nrec records, the last block may be partial).
// Every record is read or written once per launch: non-temporal accesses
// (0.86 -> 0.78 ms at 32^3 x 64).
template <bool TO_GLOBAL>
__device__ inline void tile_copy(d2* tile, d2* glob, uint32_t nrec) {
  const uint32_t nvec = nrec * 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const uint32_t c = (uint32_t)k * FORCE_REC + threadIdx.x;
    if (c < nvec) {
      if (TO_GLOBAL) __builtin_nontemporal_store(tile[c], &glob[c]);
      else tile[c] = __builtin_nontemporal_load(&glob[c]);
    }
  }
}

template <int IS>
__global__ void __launch_bounds__(FORCE_REC) k_force(double* __restrict__ grad, double* __restrict__ force,
                                                     const double* __restrict__ u, const v4u* chi, const v4u* psi,
                                                     size_t cb_vec, size_t norm_off, Geom g, int tcb) {
  __shared__ d2 tile[FORCE_REC * 9];
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int mu = (int)(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t full0 = blk * FORCE_SITES;
  const uint32_t nsite = min((uint32_t)FORCE_SITES, 2 * g.vh - full0);
  const uint32_t nrec = 4 * nsite;
  const uint32_t full = full0 + lane;
  const bool live = lane < nsite;
  const size_t rec0 = (size_t)full0 * 4 * 9;  // first vector of the block's records
  const uint32_t my = (lane * 4 + (uint32_t)mu) * 9;  // this thread's record in the tile

  C<double> G[3][3];
  if (live) {
    const int ycb = full >= g.vh ? 1 : 0;
    const uint32_t idx = full - (uint32_t)ycb * g.vh;
    const Coord c = site_coords<4>(g, idx, ycb);
    const int scb = 1 - tcb;
    const v4u* chi_t = chi + (size_t)tcb * cb_vec;
    const v4u* psi_s = psi + (size_t)scb * cb_vec;
    const float* chi_tn = (const float*)(chi_t + norm_off);
    const float* psi_sn = (const float*)(psi_s + norm_off);
    switch (mu) {
      case 0: force_link<0, IS>(g, chi_t, chi_tn, psi_s, psi_sn, tcb, ycb, idx, c, G); break;
      case 1: force_link<1, IS>(g, chi_t, chi_tn, psi_s, psi_sn, tcb, ycb, idx, c, G); break;
      case 2: force_link<2, IS>(g, chi_t, chi_tn, psi_s, psi_sn, tcb, ycb, idx, c, G); break;
      default: force_link<3, IS>(g, chi_t, chi_tn, psi_s, psi_sn, tcb, ycb, idx, c, G); break;
    }
  }
  d2 F[9];
  if (force) {
    tile_copy<false>(tile, const_cast<d2*>(reinterpret_cast<const d2*>(u)) + rec0, nrec);
    __syncthreads();
    if (live) {
      C<double> U[3][3], A[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const d2 e = tile[my + j * 3 + i];  // stored transposed: U(i,j) = stored[j][i]
          U[i][j] = {e.x, e.y};
        }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          C<double> acc{0.0, 0.0};
#pragma unroll
          for (int k = 0; k < 3; ++k) cmac(acc, U[i][k], G[k][j]);
          A[i][j] = acc;
        }
      // M = A^dagger ; F = (M - M^dagger)/2 - tr(M - M^dagger)/6
      double trim = 0.0;  // Im tr(M - M^dagger) / 2 = -Im tr(A)
#pragma unroll
      for (int i = 0; i < 3; ++i) trim += -A[i][i].im;
      trim /= 3.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          // M_ij = conj(A_ji); (M - M^dag)_ij / 2 = (conj(A_ji) - A_ij) / 2
          double re = 0.5 * (A[j][i].re - A[i][j].re);
          double im = 0.5 * (-A[j][i].im - A[i][j].im);
          if (i == j) im -= trim;
          F[i * 3 + j] = d2{re, im};
        }
    }
    __syncthreads();  // every thread has read its link record
    if (live)
#pragma unroll
      for (int k = 0; k < 9; ++k) tile[my + k] = F[k];
    __syncthreads();
    tile_copy<true>(tile, reinterpret_cast<d2*>(force) + rec0, nrec);
  }
  if (grad) {
    if (force) __syncthreads();  // the force tile has been stored
    if (live)
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) tile[my + a * 3 + b] = d2{G[a][b].re, -G[a][b].im};
    __syncthreads();
    tile_copy<true>(tile, reinterpret_cast<d2*>(grad) + rec0, nrec);
  }
}

}  // namespace wd

using namespace wd;

extern "C" int dslash_force(void* d_grad_packed, void* d_force, const void* d_u_packed, const void* d_chi,
                            const void* d_psi, const int latt[4], int isign, int cb, void* stream) {
  clear_error();
  if (int e = dslash_check_lattice(latt, 4)) return e;
  WD_REQUIRE(isign == 1 || isign == -1, "isign must be +1 or -1");
  WD_REQUIRE(cb == 0 || cb == 1, "cb (source checkerboard) must be 0 or 1");
  WD_REQUIRE(d_chi && d_psi && (d_grad_packed || d_force), "NULL pointer");
  WD_REQUIRE(!d_force || d_u_packed, "the su(3) force needs the links");
  WD_REQUIRE((((uintptr_t)d_grad_packed | (uintptr_t)d_force | (uintptr_t)d_u_packed) & 15) == 0,
             "link arrays must be 16-byte aligned");
  Geom g = make_geom(latt, 4);
  const size_t cb_bytes = spinor_cb_bytes(g.vh, F64);
  dim3 grid((2 * g.vh + FORCE_SITES - 1) / FORCE_SITES), blk(FORCE_REC);
  hipStream_t s = (hipStream_t)stream;
  const size_t cb_vec = cb_bytes / 16, norm_off = (size_t)spinor_vec(F64) * g.vh;
  if (isign > 0)
    hipLaunchKernelGGL((k_force<1>), grid, blk, 0, s, (double*)d_grad_packed, (double*)d_force,
                       (const double*)d_u_packed, (const v4u*)d_chi, (const v4u*)d_psi, cb_vec, norm_off, g, 1 - cb);
  else
    hipLaunchKernelGGL((k_force<-1>), grid, blk, 0, s, (double*)d_grad_packed, (double*)d_force,
                       (const double*)d_u_packed, (const v4u*)d_chi, (const v4u*)d_psi, cb_vec, norm_off, g, 1 - cb);
  WD_LAUNCHED();
  return 0;
}
