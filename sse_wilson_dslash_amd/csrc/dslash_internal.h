// Internal (C++) entry points shared between the translation units of
// libdslash_mi355x.so.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace wd {

// Launch shape of the single-RHS Dslash kernels for one (nd, format, link reals).
struct LaunchCfg {
  int block;  // threads per block
  int remap;  // XCD-contiguous block order
  int march;  // slices of the slowest coordinate per thread (1, 4 or 8; the x + a D psi
              // forms take 1), 16 = LDS-staged box tiles (4D FP32, FP64, sloppy
              // FP64; every form) or 32 = box tiles whose out-of-box y / z hops read
              // an LDS halo (4D FP32 and sloppy FP64; else 16)
};
constexpr int MARCH_BOX = 16, MARCH_BOX_HALO = 32;
inline bool march_ok(int m) { return m == 1 || m == 4 || m == 8 || m == MARCH_BOX || m == MARCH_BOX_HALO; }
LaunchCfg launch_cfg(int nd, int prec, int G);
void set_launch_cfg(int nd, int prec, int G, LaunchCfg c);

// res_cb = D psi_cb (xp_cb == nullptr) or res_cb = xp_cb + a * D psi_cb, all on
// one checkerboard in the device layout of `prec`; xp_cb may alias res_cb.
// With norm_partial (xpay form only) the kernel also writes one (|res|^2, 0)
// pair per block (one site per thread) or per tile (box tiles);
// apply_blocks() gives their number for the current launch shape (at most
// ceil(Vh / 64)).
int apply_cb_ex(void* res_cb, const void* psi_cb, const void* gauge_cb, const int latt[4], int nd, int prec, int G,
                int isign, int tcb, const void* xp_cb, double a, hipStream_t s, double* norm_partial = nullptr);
int apply_blocks(const int latt[4], int nd, int prec, int G);

// out = Mhat^dagger Mhat in on the even checkerboard (normal_op.hip); t_odd and
// q_even are work fields of format prec, g the bundles of target cb 0 and 1.
int apply_normal_op(void* out, const void* in, void* t_odd, void* q_even, const void* g, size_t gcb,
                    const int latt[4], double kappa, int prec, int G, hipStream_t st, double* qnorm);

}  // namespace wd
