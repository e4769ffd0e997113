// Checkerboard vector kernels of the solvers (see blas.h).  Every kernel reads
// and writes whole sites (24 reals) through the same plane loaders as the
// Dslash, so 16-bit fields get their per-site norm recomputed on every store.
// Reductions accumulate in double, write one partial per block, and a single
// block finalises the sum in a fixed order (deterministic, no atomics).
#include <hip/hip_runtime.h>

#include "blas.h"
#include "dslash_common.h"
#include "status.h"

namespace wd {

static constexpr int TPB = 256;

__device__ inline double block_sum(double v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += lds[i];
  __syncthreads();
  return t;
}

template <int P> struct VPtr {
  const v4u* v;
  const float* n;
};
template <int P> __device__ inline VPtr<P> vptr(const void* base, uint32_t vh) {
  VPtr<P> r;
  r.v = (const v4u*)base;
  r.n = (const float*)((const char*)base + (size_t)spinor_vec(P) * 16 * vh);
  return r;
}
template <int P> __device__ inline float* vnorm(void* base, uint32_t vh) {
  return (float*)((char*)base + (size_t)spinor_vec(P) * 16 * vh);
}

template <int P> using Sp = Spinor<typename Fmt<P>::compute>;

// ---------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(TPB) k_zero(void* y, uint32_t vh) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  Sp<P> z;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) z.s[s][c] = {0, 0};
  store_spinor<P, true>((v4u*)y, vnorm<P>(y, vh), i, vh, z);
}

template <int PY, int PX>
__global__ void __launch_bounds__(TPB) k_copy(void* y, const void* x, uint32_t vh) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  auto xp = vptr<PX>(x, vh);
  Sp<PX> a = load_spinor<PX>(xp.v, xp.n, i, vh);
  Sp<PY> o;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      o.s[s][c] = {(typename Fmt<PY>::compute)a.s[s][c].re, (typename Fmt<PY>::compute)a.s[s][c].im};
  store_spinor<PY, true>((v4u*)y, vnorm<PY>(y, vh), i, vh, o);
}

template <int P>
__global__ void __launch_bounds__(TPB) k_dot(const void* a, const void* b, uint32_t vh, double* partial) {
  __shared__ double lds[TPB / 64];
  auto ap = vptr<P>(a, vh);
  auto bp = vptr<P>(b, vh);
  double re = 0.0, im = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < vh; i += gridDim.x * blockDim.x) {
    Sp<P> x = load_spinor<P>(ap.v, ap.n, i, vh);
    Sp<P> y = load_spinor<P>(bp.v, bp.n, i, vh);
    typename Fmt<P>::compute sr = 0, si = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int c = 0; c < 3; ++c) {  // conj(x) * y
        sr += x.s[s][c].re * y.s[s][c].re + x.s[s][c].im * y.s[s][c].im;
        si += x.s[s][c].re * y.s[s][c].im - x.s[s][c].im * y.s[s][c].re;
      }
    re += (double)sr;
    im += (double)si;
  }
  re = block_sum(re, lds);
  im = block_sum(im, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = re;
    partial[2 * blockIdx.x + 1] = im;
  }
}

template <int P>
__global__ void __launch_bounds__(TPB) k_norm2(const void* a, uint32_t vh, double* partial) {
  __shared__ double lds[TPB / 64];
  auto ap = vptr<P>(a, vh);
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < vh; i += gridDim.x * blockDim.x) {
    Sp<P> x = load_spinor<P>(ap.v, ap.n, i, vh);
    typename Fmt<P>::compute s2 = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int c = 0; c < 3; ++c) s2 += x.s[s][c].re * x.s[s][c].re + x.s[s][c].im * x.s[s][c].im;
    acc += (double)s2;
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = acc;
    partial[2 * blockIdx.x + 1] = 0.0;
  }
}

// x += alpha p ; r -= alpha Ap ; partial |r|^2
template <int PX, int PS>
__global__ void __launch_bounds__(TPB) k_cg_update(void* x, const void* p, void* r, const void* Ap, uint32_t vh,
                                                   const double* scal, double* partial) {
  __shared__ double lds[TPB / 64];
  typedef typename Fmt<PS>::compute T;
  typedef typename Fmt<PX>::compute TX;
  const double alpha = scal[S_ALPHA];
  auto pp = vptr<PS>(p, vh);
  auto rp = vptr<PS>(r, vh);
  auto app = vptr<PS>(Ap, vh);
  auto xp = vptr<PX>(x, vh);
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < vh; i += gridDim.x * blockDim.x) {
    Sp<PS> pv = load_spinor<PS>(pp.v, pp.n, i, vh);
    Sp<PS> rv = load_spinor<PS>(rp.v, rp.n, i, vh);
    Sp<PS> av = load_spinor<PS>(app.v, app.n, i, vh);
    Sp<PX> xv = load_spinor<PX>(xp.v, xp.n, i, vh);
    const T a = (T)alpha;
    const TX ax = (TX)alpha;
    T s2 = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xv.s[s][c].re += ax * (TX)pv.s[s][c].re;
        xv.s[s][c].im += ax * (TX)pv.s[s][c].im;
        rv.s[s][c].re -= a * av.s[s][c].re;
        rv.s[s][c].im -= a * av.s[s][c].im;
        s2 += rv.s[s][c].re * rv.s[s][c].re + rv.s[s][c].im * rv.s[s][c].im;
      }
    store_spinor<PX, true>((v4u*)x, vnorm<PX>(x, vh), i, vh, xv);
    store_spinor<PS, true>((v4u*)r, vnorm<PS>(r, vh), i, vh, rv);
    acc += (double)s2;
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = acc;
    partial[2 * blockIdx.x + 1] = 0.0;
  }
}

// p <- r + beta p
template <int P>
__global__ void __launch_bounds__(TPB) k_xpay_beta(const void* r, void* p, uint32_t vh, const double* scal) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef typename Fmt<P>::compute T;
  const T beta = (T)scal[S_BETA];
  auto rp = vptr<P>(r, vh);
  auto pp = vptr<P>(p, vh);
  Sp<P> rv = load_spinor<P>(rp.v, rp.n, i, vh);
  Sp<P> pv = load_spinor<P>(pp.v, pp.n, i, vh);
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pv.s[s][c].re = rv.s[s][c].re + beta * pv.s[s][c].re;
      pv.s[s][c].im = rv.s[s][c].im + beta * pv.s[s][c].im;
    }
  store_spinor<P, true>((v4u*)p, vnorm<P>(p, vh), i, vh, pv);
}

// p <- p - rp * r  (complex rp)
template <int P>
__global__ void __launch_bounds__(TPB) k_caxpy_rp(const void* r, void* p, uint32_t vh, const double* scal) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef typename Fmt<P>::compute T;
  const T ar = (T)scal[S_RP_RE], ai = (T)scal[S_RP_IM];
  auto rp = vptr<P>(r, vh);
  auto pp = vptr<P>(p, vh);
  Sp<P> rv = load_spinor<P>(rp.v, rp.n, i, vh);
  Sp<P> pv = load_spinor<P>(pp.v, pp.n, i, vh);
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pv.s[s][c].re -= ar * rv.s[s][c].re - ai * rv.s[s][c].im;
      pv.s[s][c].im -= ar * rv.s[s][c].im + ai * rv.s[s][c].re;
    }
  store_spinor<P, true>((v4u*)p, vnorm<P>(p, vh), i, vh, pv);
}

// y += x ; x <- 0
template <int PY, int PX>
__global__ void __launch_bounds__(TPB) k_accumulate(void* y, void* x, uint32_t vh) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  auto yp = vptr<PY>(y, vh);
  auto xp = vptr<PX>(x, vh);
  Sp<PY> yv = load_spinor<PY>(yp.v, yp.n, i, vh);
  Sp<PX> xv = load_spinor<PX>(xp.v, xp.n, i, vh);
  Sp<PX> z;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      yv.s[s][c].re += (typename Fmt<PY>::compute)xv.s[s][c].re;
      yv.s[s][c].im += (typename Fmt<PY>::compute)xv.s[s][c].im;
      z.s[s][c] = {0, 0};
    }
  store_spinor<PY, true>((v4u*)y, vnorm<PY>(y, vh), i, vh, yv);
  store_spinor<PX, true>((v4u*)x, vnorm<PX>(x, vh), i, vh, z);
}

// out <- b - Ay ; partial |out|^2
template <int P>
__global__ void __launch_bounds__(TPB) k_residual(void* out, const void* b, const void* Ay, uint32_t vh,
                                                  double* partial) {
  __shared__ double lds[TPB / 64];
  auto bp = vptr<P>(b, vh);
  auto ap = vptr<P>(Ay, vh);
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < vh; i += gridDim.x * blockDim.x) {
    Sp<P> bv = load_spinor<P>(bp.v, bp.n, i, vh);
    Sp<P> av = load_spinor<P>(ap.v, ap.n, i, vh);
    typename Fmt<P>::compute s2 = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bv.s[s][c].re -= av.s[s][c].re;
        bv.s[s][c].im -= av.s[s][c].im;
        s2 += bv.s[s][c].re * bv.s[s][c].re + bv.s[s][c].im * bv.s[s][c].im;
      }
    store_spinor<P, true>((v4u*)out, vnorm<P>(out, vh), i, vh, bv);
    acc += (double)s2;
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = acc;
    partial[2 * blockIdx.x + 1] = 0.0;
  }
}

template <int P>
__global__ void __launch_bounds__(TPB) k_axpby(double a, const void* x, double b, void* y, uint32_t vh) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef typename Fmt<P>::compute T;
  auto xp = vptr<P>(x, vh);
  auto yp = vptr<P>(y, vh);
  Sp<P> xv = load_spinor<P>(xp.v, xp.n, i, vh);
  Sp<P> yv = load_spinor<P>(yp.v, yp.n, i, vh);
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      yv.s[s][c].re = (T)a * xv.s[s][c].re + (T)b * yv.s[s][c].re;
      yv.s[s][c].im = (T)a * xv.s[s][c].im + (T)b * yv.s[s][c].im;
    }
  store_spinor<P, true>((v4u*)y, vnorm<P>(y, vh), i, vh, yv);
}

__global__ void __launch_bounds__(1024) k_finalize(const double* partial, int nparts, double* scal, int op) {
  __shared__ double lds[1024 / 64];
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    re += partial[2 * i];
    im += partial[2 * i + 1];
  }
  re = block_sum(re, lds);
  im = block_sum(im, lds);
  if (threadIdx.x != 0) return;
  switch (op) {
    case FIN_ALPHA:
      scal[S_PAP] = re;
      scal[S_ALPHA] = scal[S_RR] / re;
      break;
    case FIN_RRNEW: scal[S_RRNEW] = re; break;
    case FIN_RP:
      scal[S_RP_RE] = re / scal[S_RRNEW];
      scal[S_RP_IM] = im / scal[S_RRNEW];
      break;
    case FIN_TMP:
      scal[S_TMP] = re;
      scal[S_TMP_IM] = im;
      break;
    case FIN_SET_RR: scal[S_RR] = re; break;
    case FIN_BETA:
      scal[S_BETA] = scal[S_RRNEW] / scal[S_RR];
      scal[S_RR] = scal[S_RRNEW];
      break;
    case FIN_COPY_RR: scal[S_RR] = scal[S_RRNEW]; break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline dim3 grid1(uint32_t vh) { return dim3((vh + TPB - 1) / TPB); }
static inline int red_blocks(uint32_t vh) {
  uint32_t nb = (vh + TPB - 1) / TPB;
  return (int)(nb < (uint32_t)RED_BLOCKS ? nb : (uint32_t)RED_BLOCKS);
}

#define WD_SWITCH_PREC(prec, MACRO)             \
  switch (prec) {                               \
    case F32: MACRO(F32); break;                \
    case F64: MACRO(F64); break;                \
    case F16: MACRO(F16); break;                \
    case BF16: MACRO(BF16); break;              \
    default: return set_error(DSLASH_ERR_INVALID, "unknown precision %d", prec); \
  }

static int finalize(double* partial, int nparts, double* scal, int op, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(nparts > 4 * TPB ? 1024 : TPB), 0, s, partial, nparts, scal, op);
  WD_LAUNCHED();
  return 0;
}

int blas_zero(Vec y, uint32_t vh, hipStream_t s) {
#define L_(P) hipLaunchKernelGGL((k_zero<P>), grid1(vh), dim3(TPB), 0, s, y.p, vh)
  WD_SWITCH_PREC(y.prec, L_)
#undef L_
  WD_LAUNCHED();
  return 0;
}

int blas_copy(Vec y, Vec x, uint32_t vh, hipStream_t s) {
#define L2_(PY, PX) hipLaunchKernelGGL((k_copy<PY, PX>), grid1(vh), dim3(TPB), 0, s, y.p, (const void*)x.p, vh)
#define L1_(PY)                           \
  switch (x.prec) {                       \
    case F32: L2_(PY, F32); break;        \
    case F64: L2_(PY, F64); break;        \
    case F16: L2_(PY, F16); break;        \
    default: L2_(PY, BF16); break;        \
  }
  WD_SWITCH_PREC(y.prec, L1_)
#undef L1_
#undef L2_
  WD_LAUNCHED();
  return 0;
}

int blas_dot(Vec a, Vec b, uint32_t vh, double* partial, double* scal, int op, hipStream_t s) {
  WD_REQUIRE(a.prec == b.prec, "dot of mixed formats");
  int nb = red_blocks(vh);
#define L_(P) hipLaunchKernelGGL((k_dot<P>), dim3(nb), dim3(TPB), 0, s, (const void*)a.p, (const void*)b.p, vh, partial)
  WD_SWITCH_PREC(a.prec, L_)
#undef L_
  WD_LAUNCHED();
  return finalize(partial, nb, scal, op, s);
}

int blas_norm2(Vec a, uint32_t vh, double* partial, double* scal, int op, hipStream_t s) {
  int nb = red_blocks(vh);
#define L_(P) hipLaunchKernelGGL((k_norm2<P>), dim3(nb), dim3(TPB), 0, s, (const void*)a.p, vh, partial)
  WD_SWITCH_PREC(a.prec, L_)
#undef L_
  WD_LAUNCHED();
  return finalize(partial, nb, scal, op, s);
}

int blas_cg_update(Vec x, Vec p, Vec r, Vec Ap, uint32_t vh, double* partial, double* scal, hipStream_t s) {
  WD_REQUIRE(p.prec == r.prec && r.prec == Ap.prec, "cg_update needs one sloppy format");
  WD_REQUIRE(x.prec == F32 || x.prec == F64, "solution accumulator must be F32 or F64");
  int nb = red_blocks(vh);
#define L2_(PX, PS)                                                                                              \
  hipLaunchKernelGGL((k_cg_update<PX, PS>), dim3(nb), dim3(TPB), 0, s, x.p, (const void*)p.p, r.p, (const void*)Ap.p, \
                     vh, (const double*)scal, partial)
#define L1_(PS)                       \
  if (x.prec == F32) {                \
    L2_(F32, PS);                     \
  } else {                            \
    L2_(F64, PS);                     \
  }
  WD_SWITCH_PREC(p.prec, L1_)
#undef L1_
#undef L2_
  WD_LAUNCHED();
  return finalize(partial, nb, scal, FIN_RRNEW, s);
}

int blas_xpay_beta(Vec r, Vec p, uint32_t vh, const double* scal, hipStream_t s) {
  WD_REQUIRE(r.prec == p.prec, "xpay of mixed formats");
#define L_(P) hipLaunchKernelGGL((k_xpay_beta<P>), grid1(vh), dim3(TPB), 0, s, (const void*)r.p, p.p, vh, scal)
  WD_SWITCH_PREC(r.prec, L_)
#undef L_
  WD_LAUNCHED();
  return 0;
}

int blas_caxpy_rp(Vec r, Vec p, uint32_t vh, const double* scal, hipStream_t s) {
  WD_REQUIRE(r.prec == p.prec, "caxpy of mixed formats");
#define L_(P) hipLaunchKernelGGL((k_caxpy_rp<P>), grid1(vh), dim3(TPB), 0, s, (const void*)r.p, p.p, vh, scal)
  WD_SWITCH_PREC(r.prec, L_)
#undef L_
  WD_LAUNCHED();
  return 0;
}

int blas_accumulate(Vec y, Vec x, uint32_t vh, hipStream_t s) {
  WD_REQUIRE(y.prec == F64 || y.prec == F32, "accumulator must be F32 or F64");
#define L2_(PY, PX) hipLaunchKernelGGL((k_accumulate<PY, PX>), grid1(vh), dim3(TPB), 0, s, y.p, x.p, vh)
#define L1_(PX)                 \
  if (y.prec == F64) {          \
    L2_(F64, PX);               \
  } else {                      \
    L2_(F32, PX);               \
  }
  WD_SWITCH_PREC(x.prec, L1_)
#undef L1_
#undef L2_
  WD_LAUNCHED();
  return 0;
}

int blas_residual(Vec out, Vec b, Vec Ay, uint32_t vh, double* partial, double* scal, int op, hipStream_t s) {
  WD_REQUIRE(out.prec == b.prec && b.prec == Ay.prec, "residual of mixed formats");
  int nb = red_blocks(vh);
#define L_(P)                                                                                                    \
  hipLaunchKernelGGL((k_residual<P>), dim3(nb), dim3(TPB), 0, s, out.p, (const void*)b.p, (const void*)Ay.p, vh, \
                     partial)
  WD_SWITCH_PREC(out.prec, L_)
#undef L_
  WD_LAUNCHED();
  return finalize(partial, nb, scal, op, s);
}

int blas_axpby(double a, Vec x, double b, Vec y, uint32_t vh, hipStream_t s) {
  WD_REQUIRE(x.prec == y.prec, "axpby of mixed formats");
#define L_(P) hipLaunchKernelGGL((k_axpby<P>), grid1(vh), dim3(TPB), 0, s, a, (const void*)x.p, b, y.p, vh)
  WD_SWITCH_PREC(x.prec, L_)
#undef L_
  WD_LAUNCHED();
  return 0;
}

int blas_scalar_op(double* scal, int op, hipStream_t s) { return finalize(nullptr, 0, scal, op, s); }

int blas_finalize(double* partial, int nparts, double* scal, int op, hipStream_t s) {
  return finalize(partial, nparts, scal, op, s);
}

}  // namespace wd
