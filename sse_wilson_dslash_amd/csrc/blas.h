// Checkerboard vector kernels used by the solvers: reductions with
// device-resident scalars (so a solver iteration can be captured in a hipGraph
// and replayed without host synchronisation), fused CG updates and precision
// conversions.  Host launchers are in namespace wd; all work on one
// checkerboard of Vh sites in the device layout of a storage format.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wd {

// Slots of the device scalar array used by the CG kernels.
enum ScalarSlot : int {
  S_RR = 0,      // |r|^2 of the current iterate
  S_PAP = 1,     // <p, A p>
  S_ALPHA = 2,   // rr / pAp
  S_RRNEW = 3,   // |r|^2 after the update
  S_BETA = 4,    // rr_new / rr
  S_RP_RE = 5,   // <r, p> / rr_new (real, imaginary)
  S_RP_IM = 6,
  S_TMP = 7,     // generic result
  S_TMP_IM = 8,
  S_COUNT = 16
};

// Finalisation ops of a reduction (sum of per-block partials, then a scalar step).
enum FinOp : int {
  FIN_ALPHA = 0,   // PAP = sum;  ALPHA = RR / PAP
  FIN_RRNEW = 1,   // RRNEW = sum
  FIN_RP = 2,      // RP = sum / RRNEW (complex)
  FIN_TMP = 3,     // TMP = sum (complex)
  FIN_SET_RR = 4,  // RR = sum
  FIN_BETA = 5,    // no sum: BETA = RRNEW / RR; RR = RRNEW
  FIN_COPY_RR = 6  // no sum: RR = RRNEW
};

constexpr int RED_BLOCKS = 1024;  // partial sums per reduction (<= RED_BLOCKS * 2 doubles)

struct Vec {
  void* p;    // planes (+ norm plane for 16-bit formats)
  int prec;
};

// y <- 0
int blas_zero(Vec y, uint32_t vh, hipStream_t s);
// y <- x (any formats)
int blas_copy(Vec y, Vec x, uint32_t vh, hipStream_t s);
// partial[] <- per-block sums of <a, b> (complex) ; then finalize(op)
int blas_dot(Vec a, Vec b, uint32_t vh, double* partial, double* scal, int op, hipStream_t s);
// partial[] <- per-block |a|^2 ; then finalize(op)
int blas_norm2(Vec a, uint32_t vh, double* partial, double* scal, int op, hipStream_t s);
// x += alpha p ; r -= alpha Ap ; RRNEW = |r|^2   (alpha = scal[S_ALPHA])
int blas_cg_update(Vec x, Vec p, Vec r, Vec Ap, uint32_t vh, double* partial, double* scal, hipStream_t s);
// p <- r + beta p   (beta = scal[S_BETA])
int blas_xpay_beta(Vec r, Vec p, uint32_t vh, const double* scal, hipStream_t s);
// p <- p - rp r     (rp = scal[S_RP_RE] + i scal[S_RP_IM])
int blas_caxpy_rp(Vec r, Vec p, uint32_t vh, const double* scal, hipStream_t s);
// y += x ; x <- 0
int blas_accumulate(Vec y, Vec x, uint32_t vh, hipStream_t s);
// out <- b - Ay ; RRNEW-style finalize op on |out|^2
int blas_residual(Vec out, Vec b, Vec Ay, uint32_t vh, double* partial, double* scal, int op, hipStream_t s);
// y <- a*x + b*y with host scalars
int blas_axpby(double a, Vec x, double b, Vec y, uint32_t vh, hipStream_t s);
// finalize only (FIN_BETA / FIN_COPY_RR)
int blas_scalar_op(double* scal, int op, hipStream_t s);
// sum nparts (re, im) partials written by another kernel, then apply op
int blas_finalize(double* partial, int nparts, double* scal, int op, hipStream_t s);

}  // namespace wd
