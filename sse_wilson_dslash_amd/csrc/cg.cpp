// Conjugate gradient on the even-odd preconditioned Wilson normal equations,
// with an optional low-precision inner iteration and FP64 reliable updates.
//
//   M = 1 - kappa D (D = the Wilson hopping term of dslash_apply),
//   Mhat = 1 - kappa^2 D_eo D_oe on the even checkerboard,
//   A = Mhat^dagger Mhat,  solve A x = b,  stop on ||b - A x|| / ||b|| <= tol
//   with the residual recomputed in FP64.
//
// The inner ("sloppy") iteration runs in prec_sloppy: 16-bit spinors with a
// per-site norm and 12-real 16-bit links (FP32 arithmetic), or FP32, or FP64
// (plain FP64 CG).  The solution is accumulated in FP32 (FP64 for an FP64
// inner iteration) and folded into the FP64 solution at every reliable update,
// where the FP64 residual b - A x replaces the iterated one and the search
// direction is re-orthogonalised against it (the reliable-update scheme of
// Sleijpen & van der Vorst, as used by lattice QCD solvers).  All scalars live
// in device memory, so one inner iteration is a fixed kernel sequence that is
// captured once in a hipGraph and replayed; the host reads |r|^2 once per
// iteration to decide on reliable updates and convergence.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <new>

#include "blas.h"
#include "dslash_common.h"
#include "dslash_internal.h"
#include "dslash_mi355x.h"
#include "status.h"

namespace wd {

struct CgSolver {
  int latt[4];
  uint32_t vh;
  dslash_cg_params prm;
  int ps;   // sloppy format
  int Gs;   // sloppy link reals
  int px;   // solution accumulator format
  int Gh;   // link reals of the FP64 operator
  size_t g64_cb = 0, gs_cb = 0;
  void* g64 = nullptr;  // FP64 bundles (Gh reals), both target cbs
  void* gs = nullptr;   // sloppy bundles (== g64 for an FP64 inner iteration with Gs == Gh)
  bool gs_owned = false;
  // FP64 work vectors (even cb unless noted)
  void *r_hi = nullptr, *q_hi = nullptr, *Ay_hi = nullptr, *t_hi = nullptr /* odd */;
  // sloppy work vectors
  void *r = nullptr, *p = nullptr, *Ap = nullptr, *q = nullptr, *t = nullptr /* odd */, *xlo = nullptr;
  double* scal = nullptr;
  double* partial = nullptr;
  size_t nparts = 0;         // (re, im) pairs the partial buffer holds
  double* h_scal = nullptr;  // pinned
  hipStream_t stream = nullptr;
  hipGraphExec_t g_iter = nullptr, g_pupd = nullptr;
  hipGraph_t graph_iter = nullptr, graph_pupd = nullptr;
  bool have_gauge = false;
};

static int cg_alloc(void** p, size_t n) {
  WD_HIP(hipMalloc(p, n));
  return 0;
}

static void cg_free(CgSolver* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->g_iter) (void)hipGraphExecDestroy(s->g_iter);
  if (s->g_pupd) (void)hipGraphExecDestroy(s->g_pupd);
  if (s->graph_iter) (void)hipGraphDestroy(s->graph_iter);
  if (s->graph_pupd) (void)hipGraphDestroy(s->graph_pupd);
  void* ptrs[] = {s->r_hi, s->q_hi, s->Ay_hi, s->t_hi, s->r, s->p, s->Ap, s->q, s->t, s->xlo, s->scal, s->partial};
  for (void* x : ptrs)
    if (x) (void)hipFree(x);
  if (s->g64) (void)hipFree(s->g64);
  if (s->gs_owned && s->gs) (void)hipFree(s->gs);
  if (s->h_scal) (void)hipHostFree(s->h_scal);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// out = A in = Mhat^dagger Mhat in  (even cb), in format prec with links g.
// With qnorm, the per-block partials of |q|^2 = |Mhat in|^2 = <in, A in> are written there.
static int apply_normal(const CgSolver* s, void* out, const void* in, void* t_odd, void* q_even, int prec, int G,
                        const void* g, size_t gcb, double* qnorm = nullptr) {
  return apply_normal_op(out, in, t_odd, q_even, g, gcb, s->latt, s->prm.kappa, prec, G, s->stream, qnorm);
}

// One inner iteration: Ap = A p ; alpha = rr / <p,Ap> ; x += alpha p ; r -= alpha Ap ; RRNEW = |r|^2
// <p, A p> = |Mhat p|^2 comes from the epilogue of the Dslash that forms Mhat p.
static int sloppy_iteration(CgSolver* s) {
  const int nb = apply_blocks(s->latt, 4, s->ps, s->Gs);
  WD_REQUIRE(nb >= 1 && (size_t)nb <= s->nparts, "launch shape gives %d |Mhat p|^2 partial sums; the solver holds %zu",
             nb, s->nparts);
  if (int e = apply_normal(s, s->Ap, s->p, s->t, s->q, s->ps, s->Gs, s->gs, s->gs_cb, s->partial)) return e;
  if (int e = blas_finalize(s->partial, nb, s->scal, FIN_ALPHA, s->stream)) return e;
  return blas_cg_update({s->xlo, s->px}, {s->p, s->ps}, {s->r, s->ps}, {s->Ap, s->ps}, s->vh, s->partial, s->scal,
                        s->stream);
}

// beta = RRNEW / RR ; RR = RRNEW ; p = r + beta p
static int p_update(CgSolver* s) {
  if (int e = blas_scalar_op(s->scal, FIN_BETA, s->stream)) return e;
  return blas_xpay_beta({s->r, s->ps}, {s->p, s->ps}, s->vh, s->scal, s->stream);
}

static int capture(CgSolver* s, int (*body)(CgSolver*), hipGraph_t* graph, hipGraphExec_t* exec) {
  WD_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int e = body(s);
  hipGraph_t gr = nullptr;
  hipError_t he = hipStreamEndCapture(s->stream, &gr);
  if (e) {
    if (gr) (void)hipGraphDestroy(gr);
    return e;
  }
  if (he != hipSuccess) return set_error(DSLASH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(he));
  he = hipGraphInstantiate(exec, gr, nullptr, nullptr, 0);
  if (he != hipSuccess) {
    (void)hipGraphDestroy(gr);
    return set_error(DSLASH_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(he));
  }
  *graph = gr;
  return 0;
}

static int read_scalars(CgSolver* s) {
  WD_HIP(hipMemcpyAsync(s->h_scal, s->scal, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  WD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

}  // namespace wd

using namespace wd;

extern "C" {

void* dslash_cg_create(const int latt[4], const dslash_cg_params* prm) {
  clear_error();
  if (!prm) {
    set_error(DSLASH_ERR_INVALID, "params is NULL");
    return nullptr;
  }
  if (dslash_check_lattice(latt, 4)) return nullptr;
  if (!(prm->prec_sloppy == F16 || prm->prec_sloppy == BF16 || prm->prec_sloppy == F32 || prm->prec_sloppy == F64) ||
      !(prm->gauge_reals_sloppy == 18 || prm->gauge_reals_sloppy == 12) ||
      !(prm->gauge_reals_precise == 0 || prm->gauge_reals_precise == 18 || prm->gauge_reals_precise == 12) ||
      !(prm->tol > 0) || prm->max_iter <= 0) {
    set_error(DSLASH_ERR_INVALID, "invalid CG parameters (prec %d, gauge_reals %d, tol %g, max_iter %d)",
              prm->prec_sloppy, prm->gauge_reals_sloppy, prm->tol, prm->max_iter);
    return nullptr;
  }
  CgSolver* s = new (std::nothrow) CgSolver();
  if (!s) {
    set_error(DSLASH_ERR_INVALID, "out of host memory");
    return nullptr;
  }
  for (int i = 0; i < 4; ++i) s->latt[i] = latt[i];
  s->prm = *prm;
  s->vh = (uint32_t)((uint64_t)latt[0] * latt[1] * latt[2] * latt[3] / 2);
  s->ps = prm->prec_sloppy;
  s->Gs = prm->gauge_reals_sloppy;
  s->px = s->ps == F64 ? F64 : F32;
  const size_t b64 = spinor_cb_bytes(s->vh, F64), bs = spinor_cb_bytes(s->vh, s->ps), bx = spinor_cb_bytes(s->vh, s->px);
  s->Gh = prm->gauge_reals_precise == 12 ? 12 : 18;
  s->g64_cb = gauge_cb_bytes(s->vh, 4, F64, s->Gh);
  // one partial per block (>= 64 sites) or box tile (>= 64 sites), or per reduction block
  const size_t nparts = (size_t)((s->vh + 63) / 64) > (size_t)RED_BLOCKS ? (size_t)((s->vh + 63) / 64) : (size_t)RED_BLOCKS;
  s->nparts = nparts;
  s->gs_cb = gauge_cb_bytes(s->vh, 4, s->ps, s->Gs);
  bool ok = hipStreamCreate(&s->stream) == hipSuccess && !cg_alloc(&s->g64, 2 * s->g64_cb) &&
            !cg_alloc(&s->r_hi, b64) && !cg_alloc(&s->q_hi, b64) && !cg_alloc(&s->Ay_hi, b64) &&
            !cg_alloc(&s->t_hi, b64) && !cg_alloc(&s->r, bs) && !cg_alloc(&s->p, bs) && !cg_alloc(&s->Ap, bs) &&
            !cg_alloc(&s->q, bs) && !cg_alloc(&s->t, bs) && !cg_alloc(&s->xlo, bx) &&
            !cg_alloc((void**)&s->scal, S_COUNT * sizeof(double)) &&
            !cg_alloc((void**)&s->partial, 2 * nparts * sizeof(double)) &&
            hipHostMalloc((void**)&s->h_scal, S_COUNT * sizeof(double), 0) == hipSuccess;
  if (ok) {
    if (s->ps == F64 && s->Gs == s->Gh) {
      s->gs = s->g64;
    } else {
      ok = !cg_alloc(&s->gs, 2 * s->gs_cb);
      s->gs_owned = ok;
    }
  }
  if (!ok) {
    if (!dslash_last_error()) set_error(DSLASH_ERR_HIP, "CG workspace allocation failed");
    cg_free(s);
    return nullptr;
  }
  return s;
}

int dslash_cg_set_params(void* solver, const dslash_cg_params* prm) {
  clear_error();
  CgSolver* s = (CgSolver*)solver;
  WD_REQUIRE(s && prm, "NULL argument");
  WD_REQUIRE(prm->prec_sloppy == s->prm.prec_sloppy && prm->gauge_reals_sloppy == s->prm.gauge_reals_sloppy &&
                 prm->gauge_reals_precise == s->prm.gauge_reals_precise,
             "the link and spinor formats are fixed at dslash_cg_create");
  WD_REQUIRE(prm->tol > 0 && prm->max_iter > 0, "invalid tol / max_iter");
  if (prm->kappa != s->prm.kappa) {  // kappa is baked into the captured graphs
    WD_HIP(hipStreamSynchronize(s->stream));
    if (s->g_iter) (void)hipGraphExecDestroy(s->g_iter);
    if (s->g_pupd) (void)hipGraphExecDestroy(s->g_pupd);
    if (s->graph_iter) (void)hipGraphDestroy(s->graph_iter);
    if (s->graph_pupd) (void)hipGraphDestroy(s->graph_pupd);
    s->g_iter = s->g_pupd = nullptr;
    s->graph_iter = s->graph_pupd = nullptr;
  }
  s->prm = *prm;
  return 0;
}

int dslash_cg_destroy(void* solver) {
  clear_error();
  cg_free((CgSolver*)solver);
  return 0;
}

int dslash_cg_load_gauge(void* solver, const void* d_u_packed, int prec_in, void* stream) {
  clear_error();
  CgSolver* s = (CgSolver*)solver;
  WD_REQUIRE(s && d_u_packed, "NULL argument");
  // order the packing after the caller's stream
  if (stream) WD_HIP(hipStreamSynchronize((hipStream_t)stream));
  for (int t = 0; t < 2; ++t) {
    if (int e = dslash_pack_gauge((char*)s->g64 + t * s->g64_cb, t, d_u_packed, prec_in, s->latt, 4, F64, s->Gh, nullptr,
                                  s->stream))
      return e;
    if (s->gs_owned)
      if (int e = dslash_pack_gauge((char*)s->gs + t * s->gs_cb, t, d_u_packed, prec_in, s->latt, 4, s->ps, s->Gs,
                                    nullptr, s->stream))
        return e;
  }
  WD_HIP(hipStreamSynchronize(s->stream));
  s->have_gauge = true;
  return 0;
}

int dslash_cg_apply_op(void* solver, void* d_out, const void* d_in, int sloppy) {
  clear_error();
  CgSolver* s = (CgSolver*)solver;
  WD_REQUIRE(s && d_out && d_in && d_out != d_in, "bad arguments");
  WD_REQUIRE(s->have_gauge, "dslash_cg_load_gauge first");
  int e = sloppy ? apply_normal(s, d_out, d_in, s->t, s->q, s->ps, s->Gs, s->gs, s->gs_cb)
                 : apply_normal(s, d_out, d_in, s->t_hi, s->q_hi, F64, s->Gh, s->g64, s->g64_cb);
  if (e) return e;
  WD_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int dslash_cg_solve(void* solver, void* d_x, const void* d_b, dslash_cg_result* out) {
  clear_error();
  CgSolver* s = (CgSolver*)solver;
  WD_REQUIRE(s && d_x && d_b && out, "NULL argument");
  WD_REQUIRE(s->have_gauge, "dslash_cg_load_gauge first");
  memset(out, 0, sizeof(*out));
  const uint32_t vh = s->vh;
  hipStream_t st = s->stream;
  const Vec X{d_x, F64}, B{(void*)d_b, F64}, RH{s->r_hi, F64}, AY{s->Ay_hi, F64};
  const Vec R{s->r, s->ps}, P{s->p, s->ps}, XL{s->xlo, s->px};
  auto t0 = std::chrono::steady_clock::now();

  // |b|^2
  if (int e = blas_norm2(B, vh, s->partial, s->scal, FIN_TMP, st)) return e;
  if (int e = read_scalars(s)) return e;
  const double bb = s->h_scal[S_TMP];
  WD_REQUIRE(bb > 0, "source is zero");
  const double stop2 = s->prm.tol * s->prm.tol * bb;

  // r_hi = b - A x
  if (s->prm.zero_guess) {
    if (int e = blas_zero(X, vh, st)) return e;
    if (int e = blas_copy(RH, B, vh, st)) return e;
    if (int e = blas_norm2(RH, vh, s->partial, s->scal, FIN_SET_RR, st)) return e;
  } else {
    if (int e = apply_normal(s, s->Ay_hi, d_x, s->t_hi, s->q_hi, F64, s->Gh, s->g64, s->g64_cb)) return e;
    if (int e = blas_residual(RH, B, AY, vh, s->partial, s->scal, FIN_SET_RR, st)) return e;
  }
  if (int e = blas_copy(R, RH, vh, st)) return e;
  if (int e = blas_copy(P, R, vh, st)) return e;
  if (int e = blas_zero(XL, vh, st)) return e;
  if (int e = read_scalars(s)) return e;
  double rr = s->h_scal[S_RR];
  double maxr2 = rr;  // max |r|^2 since the last reliable update
  int iters = 0, updates = 0;
  bool converged = rr <= stop2;
  const bool use_graph = s->prm.use_graph != 0;
  if (use_graph && !s->g_iter) {
    if (int e = capture(s, sloppy_iteration, &s->graph_iter, &s->g_iter)) return e;
    if (int e = capture(s, p_update, &s->graph_pupd, &s->g_pupd)) return e;
  }
  const double delta = s->prm.delta;
  while (!converged && iters < s->prm.max_iter) {
    if (use_graph) {
      WD_HIP(hipGraphLaunch(s->g_iter, st));
    } else if (int e = sloppy_iteration(s)) {
      return e;
    }
    ++iters;
    if (int e = read_scalars(s)) return e;
    const double rr_new = s->h_scal[S_RRNEW];
    if (!(rr_new == rr_new)) return set_error(DSLASH_ERR_NOT_CONVERGED, "inner iteration broke down (NaN)");
    const bool want_check = rr_new <= stop2;
    const bool reliable = want_check || (delta > 0 && rr_new < delta * delta * maxr2) || iters == s->prm.max_iter;
    if (reliable) {
      // fold the low-precision solution in, recompute the true residual in FP64
      if (int e = blas_accumulate(X, XL, vh, st)) return e;
      if (int e = apply_normal(s, s->Ay_hi, d_x, s->t_hi, s->q_hi, F64, s->Gh, s->g64, s->g64_cb)) return e;
      if (int e = blas_residual(RH, B, AY, vh, s->partial, s->scal, FIN_RRNEW, st)) return e;
      if (int e = read_scalars(s)) return e;
      ++updates;
      const double rt = s->h_scal[S_RRNEW];
      if (rt <= stop2) {
        converged = true;
        rr = rt;
        break;
      }
      maxr2 = rt;
      if (int e = blas_copy(R, RH, vh, st)) return e;
      // re-orthogonalise p against the new residual, then p = r + beta p
      if (int e = blas_dot(R, P, vh, s->partial, s->scal, FIN_RP, st)) return e;
      if (int e = blas_caxpy_rp(R, P, vh, s->scal, st)) return e;
      rr = rt;
    } else {
      rr = rr_new;
      if (rr_new > maxr2) maxr2 = rr_new;
    }
    if (use_graph) {
      WD_HIP(hipGraphLaunch(s->g_pupd, st));
    } else if (int e = p_update(s)) {
      return e;
    }
  }
  // On convergence rr is the FP64 residual of x (x already holds every update);
  // otherwise x was folded in by the last reliable update and rr is its residual.
  WD_HIP(hipStreamSynchronize(st));
  auto t1 = std::chrono::steady_clock::now();
  out->iterations = iters;
  out->reliable_updates = updates;
  out->true_residual = sqrt(rr / bb);
  out->converged = out->true_residual <= s->prm.tol;
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  if (!out->converged)
    return set_error(DSLASH_ERR_NOT_CONVERGED, "CG stopped after %d iterations at true residual %.3e", iters,
                     out->true_residual);
  return 0;
}

}  // extern "C"
