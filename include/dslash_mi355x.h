/*
 * dslash_mi355x.h -- device-resident C ABI of the MI355X Wilson Dslash.
 *
 * The legacy drop-in entry points (init_sse_su3dslash, sse_su3dslash_wilson,
 * ...) are declared in sse_dslash.h / sse_dslash_3d.h and live in
 * libsse_dslash_f32.so / libsse_dslash_f64.so.  This header is the layer those
 * libraries are built on, exported by libdslash_mi355x.so.  It works on fields
 * that already live in HBM, in the library's own layout:
 *
 *   spinor, one checkerboard:  [nvec][Vh] 16-byte vectors, the 24 reals of a
 *     site in (spin, colour, re/im) order; nvec = 6 (F32), 12 (F64,
 *     F64_SLOPPY), 3 (F16, BF16).  16-bit formats store x / max|x| per site and append one float
 *     norm per site after the vectors.  Size: dslash_spinor_cb_bytes().
 *     A full-lattice field is checkerboard 0 followed by checkerboard 1, each
 *     dslash_spinor_cb_bytes() long.
 *   gauge, one TARGET checkerboard:  for every target site x and direction mu
 *     the pair (U_mu(x), U_mu(x-mu)^dagger) as 2*G reals (G = 18 full matrix,
 *     G = 12 first two rows, third rebuilt as conj(row0 x row1)), stored as
 *     [nd][pairvec][Vh] 16-byte vectors.  Size: dslash_gauge_cb_bytes().
 *     A full gauge object is the bundle of target cb 0 followed by target cb 1.
 *
 * Site order within a checkerboard is the reference's canonical order:
 * 4D cb2 (x fastest, lib/shift_tables_scalar.c:99-123) and 3D cb3d (t fastest,
 * lib/shift_tables_scalar_3d.c:97-137).  The pack/unpack entry points convert
 * from/to the legacy host layout (full lattice, [V][4][3][3][2] transposed links
 * and [V][4][3][2] spinors, ordered by the caller's getLinearSiteIndex).
 *
 * All functions return 0 on success or a DSLASH_ERR_* code; the message of the
 * last failure on the calling thread is dslash_last_error_message().  `stream`
 * is a hipStream_t passed as void* (NULL = default stream).  Launches are
 * asynchronous with respect to the host unless stated otherwise.
 */
#ifndef DSLASH_MI355X_H
#define DSLASH_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Storage formats */
#define DSLASH_F32 0
#define DSLASH_F64 1
#define DSLASH_F16 2  /* binary16 spinors/links, per-site float norm, FP32 math */
#define DSLASH_BF16 3 /* bfloat16 spinors/links, per-site float norm, FP32 math */
/* FP64 spinors (stored exactly as DSLASH_F64) with FP32 link bundles and FP32
 * arithmetic: the device analogue of the reference's SSEDSLASH_SLOPPY FP64
 * build, whose half-spinor temporaries are float (include/types64.h:26-31;
 * accepted at 1e-7 by tests/testDslashFull.cc:50-53).  Gauge bundles of this
 * format are dslash_gauge_cb_bytes(..., DSLASH_F64_SLOPPY, G) = the F32 size. */
#define DSLASH_F64_SLOPPY 4

/* Status codes */
#define DSLASH_OK 0
#define DSLASH_ERR_INVALID 1     /* bad argument (shape, precision, isign, cb) */
#define DSLASH_ERR_HIP 2         /* a HIP runtime call or launch failed */
#define DSLASH_ERR_UNSUPPORTED 3 /* valid but not built (e.g. combination) */
#define DSLASH_ERR_NOT_INIT 4    /* legacy API used before init */
#define DSLASH_ERR_NOT_CONVERGED 5

/* ---- status ---------------------------------------------------------- */
int dslash_last_error(void);
const char* dslash_last_error_message(void);
int dslash_version(void);

/* ---- sizes (host only, no device access) ------------------------------ */
/* nd = 4 for the 4D operator, 3 for the spatial (3D) operator. */
size_t dslash_spinor_cb_bytes(const int latt[4], int nd, int prec);
size_t dslash_gauge_cb_bytes(const int latt[4], int nd, int prec, int gauge_reals);
/* 0 if latt is valid for nd (positive; dims 0..nd-1 even), else DSLASH_ERR_INVALID */
int dslash_check_lattice(const int latt[4], int nd);

/* ---- layout conversion (device pointers) ------------------------------ */
/* Legacy packed gauge [V][4][3][3][2] (prec_in F32/F64, caller order given by
 * site_table: my full index -> caller linear index; NULL = canonical order)
 * -> link bundle of one target checkerboard in prec_out / gauge_reals. */
int dslash_pack_gauge(void* d_bundle, int target_cb, const void* d_u_packed, int prec_in,
                      const int latt[4], int nd, int prec_out, int gauge_reals,
                      const int* d_site_table, void* stream);
/* Legacy full-lattice spinor [V][4][3][2] -> one checkerboard in device layout. */
int dslash_pack_spinor(void* d_cb, int cb, const void* d_aos, int prec_in,
                       const int latt[4], int nd, int prec_out,
                       const int* d_site_table, void* stream);
/* One checkerboard in device layout -> legacy AoS.  compact = 0 writes the
 * sites at their caller positions in a full-lattice array; compact = 1 writes
 * them densely in checkerboard order ([Vh][4][3][2]). */
int dslash_unpack_spinor(void* d_aos, int cb, const void* d_cb, int prec_cb,
                         const int latt[4], int nd, int prec_out,
                         const int* d_site_table, int compact, void* stream);
/* The same for the sites [site_begin, site_end) of the checkerboard only (site
 * index within the checkerboard, canonical order; other sites untouched), so a
 * caller can overlap transfers and work slab by slab. */
int dslash_pack_spinor_sites(void* d_cb, int cb, const void* d_aos, int prec_in, const int latt[4], int nd,
                             int prec_out, const int* d_site_table, size_t site_begin, size_t site_end,
                             void* stream);
int dslash_unpack_spinor_sites(void* d_aos, int cb, const void* d_cb, int prec_cb, const int latt[4], int nd,
                               int prec_out, const int* d_site_table, int compact, size_t site_begin,
                               size_t site_end, void* stream);
/* Precision conversion of one checkerboard (either direction, any formats). */
int dslash_convert_spinor(void* d_out, int prec_out, const void* d_in, int prec_in,
                          const int latt[4], int nd, void* stream);

/* ---- the hot path ----------------------------------------------------- */
/* res_cb(target) = D psi_cb(source = 1 - target_cb): the Wilson hopping term
 *   sum_mu U_mu(x)(1 - isign gamma_mu) psi(x+mu) + U_mu(x-mu)^+ (1 + isign gamma_mu) psi(x-mu)
 * over mu < nd (replaces sse_su3dslash_wilson's site loop,
 * lib/sse_su3dslash_32bit_scalar.c:185-385).  psi_cb and res_cb are single
 * checkerboards; gauge_cb is the bundle of target_cb.  Target sites are
 * overwritten.  isign must be +1 or -1. */
int dslash_apply(void* d_res_cb, const void* d_psi_cb, const void* d_gauge_cb,
                 const int latt[4], int nd, int prec, int gauge_reals,
                 int isign, int target_cb, void* stream);

/* dslash_apply on the target sites [site_begin, site_end) of the checkerboard
 * only (canonical order, one site per thread; other target sites are not
 * written).  In the canonical orders the slowest coordinate (t in 4D, z in 3D)
 * is outermost, so a range of whole slices needs only the source slices next
 * to it: the drop-in API uses this to compute slab k while slab k + 1 is
 * still arriving over PCIe. */
int dslash_apply_sites(void* d_res_cb, const void* d_psi_cb, const void* d_gauge_cb, const int latt[4], int nd,
                       int prec, int gauge_reals, int isign, int target_cb, size_t site_begin, size_t site_end,
                       void* stream);

/* 4D only: res_cb = x_cb + a * D psi_cb in one pass (x_cb may alias res_cb).
 * With a = -kappa this is the Wilson matrix M = 1 - kappa D between
 * checkerboards; the solvers use it for 1 - kappa^2 D_eo D_oe. */
int dslash_apply_xpay(void* d_res_cb, const void* d_psi_cb, const void* d_gauge_cb,
                      const void* d_x_cb, double a, const int latt[4], int prec,
                      int gauge_reals, int isign, int target_cb, void* stream);

/* Full-lattice form with the legacy meaning of cb (= SOURCE checkerboard):
 * reads psi's cb half, overwrites res's 1-cb half, leaves res's cb half
 * untouched (lib/sse_su3dslash_32bit_scalar.c:159-180).  d_gauge holds the
 * bundles of both target checkerboards. */
int dslash_apply_full(void* d_res, const void* d_psi, const void* d_gauge,
                      const int latt[4], int nd, int prec, int gauge_reals,
                      int isign, int cb, void* stream);

/* Launch shape of the single-RHS kernels for (nd, prec, gauge_reals): threads
 * per block (64, 128, 192 or 256) and whether blocks are remapped so each XCD
 * walks a contiguous range of sites (1) or not (0).  nd = 0 sets every entry.
 * dslash_tune_apply (below) measures and sets the best entry. */
int dslash_set_launch(int nd, int prec, int gauge_reals, int block_threads, int xcd_remap);
int dslash_get_launch(int nd, int prec, int gauge_reals, int* block_threads, int* xcd_remap);
/* Marching length of res = D psi for (nd, prec, gauge_reals): 1 = one target
 * site per thread; 4 or 8 = a thread walks that many consecutive slices of the
 * slowest coordinate of the site order (t in 4D, z in 3D) and keeps the
 * projected source spinor of its forward hop for the backward hop two slices
 * later; 16 = LDS-staged box tiles (4D FP32, FP64, sloppy FP64): a block of block_threads
 * (>= 128, >= Lx/2) holds whole x rows of a ty x tz box of (y, z), walks 8
 * t-slices, and stages each slice's source spinors in LDS for the x hops and
 * the y / z hops inside the box (links non-temporal); 32 = the same box tiles
 * (4D FP32, sloppy FP64) whose y / z hops that leave the box read an LDS halo: each thread
 * loads one neighbouring site of the next slice while the current one is
 * computed and stores it projected for its hop.  Any lattice extent is
 * allowed (the last chunk is shorter).  A launch that would have fewer than
 * 131072 threads at this length (small lattices, short t or z) uses the next
 * shorter one, down to 1; box tiles that cannot run (or would hold fewer than
 * 64 sites) fall back to 8, and 32 falls back to 16 for other formats or where
 * the halo would need more than one entry per thread.  The x + a D psi forms
 * run in box tiles under 16 and 32 and one site per thread otherwise.
 * dslash_set_launch leaves it unchanged; dslash_tune_apply tunes it with the
 * block shape. */
int dslash_set_launch_march(int nd, int prec, int gauge_reals, int march);
int dslash_get_launch_march(int nd, int prec, int gauge_reals, int* march);
/* The marching length res = D psi actually runs with on this lattice (the set
 * one, or shorter where it would leave fewer than 131072 threads or box tiles
 * cannot run). */
int dslash_launch_march_for(const int latt[4], int nd, int prec, int gauge_reals, int* march);

/* ---- many right-hand sides, many lattices ----------------------------- */
/* One lattice of a batched launch: nrhs single-checkerboard fields on one
 * gauge, field k at d_psi + k*field_stride bytes (likewise d_res). */
typedef struct {
  int latt[4];
  const void* d_gauge_cb; /* link bundle of target_cb */
  const void* d_psi;      /* nrhs source fields (checkerboard 1 - target_cb) */
  void* d_res;            /* nrhs target fields, must not overlap d_psi */
  size_t field_stride;    /* bytes, multiple of 16, >= dslash_spinor_cb_bytes */
  int nrhs;
  int isign;
  int target_cb;
} dslash_mrhs_problem;

/* 4D Dslash of every field of every problem (1..8 problems, F32 or F64) in ONE
 * launch.  A thread owns one site and rhs_block (1, 2 or 4) fields and reuses
 * each link for all of them.  grid_blocks = 0 launches one block per tile;
 * grid_blocks > 0 launches a persistent grid of that many blocks that strides
 * over the tiles of all lattices.  rhs_block / block_threads = 0 take the
 * tuned (or default) values; so does the tile shape (t_block below). */
int dslash_apply_mrhs(const dslash_mrhs_problem* problems, int nprob, int prec, int gauge_reals,
                      int rhs_block, int block_threads, int grid_blocks, void* stream);
/* The batched launch shape used for (prec, gauge_reals, persistent): rhs_block,
 * block_threads, grid_blocks (0 unless persistent) and t_block, the tile shape:
 *   1  one site per thread (any rhs_block);
 *   4  each thread walks 4 consecutive t-slices of one spatial site and keeps the
 *      projected t-neighbour half-spinor for the next-but-one slice (rhs_block 1, 2);
 *   8  a 128-thread block walks 8 slices of 128 spatial sites together and stages
 *      each slice's spinors in a double-buffered LDS slab for the x and y hops
 *      (F32, rhs_block 1, 2, block_threads 128; launches whose problems share one
 *      isign and have nrhs divisible by rhs_block; others fall back to 4);
 *  16  the same with box tiles: a block holds whole x rows of a ty x tz box of
 *      (y, z) (chosen per lattice to fit block_threads), so x hops and the y / z
 *      hops inside the box read the LDS slab (F32, rhs_block 1, 2, block_threads
 *      >= 128 and >= Lx/2 of every lattice, same launch conditions as 8; others
 *      fall back to 8, then 4).  With rhs_block 1, block_threads may go up to 768
 *      (wide tiles: one block per CU, larger boxes, fewer hops leaving the box;
 *      all fields of a box run back to back); a launch that cannot use box tiles
 *      then runs its fallback with 256 threads.
 * set: same arguments; the autotuner also sets them. */
int dslash_get_mrhs_launch(int prec, int gauge_reals, int persistent, int* rhs_block, int* block_threads,
                           int* grid_blocks, int* t_block);
int dslash_set_mrhs_launch(int prec, int gauge_reals, int persistent, int rhs_block, int block_threads,
                           int grid_blocks, int t_block);

/* ---- autotuning -------------------------------------------------------- */
/* Time every launch shape of the single-RHS kernel (block size x XCD remap x
 * marching length) on
 * the given full-lattice fields (dslash_apply_full, alternating isign and the
 * source checkerboard; both halves of res are overwritten), keep the fastest
 * for (nd, prec, gauge_reals) and report it. */
int dslash_tune_apply(void* d_res, const void* d_psi, const void* d_gauge, const int latt[4], int nd, int prec,
                      int gauge_reals, void* stream, int* block_threads, int* xcd_remap, float* best_ms);
/* Time rhs_block x block_threads x t_block (x persistent grid sizes when
 * persistent != 0) for a batched launch; keeps the fastest for (prec,
 * gauge_reals, persistent) (dslash_get_mrhs_launch reports t_block). */
int dslash_tune_mrhs(const dslash_mrhs_problem* problems, int nprob, int prec, int gauge_reals, int persistent,
                     void* stream, int* rhs_block, int* block_threads, int* grid_blocks, float* best_ms);
/* Tuned choices as text lines; load applies them (missing file = DSLASH_ERR_INVALID). */
int dslash_tune_save(const char* path);
int dslash_tune_load(const char* path);

/* ---- even-odd preconditioned Wilson CG (mixed precision) -------------- */
/* Solves A x = b on the even checkerboard, A = Mhat^dagger Mhat,
 * Mhat = 1 - kappa^2 D_eo D_oe, M = 1 - kappa D.  x and b are even-cb F64
 * fields in the device layout.  The inner iteration runs in prec_sloppy
 * (F16/BF16 spinors with per-site norms and 16-bit links, FP32 math; or F32;
 * or F64 = plain FP64 CG); the residual is recomputed in FP64 at reliable
 * updates and convergence is declared on that FP64 true residual. */
typedef struct {
  double kappa;           /* hopping parameter of M = 1 - kappa D */
  double tol;             /* stop when ||b - A x|| / ||b|| <= tol (FP64) */
  double delta;           /* reliable update when |r| < delta * max|r| since the last one;
                             <= 0: only to confirm convergence */
  int max_iter;           /* inner iterations */
  int prec_sloppy;        /* DSLASH_F16, DSLASH_BF16, DSLASH_F32 or DSLASH_F64 */
  int gauge_reals_sloppy; /* 18 or 12 */
  int use_graph;          /* replay the inner iteration as a captured hipGraph */
  int zero_guess;         /* start from x = 0 instead of the contents of x */
  int gauge_reals_precise;/* links of the FP64 operator (reliable updates, true
                             residual): 18 (or 0) = full matrices, 12 = two rows */
} dslash_cg_params;

typedef struct {
  int iterations;         /* inner iterations (one A application each) */
  int reliable_updates;   /* FP64 residual recomputations */
  int converged;
  double true_residual;   /* ||b - A x|| / ||b|| of the returned x, in FP64 */
  double seconds;         /* wall time of the solve, device synchronised */
} dslash_cg_result;

/* Allocates the workspace and link storage; NULL on failure (see last error). */
void* dslash_cg_create(const int latt[4], const dslash_cg_params* params);
/* Packs a legacy-layout gauge (device pointer, F32/F64, canonical cb2 order)
 * into FP64 18-real links and the sloppy links. */
int dslash_cg_load_gauge(void* solver, const void* d_u_packed, int prec_in, void* stream);
/* Synchronous.  Returns DSLASH_ERR_NOT_CONVERGED if tol was not reached. */
int dslash_cg_solve(void* solver, void* d_x, const void* d_b, dslash_cg_result* result);
/* out = A in (even cb): FP64 operator (sloppy = 0) or the sloppy one (fields in
 * the sloppy format).  Synchronous. */
int dslash_cg_apply_op(void* solver, void* d_out, const void* d_in, int sloppy);
/* Change kappa, tol, delta, max_iter, use_graph, zero_guess (not the formats). */
int dslash_cg_set_params(void* solver, const dslash_cg_params* params);
int dslash_cg_destroy(void* solver);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): out = A in with A = Mhat^dagger Mhat, Mhat = 1 - kappa^2
 * D_eo D_oe, on the even checkerboard -- what dslash_cg_apply_op computes with
 * a solver's workspace, here with no solver.  d_in / d_out: even-cb fields in
 * format prec (DSLASH_F32, F64, F16, BF16 or F64_SLOPPY); d_gauge: the bundles
 * of target cb 0 then 1 in that format (dslash_pack_gauge).  The odd-cb and
 * intermediate fields live in one fixed 64 MiB device array of this library
 * (each GPU it runs on holds its own copy until the process ends), so Vh is at
 * most ~174k sites in FP64 (e.g. 16^3 x 32, 24^4) and ~349k in FP32 (24^3 x
 * 48); larger lattices return DSLASH_ERR_INVALID.  Calls no hipMalloc/hipFree.
 * Asynchronous on `stream`.  All calls on one GPU share the array: calls that
 * overlap in time on one GPU (different streams, or host threads) race; calls
 * on one stream are ordered.  The first call on each device resolves the
 * array's address (hipGetSymbolAddress); later calls may be captured. */
int dslash_normal_op_noalloc(void* d_out, const void* d_in, const void* d_gauge, const int latt[4], double kappa,
                             int prec, int gauge_reals, void* stream);

/* ---- backward (gauge force) -------------------------------------------- */
/* Link gradient of S = Re <chi, D psi>, D from source checkerboard cb to 1-cb
 * with the given isign (4D, FP64).  chi, psi: full-lattice F64 device fields
 * (chi's target half and psi's source half are read).  Outputs, each
 * [V][4][3][3][2] doubles in canonical cb2 order (either may be NULL):
 *   d_grad_packed  dS / d(real entries of the packed, transposed link array);
 *   d_force        the su(3) projection F = TA((U G)^dagger) (row-major), with
 *                  dS = Re tr(X^dagger F) for dU = X U, X in su(3); needs
 *                  d_u_packed (F64 packed links, canonical order).
 * The three link arrays must be 16-byte aligned (as any hipMalloc pointer is).
 * The gradient with respect to psi is D^dagger chi (dslash_apply with -isign). */
int dslash_force(void* d_grad_packed, void* d_force, const void* d_u_packed, const void* d_chi,
                 const void* d_psi, const int latt[4], int isign, int cb, void* stream);

/* ---- domain-decomposition stages (the parscalar split) ------------------ */
/* The reference's multi-node Dslash (lib/sse_su3dslash_32bit_parscalar.c:108-820,
 * driven at :1185-1392) splits D into decomp / decomp_hvv (project the source
 * spinor, and for backward hops multiply by U^dagger at the source site),
 * an exchange of boundary half-spinors, and mvv_recons / recons on the
 * target.  These entry points are that split for the directions set in
 * dir_mask (bit mu = direction mu is partitioned; nd = 3 allows bits 0..2):
 *   D psi = halo_interior + halo_exterior(exchange(halo_pack)).
 * Fields are full-lattice device fields in the layout above, d_gauge is the
 * full gauge object (both target-cb bundles), cb is the SOURCE checkerboard
 * as in sse_su3dslash_wilson, prec F32 or F64.  On one GPU the exchange is the
 * identity (pass the send buffer as recv: the periodic wrap); across nodes
 * recv face (mu, forward) is the +mu neighbour's send face (mu, forward) and
 * recv face (mu, backward) the -mu neighbour's send face (mu, backward).
 *
 * Face buffer: per partitioned mu, the forward face (source sites on x_mu = 0,
 * projected for the -mu neighbour's forward hops, 12 reals/site) then the
 * backward face (source sites on x_mu = L_mu - 1, projected and multiplied by
 * U_mu^dagger, for the +mu neighbour's backward hops); each face holds
 * Vh / L_mu sites as [12 reals][site] planes.  offsets[2*mu] / offsets[2*mu+1]
 * give the byte offsets of the forward / backward face of mu, face_bytes[mu]
 * the size of one face, total_bytes the buffer size (any output may be NULL). */
int dslash_face_layout(const int latt[4], int nd, int prec, int dir_mask, size_t offsets[8],
                       size_t face_bytes[4], size_t* total_bytes);
/* decomp + decomp_hvv of the boundary source sites into d_send. */
int dslash_halo_pack(void* d_send, const void* d_psi, const void* d_gauge, const int latt[4], int nd, int prec,
                     int gauge_reals, int isign, int cb, int dir_mask, void* stream);
/* res (target cb 1-cb) = every hop of D that stays inside the local lattice
 * (overwrites, like dslash_apply; dir_mask = 0 is the whole operator). */
int dslash_halo_interior(void* d_res, const void* d_psi, const void* d_gauge, const int latt[4], int nd, int prec,
                         int gauge_reals, int isign, int cb, int dir_mask, void* stream);
/* res (target cb) += the boundary hops from d_recv: U_mu(x) h for forward
 * faces (mvv_recons), h for backward faces (recons); directions in increasing
 * mu, launched in order on the stream. */
int dslash_halo_exterior(void* d_res, const void* d_recv, const void* d_gauge, const int latt[4], int nd, int prec,
                         int gauge_reals, int isign, int cb, int dir_mask, void* stream);

/* ---- synthetic fields ------------------------------------------------- */
/* Fill a legacy-layout gauge [V][4][3][3][2] (prec F32/F64, canonical order of
 * nd) with SU(3) links: complex Gaussian 3x3 reunitarised by Gram-Schmidt,
 * stored transposed as the reference's packer does
 * (lib/qdp_packer_nopad.cc:7-19).  The link of a site depends only on
 * (seed, lexicographic site, mu). */
int dslash_gen_gauge_host(void* u_packed, int prec, const int latt[4], int nd, uint64_t seed);
int dslash_gen_gauge_device(void* d_u_packed, int prec, const int latt[4], int nd, uint64_t seed, void* stream);
/* Fill a legacy-layout full-lattice spinor with i.i.d. N(0,1) real/imag parts. */
int dslash_gen_spinor_host(void* psi, int prec, const int latt[4], int nd, uint64_t seed, int field);
int dslash_gen_spinor_device(void* d_psi, int prec, const int latt[4], int nd, uint64_t seed, int field, void* stream);

/* The raw Philox4x32-10 block function behind the generators (exposed so its
 * known-answer vectors can be checked on the host). */
void dslash_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- canonical layout callbacks --------------------------------------- */
/* What QDP++'s Layout supplies to init_sse_su3dslash in cb2 (nd = 4) or cb3d
 * (nd = 3) mode, for callers without QDP++: set the lattice once, then pass
 * these three functions as getSiteCoords / getLinearSiteIndex / nodeNumber.
 * Process-global state, like QDP++'s Layout. */
int dslash_layout_set(const int latt[4], int nd);
int dslash_layout_linear_index(const int coord[]);
void dslash_layout_site_coords(int coord[], int node, int linearsite);
int dslash_layout_node_number(const int coord[]);

/* ---- timing ----------------------------------------------------------- */
/* Launch dslash_apply n times back to back on `stream` (alternating isign and
 * target checkerboard when alternate != 0) bracketed by HIP events recorded on
 * that same stream; returns the average milliseconds per launch in *ms. */
int dslash_time_apply(void* d_res_full, const void* d_psi_full, const void* d_gauge_full,
                      const int latt[4], int nd, int prec, int gauge_reals,
                      int n, int alternate, void* stream, float* ms);

/* Streaming ceiling of the 4D FP32 operator's traffic: n launches of a kernel
 * that, per target site, reads 6 planes of d_in and 36 planes of d_links
 * (non-temporally) and writes 6 planes of d_out, all at the site's own index
 * (no gathers), in blocks of block_threads (64..256) with (xcd_remap = 1) or
 * without the XCD-contiguous block order; 768 bytes per site as the Dslash's
 * compulsory traffic.  Buffers: d_in and d_out >= 96 * Vh bytes (an FP32
 * checkerboard), d_links >= 576 * Vh bytes (an FP32 18-real bundle).  Returns
 * the average milliseconds per launch (HIP events on `stream`). */
int dslash_time_ceiling(void* d_out, const void* d_in, const void* d_links, const int latt[4], int block_threads,
                        int xcd_remap, int n, void* stream, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* DSLASH_MI355X_H */
