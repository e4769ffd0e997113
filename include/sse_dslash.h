/*
 * sse_dslash.h -- drop-in 4D Wilson Dslash API, served by the MI355X GPU.
 *
 * Same entry points, signatures and semantics as the reference's
 * include/sse_dslash.h:13-36, so Chroma/QDP++ code that calls the SSE library
 * links against libsse_dslash_f32.so or libsse_dslash_f64.so unchanged.
 * Host pointers are full-lattice arrays ordered by the caller's
 * getLinearSiteIndex; the gauge is the packed [V][4][3][3][2] array of
 * transposed links (lib/qdp_packer_nopad.cc:7-19).  Each call uploads u and the
 * source checkerboard of psi, runs the fused kernel and writes back the target
 * checkerboard of res; see the *_resident / *_device extensions below for the
 * paths that keep fields in HBM.
 */
#ifndef SSE_DSLASH_H
#define SSE_DSLASH_H

#include "sse_config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces init_sse_su3dslash (include/sse_dslash.h:13-17,
 * lib/sse_su3dslash_32bit_scalar.c:84-121).  Reference-counted: a second init
 * only increments the count (the lattice size is then ignored).  Every global
 * dimension must be even, otherwise a message goes to stderr and the process
 * exits with status 1.  The callbacks give the caller's site order; nodeNumber
 * is not used on one node. */
void init_sse_su3dslash(const int* latt_size,
                        void (*getSiteCoords)(int coord[], int node, int linearsite),
                        int (*getLinearSiteIndex)(const int coord[]),
                        int (*nodeNumber)(const int coord[]));

/* Replaces free_sse_su3dslash (include/sse_dslash.h:20,
 * lib/sse_su3dslash_32bit_scalar.c:124-140): decrements the count and releases
 * tables and device buffers when it reaches zero. */
void free_sse_su3dslash(void);

/* Replaces sse_su3dslash_wilson (include/sse_dslash.h:32,
 * lib/sse_su3dslash_32bit_scalar.c:159-180).  cb is the SOURCE checkerboard:
 * res on checkerboard 1-cb is overwritten with D psi (isign = +1) or
 * D^dagger psi (isign = -1); res on checkerboard cb is not touched; psi and res
 * may be the same array.  isign other than +-1 does nothing. */
void sse_su3dslash_wilson(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb);

/* Replaces sse_su3dslash_prepost_receives (include/sse_dslash.h:36): no-op on
 * one node, as lib/sse_su3dslash_32bit_scalar.c:79-82. */
void sse_su3dslash_prepost_receives(void);

/* ---- extensions (no counterpart in the reference) ---------------------- */

/* 0 or the DSLASH_ERR_* code of the last failed call of this library. */
int sse_su3dslash_last_error(void);

/* Upload and pack the gauge once (gauge_reals = 18, or 12 for SU(3) links),
 * keeping it resident in HBM until released or free_sse_su3dslash. */
int sse_su3dslash_gauge_upload(SSEREAL* u, int gauge_reals);
int sse_su3dslash_gauge_release(void);

/* As sse_su3dslash_wilson with the resident gauge: only spinors cross PCIe. */
int sse_su3dslash_wilson_resident(SSEREAL* psi, SSEREAL* res, int isign, int cb);

/* Fully device-resident: d_psi/d_res are full-lattice fields in the device
 * layout of dslash_mi355x.h in this library's precision; stream is a
 * hipStream_t.  Asynchronous.  A later gauge upload, gauge reuse rebuild or
 * free_sse_su3dslash waits for the kernels these calls queued on `stream`
 * before it rewrites or frees the resident links. */
int sse_su3dslash_wilson_device(void* d_res, const void* d_psi, int isign, int cb, void* stream);

/* Gauge reuse for sse_su3dslash_wilson (opt-in; the reference uploads nothing
 * and so has no such switch).  mode 0 (default): the gauge is uploaded on every
 * call, as the reference reads it on every call.  mode 1: the packed links of
 * both checkerboards stay in HBM and are uploaded again only when u is a
 * different pointer or a 64-bit checksum of the whole array (computed on the
 * host threads each call) differs from the last upload's.  mode 2: reuse
 * whenever u is the same pointer; the caller promises not to change the array
 * in place, or calls sse_su3dslash_gauge_release() after changing it.  Either
 * mode replaces a gauge set with sse_su3dslash_gauge_upload. */
int sse_su3dslash_set_gauge_reuse(int mode);

/* Sloppy FP64 (FP64 library only; DSLASH_ERR_UNSUPPORTED in the FP32 one):
 * on = 1 keeps FP64 spinors at the interface but packs links in FP32 and runs
 * the kernel in FP32 arithmetic (DSLASH_F64_SLOPPY), the run-time counterpart
 * of the reference's SSEDSLASH_SLOPPY build (include/types64.h:26-31, accepted
 * at 1e-7 by tests/testDslashFull.cc:50-53).  A library built with
 * -DSSEDSLASH_SLOPPY starts in this mode.  Switching drops the resident gauge
 * (upload it again). */
int sse_su3dslash_set_sloppy(int on);

/* CPU build (libsse_dslash_cpu_f32/f64.so): the OpenMP threads a call uses
 * (OMP_NUM_THREADS or the runtime's default); the GPU build returns 0. */
int sse_su3dslash_cpu_threads(void);

/* The reference's C gauge packer (lib/packer_nopad.c:29-45, declared in no
 * reference header): reorders [mu][cb][Vh] links into the packed
 * [cb][Vh][mu] order, without transposing.  volume = V/2. */
typedef SSEREAL sse_u_mat_array[3][3][2];
void pack_gauge_field(int volume, sse_u_mat_array* u, sse_u_mat_array* u_tmp);
/* The reference's unpack_gauge_field (lib/packer_nopad.c:47-69), byte for byte:
 * it copies u_tmp to u position by position in [mu][cb][volume] order, so u
 * becomes a copy of u_tmp (it does NOT undo pack_gauge_field). */
void unpack_gauge_field(int volume, sse_u_mat_array* u_tmp, sse_u_mat_array* u);
/* Extension (no counterpart in the reference): the true inverse of
 * pack_gauge_field, [cb][volume][mu] back to [mu][cb][volume]. */
void unpack_gauge_field_inverse(int volume, sse_u_mat_array* u_tmp, sse_u_mat_array* u);

#ifdef __cplusplus
}
#endif

#endif /* SSE_DSLASH_H */
