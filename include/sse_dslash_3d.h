/*
 * sse_dslash_3d.h -- drop-in spatial (3D) Wilson Dslash API on the MI355X.
 *
 * Same entry points as the reference's include/sse_dslash_3d.h:13-33: the sum
 * runs over mu = 0,1,2 only, the checkerboard is cb3 = (x+y+z)&1 and the
 * canonical order is cb3d with t fastest (lib/shift_tables_scalar_3d.c:97-137).
 * The packed gauge still carries 4 directions per site
 * (lib/qdp_packer_3d.cc:11-25); direction 3 is not read.
 */
#ifndef SSE_DSLASH3D_H
#define SSE_DSLASH3D_H

#include "sse_config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces init_sse_su3dslash_3d (include/sse_dslash_3d.h:13-17,
 * lib/sse_su3dslash_32bit_scalar_3d.c:79-118).  Dimensions 0..2 must be even
 * (stderr message and exit(1) otherwise); refcounted like the 4D version. */
void init_sse_su3dslash_3d(const int* latt_size,
                           void (*getSiteCoords)(int coord[], int node, int linearsite),
                           int (*getLinearSiteIndex)(const int coord[]),
                           int (*nodeNumber)(const int coord[]));

/* Replaces free_sse_su3dslash_3d (include/sse_dslash_3d.h:21). */
void free_sse_su3dslash_3d(void);

/* Replaces sse_su3dslash_wilson_3d (include/sse_dslash_3d.h:32,
 * lib/sse_su3dslash_32bit_scalar_3d.c:150-170); cb = source cb3. */
void sse_su3dslash_wilson_3d(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb);

/* ---- extensions (no counterpart in the reference) ---------------------- */
int sse_su3dslash_3d_last_error(void);
int sse_su3dslash_3d_gauge_upload(SSEREAL* u, int gauge_reals);
int sse_su3dslash_3d_gauge_release(void);
int sse_su3dslash_wilson_3d_resident(SSEREAL* psi, SSEREAL* res, int isign, int cb);
int sse_su3dslash_wilson_3d_device(void* d_res, const void* d_psi, int isign, int cb, void* stream);
/* Sloppy FP64 (FP64 library only), as sse_su3dslash_set_sloppy. */
int sse_su3dslash_3d_set_sloppy(int on);
/* Gauge reuse for sse_su3dslash_wilson_3d, as sse_su3dslash_set_gauge_reuse. */
int sse_su3dslash_3d_set_gauge_reuse(int mode);

#ifdef __cplusplus
}
#endif

#endif /* SSE_DSLASH3D_H */
