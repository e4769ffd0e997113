/*
 * sse_dslash_qdp_packer_3d.h -- QDP++-facing gauge packer of the 3D operator
 * (header-only).
 *
 * Same names and signature as the reference's include/sse_dslash_qdp_packer_3d.h:14-16
 * (SSEDslash3D::PrimitiveSU3Matrix, SSEDslash3D::qdp_pack_gauge_3d).  All four
 * directions are packed, as lib/qdp_packer_3d.cc:11-25 does: the 3D Dslash
 * reads mu = 0..2 of the same [V][4][3][3][2] transposed-link array.  Requires
 * QDP++ (see sse_dslash_qdp_packer.h).
 */
#ifndef SSE_DSLASH_QDP_PACKER_3D_H
#define SSE_DSLASH_QDP_PACKER_3D_H

#include "sse_config.h"
#include "sse_dslash_3d.h"

#ifndef QDP_INCLUDE
#include "qdp.h"
#endif

using namespace QDP;

namespace SSEDslash3D {

typedef PColorMatrix<RComplex<REAL>, Nc> PrimitiveSU3Matrix;

/* packed[4*site + mu] = transpose(u[mu] at site), mu = 0..3. */
inline void qdp_pack_gauge_3d(const multi1d<LatticeColorMatrix>& u, multi1d<PrimitiveSU3Matrix>& packed) {
  const int nsites = Layout::sitesOnNode();
  if (packed.size() < 4 * nsites) packed.resize(4 * nsites);
  for (int site = 0; site < nsites; ++site)
    for (int mu = 0; mu < 4; ++mu) packed[4 * site + mu] = transpose(u[mu].elem(site).elem());
}

/* Extension: pack and keep resident for sse_su3dslash_wilson_3d_resident/_device. */
inline int qdp_gauge_upload_3d(const multi1d<LatticeColorMatrix>& u, int gauge_reals) {
  static_assert(sizeof(PrimitiveSU3Matrix) == 18 * sizeof(SSEREAL),
                "QDP++ REAL and SSE_PRECISION must agree (link = 18 reals)");
  multi1d<PrimitiveSU3Matrix> packed;
  qdp_pack_gauge_3d(u, packed);
  return sse_su3dslash_3d_gauge_upload((SSEREAL*)&packed[0], gauge_reals);
}

}  // namespace SSEDslash3D

#endif /* SSE_DSLASH_QDP_PACKER_3D_H */
