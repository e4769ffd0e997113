/*
 * sse_dslash_qdp_packer.h -- QDP++-facing gauge packer (header-only).
 *
 * Same names and signatures as the reference's include/sse_dslash_qdp_packer.h:15-19
 * (SSEDslash::PrimitiveSU3Matrix, SSEDslash::qdp_pack_gauge,
 * SSEDslash::qdp_pack_gauge3d), so Chroma code that packs its gauge field for
 * the SSE library compiles unchanged against this tree.  The packed array is
 * the host wire format sse_su3dslash_wilson reads: entry 4*site + mu holds the
 * TRANSPOSE of U_mu at the node's linear site `site` (QDP++'s site order, the
 * one its getLinearSiteIndex callback reports), i.e. the [V][4][3][3][2] array
 * of lib/qdp_packer_nopad.cc:7-19.
 *
 * Requires QDP++ (qdp.h on the include path; define QDP_INCLUDE if it is
 * already included).  Header-only: none of this tree's libraries depends on
 * QDP++.
 */
#ifndef SSE_DSLASH_QDP_PACKER_H
#define SSE_DSLASH_QDP_PACKER_H

#include "sse_config.h"
#include "sse_dslash.h"

#ifndef QDP_INCLUDE
#include "qdp.h"
#endif

using namespace QDP;

namespace SSEDslash {

/* One link in QDP++'s primitive colour-matrix type (9 complex REALs). */
typedef PColorMatrix<RComplex<REAL>, Nc> PrimitiveSU3Matrix;

/* packed[4*site + mu] = transpose(u[mu] at site), for every site on the node.
 * packed is resized to 4 * Layout::sitesOnNode() if it is shorter. */
inline void qdp_pack_gauge(const multi1d<LatticeColorMatrix>& u, multi1d<PrimitiveSU3Matrix>& packed) {
  const int nsites = Layout::sitesOnNode();
  if (packed.size() < 4 * nsites) packed.resize(4 * nsites);
  for (int site = 0; site < nsites; ++site)
    for (int mu = 0; mu < 4; ++mu) packed[4 * site + mu] = transpose(u[mu].elem(site).elem());
}

/* Declared by the reference header but defined nowhere in the reference; here
 * it is the 3D operator's packer (all four directions packed, the 3D Dslash
 * reads mu = 0..2), the same layout as SSEDslash3D::qdp_pack_gauge_3d. */
inline void qdp_pack_gauge3d(const multi1d<LatticeColorMatrix>& u, multi1d<PrimitiveSU3Matrix>& packed) {
  qdp_pack_gauge(u, packed);
}

/* Extension: pack and upload in one call, keeping the gauge resident in HBM
 * for sse_su3dslash_wilson_resident / _device (gauge_reals 18, or 12 for
 * SU(3) links).  Returns 0 or a DSLASH_ERR_* code. */
inline int qdp_gauge_upload(const multi1d<LatticeColorMatrix>& u, int gauge_reals) {
  static_assert(sizeof(PrimitiveSU3Matrix) == 18 * sizeof(SSEREAL),
                "QDP++ REAL and SSE_PRECISION must agree (link = 18 reals)");
  multi1d<PrimitiveSU3Matrix> packed;
  qdp_pack_gauge(u, packed);
  return sse_su3dslash_gauge_upload((SSEREAL*)&packed[0], gauge_reals);
}

}  // namespace SSEDslash

#endif /* SSE_DSLASH_QDP_PACKER_H */
