// Compiles the QDP++-facing packer headers against tests/qdp_stub and checks
// the packed array against the wire format sse_su3dslash_wilson reads:
// reals[((site*4 + mu)*3 + a)*3 + b][c] = U_mu(site)(b, a).{re,im}
// (lib/qdp_packer_nopad.cc:7-19, lib/qdp_packer_3d.cc:11-25).
#include <stdio.h>

#include "sse_dslash_qdp_packer.h"
#include "sse_dslash_qdp_packer_3d.h"

static double value(int site, int mu, int i, int j, int c) { return site * 1000.0 + mu * 100 + i * 10 + j + 0.5 * c; }

int main() {
  const int V = 37;  // odd on purpose: no assumption of a checkerboard split
  QDP::Layout::stub_sites() = V;
  multi1d<LatticeColorMatrix> u(4);
  for (int mu = 0; mu < 4; ++mu) {
    u[mu].F.resize(V);
    for (int s = 0; s < V; ++s)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          u[mu].elem(s).elem().elem(i, j).real() = (REAL)value(s, mu, i, j, 0);
          u[mu].elem(s).elem().elem(i, j).imag() = (REAL)value(s, mu, i, j, 1);
        }
  }
  int bad = 0;
  for (int variant = 0; variant < 3; ++variant) {
    multi1d<SSEDslash::PrimitiveSU3Matrix> p4;  // empty: the packer sizes it
    multi1d<SSEDslash3D::PrimitiveSU3Matrix> p3(4 * V);
    const SSEREAL* r;
    if (variant == 0) {
      SSEDslash::qdp_pack_gauge(u, p4);
      r = (const SSEREAL*)&p4[0];
      if (p4.size() != 4 * V) ++bad;
    } else if (variant == 1) {
      SSEDslash::qdp_pack_gauge3d(u, p4);
      r = (const SSEREAL*)&p4[0];
    } else {
      SSEDslash3D::qdp_pack_gauge_3d(u, p3);
      r = (const SSEREAL*)&p3[0];
    }
    for (int s = 0; s < V; ++s)
      for (int mu = 0; mu < 4; ++mu)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 2; ++c) {
              SSEREAL got = r[((((size_t)s * 4 + mu) * 3 + a) * 3 + b) * 2 + c];
              if (got != (SSEREAL)value(s, mu, b, a, c)) ++bad;
            }
  }
  static_assert(sizeof(SSEDslash::PrimitiveSU3Matrix) == 18 * sizeof(SSEREAL), "link size");
  printf("%s %d\n", bad ? "FAIL" : "OK", bad);
  return bad ? 1 : 0;
}
