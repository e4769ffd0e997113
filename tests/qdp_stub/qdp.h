// Minimal model of the QDP++ types the packer headers touch (OLattice /
// PScalar / PColorMatrix / RComplex element access, transpose, multi1d,
// Layout::sitesOnNode), used only to compile and run this project's
// include/sse_dslash_qdp_packer*.h in a CPU test.  It is not QDP++ and holds
// none of its code; REAL follows SSE_PRECISION as a QDP++ build of the same
// precision would.
#pragma once
#include <vector>

namespace QDP {

#if SSE_PRECISION == 32
typedef float REAL;
#else
typedef double REAL;
#endif
const int Nc = 3;

template <typename T> struct RComplex {
  T re, im;
  T& real() { return re; }
  T& imag() { return im; }
};

template <typename T, int N> struct PColorMatrix {
  T F[N][N];
  T& elem(int i, int j) { return F[i][j]; }
  const T& elem(int i, int j) const { return F[i][j]; }
};

template <typename T, int N> PColorMatrix<T, N> transpose(const PColorMatrix<T, N>& m) {
  PColorMatrix<T, N> r;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) r.F[i][j] = m.F[j][i];
  return r;
}

template <typename T> struct PScalar {
  T F;
  T& elem() { return F; }
  const T& elem() const { return F; }
};

template <typename T> struct OLattice {
  std::vector<T> F;
  T& elem(int i) { return F[i]; }
  const T& elem(int i) const { return F[i]; }
};

typedef OLattice<PScalar<PColorMatrix<RComplex<REAL>, Nc>>> LatticeColorMatrix;

template <typename T> struct multi1d {
  std::vector<T> v;
  multi1d() {}
  explicit multi1d(int n) : v(n) {}
  int size() const { return (int)v.size(); }
  void resize(int n) { v.resize(n); }
  T& operator[](int i) { return v[i]; }
  const T& operator[](int i) const { return v[i]; }
};

namespace Layout {
inline int& stub_sites() {
  static int n = 0;
  return n;
}
inline int sitesOnNode() { return stub_sites(); }
}  // namespace Layout

}  // namespace QDP
