// CPU check of the drop-in API's staging pool (sse_wilson_dslash_amd/csrc/stager.h)
// against tests/hip_stub: round trips of sizes below, at and above the
// pipelining threshold (partial last chunks), repeated start / finish cycles
// of one object (a restarted pool must not replay the previous job), a pool
// destroyed while started (workers stopped, not terminate()), built
// with -fsanitize=thread or =address by tests/test_library_cpu.py.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "stager.h"

int main() {
  const size_t N = (size_t)96 << 20;
  std::vector<char> h(N), d(N), back(N);
  for (size_t i = 0; i < N; ++i) h[i] = (char)(i * 131 + 7);
  wd::Stager st;
  for (int cycle = 0; cycle < 3; ++cycle) {
    st.start();
    const size_t sizes[] = {N, N - 12345, (size_t)64 << 20, ((size_t)64 << 20) - 1, (size_t)1 << 20, 0};
    for (size_t n : sizes) {
      memset(d.data(), 0, N);
      memset(back.data(), 0, N);
      if (st.h2d(d.data(), h.data(), n, nullptr) != hipSuccess || st.d2h(back.data(), d.data(), n, nullptr) != hipSuccess) {
        printf("error: cycle %d size %zu\n", cycle, n);
        return 1;
      }
      if (memcmp(d.data(), h.data(), n) != 0 || memcmp(back.data(), h.data(), n) != 0) {
        printf("mismatch: cycle %d size %zu\n", cycle, n);
        return 1;
      }
    }
    // the split download used by the slab pipeline: pieces through
    // alternating slots, each finished after the next one is issued
    {
      const size_t piece = ((size_t)5 << 20) + 123;
      const size_t npieces = N / piece;
      memset(back.data(), 0, N);
      for (size_t k = 0; k <= npieces; ++k) {
        if (k < npieces && st.d2h_issue((int)(k & 1), h.data() + k * piece, piece, nullptr) != hipSuccess) return 1;
        if (k > 0 && st.d2h_finish(back.data() + (k - 1) * piece, (int)((k - 1) & 1), piece) != hipSuccess) return 1;
      }
      if (memcmp(back.data(), h.data(), npieces * piece) != 0) {
        printf("split download mismatch\n");
        return 1;
      }
      if (st.d2h_issue(2, h.data(), 16, nullptr) == hipSuccess ||
          st.d2h_issue(0, h.data(), wd::Stager::CHUNK + 1, nullptr) == hipSuccess) {
        printf("split download accepted a bad slot or size\n");
        return 1;
      }
      // uploads of pieces smaller than a chunk, back to back (buffers alternate across calls)
      memset(d.data(), 0, N);
      for (size_t k = 0; k < npieces; ++k)
        if (st.h2d(d.data() + k * piece, h.data() + k * piece, piece, nullptr) != hipSuccess) return 1;
      if (memcmp(d.data(), h.data(), npieces * piece) != 0) {
        printf("piecewise upload mismatch\n");
        return 1;
      }
    }
    // checksum: reproducible, and any single changed byte changes it (first,
    // part boundaries, last full word, the tail), for the parallel path (n >=
    // 64 MiB) and the single-thread one
    for (size_t n : {N - 3, (size_t)1 << 20}) {
      const uint64_t c0 = st.checksum(h.data(), n);
      if (st.checksum(h.data(), n) != c0) {
        printf("checksum not reproducible: size %zu\n", n);
        return 1;
      }
      const size_t at[] = {0, n / 8, n / 2 + 1, n - 9, n - 1};
      for (size_t i : at) {
        h[i] ^= 0x10;
        const uint64_t c1 = st.checksum(h.data(), n);
        h[i] ^= 0x10;
        if (c1 == c0) {
          printf("checksum blind to byte %zu of %zu\n", i, n);
          return 1;
        }
      }
      if (st.checksum(h.data(), n - 1) == c0) {
        printf("checksum blind to length\n");
        return 1;
      }
    }
    st.finish();
  }
#ifndef __SANITIZE_ADDRESS__
  // a started pool destroyed without finish() (a caller that never frees the
  // library) must stop its workers instead of terminating the process; its
  // staging buffer is left to the exiting process on purpose, so this part runs
  // in the ThreadSanitizer build only (LeakSanitizer would report that buffer)
  {
    wd::Stager leaked;
    leaked.start();
    if (leaked.h2d(d.data(), h.data(), (size_t)70 << 20, nullptr) != hipSuccess) return 1;
  }
#endif
  puts("stager ok");
  return 0;
}
