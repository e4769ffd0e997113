// Pipelined host <-> device copies through pinned staging buffers, used by the
// drop-in API for the caller's pageable host arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <emmintrin.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace wd {

// Host <-> device copies of the legacy host arrays through two pinned staging
// buffers: host threads copy chunk i between the caller's (pageable) array and
// one buffer while the DMA engine moves chunk i - 1 through the other.  On
// MI355X this moves ~53 GB/s host-to-device and ~55 GB/s device-to-host, where
// hipMemcpy from pageable memory gives 19-29 and 17 GB/s.  Without pinned
// memory (allocation failed) it falls back to plain copies.
class Stager {
 public:
  static constexpr size_t CHUNK = (size_t)32 << 20;
  static constexpr size_t SMALL = (size_t)1 << 20;  // smaller uploads go straight to hipMemcpyAsync

  Stager() = default;
  Stager(const Stager&) = delete;
  Stager& operator=(const Stager&) = delete;
  // A caller that never frees the library still exits cleanly: the workers are
  // stopped and joined (joinable std::threads would terminate the process).  No
  // HIP call here, the runtime may already be gone at static destruction.
  ~Stager() { stop_workers(); }

  // share = how many pools run copies at the same time: each pool takes that
  // fraction of the host's threads (leaving one per pool for other work), at
  // most 7 workers besides the calling thread.
  void start(unsigned share = 1) {
    if (pinned) return;
    if (hipHostMalloc((void**)&pinned, 2 * CHUNK, 0) != hipSuccess) {
      pinned = nullptr;
      return;
    }
    for (hipEvent_t& e : ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency()) / std::max(1u, share);
    const int n = (int)std::min(7u, hw > 2 ? hw - 2 : 0u);
    stop = false;
    // workers wait for the NEXT job: a generation count left from an earlier
    // start/finish cycle must not look like a pending job
    const unsigned g0 = gen;
    for (int i = 0; i < n; ++i) workers.emplace_back([this, i, g0] { work(i + 1, g0); });
  }

  void finish() {
    // the last uploads may still be reading the buffers
    (void)reuse(0);
    (void)reuse(1);
    stop_workers();
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    busy[0] = busy[1] = false;
    next = 0;
  }

  // Enqueues the upload on s; returns once every byte of h has been read (the
  // last DMAs may still be running from the staging buffers).  Chunks alternate
  // between the two buffers across calls too, so the host copy of one call's
  // first chunk overlaps the DMA of the previous call's last one.
  hipError_t h2d(void* d, const void* h, size_t n, hipStream_t s) {
    if (!usable() || n < SMALL) return hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s);
    for (size_t off = 0; off < n; off += CHUNK) {
      const size_t len = std::min(CHUNK, n - off);
      const int b = next;
      next ^= 1;
      if (hipError_t e = reuse(b)) return e;
      copy((char*)pinned + b * CHUNK, (const char*)h + off, len);
      if (hipError_t e = hipMemcpyAsync((char*)d + off, pinned + b * CHUNK, len, hipMemcpyHostToDevice, s)) return e;
      if (hipError_t e = hipEventRecord(ev[b], s)) return e;
      busy[b] = true;
    }
    return hipSuccess;
  }

  // 64-bit checksum of n host bytes, computed by the calling thread and the
  // workers in parallel (a change detector for a host array, not a
  // cryptographic hash).  The value depends on the worker count, which is
  // fixed between start() and finish().
  uint64_t checksum(const void* h, size_t n) {
    if (workers.empty() || n < 2 * CHUNK) return digest((const char*)h, n);
    run(nullptr, (const char*)h, n);
    uint64_t r = (uint64_t)job_parts;
    for (int k = 0; k < job_parts; ++k) r = (r ^ part_sum[k]) * 0x9E3779B97F4A7C15ull, r ^= r >> 31;
    return r;
  }

  // Split download for pipelines: d2h_issue starts the DMA of n <= CHUNK bytes
  // into staging buffer `slot` (0 or 1) after the work already on s;
  // d2h_finish waits for it and copies the bytes to h.  A slot is reused only
  // after its finish.
  hipError_t d2h_issue(int slot, const void* d, size_t n, hipStream_t s) {
    if (!usable() || n > CHUNK || (slot & ~1)) return hipErrorInvalidValue;
    if (hipError_t e = reuse(slot)) return e;
    if (hipError_t e = hipMemcpyAsync(pinned + slot * CHUNK, d, n, hipMemcpyDeviceToHost, s)) return e;
    if (hipError_t e = hipEventRecord(ev[slot], s)) return e;
    busy[slot] = true;
    return hipSuccess;
  }
  hipError_t d2h_finish(void* h, int slot, size_t n) {
    if (!usable() || n > CHUNK || (slot & ~1)) return hipErrorInvalidValue;
    if (hipError_t e = reuse(slot)) return e;
    copy((char*)h, pinned + slot * CHUNK, n);
    return hipSuccess;
  }
  bool pipelined() const { return usable(); }

  // Downloads after the work already on s; returns with h complete.
  hipError_t d2h(void* h, const void* d, size_t n, hipStream_t s) {
    if (!usable() || n < 2 * CHUNK) {
      if (hipError_t e = hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s)) return e;
      return hipStreamSynchronize(s);
    }
    const size_t nch = (n + CHUNK - 1) / CHUNK;
    auto issue = [&](size_t i) -> hipError_t {
      const int b = (int)(i & 1);
      const size_t off = i * CHUNK, len = std::min(CHUNK, n - off);
      if (hipError_t e = reuse(b)) return e;
      if (hipError_t e = hipMemcpyAsync(pinned + b * CHUNK, (const char*)d + off, len, hipMemcpyDeviceToHost, s))
        return e;
      if (hipError_t e = hipEventRecord(ev[b], s)) return e;
      busy[b] = true;
      return hipSuccess;
    };
    for (size_t i = 0; i < std::min<size_t>(2, nch); ++i)
      if (hipError_t e = issue(i)) return e;
    for (size_t i = 0; i < nch; ++i) {
      const int b = (int)(i & 1);
      const size_t off = i * CHUNK, len = std::min(CHUNK, n - off);
      if (hipError_t e = reuse(b)) return e;
      copy((char*)h + off, pinned + b * CHUNK, len);
      if (i + 2 < nch)
        if (hipError_t e = issue(i + 2)) return e;
    }
    return hipSuccess;
  }

 private:
  bool usable() const { return pinned && ev[0] && ev[1]; }

  void stop_workers() {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv.notify_all();
    for (std::thread& t : workers) t.join();
    workers.clear();
  }

  hipError_t reuse(int b) {  // wait until the DMA that last used buffer b is done
    if (!busy[b]) return hipSuccess;
    busy[b] = false;
    return hipEventSynchronize(ev[b]);
  }

  // memcpy split over the workers and the calling thread
  void copy(char* dst, const char* src, size_t n) { run(dst, src, n); }

  // One job (copy when dst != nullptr, else checksum) split over the workers
  // and the calling thread.
  void run(char* dst, const char* src, size_t n) {
    const int parts = (int)workers.size() + 1;
    {
      std::lock_guard<std::mutex> lk(m);
      job_dst = dst;
      job_src = src;
      job_n = n;
      job_parts = parts;
      pending = parts - 1;
      ++gen;
    }
    cv.notify_all();
    part(0);
    std::unique_lock<std::mutex> lk(m);
    done.wait(lk, [this] { return pending == 0; });
  }

  void part(int k) {
    const size_t a = job_n * k / job_parts, b = job_n * (k + 1) / job_parts;
    if (job_dst)
      stream_copy(job_dst + a, job_src + a, b - a);
    else
      part_sum[k] = digest(job_src + a, b - a);
  }

  // memcpy with non-temporal 16-byte stores: the staging copies are a few MB
  // per thread, below the size where memcpy itself stops reading the
  // destination lines before writing them, and neither side is read again by
  // this thread, so streaming stores save that read of host memory (the
  // bound of the host-pointer path).  The fence makes the stores visible before
  // the copy is reported done.
  static void stream_copy(char* dst, const char* src, size_t n) {
    if (n < ((size_t)64 << 10)) {
      memcpy(dst, src, n);
      return;
    }
    const size_t head = (16 - ((uintptr_t)dst & 15)) & 15;
    memcpy(dst, src, head);
    dst += head;
    src += head;
    n -= head;
    const size_t body = n & ~(size_t)63;
    for (size_t i = 0; i < body; i += 64) {
      const __m128i v0 = _mm_loadu_si128((const __m128i*)(src + i));
      const __m128i v1 = _mm_loadu_si128((const __m128i*)(src + i + 16));
      const __m128i v2 = _mm_loadu_si128((const __m128i*)(src + i + 32));
      const __m128i v3 = _mm_loadu_si128((const __m128i*)(src + i + 48));
      _mm_stream_si128((__m128i*)(dst + i), v0);
      _mm_stream_si128((__m128i*)(dst + i + 16), v1);
      _mm_stream_si128((__m128i*)(dst + i + 32), v2);
      _mm_stream_si128((__m128i*)(dst + i + 48), v3);
    }
    memcpy(dst + body, src + body, n - body);
    _mm_sfence();
  }

  // 64-bit digest of a byte range: four independent lanes, each folding one
  // 8-byte word per step through xor, a multiply by an odd constant and an
  // xor-shift (each a bijection of the lane state), then a tail and length fold.
  static uint64_t digest(const char* p, size_t n) {
    uint64_t h[4] = {0x243F6A8885A308D3ull, 0x13198A2E03707344ull, 0xA4093822299F31D0ull, 0x082EFA98EC4E6C89ull};
    auto mix = [](uint64_t x, uint64_t w) {
      x = (x ^ w) * 0x9E3779B97F4A7C15ull;
      return x ^ (x >> 29);
    };
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
      uint64_t w[4];
      memcpy(w, p + i, 32);
      h[0] = mix(h[0], w[0]);
      h[1] = mix(h[1], w[1]);
      h[2] = mix(h[2], w[2]);
      h[3] = mix(h[3], w[3]);
    }
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      memcpy(&w, p + i, 8);
      h[0] = mix(h[0], w);
    }
    uint64_t t = 0;
    for (int j = 0; i < n; ++i, ++j) t |= (uint64_t)(unsigned char)p[i] << (8 * j);
    return mix(mix(mix(mix(mix(h[0], h[1]), h[2]), h[3]), t), (uint64_t)n);
  }

  void work(int k, unsigned seen) {
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return stop || gen != seen; });
        if (stop) return;
        seen = gen;
      }
      part(k);
      {
        std::lock_guard<std::mutex> lk(m);
        --pending;
      }
      done.notify_one();
    }
  }

  char* pinned = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool busy[2] = {false, false};
  int next = 0;  // staging buffer of the next upload chunk
  std::vector<std::thread> workers;
  std::mutex m;
  std::condition_variable cv, done;
  bool stop = false;
  unsigned gen = 0;
  int pending = 0;
  char* job_dst = nullptr;
  const char* job_src = nullptr;
  size_t job_n = 0;
  int job_parts = 1;
  uint64_t part_sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // one per part (at most 7 workers + caller)
};

}  // namespace wd
