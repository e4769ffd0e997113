// The reference's drop-in C API (include/sse_dslash.h, include/sse_dslash_3d.h)
// on top of libdslash_mi355x.so.  Built twice: with -DSSE_PRECISION=32 into
// libsse_dslash_f32.so and with -DSSE_PRECISION=64 into libsse_dslash_f64.so,
// so both export the reference's symbol names with the matching SSEREAL.
//
// Host state mirrors the reference's static tables: a reference count, the
// lattice size and the site table (my checkerboard index -> caller linear
// index, built through the caller's callbacks exactly as
// lib/shift_tables_scalar.c:236-263 and lib/shift_tables_scalar_3d.c:256-278
// do).  Neighbour tables are not needed: the kernel computes neighbours from
// coordinates, and a non-canonical caller order is applied through the site
// table when packing and unpacking on the device.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "dslash_mi355x.h"
#include "sse_dslash.h"
#include "sse_dslash_3d.h"
#include "stager.h"

#if SSE_PRECISION == 32
#define LEG_PREC DSLASH_F32
#else
#define LEG_PREC DSLASH_F64
#endif
// A build with -DSSEDSLASH_SLOPPY (the reference's --enable-sloppy) starts the
// FP64 library in sloppy mode; sse_su3dslash_set_sloppy switches at run time.
#if SSE_PRECISION == 64 && defined(SSEDSLASH_SLOPPY)
#define LEG_DEV_PREC DSLASH_F64_SLOPPY
#else
#define LEG_DEV_PREC LEG_PREC
#endif

namespace {

const size_t RS = sizeof(SSEREAL);

struct Legacy {
  const int nd;
  const char* name;
  int refcount = 0;
  int last_error = 0;
  int latt[4] = {0, 0, 0, 0};
  size_t V = 0, Vh = 0;
  std::vector<int> site_table;
  bool canonical = true;
  int* d_site_table = nullptr;
  void* d_u_aos = nullptr;
  void* d_psi_aos = nullptr;
  void* d_res_aos = nullptr;
  void* d_psi_cb = nullptr;
  void* d_res_cb = nullptr;
  void* d_bundle = nullptr;       // resident bundles of target cb 0 and 1
  void* d_bundle_call = nullptr;  // per-call bundle of one target cb
  size_t sb = 0, gb18 = 0;
  int resident_G = 0;        // 0: nothing resident; 18/12: both bundles hold the uploaded gauge
  int prec = LEG_DEV_PREC;   // device format of bundles and kernel (LEG_PREC or F64_SLOPPY)
  hipStream_t stream = nullptr;
  // *_device calls run on the caller's stream and read d_bundle after the call
  // returns; dev_done marks the last of them, and everything that rewrites or
  // frees d_bundle waits for it first.
  hipEvent_t dev_done = nullptr;
  bool dev_pending = false;
  // Gauge reuse for the host-pointer call (0 off, 1 pointer + checksum, 2 pointer):
  // the resident bundles then hold the gauge last uploaded from reuse_ptr.
  int reuse_mode = 0;
  bool reuse_valid = false;
  const SSEREAL* reuse_ptr = nullptr;
  uint64_t reuse_sum = 0;
  std::vector<SSEREAL> h_stage;
  wd::Stager stager;     // uploads (and everything that is not pipelined)
  wd::Stager stager_dn;  // downloads of the slab pipeline, so both directions move at once
  // Slab pipeline of the host-pointer call (canonical orders): the source
  // checkerboard moves up in slabs of whole slices of the slowest coordinate
  // on s_up while the compute stream works on the slabs already there and s_dn
  // brings finished target slabs back.
  static constexpr int MAX_SLABS = 32;
  static constexpr int SLABS = 8;  // default slab count
  hipStream_t s_up = nullptr, s_dn = nullptr;
  hipEvent_t ev_up[MAX_SLABS] = {}, ev_done[MAX_SLABS] = {};
  int nslab = 0;

  explicit Legacy(int nd_, const char* name_) : nd(nd_), name(name_) {}

  int fail(int code, const char* what) {
    last_error = code;
    fprintf(stderr, "%s: %s (%s)\n", name, what, dslash_last_error_message());
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    last_error = DSLASH_ERR_HIP;
    fprintf(stderr, "%s: %s: %s\n", name, what, hipGetErrorString(e));
    return DSLASH_ERR_HIP;
  }

  // My canonical full index of a site (cb2 for 4D, cb3d for 3D).
  size_t my_index(const int c[4]) const {
    int lxh = latt[0] / 2;
    int cb = nd == 4 ? ((c[0] + c[1] + c[2] + c[3]) & 1) : ((c[0] + c[1] + c[2]) & 1);
    size_t idx;
    if (nd == 4)
      idx = (size_t)(c[0] / 2) + (size_t)lxh * ((size_t)c[1] + (size_t)latt[1] * ((size_t)c[2] + (size_t)latt[2] * (size_t)c[3]));
    else
      idx = (size_t)c[3] + (size_t)latt[3] * ((size_t)(c[0] / 2) + (size_t)lxh * ((size_t)c[1] + (size_t)latt[1] * (size_t)c[2]));
    return (size_t)cb * Vh + idx;
  }

  void init(const int* ls, void (*getSiteCoords)(int[], int, int), int (*getLinearSiteIndex)(const int[])) {
    if (refcount > 0) {  // lib/sse_su3dslash_32bit_scalar.c:95-99
      ++refcount;
      return;
    }
    for (int mu = 0; mu < nd; ++mu)
      if (ls[mu] % 2 != 0) {  // lib/sse_su3dslash_32bit_scalar.c:103-110
        if (nd == 4)
          fprintf(stderr,
                  "This is a Dslash with checkerboarding in 4 dimensions. Each GLOBAL dimension must be even,  "
                  "Your lattice is not like this: latt_size[%d]=%d\n",
                  mu, ls[mu]);
        else
          fprintf(stderr,
                  "This is a Dslash with checkerboarding in 3 dimensions. GLOBAL dimension 0,1,2 (corresponding to "
                  "x,y,z) must be even,  Your lattice is not like this: latt_size[%d]=%d\n",
                  mu, ls[mu]);
        exit(1);
      }
    for (int mu = 0; mu < 4; ++mu) latt[mu] = ls[mu];
    if (dslash_check_lattice(latt, nd) != 0) {
      fprintf(stderr, "%s: %s\n", name, dslash_last_error_message());
      exit(1);
    }
    V = (size_t)latt[0] * latt[1] * latt[2] * latt[3];
    Vh = V / 2;
    site_table.assign(V, -1);
    // Visit sites in the reference's loop order and record caller indices.
    int c[4];
    if (nd == 4) {
      for (int p = 0; p < 2; ++p)
        for (int t = 0; t < latt[3]; ++t)
          for (int z = 0; z < latt[2]; ++z)
            for (int y = 0; y < latt[1]; ++y)
              for (int x = 0; x < latt[0] / 2; ++x) {
                c[0] = 2 * x + p; c[1] = y; c[2] = z; c[3] = t;
                site_table[my_index(c)] = getLinearSiteIndex(c);
              }
    } else {
      for (int p = 0; p < 2; ++p)
        for (int z = 0; z < latt[2]; ++z)
          for (int y = 0; y < latt[1]; ++y)
            for (int x = 0; x < latt[0] / 2; ++x)
              for (int t = 0; t < latt[3]; ++t) {
                c[0] = 2 * x + p; c[1] = y; c[2] = z; c[3] = t;
                site_table[my_index(c)] = getLinearSiteIndex(c);
              }
    }
    // The caller's order must be a permutation whose coordinates round-trip
    // through getSiteCoords (the reference builds its neighbour table from
    // them, lib/shift_tables_scalar.c:275-302).
    std::vector<char> seen(V, 0);
    canonical = true;
    for (size_t i = 0; i < V; ++i) {
      int q = site_table[i];
      if (q < 0 || (size_t)q >= V || seen[q]) {
        fprintf(stderr, "%s: getLinearSiteIndex is not a permutation of 0..V-1 (site %zu -> %d)\n", name, i, q);
        exit(1);
      }
      seen[q] = 1;
      if ((size_t)q != i) canonical = false;
      if (getSiteCoords) {
        int cc[4] = {0, 0, 0, 0};
        getSiteCoords(cc, 0, q);
        if (my_index(cc) != i) {
          fprintf(stderr, "%s: getSiteCoords(%d) is inconsistent with getLinearSiteIndex\n", name, q);
          exit(1);
        }
      }
    }
    sb = dslash_spinor_cb_bytes(latt, nd, LEG_PREC);
    gb18 = dslash_gauge_cb_bytes(latt, nd, LEG_PREC, 18);
    hipError_t e;
    if ((e = hipStreamCreate(&stream)) != hipSuccess || (e = hipStreamCreate(&s_up)) != hipSuccess ||
        (e = hipStreamCreate(&s_dn)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&dev_done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipMalloc(&d_u_aos, V * 4 * 18 * RS)) != hipSuccess ||
        (e = hipMalloc(&d_psi_aos, V * 24 * RS)) != hipSuccess ||
        (e = hipMalloc(&d_res_aos, V * 24 * RS)) != hipSuccess ||
        (e = hipMalloc(&d_psi_cb, sb)) != hipSuccess ||
        (e = hipMalloc(&d_res_cb, sb)) != hipSuccess ||
        (e = hipMalloc(&d_bundle, 2 * gb18)) != hipSuccess ||
        (e = hipMalloc(&d_bundle_call, gb18)) != hipSuccess) {
      fprintf(stderr, "%s: device allocation failed: %s\n", name, hipGetErrorString(e));
      exit(1);
    }
    if (!canonical) {
      if ((e = hipMalloc(&d_site_table, V * sizeof(int))) != hipSuccess ||
          (e = hipMemcpy(d_site_table, site_table.data(), V * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess) {
        fprintf(stderr, "%s: site table upload failed: %s\n", name, hipGetErrorString(e));
        exit(1);
      }
      h_stage.resize(Vh * 24);
    }
    nslab = 0;
    if (canonical) {
      // slabs of whole slices, each at most one staging chunk: the first target
      // slab waits for three source slabs and the last two downloads follow the
      // last upload, which favours short slabs; each slab costs a few host
      // synchronisations and launches, which favours few.  SSE_DSLASH_SLABS
      // overrides the count.
      const size_t L = (size_t)latt[nd - 1], slice_bytes = Vh / L * 24 * RS;
      size_t S = std::max<size_t>(SLABS, (Vh * 24 * RS + wd::Stager::CHUNK - 1) / wd::Stager::CHUNK);
      if (const char* env = getenv("SSE_DSLASH_SLABS")) S = (size_t)std::max(0, atoi(env));
      S = std::min<size_t>(S, std::min<size_t>(L, MAX_SLABS));
      if (S >= 3 && (L + S - 1) / S * slice_bytes <= wd::Stager::CHUNK) nslab = (int)S;
    }
    // with the slab pipeline two pools copy at the same time (uploads, downloads)
    stager.start(nslab ? 2 : 1);
    if (nslab) {
      for (int k = 0; k < nslab; ++k)
        if ((e = hipEventCreateWithFlags(&ev_up[k], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&ev_done[k], hipEventDisableTiming)) != hipSuccess) {
          fprintf(stderr, "%s: event creation failed: %s\n", name, hipGetErrorString(e));
          exit(1);
        }
      stager_dn.start(2);
    }
    resident_G = 0;
    dev_pending = false;
    reuse_valid = false;
    refcount = 1;
    last_error = 0;
  }

  void release() {
    if (refcount == 0) return;  // lib/sse_su3dslash_32bit_scalar.c:127-129
    if (--refcount > 0) return;
    if (stream) (void)hipStreamSynchronize(stream);
    if (dev_pending) (void)hipEventSynchronize(dev_done);
    dev_pending = false;
    if (s_up) (void)hipStreamSynchronize(s_up);
    if (s_dn) (void)hipStreamSynchronize(s_dn);
    stager.finish();
    stager_dn.finish();
    for (int k = 0; k < MAX_SLABS; ++k) {
      if (ev_up[k]) (void)hipEventDestroy(ev_up[k]);
      if (ev_done[k]) (void)hipEventDestroy(ev_done[k]);
      ev_up[k] = ev_done[k] = nullptr;
    }
    nslab = 0;
    void* ptrs[] = {d_u_aos, d_psi_aos, d_res_aos, d_psi_cb, d_res_cb, d_bundle, d_bundle_call, d_site_table};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    d_u_aos = d_psi_aos = d_res_aos = d_psi_cb = d_res_cb = d_bundle = d_bundle_call = nullptr;
    d_site_table = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    if (s_up) (void)hipStreamDestroy(s_up);
    if (s_dn) (void)hipStreamDestroy(s_dn);
    stream = s_up = s_dn = nullptr;
    if (dev_done) (void)hipEventDestroy(dev_done);
    dev_done = nullptr;
    site_table.clear();
    h_stage.clear();
    resident_G = 0;
    reuse_valid = false;
  }

  // Order the internal stream after every *_device launch issued so far: the
  // next kernel that rewrites the resident bundles must not overtake a kernel
  // on the caller's stream that still reads them.
  int after_device_calls() {
    if (!dev_pending) return 0;
    hipError_t e = hipStreamWaitEvent(stream, dev_done, 0);
    if (e != hipSuccess) return hip_fail(e, "wait for device-resident calls");
    dev_pending = false;
    return 0;
  }

  int check_ready() {
    if (refcount == 0) {
      last_error = DSLASH_ERR_NOT_INIT;
      fprintf(stderr, "%s: called before init\n", name);
      return DSLASH_ERR_NOT_INIT;
    }
    return 0;
  }

  // Upload the packed gauge and build the bundles of the listed target cbs.
  int upload_gauge(const SSEREAL* u, int G, void* const slot[2]) {
    hipError_t e = stager.h2d(d_u_aos, u, V * 4 * 18 * RS, stream);
    if (e != hipSuccess) return hip_fail(e, "gauge upload");
    if (int r = after_device_calls()) return r;
    for (int t = 0; t < 2; ++t)
      if (slot[t] && dslash_pack_gauge(slot[t], t, d_u_aos, LEG_PREC, latt, nd, prec, G, d_site_table, stream))
        return fail(dslash_last_error(), "gauge pack");
    return 0;
  }

  // First site (within a checkerboard) of slab k of the slab pipeline.
  size_t slab_begin(int k) const {
    const size_t L = (size_t)latt[nd - 1];
    return L * (size_t)k / (size_t)nslab * (Vh / L);
  }

  // The host-pointer apply as a pipeline over slabs of whole slices of the
  // slowest coordinate (t in 4D, z in 3D; canonical orders).  Target slab k
  // needs source slabs k - 1, k, k + 1 (periodic), so the source slabs go up
  // in the order S-1, 0, 1, ..., S-2 on s_up (a host thread, through the
  // upload staging buffers), the compute stream packs each source slab once it
  // has arrived and computes target slab k as soon as its three source slabs
  // are in, and s_dn brings target slab k back (its own staging buffers)
  // while slab k + 1 is computed and later source slabs still go up.
  int apply_host_slabs(SSEREAL* psi, SSEREAL* res, int isign, int cb, const void* bundle_t, int G) {
    const int tcb = 1 - cb, S = nslab;
    const size_t site_bytes = 24 * RS;
    std::mutex m;
    std::condition_variable cv;
    int uploaded = 0;  // slabs whose upload event is recorded
    hipError_t up_err = hipSuccess;
    std::thread up([&] {
      for (int step = 0; step < S; ++step) {
        const int j = step == 0 ? S - 1 : step - 1;
        const size_t a = slab_begin(j), b = slab_begin(j + 1);
        const size_t off = ((size_t)cb * Vh + a) * 24;
        hipError_t e = stager.h2d((SSEREAL*)d_psi_aos + off, psi + off, (b - a) * site_bytes, s_up);
        if (e == hipSuccess) e = hipEventRecord(ev_up[j], s_up);
        std::lock_guard<std::mutex> lk(m);
        if (e != hipSuccess) {
          up_err = e;
          uploaded = S;  // no further progress: let the consumer stop
        } else {
          uploaded = step + 1;
        }
        cv.notify_one();
        if (e != hipSuccess) return;
      }
    });
    auto step_of = [S](int j) { return j == S - 1 ? 0 : j + 1; };
    std::vector<char> packed(S, 0);
    int rc = 0;
    hipError_t e = hipSuccess;
    for (int k = 0; k < S && !rc; ++k) {
      const int nb[3] = {(k + S - 1) % S, k, (k + 1) % S};
      int need = 0;
      for (int j : nb) need = std::max(need, step_of(j) + 1);
      {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return uploaded >= need; });
        if (up_err != hipSuccess) break;
      }
      for (int j : nb) {
        if (packed[j]) continue;
        if ((e = hipStreamWaitEvent(stream, ev_up[j], 0)) != hipSuccess) break;
        if (dslash_pack_spinor_sites(d_psi_cb, cb, d_psi_aos, LEG_PREC, latt, nd, prec, nullptr, slab_begin(j),
                                     slab_begin(j + 1), stream)) {
          rc = fail(dslash_last_error(), "spinor pack");
          break;
        }
        packed[j] = 1;
      }
      if (e != hipSuccess || rc) break;
      const size_t a = slab_begin(k), b = slab_begin(k + 1);
      if (dslash_apply_sites(d_res_cb, d_psi_cb, bundle_t, latt, nd, prec, G, isign, tcb, a, b, stream)) {
        rc = fail(dslash_last_error(), "dslash");
        break;
      }
      if (dslash_unpack_spinor_sites(d_res_aos, tcb, d_res_cb, prec, latt, nd, LEG_PREC, nullptr, 0, a, b, stream)) {
        rc = fail(dslash_last_error(), "spinor unpack");
        break;
      }
      const size_t off = ((size_t)tcb * Vh + a) * 24;
      if ((e = hipEventRecord(ev_done[k], stream)) != hipSuccess ||
          (e = hipStreamWaitEvent(s_dn, ev_done[k], 0)) != hipSuccess ||
          (e = stager_dn.d2h_issue(k & 1, (const SSEREAL*)d_res_aos + off, (b - a) * site_bytes, s_dn)) !=
              hipSuccess)
        break;
      if (k > 0) {
        const size_t a0 = slab_begin(k - 1), off0 = ((size_t)tcb * Vh + a0) * 24;
        if ((e = stager_dn.d2h_finish(res + off0, (k - 1) & 1, (a - a0) * site_bytes)) != hipSuccess) break;
      }
    }
    if (!rc && e == hipSuccess && up_err == hipSuccess) {
      const size_t a0 = slab_begin(S - 1), off0 = ((size_t)tcb * Vh + a0) * 24;
      e = stager_dn.d2h_finish(res + off0, (S - 1) & 1, (Vh - a0) * site_bytes);
    }
    up.join();
    // leave no stream work behind that could still touch the caller's arrays
    (void)hipStreamSynchronize(s_up);
    (void)hipStreamSynchronize(s_dn);
    (void)hipStreamSynchronize(stream);
    if (rc) return rc;
    if (up_err != hipSuccess) return hip_fail(up_err, "spinor upload");
    if (e != hipSuccess) return hip_fail(e, "slab pipeline");
    last_error = 0;
    return 0;
  }

  // Shared tail of the host-pointer paths: psi's source cb in, res target cb out.
  int apply_host(SSEREAL* psi, SSEREAL* res, int isign, int cb, const void* bundle_t, int G) {
    if (nslab && stager.pipelined() && stager_dn.pipelined())
      return apply_host_slabs(psi, res, isign, cb, bundle_t, G);
    const int tcb = 1 - cb;
    hipError_t e;
    if (canonical)
      e = stager.h2d((char*)d_psi_aos + (size_t)cb * Vh * 24 * RS, psi + (size_t)cb * Vh * 24, Vh * 24 * RS,
                     stream);
    else
      e = stager.h2d(d_psi_aos, psi, V * 24 * RS, stream);
    if (e != hipSuccess) return hip_fail(e, "spinor upload");
    if (dslash_pack_spinor(d_psi_cb, cb, d_psi_aos, LEG_PREC, latt, nd, prec, d_site_table, stream))
      return fail(dslash_last_error(), "spinor pack");
    if (dslash_apply(d_res_cb, d_psi_cb, bundle_t, latt, nd, prec, G, isign, tcb, stream))
      return fail(dslash_last_error(), "dslash");
    if (canonical) {
      if (dslash_unpack_spinor(d_res_aos, tcb, d_res_cb, prec, latt, nd, LEG_PREC, nullptr, 0, stream))
        return fail(dslash_last_error(), "spinor unpack");
      e = stager.d2h(res + (size_t)tcb * Vh * 24, (char*)d_res_aos + (size_t)tcb * Vh * 24 * RS, Vh * 24 * RS,
                     stream);
      if (e != hipSuccess) return hip_fail(e, "result download");
    } else {
      if (dslash_unpack_spinor(d_res_aos, tcb, d_res_cb, prec, latt, nd, LEG_PREC, nullptr, 1, stream))
        return fail(dslash_last_error(), "spinor unpack");
      e = stager.d2h(h_stage.data(), d_res_aos, Vh * 24 * RS, stream);
      if (e != hipSuccess) return hip_fail(e, "result download");
      for (size_t i = 0; i < Vh; ++i)
        memcpy(res + (size_t)site_table[(size_t)tcb * Vh + i] * 24, &h_stage[i * 24], 24 * RS);
    }
    last_error = 0;
    return 0;
  }

  int wilson(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb) {
    if (isign != 1 && isign != -1) return 0;  // lib/sse_su3dslash_32bit_scalar.c:162-179 does nothing
    if (int r = check_ready()) return r;
    if (cb != 0 && cb != 1) return fail(DSLASH_ERR_INVALID, "cb must be 0 or 1");
    if (reuse_mode) return wilson_reuse(u, psi, res, isign, cb);
    // The gauge may change between calls, so it is uploaded every call and
    // only the target checkerboard's bundle is built.
    const int tcb = 1 - cb;
    void* slots[2] = {tcb == 0 ? d_bundle_call : nullptr, tcb == 1 ? d_bundle_call : nullptr};
    if (int r = upload_gauge(u, 18, slots)) return r;
    return apply_host(psi, res, isign, cb, d_bundle_call, 18);
  }

  int gauge_upload(const SSEREAL* u, int G) {
    if (int r = check_ready()) return r;
    if (G != 18 && G != 12) return fail(DSLASH_ERR_INVALID, "gauge_reals must be 18 or 12");
    size_t gb = dslash_gauge_cb_bytes(latt, nd, prec, G);
    void* slots[2] = {d_bundle, (char*)d_bundle + gb};
    resident_G = 0;
    reuse_valid = false;
    if (int r = upload_gauge(u, G, slots)) return r;
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "synchronize");
    resident_G = G;
    last_error = 0;
    return 0;
  }

  int gauge_release() {
    resident_G = 0;
    reuse_valid = false;
    return 0;
  }

  // Host-pointer call with gauge reuse: the bundles of both target
  // checkerboards stay resident and are rebuilt only when u is a different
  // pointer than last time, or (mode 1) its checksum changed.
  int wilson_reuse(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb) {
    const size_t ub = V * 4 * 18 * RS;
    uint64_t sum = 0;
    if (reuse_mode == 1 && nslab) {  // both staging pools hash half of the array each
      const size_t h1 = ub / 2;
      uint64_t s2 = 0;
      std::thread t2([&] { s2 = stager_dn.checksum((const char*)u + h1, ub - h1); });
      const uint64_t s1 = stager.checksum(u, h1);
      t2.join();
      sum = (s1 * 0x9E3779B97F4A7C15ull) ^ s2;
    } else if (reuse_mode == 1) {
      sum = stager.checksum(u, ub);
    }
    if (!(reuse_valid && resident_G == 18 && u == reuse_ptr && (reuse_mode == 2 || sum == reuse_sum))) {
      reuse_valid = false;
      resident_G = 0;
      const size_t gb = dslash_gauge_cb_bytes(latt, nd, prec, 18);
      void* slots[2] = {d_bundle, (char*)d_bundle + gb};
      if (int r = upload_gauge(u, 18, slots)) return r;
      resident_G = 18;
      reuse_valid = true;
      reuse_ptr = u;
      reuse_sum = sum;
    }
    const size_t gb = dslash_gauge_cb_bytes(latt, nd, prec, 18);
    return apply_host(psi, res, isign, cb, (char*)d_bundle + (size_t)(1 - cb) * gb, 18);
  }

  int set_gauge_reuse(int mode) {
    if (mode < 0 || mode > 2) return fail(DSLASH_ERR_INVALID, "gauge reuse mode must be 0, 1 or 2");
    if (mode != reuse_mode) reuse_valid = false;
    reuse_mode = mode;
    last_error = 0;
    return 0;
  }

  // Sloppy FP64: FP64 spinors, FP32 links and arithmetic.  The bundles change
  // format, so a resident gauge is dropped and must be uploaded again.
  int set_sloppy(int on) {
    if (on && LEG_PREC != DSLASH_F64) return fail(DSLASH_ERR_UNSUPPORTED, "sloppy mode exists for the FP64 library only");
    int p = on ? DSLASH_F64_SLOPPY : LEG_PREC;
    if (p != prec) resident_G = 0, reuse_valid = false;
    prec = p;
    last_error = 0;
    return 0;
  }

  int wilson_resident(SSEREAL* psi, SSEREAL* res, int isign, int cb) {
    if (int r = check_ready()) return r;
    if (isign != 1 && isign != -1) return fail(DSLASH_ERR_INVALID, "isign must be +1 or -1");
    if (cb != 0 && cb != 1) return fail(DSLASH_ERR_INVALID, "cb must be 0 or 1");
    if (!resident_G) return fail(DSLASH_ERR_NOT_INIT, "no resident gauge: call *_gauge_upload first");
    size_t gb = dslash_gauge_cb_bytes(latt, nd, prec, resident_G);
    return apply_host(psi, res, isign, cb, (char*)d_bundle + (size_t)(1 - cb) * gb, resident_G);
  }

  int wilson_device(void* d_res, const void* d_psi, int isign, int cb, void* s) {
    if (int r = check_ready()) return r;
    if (!resident_G) return fail(DSLASH_ERR_NOT_INIT, "no resident gauge: call *_gauge_upload first");
    if (dslash_apply_full(d_res, d_psi, d_bundle, latt, nd, prec, resident_G, isign, cb, s))
      return fail(dslash_last_error(), "dslash");
    hipError_t e = hipEventRecord(dev_done, (hipStream_t)s);
    if (e != hipSuccess) return hip_fail(e, "record device-call event");
    dev_pending = true;
    last_error = 0;
    return 0;
  }
};

Legacy g4(4, "sse_su3dslash");
Legacy g3(3, "sse_su3dslash_3d");

}  // namespace

extern "C" {

void init_sse_su3dslash(const int* latt_size, void (*getSiteCoords)(int coord[], int node, int linearsite),
                        int (*getLinearSiteIndex)(const int coord[]), int (*nodeNumber)(const int coord[])) {
  (void)nodeNumber;
  g4.init(latt_size, getSiteCoords, getLinearSiteIndex);
}
void free_sse_su3dslash(void) { g4.release(); }
void sse_su3dslash_wilson(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  (void)g4.wilson(u, psi, res, isign, cb);
}
void sse_su3dslash_prepost_receives(void) {}
// The CPU build's thread query: no host threads compute here.
int sse_su3dslash_cpu_threads(void) { return 0; }
int sse_su3dslash_last_error(void) { return g4.last_error; }
int sse_su3dslash_gauge_upload(SSEREAL* u, int gauge_reals) { return g4.gauge_upload(u, gauge_reals); }
int sse_su3dslash_gauge_release(void) { return g4.gauge_release(); }
int sse_su3dslash_wilson_resident(SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  return g4.wilson_resident(psi, res, isign, cb);
}
int sse_su3dslash_wilson_device(void* d_res, const void* d_psi, int isign, int cb, void* stream) {
  return g4.wilson_device(d_res, d_psi, isign, cb, stream);
}
int sse_su3dslash_set_sloppy(int on) { return g4.set_sloppy(on); }
int sse_su3dslash_set_gauge_reuse(int mode) { return g4.set_gauge_reuse(mode); }

void init_sse_su3dslash_3d(const int* latt_size, void (*getSiteCoords)(int coord[], int node, int linearsite),
                           int (*getLinearSiteIndex)(const int coord[]), int (*nodeNumber)(const int coord[])) {
  (void)nodeNumber;
  g3.init(latt_size, getSiteCoords, getLinearSiteIndex);
}
void free_sse_su3dslash_3d(void) { g3.release(); }
void sse_su3dslash_wilson_3d(SSEREAL* u, SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  (void)g3.wilson(u, psi, res, isign, cb);
}
int sse_su3dslash_3d_last_error(void) { return g3.last_error; }
int sse_su3dslash_3d_gauge_upload(SSEREAL* u, int gauge_reals) { return g3.gauge_upload(u, gauge_reals); }
int sse_su3dslash_3d_gauge_release(void) { return g3.gauge_release(); }
int sse_su3dslash_wilson_3d_resident(SSEREAL* psi, SSEREAL* res, int isign, int cb) {
  return g3.wilson_resident(psi, res, isign, cb);
}
int sse_su3dslash_wilson_3d_device(void* d_res, const void* d_psi, int isign, int cb, void* stream) {
  return g3.wilson_device(d_res, d_psi, isign, cb, stream);
}
int sse_su3dslash_3d_set_sloppy(int on) { return g3.set_sloppy(on); }
int sse_su3dslash_3d_set_gauge_reuse(int mode) { return g3.set_gauge_reuse(mode); }

#include "gauge_packers.h"

}  // extern "C"
