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
  wd::Stager stager;

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
    if ((e = hipStreamCreate(&stream)) != hipSuccess ||
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
    stager.start();
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
    stager.finish();
    void* ptrs[] = {d_u_aos, d_psi_aos, d_res_aos, d_psi_cb, d_res_cb, d_bundle, d_bundle_call, d_site_table};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    d_u_aos = d_psi_aos = d_res_aos = d_psi_cb = d_res_cb = d_bundle = d_bundle_call = nullptr;
    d_site_table = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
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

  // Shared tail of the host-pointer paths: psi's source cb in, res target cb out.
  int apply_host(SSEREAL* psi, SSEREAL* res, int isign, int cb, const void* bundle_t, int G) {
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
    const uint64_t sum = reuse_mode == 1 ? stager.checksum(u, ub) : 0;
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
