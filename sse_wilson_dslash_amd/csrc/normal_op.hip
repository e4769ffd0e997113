// The even-odd preconditioned normal operator A = Mhat^dagger Mhat,
// Mhat = 1 - kappa^2 D_eo D_oe (M = 1 - kappa D), on the even checkerboard:
// the kernel sequence the CG applies, and dslash_normal_op_noalloc, the form of
// it that allocates nothing.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dslash_common.h"
#include "dslash_internal.h"
#include "dslash_mi355x.h"
#include "status.h"

namespace wd {

// out = A in.  t_odd (odd cb) and q_even (even cb) are work fields in format
// prec; g holds the link bundles of target cb 0 then cb 1, gcb bytes each.
// With qnorm, the per-block partials of |q|^2 = |Mhat in|^2 = <in, A in> are
// written there (the fused <p, A p> of the CG).
int apply_normal_op(void* out, const void* in, void* t_odd, void* q_even, const void* g, size_t gcb,
                    const int latt[4], double kappa, int prec, int G, hipStream_t st, double* qnorm) {
  const double k2 = -kappa * kappa;
  const char* g0 = (const char*)g;
  const char* g1 = g0 + gcb;
  // q = in - kappa^2 D_eo D_oe in
  if (int e = apply_cb_ex(t_odd, in, g1, latt, 4, prec, G, +1, 1, nullptr, 0.0, st)) return e;
  if (int e = apply_cb_ex(q_even, t_odd, g0, latt, 4, prec, G, +1, 0, in, k2, st, qnorm)) return e;
  // out = q - kappa^2 (D^dag)_eo (D^dag)_oe q
  if (int e = apply_cb_ex(t_odd, q_even, g1, latt, 4, prec, G, -1, 1, nullptr, 0.0, st)) return e;
  if (int e = apply_cb_ex(out, t_odd, g0, latt, 4, prec, G, -1, 0, q_even, k2, st)) return e;
  return 0;
}

}  // namespace wd

using namespace wd;

namespace {

// Work fields of dslash_normal_op_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never uses its name as an
// address: the device address comes from hipGetSymbolAddress, once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES];
hipError_t g_scratch_err[NOALLOC_MAX_DEVICES];
std::once_flag g_scratch_once[NOALLOC_MAX_DEVICES];

}  // namespace

extern "C" {

int dslash_normal_op_noalloc(void* d_out, const void* d_in, const void* d_gauge, const int latt[4], double kappa,
                             int prec, int gauge_reals, void* stream) {
  clear_error();
  WD_REQUIRE(latt, "latt is NULL");
  if (int e = dslash_check_lattice(latt, 4)) return e;
  WD_REQUIRE(prec == F32 || prec == F64 || prec == F16 || prec == BF16 || prec == F64S, "unknown format %d", prec);
  WD_REQUIRE(gauge_reals == 18 || gauge_reals == 12, "gauge_reals must be 18 or 12 (got %d)", gauge_reals);
  WD_REQUIRE(d_out && d_in && d_gauge, "NULL field pointer");
  WD_REQUIRE(d_out != d_in, "out and in must be different fields");
  const uint32_t vh = (uint32_t)((uint64_t)latt[0] * latt[1] * latt[2] * latt[3] / 2);
  const size_t sb = (spinor_cb_bytes(vh, prec) + 255) / 256 * 256;
  WD_REQUIRE(2 * sb <= NOALLOC_SCRATCH_BYTES,
             "dslash_normal_op_noalloc: %d x %d x %d x %d needs %zu bytes of work fields in this format, the "
             "library's fixed array holds %zu (use dslash_cg_apply_op for larger lattices)",
             latt[0], latt[1], latt[2], latt[3], 2 * sb, NOALLOC_SCRATCH_BYTES);
  int dev = 0;
  WD_HIP(hipGetDevice(&dev));
  WD_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device %d beyond the %d this entry point serves", dev,
             NOALLOC_MAX_DEVICES);
  // set-up on the first call on this device (with it current): the array's
  // device address
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
    return set_error(DSLASH_ERR_HIP, "hipGetSymbolAddress of the work array on device %d: %s", dev,
                     hipGetErrorString(g_scratch_err[dev]));
  char* base = (char*)g_scratch_ptr[dev];
  const size_t gcb = gauge_cb_bytes(vh, 4, prec, gauge_reals);
  return apply_normal_op(d_out, d_in, base, base + sb, d_gauge, gcb, latt, kappa, prec, gauge_reals,
                         (hipStream_t)stream, nullptr);
}

}  // extern "C"
