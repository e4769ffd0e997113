"""ctypes binding of the C ABI in include/dslash_mi355x.h (libdslash_mi355x.so).

The shared libraries are built in-tree by ``sse_wilson_dslash_amd.build``
(hipcc, gfx950) into ``sse_wilson_dslash_amd/_lib/``.  Nothing here falls back
to a CPU implementation: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
CORE = "libdslash_mi355x.so"
LEGACY = {32: "libsse_dslash_f32.so", 64: "libsse_dslash_f64.so"}
# the same drop-in API on the host CPU (OpenMP), for nodes without a GPU
LEGACY_CPU = {32: "libsse_dslash_cpu_f32.so", 64: "libsse_dslash_cpu_f64.so"}

F32, F64, F16, BF16, F64S = 0, 1, 2, 3, 4
PREC_NAMES = {F32: "f32", F64: "f64", F16: "f16", BF16: "bf16", F64S: "f64_sloppy"}
OK, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_NOT_INIT, ERR_NOT_CONVERGED = range(6)

_vp, _i, _sz, _u64, _f = C.c_void_p, C.c_int, C.c_size_t, C.c_uint64, C.c_float
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)

class MrhsProblem(C.Structure):
    """dslash_mrhs_problem"""
    _fields_ = [("latt", C.c_int * 4), ("d_gauge_cb", C.c_void_p), ("d_psi", C.c_void_p), ("d_res", C.c_void_p),
                ("field_stride", C.c_size_t), ("nrhs", C.c_int), ("isign", C.c_int), ("target_cb", C.c_int)]


class CgParams(C.Structure):
    """dslash_cg_params"""
    _fields_ = [("kappa", C.c_double), ("tol", C.c_double), ("delta", C.c_double), ("max_iter", C.c_int),
                ("prec_sloppy", C.c_int), ("gauge_reals_sloppy", C.c_int), ("use_graph", C.c_int),
                ("zero_guess", C.c_int), ("gauge_reals_precise", C.c_int)]


class CgResult(C.Structure):
    """dslash_cg_result"""
    _fields_ = [("iterations", C.c_int), ("reliable_updates", C.c_int), ("converged", C.c_int),
                ("true_residual", C.c_double), ("seconds", C.c_double)]


# name -> (restype, argtypes); mirrors include/dslash_mi355x.h
PROTOTYPES = {
    "dslash_cg_create": (_vp, [_ip, C.POINTER(CgParams)]),
    "dslash_cg_load_gauge": (_i, [_vp, _vp, _i, _vp]),
    "dslash_cg_solve": (_i, [_vp, _vp, _vp, C.POINTER(CgResult)]),
    "dslash_cg_apply_op": (_i, [_vp, _vp, _vp, _i]),
    "dslash_cg_set_params": (_i, [_vp, C.POINTER(CgParams)]),
    "dslash_cg_destroy": (_i, [_vp]),
    "dslash_normal_op_noalloc": (_i, [_vp, _vp, _vp, _ip, C.c_double, _i, _i, _vp]),
    "dslash_time_ceiling": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _vp, _fp]),
    "dslash_pack_spinor_sites": (_i, [_vp, _i, _vp, _i, _ip, _i, _i, _vp, _sz, _sz, _vp]),
    "dslash_unpack_spinor_sites": (_i, [_vp, _i, _vp, _i, _ip, _i, _i, _vp, _i, _sz, _sz, _vp]),
    "dslash_apply_sites": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _i, _i, _sz, _sz, _vp]),
    "dslash_last_error": (_i, []),
    "dslash_last_error_message": (C.c_char_p, []),
    "dslash_version": (_i, []),
    "dslash_spinor_cb_bytes": (_sz, [_ip, _i, _i]),
    "dslash_gauge_cb_bytes": (_sz, [_ip, _i, _i, _i]),
    "dslash_check_lattice": (_i, [_ip, _i]),
    "dslash_pack_gauge": (_i, [_vp, _i, _vp, _i, _ip, _i, _i, _i, _vp, _vp]),
    "dslash_pack_spinor": (_i, [_vp, _i, _vp, _i, _ip, _i, _i, _vp, _vp]),
    "dslash_unpack_spinor": (_i, [_vp, _i, _vp, _i, _ip, _i, _i, _vp, _i, _vp]),
    "dslash_convert_spinor": (_i, [_vp, _i, _vp, _i, _ip, _i, _vp]),
    "dslash_apply": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _i, _i, _vp]),
    "dslash_apply_full": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _i, _i, _vp]),
    "dslash_apply_xpay": (_i, [_vp, _vp, _vp, _vp, C.c_double, _ip, _i, _i, _i, _i, _vp]),
    "dslash_set_launch": (_i, [_i, _i, _i, _i, _i]),
    "dslash_get_launch": (_i, [_i, _i, _i, _ip, _ip]),
    "dslash_set_launch_march": (_i, [_i, _i, _i, _i]),
    "dslash_get_launch_march": (_i, [_i, _i, _i, _ip]),
    "dslash_launch_march_for": (_i, [_ip, _i, _i, _i, _ip]),
    "dslash_apply_mrhs": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp]),
    "dslash_tune_apply": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _vp, _ip, _ip, _fp]),
    "dslash_tune_mrhs": (_i, [_vp, _i, _i, _i, _i, _vp, _ip, _ip, _ip, _fp]),
    "dslash_force": (_i, [_vp, _vp, _vp, _vp, _vp, _ip, _i, _i, _vp]),
    "dslash_get_mrhs_launch": (_i, [_i, _i, _i, _ip, _ip, _ip, _ip]),
    "dslash_set_mrhs_launch": (_i, [_i, _i, _i, _i, _i, _i, _i]),
    "dslash_face_layout": (_i, [_ip, _i, _i, _i, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)]),
    "dslash_halo_pack": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _i, _i, _i, _vp]),
    "dslash_halo_interior": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _i, _i, _i, _vp]),
    "dslash_halo_exterior": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _i, _i, _i, _vp]),
    "dslash_tune_save": (_i, [C.c_char_p]),
    "dslash_tune_load": (_i, [C.c_char_p]),
    "dslash_gen_gauge_host": (_i, [_vp, _i, _ip, _i, _u64]),
    "dslash_gen_gauge_device": (_i, [_vp, _i, _ip, _i, _u64, _vp]),
    "dslash_gen_spinor_host": (_i, [_vp, _i, _ip, _i, _u64, _i]),
    "dslash_gen_spinor_device": (_i, [_vp, _i, _ip, _i, _u64, _i, _vp]),
    "dslash_time_apply": (_i, [_vp, _vp, _vp, _ip, _i, _i, _i, _i, _i, _vp, _fp]),
    "dslash_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "dslash_layout_set": (_i, [_ip, _i]),
    "dslash_layout_linear_index": (_i, [_ip]),
    "dslash_layout_site_coords": (None, [_ip, _i, _i]),
    "dslash_layout_node_number": (_i, [_ip]),
}

_core = None


class DslashError(RuntimeError):
    """A C ABI call returned a non-zero DSLASH_ERR_* status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[dslash error {code}] {msg}")
        self.code = code


def lib_path(name: str = CORE) -> str:
    return os.path.join(LIB_DIR, name)


def core() -> C.CDLL:
    """Load libdslash_mi355x.so (once) and declare its prototypes."""
    global _core
    if _core is None:
        path = lib_path(CORE)
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _core = lib
    return _core


def check(status: int) -> None:
    """Raise DslashError for a non-zero status of a core call."""
    if status != 0:
        msg = core().dslash_last_error_message().decode(errors="replace")
        raise DslashError(status, msg)


def latt_arg(latt) -> C.Array:
    if len(latt) != 4:
        raise ValueError("lattice size must have 4 entries")
    return (C.c_int * 4)(*[int(v) for v in latt])


def ptr(x) -> int | None:
    """Device or host address of a torch tensor / numpy array / int / None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(f"cannot take the address of {type(x)!r}")
