#!/usr/bin/env python3
"""Benchmark of the MI355X Wilson Dslash (one GPU).

Headline: FP32 4D Wilson Dslash on a 32^3 x 64 lattice, GFLOP/s at 1320
flop/site, one "step" = one application over one checkerboard (Vh = 1,048,576
target sites), steps alternating isign = +-1 and the source checkerboard, with
gauge and spinors resident in HBM.  Inputs are synthetic: SU(3) links from
Gram-Schmidt-reunitarised complex Gaussians (seed 1) and N(0,1) spinors,
generated on the device.

Also reported (same JSON line): the roofline of the dominant kernel measured
with HIP events on its launch stream, the CPU baseline (the host-CPU build of
the same drop-in API, libsse_dslash_cpu_f32.so, on all usable host cores and
on one core; a bounded sample of the same lattice), and secondary figures
(FP64, 3D, 12-real links, 16-bit and sloppy formats, batched launches, the
domain-decomposition split, the backward kernel, the CG, the drop-in API).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--cpu-seconds S]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E nominal
LATT = (32, 32, 32, 64)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def fp64_result(latt, nd: int, seed: int, src, isign: int, src_cb: int):
    """D psi of the FP64 operator (equal to the oracle to 1e-13 in the GPU tests) on
    the seeded gauge for a host spinor src, as a full-lattice float64 array with the
    target checkerboard filled."""
    import numpy as np
    import torch
    from sse_wilson_dslash_amd import F64
    from sse_wilson_dslash_amd.device import WilsonDslash

    ref = WilsonDslash(latt, nd, F64, 18)
    ref.gen_gauge(seed=seed)
    psi = ref.pack_spinor(np.ascontiguousarray(src, dtype=np.float64))
    res = ref.new_spinor(zero=True)
    ref.apply(res, psi, isign, src_cb)
    want = ref.unpack_spinor(res, dtype=np.float64, cbs=(1 - src_cb,))
    del ref, psi, res
    torch.cuda.empty_cache()
    return want


def rel_err(got, want) -> float:
    import numpy as np
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


def gpu_leg(steps: int, warmup: int, prec_name: str, nd: int, G: int, verbose=True, check_err=True):
    import torch
    from sse_wilson_dslash_amd import F32, F64
    from sse_wilson_dslash_amd.device import WilsonDslash

    prec = {"f32": F32, "f64": F64}[prec_name]
    from sse_wilson_dslash_amd.batch import tune_apply

    op = WilsonDslash(LATT, nd=nd, prec=prec, gauge_reals=G)
    op.gen_gauge(seed=1)
    psi = op.gen_spinor(seed=1, field=0)
    res = op.new_spinor(zero=True)
    torch.cuda.synchronize()
    shape = tune_apply(op, res, psi)  # picks threads/block, XCD order and marching length
    op.time_apply(res, psi, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = op.time_apply(res, psi, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    flops = op.flops_per_call()
    byts = op.bytes_per_call()
    out = {
        "ms": ms,
        "gflops": flops / (ms * 1e-3) / 1e9,
        "gbs": byts / (ms * 1e-3) / 1e9,
        "bytes_per_call": byts,
        "wall_s": wall,
        "launch": {"block_threads": shape["block_threads"], "xcd_remap": shape["xcd_remap"], "march": shape["march"]},
    }
    err = ""
    if prec == F32 and check_err:  # error of the timed launch shape against the FP64 operator, same fields
        import numpy as np
        src = op.unpack_spinor(psi, dtype=np.float64, cbs=(0,))
        op.apply(res, psi, 1, 0)
        out["rel_err_vs_fp64"] = rel_err(op.unpack_spinor(res, dtype=np.float64, cbs=(1,)),
                                         fp64_result(LATT, nd, 1, src, 1, 0))
        err = f", rel. error vs FP64 operator {out['rel_err_vs_fp64']:.2e}"
    if verbose:
        log(f"[gpu] {prec_name} nd={nd} G={G}: {ms * 1e3:.1f} us/call, {out['gflops']:.0f} GFLOP/s, "
            f"{out['gbs']:.0f} GB/s ({out['gbs'] / PEAK_HBM_GBS:.1%} of HBM peak), launch {out['launch']}{err}")
    del op, psi, res
    torch.cuda.empty_cache()
    return out


def host_cpu_info() -> dict:
    """Host CPUs this process may use: the affinity set, its physical cores
    (distinct (package, core) pairs in /proc/cpuinfo) and the cgroup CPU quota."""
    aff = sorted(os.sched_getaffinity(0))
    info = {"affinity_cpus": len(aff)}
    try:
        cores, model, cur = set(), None, {}
        for line in open("/proc/cpuinfo"):
            if not line.strip():
                if cur.get("processor") in aff:
                    cores.add((cur.get("physical id", 0), cur.get("core id", cur.get("processor"))))
                cur = {}
                continue
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("processor", "physical id", "core id"):
                cur[k] = int(v)
            elif k == "model name" and model is None:
                model = v
        if cur.get("processor") in aff:
            cores.add((cur.get("physical id", 0), cur.get("core id", cur.get("processor"))))
        info["physical_cores"] = len(cores)
        info["model"] = model
    except OSError:
        pass
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_cpu_quota"] = round(int(q) / int(period), 2)
    except (OSError, ValueError):
        pass
    return info


def cpu_child(prec: int, seconds: float, nd: int = 4) -> None:
    """Runs in a child process (OMP_* set by the parent): times the host-CPU
    library (libsse_dslash_cpu_f{32,64}.so, the drop-in API computed with
    OpenMP) at 32^3 x 64 and prints one JSON line."""
    # no torch in this process: its OpenMP runtime would set its own thread count
    import ctypes as C
    import numpy as np
    from sse_wilson_dslash_amd import _lib
    from sse_wilson_dslash_amd.legacy import SSEDslash

    lg = SSEDslash(prec, backend="cpu")
    core = _lib.core()
    V = int(np.prod(LATT))
    pc = _lib.F32 if prec == 32 else _lib.F64
    u = np.empty((V, 4, 3, 3, 2), dtype=lg.dtype)
    psi = np.empty((V, 4, 3, 2), dtype=lg.dtype)
    _lib.check(core.dslash_gen_gauge_host(u.ctypes.data, pc, _lib.latt_arg(LATT), nd, 1))
    _lib.check(core.dslash_gen_spinor_host(psi.ctypes.data, pc, _lib.latt_arg(LATT), nd, 1, 0))
    res = np.zeros_like(psi)
    lg.lib.sse_su3dslash_cpu_threads.restype = C.c_int
    threads = lg.lib.sse_su3dslash_cpu_threads()
    init, apply = ((lg.init_sse_su3dslash, lg.sse_su3dslash_wilson) if nd == 4 else
                   (lg.init_sse_su3dslash_3d, lg.sse_su3dslash_wilson_3d))
    init(LATT)
    for k in range(3):
        apply(u, psi, res, 1 if k % 4 < 2 else -1, k % 2)
    times = []
    t_start = time.perf_counter()
    k = 0
    while k < 20 or (time.perf_counter() - t_start < seconds and k < 2000):
        t0 = time.perf_counter()
        apply(u, psi, res, 1 if k % 4 < 2 else -1, k % 2)
        times.append(time.perf_counter() - t0)
        k += 1
    times.sort()
    med = times[len(times) // 2]
    vh = int(np.prod(LATT)) // 2
    fl = (1320 if nd == 4 else 984) * vh
    print(json.dumps({"ms": med * 1e3, "gflops": fl / med / 1e9, "calls": k, "threads": threads,
                      "wall_s": time.perf_counter() - t_start}), flush=True)


def cpu_leg(seconds: float, threads: int):
    """The host-CPU build of the drop-in API (csrc/cpu_dslash.cpp: OpenMP, AVX2,
    projector-specialised hops, blocked traversal) at 32^3 x 64, FP32 and FP64,
    on all usable host cores and on one core, each in a child process with
    OMP_NUM_THREADS, OMP_PROC_BIND=spread and OMP_PLACES=cores."""
    import subprocess
    host = host_cpu_info()
    out = {"host": host, "threads_all": threads}
    for prec in (32, 64):
        for n in (threads, 1):
            env = dict(os.environ, OMP_NUM_THREADS=str(n), OMP_PROC_BIND="spread", OMP_PLACES="cores")
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", str(prec),
                                "--cpu-seconds", str(seconds if n > 1 else seconds / 2)],
                               env=env, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"CPU leg failed: {r.stderr[-2000:]}")
            d = json.loads(r.stdout.strip().splitlines()[-1])
            if d["threads"] != n:
                raise RuntimeError(f"CPU leg asked for {n} OpenMP threads, the library runs {d['threads']}")
            key = f"f{prec}_{'all' if n > 1 else '1'}_thread{'s' if n > 1 else ''}"
            out[key] = {"ms_per_call": round(d["ms"], 2), "gflops": round(d["gflops"], 2), "calls": d["calls"],
                        "threads": n}
            log(f"[cpu] libsse_dslash_cpu_f{prec} on {n} threads: {d['ms']:.2f} ms/call, {d['gflops']:.1f} GFLOP/s "
                f"({d['calls']} calls, {d['wall_s']:.1f} s)")
    return out


def half_leg(steps: int, prec_name: str, G: int = 12):
    """Reduced-precision Dslash at 32^3x64: 16-bit storage (per-site norm, 16-bit links,
    FP32 math) or sloppy FP64 (FP64 spinors, FP32 links and math); time, and the error
    against the FP64 operator on the same seeded fields."""
    import numpy as np
    import torch
    from sse_wilson_dslash_amd import BF16, F16, F64, F64S
    from sse_wilson_dslash_amd.batch import tune_apply
    from sse_wilson_dslash_amd.device import WilsonDslash

    prec = {"f16": F16, "bf16": BF16, "f64_sloppy": F64S}[prec_name]
    ref = WilsonDslash(LATT, 4, F64, 18)
    ref.gen_gauge(seed=1)
    psi64 = ref.gen_spinor(seed=1)
    res64 = ref.new_spinor(zero=True)
    ref.apply(res64, psi64, 1, 0)
    want = ref.unpack_spinor(res64, dtype=np.float64, cbs=(1,))
    src = ref.unpack_spinor(psi64, dtype=np.float64, cbs=(0,))
    del ref, psi64, res64
    torch.cuda.empty_cache()
    op = WilsonDslash(LATT, 4, prec, G)
    op.gen_gauge(seed=1)
    psi = op.pack_spinor(src)
    res = op.new_spinor(zero=True)
    op.apply(res, psi, 1, 0)
    got = op.unpack_spinor(res, dtype=np.float64, cbs=(1,))
    err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    shape = tune_apply(op, res, psi)
    op.time_apply(res, psi, 20)
    ms = op.time_apply(res, psi, steps)
    gbs = op.bytes_per_call() / (ms * 1e-3) / 1e9
    log(f"[gpu] {prec_name} nd=4 G={G}: {ms * 1e3:.1f} us/call, {op.flops_per_call() / (ms * 1e-3) / 1e9:.0f} GFLOP/s, "
        f"{gbs:.0f} GB/s ({gbs / PEAK_HBM_GBS:.1%}), rel. error vs FP64 operator {err:.2e}")
    out = {"gflops": round(op.flops_per_call() / (ms * 1e-3) / 1e9, 1), "us_per_call": round(ms * 1e3, 2),
           "hbm_gbs": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM_GBS, 4), "rel_err_vs_fp64": err,
           "launch": shape}
    del op, psi, res
    torch.cuda.empty_cache()
    return out


def halo_leg(reps: int = 200):
    """Domain-decomposition split (face pack, interior, exterior) with T partitioned, at the
    local shape of an 8-way T split of 32^3x64 (32^3x8) and at 32^3x64; one GPU, so the
    exchange is the identity.  Event timing on the launch stream; the faces' cost is
    split - interior."""
    import torch
    from sse_wilson_dslash_amd import F32
    from sse_wilson_dslash_amd.device import WilsonDslash
    from sse_wilson_dslash_amd.halo import HaloDslash

    def timed(fn):
        for _ in range(10):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    out = {}
    for latt in ((32, 32, 32, 8), LATT):
        op = WilsonDslash(latt, 4, F32, 18)
        op.gen_gauge(seed=1)
        psi = op.gen_spinor(seed=1)
        res = op.new_spinor(zero=True)
        h = HaloDslash(op, 0b1000)
        faces = h.new_faces()
        t_fused = timed(lambda: op.apply(res, psi, 1, 0))
        t_int = timed(lambda: h.interior(res, psi, 1, 0))
        t_split = timed(lambda: h.apply(res, psi, 1, 0, faces))
        key = "x".join(str(v) for v in latt)
        out[key] = {"fused_us": round(t_fused, 2), "interior_us": round(t_int, 2), "split_us": round(t_split, 2),
                    "face_bytes_per_direction": 2 * h.face_bytes[3]}
        log(f"[gpu] halo split (T partitioned) {key}: fused {t_fused:.1f} us, interior {t_int:.1f} us, "
            f"pack+interior+exterior {t_split:.1f} us, faces {2 * h.face_bytes[3] / 1024:.0f} KiB")
        del op, psi, res, faces
        torch.cuda.empty_cache()
    return out


def force_leg(reps: int = 50):
    """Backward kernel (link gradient and su(3) force of Re<chi, D psi>, FP64) at
    16^3x32 (the backward config) and 32^3x64.  Algorithmic bytes per link: the
    two spinor fields read once (2 x 192 B per site pair = 48 B per link), the link
    read once (144 B), both outputs written once (2 x 144 B)."""
    import torch
    from sse_wilson_dslash_amd import F64
    from sse_wilson_dslash_amd._lib import check, latt_arg, ptr
    from sse_wilson_dslash_amd.device import WilsonDslash, current_stream

    out = {}
    for latt in ((16, 16, 16, 32), LATT):
        op = WilsonDslash(latt, 4, F64, 18)
        d_u = torch.empty(op.V * 72, dtype=torch.float64, device="cuda")
        check(op.lib.dslash_gen_gauge_device(ptr(d_u), F64, latt_arg(latt), 4, 1, current_stream()))
        chi = op.gen_spinor(seed=2, field=0)
        psi = op.gen_spinor(seed=2, field=1)
        grad = torch.empty_like(d_u)
        force = torch.empty_like(d_u)

        def call(k):
            check(op.lib.dslash_force(ptr(grad), ptr(force), ptr(d_u), ptr(chi), ptr(psi), latt_arg(latt),
                                      1 if k % 2 == 0 else -1, k % 2, current_stream()))

        for k in range(5):
            call(k)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(reps):
            call(k)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        links = 4 * op.V
        gbs = links * 480 / (ms * 1e-3) / 1e9
        key = "x".join(str(v) for v in latt)
        out[key] = {"us_per_call": round(ms * 1e3, 2), "links": links, "hbm_gbs": round(gbs, 1),
                    "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}
        log(f"[force] {key} FP64 gradient + su(3) force: {ms * 1e3:.1f} us/call, {gbs:.0f} GB/s "
            f"({gbs / PEAK_HBM_GBS:.1%} of HBM peak, 480 B/link)")
        del op, d_u, chi, psi, grad, force
        torch.cuda.empty_cache()
    return out


def copy_leg(gib: float = 2.0, reps: int = 20):
    """Device-to-device copy bandwidth (read + write bytes / time): the practical HBM ceiling."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 2 * n * 4 / (ms * 1e-3) / 1e9
    log(f"[copy] {gib:.1f} GiB device copy: {ms:.3f} ms, {gbs:.0f} GB/s (read+write)")
    del a, b
    torch.cuda.empty_cache()
    return {"gbs": round(gbs, 1), "frac_of_peak": round(gbs / PEAK_HBM_GBS, 4), "bytes": 2 * n * 4}


def ceiling_leg(reps: int = 200):
    """The streaming ceiling of the headline kernel's traffic (dslash_time_ceiling:
    the FP32 operator's 42 read / 6 written planes per site, links non-temporal,
    no gathers): at 32^3 x 64 over block sizes with and without the XCD block
    order, re-reading the same 805 MB every launch (which the 256 MB MALL partly
    serves); alternating two buffer sets as the headline alternates its two
    checkerboards' link bundles (footprint beyond the MALL); and at 48^3 x 96,
    the CG's lattice."""
    import ctypes as C
    import torch
    from sse_wilson_dslash_amd._lib import check, core, latt_arg, ptr
    from sse_wilson_dslash_amd.device import current_stream

    lib = core()
    ms = C.c_float()

    def bufs(latt):
        vh = latt[0] * latt[1] * latt[2] * latt[3] // 2
        return (torch.ones(96 * vh, dtype=torch.uint8, device="cuda"), torch.empty(96 * vh, dtype=torch.uint8,
                device="cuda"), torch.ones(576 * vh, dtype=torch.uint8, device="cuda"), vh)

    def run(b, latt, block, remap, n):
        check(lib.dslash_time_ceiling(ptr(b[1]), ptr(b[0]), ptr(b[2]), latt_arg(latt), block, remap, n,
                                      current_stream(), C.byref(ms)))
        return ms.value

    b1 = bufs(LATT)
    vh = b1[3]
    res = {}
    for block, remap in ((64, 1), (128, 1), (256, 1), (64, 0), (256, 0)):
        run(b1, LATT, block, remap, 20)
        res[f"{block}{'_xcd' if remap else ''}"] = round(run(b1, LATT, block, remap, reps) * 1e3, 2)
    best_key = min(res, key=res.get)
    best = res[best_key]
    blk, rm = int(best_key.split("_")[0]), 1 if best_key.endswith("xcd") else 0
    b2 = bufs(LATT)
    tot = 0.0
    for k in range(2 * (reps // 2)):
        tot += run(b2 if k % 2 else b1, LATT, blk, rm, 1)
    alt = tot / (2 * (reps // 2)) * 1e3
    del b1, b2
    torch.cuda.empty_cache()
    big = (48, 48, 48, 96)
    b3 = bufs(big)
    run(b3, big, blk, rm, 5)
    t48 = run(b3, big, blk, rm, 30) * 1e3
    vh48 = b3[3]
    del b3
    torch.cuda.empty_cache()
    gbs = 768 * vh / (best * 1e-6) / 1e9
    gbs_alt = 768 * vh / (alt * 1e-6) / 1e9
    gbs48 = 768 * vh48 / (t48 * 1e-6) / 1e9
    log(f"[ceiling] 42 read + 6 written planes/site, 32^3x64: same data {best:.1f} us = {gbs:.0f} GB/s "
        f"({gbs / PEAK_HBM_GBS:.1%}); two alternating buffer sets {alt:.1f} us = {gbs_alt:.0f} GB/s "
        f"({gbs_alt / PEAK_HBM_GBS:.1%}); 48^3x96 {t48:.1f} us = {gbs48:.0f} GB/s ({gbs48 / PEAK_HBM_GBS:.1%}); "
        f"shapes (us): {res}")
    return {"us_per_launch": best, "gbs": round(gbs, 1), "frac_of_peak": round(gbs / PEAK_HBM_GBS, 4),
            "bytes_per_launch": 768 * vh, "shapes_us": res,
            "alternating_sets": {"us_per_launch": round(alt, 2), "gbs": round(gbs_alt, 1),
                                 "frac_of_peak": round(gbs_alt / PEAK_HBM_GBS, 4)},
            "lattice_48x48x48x96": {"us_per_launch": round(t48, 2), "gbs": round(gbs48, 1),
                                    "frac_of_peak": round(gbs48 / PEAK_HBM_GBS, 4)}}


def cg_leg(latt=(48, 48, 48, 96), kappa=0.12, tol=1e-10):
    """Mixed-precision (F16 spinors + norms, 12-real F16 links, FP64 reliable
    updates) vs plain FP64 even-odd CG on Mhat^dag Mhat, same Gaussian source,
    time to FP64 true residual tol."""
    import torch
    from sse_wilson_dslash_amd import BF16, F16, F64
    from sse_wilson_dslash_amd._lib import check, latt_arg, ptr
    from sse_wilson_dslash_amd.device import WilsonDslash, current_stream
    from sse_wilson_dslash_amd.solver import WilsonCG

    op = WilsonDslash(latt, 4, F64, 18)
    V = op.V
    d_u = torch.empty(V * 72, dtype=torch.float64, device="cuda")
    check(op.lib.dslash_gen_gauge_device(ptr(d_u), F64, latt_arg(latt), 4, 1, current_stream()))
    src = op.gen_spinor(seed=1, field=7)
    b = op.cb_view(src, 0).clone()
    del src
    out = {}
    # plain FP64 CG: no reliable updates except the one confirming convergence (delta = 0);
    # mixed: reliable update when |r| has dropped by delta = 3e-3 since the last one.
    # Each pair shares the FP64 operator's links (18 reals, or 12 reals for SU(3)).
    # BF16 keeps 8 significant bits, so its reliable updates come sooner (delta = 1e-2)
    runs = (("fp64", F64, 18, 0.0, 18), ("mixed_f16", F16, 12, 3e-3, 18), ("mixed_bf16", BF16, 12, 1e-2, 18),
            ("fp64_g12", F64, 12, 0.0, 12), ("mixed_f16_g12", F16, 12, 3e-3, 12))
    for name, prec, G, delta, Gh in runs:
        cg = WilsonCG(latt, kappa=kappa, tol=tol, prec_sloppy=prec, gauge_reals_sloppy=G, delta=delta,
                      gauge_reals_precise=Gh)
        cg.load_gauge(d_u)
        x = op.new_spinor(full=False)
        cg.solve(x, b)  # warm-up (graph capture, first-touch)
        r = cg.solve(x, b)
        log(f"[cg] {name}: {r['seconds'] * 1e3:.1f} ms, {r['iterations']} iterations, "
            f"{r['reliable_updates']} reliable updates, true residual {r['true_residual']:.3e}")
        out[name] = {"ms": round(r["seconds"] * 1e3, 2), "iterations": r["iterations"],
                     "reliable_updates": r["reliable_updates"], "true_residual": r["true_residual"],
                     "converged": r["converged"], "delta": delta,
                     "sloppy": {F64: "f64", F16: "f16", BF16: "bf16"}[prec] + f" spinors, {G}-real links",
                     "fp64_operator_links": Gh}
        cg.close()
        del x
        torch.cuda.empty_cache()
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed_f16"]["ms"], 3)
    out["speedup_mixed_bf16_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed_bf16"]["ms"], 3)
    out["speedup_mixed_vs_fp64_g12_links"] = round(out["fp64_g12"]["ms"] / out["mixed_f16_g12"]["ms"], 3)
    out["config"] = {"lattice": list(latt), "kappa": kappa, "tol": tol, "operator": "Mhat^dag Mhat, even-odd",
                     "source": "Gaussian, even checkerboard, seed 1"}
    log(f"[cg] mixed vs FP64 speedup: {out['speedup_mixed_vs_fp64']:.2f}x (18-real FP64 links; BF16 "
        f"{out['speedup_mixed_bf16_vs_fp64']:.2f}x), "
        f"{out['speedup_mixed_vs_fp64_g12_links']:.2f}x (12-real FP64 links)")
    del op, d_u, b
    torch.cuda.empty_cache()
    return out


def half_pmc_workload(reps: int = 20):
    """For rocprofv3 --pmc runs: the 16-bit operators at 32^3 x 64 (F16 with 12- and
    18-real links, BF16 with 12-real links; one site per thread, 64-thread blocks,
    XCD order) and the FP32 one-site kernel for comparison, reps launches each with
    fixed shapes (no tuning)."""
    import ctypes as C
    import torch
    from sse_wilson_dslash_amd import BF16, F16, F32
    from sse_wilson_dslash_amd._lib import check
    from sse_wilson_dslash_amd.device import WilsonDslash

    for prec, G in ((F16, 12), (F16, 18), (BF16, 12), (F32, 18)):
        op = WilsonDslash(LATT, 4, prec, G)
        b, r, m = C.c_int(), C.c_int(), C.c_int()
        check(op.lib.dslash_get_launch(4, prec, G, C.byref(b), C.byref(r)))
        check(op.lib.dslash_get_launch_march(4, prec, G, C.byref(m)))
        check(op.lib.dslash_set_launch(4, prec, G, 64, 1))
        check(op.lib.dslash_set_launch_march(4, prec, G, 1))
        op.gen_gauge(seed=1)
        psi = op.gen_spinor(seed=1)
        res = op.new_spinor(zero=True)
        ms = op.time_apply(res, psi, reps)
        log(f"[pmc-half] prec={prec} G={G}: {ms * 1e3:.1f} us/launch, {op.bytes_per_call()} B/launch")
        check(op.lib.dslash_set_launch(4, prec, G, b.value, r.value))
        check(op.lib.dslash_set_launch_march(4, prec, G, m.value))
        del op, psi, res
        torch.cuda.empty_cache()


def cg_lattice_pmc_workload(reps: int = 10):
    """For rocprofv3 --pmc runs: the CG's operators at 48^3 x 96 in their default
    launch shapes, reps launches each: FP64 18-real D psi (box tiles), the FP64
    x + a D psi form the CG uses, and F16 12-real D psi (one site per thread); then
    the streaming ceiling kernel at the same lattice."""
    import ctypes as C
    import torch
    from sse_wilson_dslash_amd import F16, F64
    from sse_wilson_dslash_amd._lib import check, core, latt_arg, ptr
    from sse_wilson_dslash_amd.device import WilsonDslash, current_stream

    latt = (48, 48, 48, 96)
    for prec, G in ((F64, 18), (F16, 12)):
        op = WilsonDslash(latt, 4, prec, G)
        op.gen_gauge(seed=1)
        psi = op.gen_spinor(seed=1)
        res = op.new_spinor(zero=True)
        op.time_apply(res, psi, 3)  # first touch of the fresh fields
        ms = op.time_apply(res, psi, reps)
        log(f"[pmc-48] prec={prec} G={G}: {ms * 1e3:.1f} us/launch, {op.bytes_per_call()} B/launch")
        if prec == F64:
            x = op.cb_view(psi, 0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                check(op.lib.dslash_apply_xpay(ptr(op.cb_view(res, 0)), ptr(op.cb_view(psi, 1)),
                                               ptr(op.gauge), ptr(x), -0.0144, op._L, F64, 18, 1, 0,
                                               current_stream()))
            e1.record()
            e1.synchronize()
            log(f"[pmc-48] FP64 x + a D psi: {e0.elapsed_time(e1) / reps * 1e3:.1f} us/launch")
        del op, psi, res
        torch.cuda.empty_cache()
    vh = 48 * 48 * 48 * 96 // 2
    a = torch.ones(96 * vh, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    links = torch.ones(576 * vh, dtype=torch.uint8, device="cuda")
    msf = C.c_float()
    check(core().dslash_time_ceiling(ptr(b), ptr(a), ptr(links), latt_arg(latt), 64, 1, reps, current_stream(),
                                     C.byref(msf)))
    log(f"[pmc-48] ceiling: {msf.value * 1e3:.1f} us/launch, {768 * vh} B/launch")


def operators_pmc_workload(reps: int = 12):
    """For rocprofv3 --pmc runs: single-RHS operators at 32^3 x 64 with 18-real links
    in their default launch shapes: 3D FP32, FP64 and sloppy FP64 (z-marching), then
    4D FP64 (box tiles) and 4D sloppy FP64 (box tiles with the LDS halo); 3
    first-touch launches and then reps launches each."""
    import ctypes as C
    import torch
    from sse_wilson_dslash_amd import F32, F64, F64S
    from sse_wilson_dslash_amd._lib import check
    from sse_wilson_dslash_amd.device import WilsonDslash

    for nd, prec in ((3, F32), (3, F64), (3, F64S), (4, F64), (4, F64S)):
        op = WilsonDslash(LATT, nd, prec, 18)
        b, r, m = C.c_int(), C.c_int(), C.c_int()
        check(op.lib.dslash_get_launch(nd, prec, 18, C.byref(b), C.byref(r)))
        check(op.lib.dslash_launch_march_for(op._L, nd, prec, 18, C.byref(m)))
        op.gen_gauge(seed=1)
        psi = op.gen_spinor(seed=1)
        res = op.new_spinor(zero=True)
        op.time_apply(res, psi, 3)
        ms = op.time_apply(res, psi, reps)
        log(f"[pmc-ops] nd={nd} prec={prec}: {ms * 1e3:.1f} us/launch, {op.bytes_per_call()} B/launch, "
            f"{b.value} threads, march {m.value}")
        del op, psi, res
        torch.cuda.empty_cache()


def mrhs_pmc_workload(warmup: int = 8, reps: int = 30):
    """For rocprofv3 --pmc runs: the 32-RHS 24^3x48 launch with the LDS-staged
    shapes fixed (128-site rows; 12x4x4 boxes at 192 threads, two fields per tile;
    12x8x8 wide boxes at 768 threads, one field per tile), each warmup + reps
    launches, no tuning."""
    import torch
    from sse_wilson_dslash_amd import F32
    from sse_wilson_dslash_amd.batch import BatchItem, RhsBatch, apply_batch, get_launch, set_launch
    from sse_wilson_dslash_amd.device import WilsonDslash

    op = WilsonDslash((24, 24, 24, 48), 4, F32, 18)
    op.gen_gauge(seed=1)
    psi = RhsBatch(op, 32)
    for k in range(32):
        psi.field(k).copy_(op.cb_view(op.gen_spinor(seed=1, field=k), 1))
    items = [BatchItem(op, psi, RhsBatch(op, 32, zero=True), 1, 0)]
    saved = get_launch(F32, 18, False)
    try:
        for R, block, t_block in ((2, 128, 8), (2, 192, 16), (1, 768, 16)):
            set_launch(F32, 18, False, R, block, 0, t_block)
            for _ in range(warmup):
                apply_batch(items)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                apply_batch(items)
            e1.record()
            e1.synchronize()
            log(f"[mrhs pmc] rhs_block {R}, {block} threads, t_block {t_block}: "
                f"{e0.elapsed_time(e1) / reps * 1e3:.1f} us/launch")
    finally:
        set_launch(F32, 18, False, saved["rhs_block"], saved["block_threads"], saved["grid_blocks"],
                   saved["t_block"])


def mrhs_leg(reps: int = 50):
    """32 right-hand sides on one 24^3x48 gauge, and one persistent launch over
    8^4, 16^3x32, 24^3x48 with 4 RHS each (FP32, 18-real links), tuned shapes."""
    import torch
    from sse_wilson_dslash_amd import F32
    from sse_wilson_dslash_amd.batch import (BatchItem, RhsBatch, apply_batch, compulsory_bytes, flops,
                                            tune_batch)
    from sse_wilson_dslash_amd.device import WilsonDslash

    import numpy as np

    def item(latt, nrhs, seed):
        op = WilsonDslash(latt, 4, F32, 18)
        op.gen_gauge(seed=seed)
        psi = RhsBatch(op, nrhs)
        for k in range(nrhs):
            psi.field(k).copy_(op.cb_view(op.gen_spinor(seed=seed, field=k), 1))
        return BatchItem(op, psi, RhsBatch(op, nrhs, zero=True), 1, 0)

    def timed(items, grid):
        apply_batch(items, grid_blocks=grid)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            apply_batch(items, grid_blocks=grid)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    items = [item((24, 24, 24, 48), 32, 1)]
    t = tune_batch(items, persistent=False)
    ms = timed(items, 0)
    out["rhs32_24c48"] = {"ms": round(ms, 4), "gflops": round(flops(items) / (ms * 1e-3) / 1e9, 1),
                          "hbm_gbs": round(compulsory_bytes(items) / (ms * 1e-3) / 1e9, 1), "tuned": t}
    errs = []
    for k in (0, 31):  # fields of the timed launch against the FP64 operator on the same sources
        op, psi, res = items[0].op, items[0].psi, items[0].res
        src = op.unpack_cb(psi.field(k), 1, dtype=np.float64)
        errs.append(rel_err(op.unpack_cb(res.field(k), 0, dtype=np.float64),
                            fp64_result((24, 24, 24, 48), 4, 1, src, 1, 1)))
    out["rhs32_24c48"]["rel_err_vs_fp64"] = max(errs)
    log(f"[mrhs] 32 RHS 24^3x48: {ms * 1e3:.1f} us/launch, {out['rhs32_24c48']['gflops']:.0f} GFLOP/s, tuned {t}, "
        f"rel. error vs FP64 operator (fields 0, 31) {max(errs):.2e}")
    del items
    items = [item((8, 8, 8, 8), 4, 2), item((16, 16, 16, 32), 4, 3), item((24, 24, 24, 48), 4, 4)]
    t = tune_batch(items, persistent=True)
    ms = timed(items, t["grid_blocks"])
    out["persistent_uneven"] = {"ms": round(ms, 4), "gflops": round(flops(items) / (ms * 1e-3) / 1e9, 1),
                                "hbm_gbs": round(compulsory_bytes(items) / (ms * 1e-3) / 1e9, 1), "tuned": t}
    log(f"[mrhs] persistent 8^4+16^3x32+24^3x48 x4 RHS: {ms * 1e3:.1f} us/launch, "
        f"{out['persistent_uneven']['gflops']:.0f} GFLOP/s, tuned {t}")
    del items
    torch.cuda.empty_cache()
    return out


def legacy_leg(calls: int = 12):
    """The drop-in host-pointer API (sse_su3dslash_wilson, FP32, 32^3x64): every call
    uploads the 604 MB gauge and the source half of psi and downloads the target
    half of res over PCIe; with the opt-in gauge reuse (checksum, pointer); and the
    resident-gauge extension, where only spinors move.  Median of `calls` calls
    after one untimed call, alternating isign and cb; PCIe-inclusive."""
    import numpy as np
    from sse_wilson_dslash_amd.device import gen_gauge_host, gen_spinor_host
    from sse_wilson_dslash_amd.legacy import SSEDslash

    lg = SSEDslash(32)
    u = gen_gauge_host(LATT, 4, seed=1, dtype=np.float32)
    psi = gen_spinor_host(LATT, 4, seed=1, dtype=np.float32)
    res = np.zeros_like(psi)
    lg.init_sse_su3dslash(LATT)
    out = {}
    fl = 1320 * (int(np.prod(LATT)) // 2)

    def timed(call, key):
        call(1, 0)
        ts = []
        for k in range(calls):
            t0 = time.perf_counter()
            call(1 if k % 4 < 2 else -1, k % 2)
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts)) * 1e3
        out[key + "_ms_per_call"] = round(ms, 2)
        out[key + "_gflops"] = round(fl / (ms * 1e-3) / 1e9, 1)

    try:
        timed(lambda isign, cb: lg.sse_su3dslash_wilson(u, psi, res, isign, cb), "host_pointers")
        # opt-in gauge reuse (sse_su3dslash_set_gauge_reuse): 1 = pointer + checksum,
        # 2 = pointer only; the untimed first call uploads, the timed ones reuse
        for mode, key in ((1, "gauge_reuse_checksum"), (2, "gauge_reuse_pointer")):
            assert lg.set_gauge_reuse(mode) == 0
            timed(lambda isign, cb: lg.sse_su3dslash_wilson(u, psi, res, isign, cb), key)
        assert lg.set_gauge_reuse(0) == 0
        assert lg.gauge_upload(u, 18) == 0
        timed(lambda isign, cb: lg.wilson_resident(psi, res, isign, cb), "resident_gauge")
    finally:
        lg.free_sse_su3dslash()
    log(f"[legacy] host-pointer API {out['host_pointers_ms_per_call']} ms/call "
        f"({out['host_pointers_gflops']} GFLOP/s, PCIe-inclusive); gauge reuse "
        f"{out['gauge_reuse_checksum_ms_per_call']} ms/call (checksum), "
        f"{out['gauge_reuse_pointer_ms_per_call']} ms/call (pointer); resident gauge "
        f"{out['resident_gauge_ms_per_call']} ms/call ({out['resident_gauge_gflops']} GFLOP/s)")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=6.0)
    ap.add_argument("--cpu-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip FP64 / 3D / 12-real variants")
    ap.add_argument("--no-cg", action="store_true", help="skip the 48^3x96 mixed vs FP64 CG")
    ap.add_argument("--no-mrhs", action="store_true", help="skip the batched Dslash legs")
    ap.add_argument("--pmc-mrhs", action="store_true",
                    help="only the 32-RHS launch in two fixed LDS-staged shapes (for rocprofv3 --pmc runs)")
    ap.add_argument("--pmc-cg-lattice", action="store_true",
                    help="only the CG's operators and the ceiling kernel at 48^3x96 (for rocprofv3 --pmc runs)")
    ap.add_argument("--pmc-half", action="store_true",
                    help="only the 16-bit and FP32 one-site kernels in fixed shapes (for rocprofv3 --pmc runs)")
    ap.add_argument("--pmc-ops", action="store_true",
                    help="only the 3D FP32 / FP64 / sloppy-FP64 and 4D FP64 / sloppy-FP64 operators in their default "
                         "shapes (for rocprofv3 --pmc runs)")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="only a 2 GiB calibration copy and the headline kernel (for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args.cpu_child, args.cpu_seconds)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the Dslash here runs on one GPU only "
              "(multi-GPU halo exchange is not built); run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2

    log(f"[bench] device {torch.cuda.get_device_name(0)}")
    if args.pmc_mrhs:
        mrhs_pmc_workload()
        return 0
    if args.pmc_half:
        half_pmc_workload()
        return 0
    if args.pmc_cg_lattice:
        cg_lattice_pmc_workload()
        return 0
    if args.pmc_ops:
        operators_pmc_workload()
        return 0
    if args.pmc_workload:
        copy_leg(2.0, 3)
        gpu_leg(args.steps, args.warmup, "f32", 4, 18, check_err=False)
        return 0
    head = gpu_leg(args.steps, args.warmup, "f32", 4, 18)
    extra = {"copy": copy_leg(), "ceiling": ceiling_leg(),
             "f32_4d": {"rel_err_vs_fp64": head.get("rel_err_vs_fp64"), "launch": head["launch"]}}
    extra["f32_4d"]["frac_of_ceiling"] = round(extra["ceiling"]["us_per_launch"] / (head["ms"] * 1e3), 4)
    if not args.no_extra:
        n = max(200, args.steps // 4)
        for key, (p, nd, G) in {"f64_4d": ("f64", 4, 18), "f32_4d_g12": ("f32", 4, 12),
                                "f32_3d": ("f32", 3, 18), "f64_3d": ("f64", 3, 18)}.items():
            r = gpu_leg(n, 20, p, nd, G)
            extra[key] = {"gflops": round(r["gflops"], 1), "us_per_call": round(r["ms"] * 1e3, 2),
                          "hbm_gbs": round(r["gbs"], 1), "hbm_frac": round(r["gbs"] / PEAK_HBM_GBS, 4),
                          "launch": r["launch"]}
            if "rel_err_vs_fp64" in r:
                extra[key]["rel_err_vs_fp64"] = r["rel_err_vs_fp64"]
        for p in ("f16", "bf16"):
            extra[f"{p}_4d_g12"] = half_leg(n, p)
        extra["f64_sloppy_4d"] = half_leg(n, "f64_sloppy", 18)
    if not args.no_mrhs:
        extra["mrhs"] = mrhs_leg()
    if not args.no_extra:
        extra["halo_split"] = halo_leg()
        extra["force"] = force_leg()
    if not args.no_cg:
        extra["cg_48c96"] = cg_leg()
    if not args.no_extra:
        extra["legacy_api_f32_32c64"] = legacy_leg()

    cpu = None
    if not args.no_cpu:
        host = host_cpu_info()
        threads = args.cpu_threads or host["affinity_cpus"]
        if host.get("cgroup_cpu_quota"):  # more threads than the CPU share only queue up
            threads = max(1, min(threads, int(host["cgroup_cpu_quota"])))
        c = cpu_leg(args.cpu_seconds, threads)
        extra["cpu_dslash"] = c
        a = c["f32_all_threads"]
        cpu = {"value": a["gflops"], "unit": "GFLOP/s", "cores": threads, "kind": "port",
               "sample": f"{a['calls']} calls of the host-CPU drop-in library (FP32 4D, 32x32x32x64, one "
                         f"checkerboard each, alternating isign and cb) on {threads} threads of "
                         f"{host.get('physical_cores')} physical cores ({host.get('model')}), median "
                         f"{a['ms_per_call']} ms/call; 1 thread: {c['f32_1_thread']['ms_per_call']} ms/call"}
        extra["gpu_vs_cpu_all_threads"] = round(head["gflops"] / a["gflops"], 1)

    gbs = head["gbs"]
    traffic = measured_traffic("dslash_wilson_f32_32c64_pmc.txt")
    line = {
        "metric": "Dslash GFLOP/s + % HBM roofline, 32^3x64; mixed vs FP64 CG time-to-1e-10",
        "value": round(head["gflops"], 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SU(3) gauge, Gaussian spinors, generated on device)",
        "config": {"workload": "wilson_dslash_4d_fp32_32x32x32x64", "lattice": list(LATT),
                   "sites_per_step": int(32 * 32 * 32 * 64 // 2), "flop_per_site": 1320,
                   "gauge_reals": 18, "steps_alternate": "isign +-1, source cb 0/1"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
