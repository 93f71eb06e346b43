#!/usr/bin/env python
"""Benchmark: FDK cone-beam backprojection (cbct_back 'fdk,ts,back') on one MI355X.

Default workload (the one the headline metric is quoted on), ``clinical_arc``:
arc detector 888 x 64 at ds = 1.0239 mm, dt = 1.0964 mm, quarter-detector
offset 1.25, dso = 541 mm, dsd = 949 mm, 984 views over 360 deg; image
512 x 512 x 256 at 0.7 mm in-plane (dy = -dx) and 0.625 mm in z; 30 seeded
ellipsoids, analytic line integrals, Poisson noise at I0 = 2e5, then the FDK
cosine weight and ramp filter.  The filtered views are stored on the device
in FP16 with a per-view scale (FP32 and BF16 are timed beside it).

One step = one backprojection of all views into the whole volume (the plan's
graph: a zero fill of the slices no view reaches, the zero-padded copy of the
views and the tiled kernel), projections and image resident in HBM.
value = nx*ny*nz*na / t.  The roofline entry is for the backprojection kernel
alone, timed with HIP events on its own stream (fdkhip_plan_profile).

``--workload uneven_scans`` times 32 seeded scans of uneven size in one batch
execute (FdkBatch, autotuned: the persistent kernel or one launch per scan,
replayed from a HIP graph) against the same scans run one untuned plan after
another; other workloads are listed in mirt_amd/workloads.py.

Prints ONE JSON line.  The CPU baseline is the reference C (oracle/_ref,
fdk-helix axial path, pthreads) timed on this host's cores on a stated sample
of the same workload; relative errors against the FP64 oracle are measured on
a seeded sample of columns (all slices, all views).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector peak, AMD spec
FLOP_PER_UPDATE = 14         # wu, wd, 2x(2 mul + add), vertical lerp (3), x w2, +=, t_bin +=
HEADLINE_METRIC = ("voxel-view updates/s (GUPS), 512²×256 ×984 views; "
                   "fp16-stored vs fp32 at rel-L2")

# storage/math path timed as the headline of each workload
HEADLINE = {"clinical_arc": "f16", "large_flat": "f16", "plain_flat": "f32_exact",
            "cpu_small": "f32_exact", "uneven_scans": "f16"}
PATHS = {"f32_exact": ("f32", "exact"), "f32_fast": ("f32", "fast"),
         "f16": ("f16", "fast"), "bf16": ("bf16", "fast")}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_threads():
    """Threads for the CPU reference: every core this process may run on,
    limited by OMP_NUM_THREADS where the host sets it (its CPU share)."""
    n = host_cores()
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def oracle_geom(cg, ig):
    from oracle import oracle as O
    return O.geom((ig.nx, ig.ny, ig.nz), (ig.dx, ig.dy, ig.dz),
                  (ig.offset_x, ig.offset_y, ig.offset_z), cg.dso, cg.dsd, cg.dfs,
                  cg.ns, cg.nt, (cg.ds, cg.dt), (cg.offset_s, cg.offset_t))


def profiled_traffic(workload, fmt, cfg, kernel="back_kernel"):
    """HBM bytes per launch of the backprojection kernel from the committed
    rocprofv3 summary for this workload and path (FETCH_SIZE and WRITE_SIZE
    passes, scaled by the copy-kernel calibration), if the summary's launch
    shape (waves, slices per lane) matches the one timed here; else None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_{fmt}.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
    if m and cfg["waves"] == 0:
        # a batch in one launch per scan: the summary covers one execute
        return float(m.group(1)) if re.search(r"^launch:\s*per_scan", text, re.M) else None
    shape = re.search(r"^launch:\s*waves=(\d+) zper=(\d+) vstage=(\d+)", text, re.M)
    if not m or not shape:
        return None
    waves, zper, _vstage = (int(x) for x in shape.groups())
    if (waves, zper) != (cfg["waves"], cfg["zper"]):
        return None
    return float(m.group(1))


def profiled_valu_issue(name):
    """VALU issue utilisation (fraction of the CU's wave64 VALU slots used)
    stated in the committed rocprofv3 summary profiles/<name>.md, or None."""
    path = os.path.join(ROOT, "profiles", f"{name}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"(\d+) % VALU issue utilisation", open(path).read())
    return int(m.group(1)) / 100 if m else None


def prepare(tsa, fmt):
    from mirt_amd import quantize_views
    if fmt == "f32":
        return (tsa,)
    return quantize_views(tsa, fmt)


def time_steps(fn, steps, warmup):
    import torch
    s = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps      # ms per step


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_reference_rate(g, tsa_host, beta, mask, nx_ny_nz, target_s=15.0, nthread=None):
    """Reference C on this host on an evenly spaced sample of views."""
    from oracle import oracle as O
    if not O.have_reference("fixed"):
        return None
    if nthread is None:
        nthread = cpu_threads()
    na = tsa_host.shape[0]
    probe = np.arange(0, na, max(1, na // 4))[:4]
    O.ref_back_ts("fixed", g, tsa_host[probe], beta[probe], mask, nthread=nthread)  # warm-up
    _, tp = O.ref_back_ts("fixed", g, tsa_host[probe], beta[probe], mask,
                          nthread=nthread, timed=True)
    per_view = tp / len(probe)
    nv = int(min(na, max(4, target_s / per_view)))
    step = max(1, na // nv)
    views = np.arange(0, na, step)[:nv]
    times = []
    for _ in range(3):
        _, t = O.ref_back_ts("fixed", g, tsa_host[views], beta[views], mask,
                             nthread=nthread, timed=True)
        times.append(t)
        if sum(times) > 2.5 * target_s:
            break
    t = float(np.median(times))
    nx, ny, nz = nx_ny_nz
    return {"value": nx * ny * nz * len(views) / t, "unit": "voxel-view updates/s",
            "cores": nthread, "kind": "reference",
            "sample": f"reference C (fdk-helix fdk_ts_back_t, gcc -O3, {nthread} pthreads) "
                      f"on {len(views)} of {na} views (every {step}th), full "
                      f"{nx}x{ny}x{nz} volume; median of {len(times)} runs of {t:.2f} s; "
                      f"host CPU {cpu_model()}, {host_cores()} cores available",
            "seconds": t}


def end_to_end_ms(g_mex, tsa_host, beta, reps=2):
    """Host buffers in, host image out through the drop-in C-ABI entry
    (fdkhip_fdk_ts_back: views to the device through pinned chunks, FP32
    EXACT kernel, image back the same way; the plan and device buffers of
    the first call are reused by the next): wall milliseconds, best of
    `reps`."""
    from mirt_amd import fdk_mex
    ptsa = np.asfortranarray(tsa_host.transpose(2, 1, 0))   # [nt ns na]
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        fdk_mex("fdk,ts,back", *g_mex, ptsa, beta, np.array([1], np.int32))
        t = (time.perf_counter() - t0) * 1e3
        best = t if best is None else min(best, t)
    return best


def column_errors(g, tsa_host, beta, mask, images, ncol=1024, seed=7):
    """Relative L2 of each GPU image against the FP64 oracle and the FP32
    restatement of the reference kernel, on seeded columns."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    ncol = min(ncol, g.nx * g.ny)
    cols = np.sort(rng.choice(g.nx * g.ny, ncol, replace=False)).astype(np.int32)
    r64 = O.back_ts_f64(g, tsa_host, beta, mask, cols=cols)
    r32 = O.back_ts_f32(g, tsa_host, beta, mask, cols=cols)
    errs = {"columns_sampled": int(cols.size)}
    for name, img in images.items():
        sub = img.reshape(-1, g.nz)[cols]
        errs[f"relL2_{name}_vs_fp64"] = O.rel_l2(sub, r64)
        errs[f"relL2_{name}_vs_reference_c"] = O.rel_l2(sub, r32)
        errs[f"maxabs_{name}_vs_reference_c"] = float(np.abs(sub - r32).max())
    errs["relL2_reference_c_vs_fp64"] = O.rel_l2(r32, r64)
    return errs


def filter_times(args, cg, dev):
    """The FDK weight + ramp filter before the backprojection, on synthetic
    FP32 views of the workload's size: fdkhip_fdk_filter_views (one fused
    kernel: weight, FFT of the zero-padded rows in LDS, response, inverse
    FFT, transpose) against torch's FP32 FFT route (feldkamp_weight1 +
    fdk_filter + the permute to [nt ns na]).  Bound: HBM, one read of the
    views and one write of the result."""
    import torch
    from mirt_amd.fdk import fdk_filter, feldkamp_weight1, weight_filter_views
    ns, nt, na = cg.ns, cg.nt, cg.na
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    sta = torch.randn((na, nt, ns), generator=g, device=dev, dtype=torch.float32)
    native = time_steps(lambda: weight_filter_views(sta, cg), args.steps, args.warmup)
    p = sta.permute(2, 1, 0)                                   # (ns, nt, na) view
    fft = time_steps(lambda: fdk_filter(feldkamp_weight1(p, cg), cg).permute(2, 0, 1)
                     .contiguous(), args.steps, args.warmup)
    nbytes = 8.0 * ns * nt * na
    gbs = nbytes / (native * 1e-3) / 1e9
    log(f"fdk filter: {native:.3f} ms fused kernel ({gbs:.0f} GB/s), "
        f"{fft:.3f} ms torch FP32 FFT")
    return {"ms": native, "path": "fdkhip_fdk_filter_views (fused weight + FFT filter + "
                                  "transpose)", "torch_fft_f32_ms": fft,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                         "frac": gbs / 8000.0, "algorithmic_bytes": nbytes},
            "note": f"{ns}x{nt}x{na} FP32 views; algorithmic bytes = 4 B read + 4 B written "
                    "per sample; time includes the output allocation"}


def run_single(args, dev):
    import torch
    from mirt_amd import FdkPlan, geom_of
    from mirt_amd.cbct_back import _views
    from mirt_amd.workloads import WORKLOADS, filtered_views

    scan = WORKLOADS[args.workload]()
    cg, ig = scan.cg, scan.ig
    t0 = time.perf_counter()
    tsa = filtered_views(scan, dev)
    torch.cuda.synchronize()
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s")
    beta = _views(cg, cg.na)
    geom = geom_of(cg, ig)
    mask = ig.mask_or
    head = args.path or HEADLINE[args.workload]
    order = [head] + ([] if args.fmt_only else [p for p in PATHS if p != head])

    U = ig.nx * ig.ny * ig.nz * cg.na
    paths, images, kernel = {}, {}, {}
    U_eff = cfg = None
    for name in order:
        fmt, mm = PATHS[name]
        plan = FdkPlan(geom, beta, mask, fmt=fmt, math_mode=mm)
        a = prepare(tsa, fmt)
        out = torch.empty((ig.ny, ig.nx, ig.nz), dtype=torch.float32, device=dev)
        if args.launch:
            plan.set_config(*args.launch[:3])
            if len(args.launch) > 3:
                plan.set_occupancy(args.launch[3])
        elif not args.no_autotune:
            t0 = time.perf_counter()
            best = plan.autotune(*a, out=out, reps=2)
            log(f"{name}: autotune {time.perf_counter() - t0:.1f} s -> {plan.config}, "
                f"{plan.occupancy} waves/SIMD ({best:.3f} ms)")
        paths[name] = time_steps(lambda: plan(*a, out=out), args.steps, args.warmup)
        kernel[name] = plan.profile(*a, out=out, reps=max(2, args.steps))
        images[name] = out
        if name == head:
            U_eff = plan.count_updates()
            cfg = dict(plan.config, occupancy=plan.occupancy)
        log(f"{name}: {paths[name]:.3f} ms/step, kernel {kernel[name]['back']:.3f} ms, "
            f"{U / paths[name] * 1e3:.3e} updates/s, config {plan.config}")
        del plan, a

    ms = paths[head]
    kms = kernel[head]["back"]
    achieved = FLOP_PER_UPDATE * U_eff / (kms * 1e-3) / 1e12
    traffic = profiled_traffic(args.workload, head, cfg)
    valu_issue = (profiled_valu_issue(f"back_kernel_{args.workload}_{head}")
                  if traffic is not None else None)
    det = "arc" if cg.dfs == 0 else "flat"
    result = {
        "metric": HEADLINE_METRIC if args.workload == "clinical_arc"
        else f"voxel-view updates/s (GUPS), {args.workload}",
        "value": U / (ms * 1e-3),
        "unit": "voxel-view updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: {len(scan.ellipsoids)} seeded ellipsoids, analytic {det} "
                f"cone-beam line integrals"
                + (f", Poisson I0={scan.i0:g}" if scan.i0 else ", noise-free")
                + f", FDK weight + ramp filter, {head} storage"
                + (" with per-view scale" if head in ("f16", "bf16") else ""),
        "config": {"workload": f"{args.workload}: cbct_back fdk,ts,back, {det} "
                               f"{cg.ns}x{cg.nt} det, {cg.na} views/360 deg, dso "
                               f"{cg.dso:g} dsd {cg.dsd:g} -> {ig.nx}x{ig.ny}x{ig.nz} "
                               f"({ig.dx:g}/{ig.dz:g} mm)",
                   "nx": ig.nx, "ny": ig.ny, "nz": ig.nz, "na": cg.na, "ns": cg.ns,
                   "nt": cg.nt, "storage": head, "accumulate": "f32", "launch": cfg},
        "gups_nominal": U / (ms * 1e-3) / 1e9,
        "gups_effective": U_eff / (ms * 1e-3) / 1e9,
        "updates_nominal": U,
        "updates_effective": U_eff,
        "paths_ms_per_step": paths,
        "kernel_ms": kernel,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic, "valu_issue": valu_issue,
                     "note": f"{FLOP_PER_UPDATE} FP32 flop x {U_eff} in-support updates per "
                             f"launch / {kms:.3f} ms back_kernel (HIP events on its "
                             "stream); traffic = HBM bytes per launch from the "
                             "rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/, "
                             "valu_issue = SQ_INSTS_VALU / (CUs x GRBM_GUI_ACTIVE cycles) "
                             "from the same summary"},
    }
    if not args.fmt_only:
        result["fdk_filter"] = filter_times(args, cg, dev)
    if args.slabs and args.slabs > 1:
        result["slabs"] = slab_times(args, geom, beta, mask, ig, tsa, head, cfg, "z")
        result["strips"] = slab_times(args, geom, beta, mask, ig, tsa, head, cfg, "y")
    ref = images.get("f32_exact")
    if ref is not None:
        for name in images:
            if name != "f32_exact":
                d = (images[name] - ref).double()
                result[f"relL2_{name}_vs_f32_exact_gpu"] = float(
                    d.norm() / ref.double().norm())
    result["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        g = oracle_geom(cg, ig)
        tsa_host = tsa.cpu().numpy()
        m = np.ascontiguousarray(np.asarray(mask).T).astype(np.uint8)
        t0 = time.perf_counter()
        result["cpu_baseline"] = cpu_reference_rate(g, tsa_host, beta, m,
                                                    (ig.nx, ig.ny, ig.nz))
        one = cpu_reference_rate(g, tsa_host, beta, m, (ig.nx, ig.ny, ig.nz),
                                 target_s=5.0, nthread=1)
        if result["cpu_baseline"] is not None and one is not None:
            result["cpu_baseline"]["value_1thread"] = one["value"]
            result["cpu_baseline"]["sample_1thread"] = one["sample"]
        g_mex = (np.array([ig.nx, ig.ny, ig.nz], np.int32),
                 np.array([ig.dx, ig.dy, ig.dz], np.float64),
                 np.array([ig.offset_x, ig.offset_y, ig.offset_z], np.float64),
                 np.asfortranarray(np.asarray(mask) != 0).astype(np.uint8),
                 float(cg.dso), float(cg.dsd), float(cg.dfs),
                 np.array([cg.ds, cg.dt], np.float64),
                 np.array([cg.offset_s, cg.offset_t], np.float64))
        e2e = end_to_end_ms(g_mex, tsa_host, np.ascontiguousarray(beta, np.float64))
        result["end_to_end"] = {
            "ms": e2e, "value": U / (e2e * 1e-3), "unit": "voxel-view updates/s",
            "path": "host buffers through fdkhip_fdk_ts_back (views in and image out "
                    "through pinned chunks, FP32 EXACT kernel; the first call's plan and "
                    "device buffers reused, best of 2 calls); not the headline value"}
        log(f"end to end (host buffers, FP32 EXACT): {e2e:.1f} ms")
        imgs = {k: v.cpu().numpy() for k, v in images.items()}
        result["errors"] = column_errors(g, tsa_host, beta, m, imgs)
        if not args.no_full_error:
            result["errors"].update(volume_errors(g, tsa_host, beta, m, imgs, head))
        log(f"cpu baseline + oracle errors in {time.perf_counter() - t0:.1f} s")
    if args.workload == "clinical_arc" and not args.fmt_only and not args.no_batch_leg:
        result["uneven_scans"] = batch_leg(args, dev)
    return result


def volume_errors(g, tsa_host, beta, mask, images, head):
    """Relative L2 and max-abs of the headline image over the whole volume
    against the FP64 oracle on the same FP32 views (multithreaded C)."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    r64 = O.back_ts_f64(g, tsa_host, beta, mask)
    t = time.perf_counter() - t0
    out = {"whole_volume_seconds_fp64": t}
    for name, img in images.items():
        out[f"relL2_{name}_vs_fp64_whole_volume"] = O.rel_l2(img, r64)
        out[f"maxabs_{name}_vs_fp64_whole_volume"] = float(np.abs(img - r64).max())
    out["maxabs_fp64_whole_volume"] = float(np.abs(r64).max())
    log(f"whole-volume FP64 oracle in {t:.1f} s: {head} relL2 "
        f"{out[f'relL2_{head}_vs_fp64_whole_volume']:.3e}")
    return out


# per in-range voxel-view update of cbct_back_mat, in MATLAB's expression
# order: tprime (sub, mul), bt (div, add), wu, wd, p1 (2 mul + add), p2 (same),
# p0 (2 mul + add), the weight (1 mul flat / mul + div arc), img2 += p0
MAT_FLOP_PER_UPDATE = {"flat": 17, "arc": 18}
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector peak, AMD spec


def mat_cpu_rate(g, proj_sta, beta, mask, target_s=15.0):
    """The oracle's numpy restatement of cbct_back_mat (one thread) on a slab
    of slices and a sample of views: voxel-view updates/s."""
    from oracle import oracle as O

    def slab(nzs):
        gs = O.geom((g.nx, g.ny, nzs), (g.dx, g.dy, g.dz), (g.offset_x, g.offset_y, 0.0),
                    g.dso, g.dsd, g.dfs, g.ns, g.nt, (g.ds, g.dt), (g.offset_s, g.offset_t))
        return gs

    t0 = time.perf_counter()
    O.back_mat_f64(slab(4), proj_sta[:, :, :1], beta[:1], mask=mask)
    per = (time.perf_counter() - t0) / 4                    # s per slice-view
    nzs = int(max(4, min(g.nz, target_s / per / 4)))
    nv = int(max(1, min(proj_sta.shape[2], target_s / per / nzs)))
    views = np.linspace(0, proj_sta.shape[2] - 1, nv).astype(int)
    t0 = time.perf_counter()
    O.back_mat_f64(slab(nzs), proj_sta[:, :, views], beta[views], mask=mask)
    t = time.perf_counter() - t0
    return {"value": g.nx * g.ny * nzs * nv / t, "unit": "voxel-view updates/s",
            "cores": 1, "kind": "port",
            "sample": f"numpy restatement of cbct_back_mat (oracle.back_mat_f64, one thread) "
                      f"on {nv} of {proj_sta.shape[2]} views x a {g.nx}x{g.ny}x{nzs} slab, "
                      f"{t:.1f} s; host CPU {cpu_model()}",
            "seconds": t}


def run_mat(args, dev):
    """cbct_back_mat (use_mex = 0) on the workload's filtered views: the FP64
    MATLAB-rule backprojector, views [ns nt na] float32 and the float64 image
    resident on the device (fdkhip_cbct_back_mat_direct)."""
    import ctypes as C
    import torch
    from mirt_amd import _lib, geom_of
    from mirt_amd.cbct_back import _views
    from mirt_amd.workloads import WORKLOADS, filtered_views

    lib = _lib.load()
    scan = WORKLOADS[args.workload]()
    cg, ig = scan.cg, scan.ig
    tsa = filtered_views(scan, dev)                       # (na, ns, nt)
    sta = tsa.permute(0, 2, 1).contiguous()               # (na, nt, ns) = [ns nt na]
    del tsa
    beta = np.ascontiguousarray(_views(cg, cg.na), np.float64)
    mask = np.asfortranarray(ig.mask_or != 0).astype(np.uint8, order="F")
    img = torch.empty((ig.nz, ig.ny, ig.nx), dtype=torch.float64, device=dev)
    geom = geom_of(cg, ig)
    s = torch.cuda.current_stream()
    kms = []

    def step():
        ms = C.c_float(0)
        _lib.check(lib.fdkhip_cbct_back_mat_direct(
            C.byref(geom), C.c_void_p(sta.data_ptr()), 0, cg.na, beta.ctypes.data_as(C.c_void_p),
            None, 1, 0.0, mask.ctypes.data_as(C.c_void_p), 1.0, C.c_void_p(img.data_ptr()),
            C.c_void_p(s.cuda_stream), C.byref(ms)))
        kms.append(ms.value)

    ms = time_steps(step, args.steps, args.warmup)
    kernel_ms = float(np.mean(kms[args.warmup:]))
    U = ig.nx * ig.ny * ig.nz * cg.na
    nmask = int(mask.sum())
    U_mask = nmask * ig.nz * cg.na                        # cbct_back_mat visits masked voxels
    det = "arc" if cg.dfs == 0 else "flat"
    achieved = MAT_FLOP_PER_UPDATE[det] * U_mask / (kernel_ms * 1e-3) / 1e12
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"mat_kernel_{args.workload}_f64.md")
    if os.path.exists(prof):
        m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(prof).read(), re.M)
        traffic = float(m.group(1)) if m else None
    log(f"mat: {ms:.3f} ms/step, kernel {kernel_ms:.3f} ms, {U / ms * 1e3:.3e} updates/s")
    result = {
        "metric": f"voxel-view updates/s (GUPS), {args.workload}, cbct_back_mat "
                  "(use_mex=0, FP64 MATLAB backprojector)",
        "value": U / (ms * 1e-3), "unit": "voxel-view updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": f"synthetic: {args.workload} FDK-filtered phantom views, float32 [ns nt na]",
        "config": {"workload": f"{args.workload} through cbct_back_mat: {det} "
                               f"{cg.ns}x{cg.nt} det, {cg.na} views -> "
                               f"{ig.nx}x{ig.ny}x{ig.nz}, every voxel x every view",
                   "nx": ig.nx, "ny": ig.ny, "nz": ig.nz, "na": cg.na},
        "kernel_ms": kernel_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic,
                     "note": f"{MAT_FLOP_PER_UPDATE[det]} FP64 flop x {U_mask} masked "
                             f"voxel-view updates / {kernel_ms:.3f} ms mat_kernel (HIP "
                             "events on its stream)"},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        g = oracle_geom(cg, ig)
        proj_sta = sta.cpu().numpy().transpose(2, 1, 0)   # (ns, nt, na) view
        result["cpu_baseline"] = mat_cpu_rate(g, proj_sta, beta, ig.mask_or)
    return result


def helix_update_count(cg, ig, beta, zs, izb, dev, z_range=None):
    """Voxel-view updates of the reference's helical loop (fdk,ts.c:112-245:
    columns whose s bin is in [-1, ns), slices [iz_min, iz_max) inside the
    view window), evaluated in float32 with torch on the device -- the count
    behind the updates/s and the roofline of the helical runs (float rounding
    of the slice bounds may move it by a few 1e-6)."""
    import torch
    f = dict(device=dev, dtype=torch.float32)
    nx, ny, nz = ig.nx, ig.ny, ig.nz
    wx = (nx - 1) / 2 + ig.offset_x
    wy = (ny - 1) / 2 + ig.offset_y
    wz = (nz - 1) / 2 + ig.offset_z
    ws = (cg.ns - 1) / 2 + cg.offset_s
    wt = (cg.nt - 1) / 2 + cg.offset_t
    xx = (ig.dx * (torch.arange(nx, **f) - wx))[None, None, :]
    yy = (ig.dy * (torch.arange(ny, **f) - wy))[None, :, None]
    m = torch.as_tensor(np.ascontiguousarray(np.asarray(ig.mask_or).T), device=dev)
    lo_z, hi_z = z_range or (0, nz)
    total = 0
    for v0 in range(0, len(beta), 32):
        b = torch.as_tensor(beta[v0:v0 + 32], **f)[:, None, None]
        z = torch.as_tensor(zs[v0:v0 + 32], **f)[:, None, None]
        i0 = torch.as_tensor(izb[0][v0:v0 + 32], device=dev)[:, None, None]
        i1 = torch.as_tensor(izb[1][v0:v0 + 32], device=dev)[:, None, None]
        cb, sb = torch.cos(b), torch.sin(b)
        xb = xx * cb + yy * sb
        yb = cg.dso - (-xx * sb + yy * cb)
        mag = cg.dsd / yb
        ss = cg.dsd * torch.atan2(xb, yb) if cg.dfs == 0 else mag * xb
        isb = torch.floor(ss / cg.ds + ws)
        ok = (isb >= -1) & (isb < cg.ns) & (m[None] != 0)
        t_inc = mag * ig.dz / cg.dt
        wzb = wz + z / ig.dz
        izmin = torch.floor(wzb - (wt + 1) / t_inc) + 1
        izmax = torch.ceil(wzb + (cg.nt - wt) / t_inc)
        izmin = torch.maximum(torch.maximum(izmin, i0.to(izmin.dtype)),
                              torch.tensor(float(lo_z), **f))
        izmax = torch.minimum(torch.minimum(izmax, i1.to(izmax.dtype)),
                              torch.tensor(float(hi_z), **f))
        total += int(torch.where(ok, (izmax - izmin).clamp(min=0), 0).sum().item())
    return total


def helix_windows(cg, ig, na1_half):
    """iz_min_beta / iz_max_beta of every view (fdk,ts,t.c:78-85)."""
    zs = np.asarray(cg.source_zs, np.float64)
    na = len(zs)
    wz = np.float32((ig.nz - 1) / 2 + np.float32(ig.offset_z))
    dz = np.float32(ig.dz)
    a0 = np.maximum(np.arange(na) - na1_half, 0)
    a1 = np.minimum(np.arange(na) + na1_half, na - 1)
    c0 = np.ceil((wz + zs[a0] / dz).astype(np.float32))
    c1 = np.ceil((wz + zs[a1] / dz).astype(np.float32))
    return np.maximum(c0, 0).astype(np.int64), np.minimum(c1, ig.nz).astype(np.int64)


def run_helix(args, dev):
    """The helical 'fdk,ts,back' of the fdk-helix kernel on a helical scan
    (clinical_helix: 3 turns of the clinical_arc detector) in native helical
    geometry: views [nt ns na] FP32 and the [nz nx ny] image resident on the
    device, a helical plan (fdkhip_plan_create_helix: the axial pipeline with
    per-view source heights and slice windows, EXACT arithmetic), autotuned,
    replayed from its HIP graph."""
    import torch
    from mirt_amd import FdkPlan
    from mirt_amd.cbct_back import _views
    from mirt_amd.cbct_back_helix import helix_operands
    from mirt_amd.workloads import WORKLOADS, filtered_views

    name = args.workload if args.workload != "clinical_arc" else "clinical_helix"
    scan = WORKLOADS[name]()
    cg, ig = scan.cg, scan.ig
    t0 = time.perf_counter()
    tsa = filtered_views(scan, dev)                       # (na, ns, nt)
    torch.cuda.synchronize()
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s")
    beta = np.ascontiguousarray(_views(cg, cg.na), np.float64)
    zs = np.ascontiguousarray(cg.source_zs, np.float64)
    na1_half, dpr = helix_operands(cg, 1)
    mask = np.ascontiguousarray(np.asarray(ig.mask_or).T).astype(np.uint8)
    img = torch.empty((ig.ny, ig.nx, ig.nz), dtype=torch.float32, device=dev)
    plan = FdkPlan.helix_from_geoms(cg, ig)
    if args.launch:
        plan.set_config(*args.launch[:3])
        if len(args.launch) > 3:
            plan.set_occupancy(args.launch[3])
    elif not args.no_autotune:
        t0 = time.perf_counter()
        plan.autotune(tsa, out=img, reps=2)
        log(f"helix: autotune {time.perf_counter() - t0:.1f} s -> {plan.config}")
    cfg = plan.config
    launch = f"waves={cfg['waves']} zper={cfg['zper']} vstage={cfg['vstage']}"
    ms = time_steps(lambda: plan(tsa, out=img), args.steps, args.warmup)
    prof_ms = plan.profile(tsa, out=img, reps=max(2, args.steps))
    kernel_ms = prof_ms["back"]
    izb = helix_windows(cg, ig, na1_half)
    U = helix_update_count(cg, ig, beta, zs, izb, dev)
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"back_kernel_{name}_f32.md")
    if os.path.exists(prof):
        text = open(prof).read()
        m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
        if m and re.search(r"^launch: " + re.escape(launch) + "$", text, re.M):
            traffic = float(m.group(1))
    U_nom = ig.nx * ig.ny * ig.nz * cg.na
    achieved = FLOP_PER_UPDATE * U / (kernel_ms * 1e-3) / 1e12
    log(f"helix: {ms:.3f} ms/step, back_kernel {kernel_ms:.3f} ms ({launch}), "
        f"pad {prof_ms['pad']:.3f} ms, {U:.4e} updates ({U / ms * 1e3:.3e}/s); "
        f"plan counts {plan.count_updates():.4e}")
    result = {
        "metric": f"voxel-view updates/s (in-window updates), {name}, helical fdk,ts,back",
        "value": U / (ms * 1e-3), "unit": "voxel-view updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: {len(scan.ellipsoids)} seeded ellipsoids, analytic helical arc "
                f"cone-beam line integrals, Poisson I0={scan.i0:g}, FDK weight + ramp filter",
        "config": {"workload": f"{name}: arc {cg.ns}x{cg.nt} det, {cg.na} views over "
                               f"{cg.orbit:g} deg, pitch {cg.pitch:g} -> "
                               f"{ig.nx}x{ig.ny}x{ig.nz} ({ig.dx:g}/{ig.dz:g} mm)",
                   "nx": ig.nx, "ny": ig.ny, "nz": ig.nz, "na": cg.na, "na1_half": na1_half},
        "updates_in_window": U, "updates_nominal": U_nom,
        "gups_nominal": U_nom / (ms * 1e-3) / 1e9,
        "kernel_ms": kernel_ms, "launch": launch, "pad_ms": prof_ms["pad"],
        "path": "helical plan (fdkhip_plan_create_helix), EXACT",
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic,
                     "valu_issue": (profiled_valu_issue(f"back_kernel_{name}_f32")
                                    if traffic is not None else None),
                     "note": f"{FLOP_PER_UPDATE} FP32 flop x {U} in-window updates / "
                             f"{kernel_ms:.3f} ms back_kernel (HIP events on its stream)"},
        "cpu_baseline": None,
    }
    # the other arithmetic on the same helical plans: FAST FP32 and FP16/BF16
    # storage, each against the EXACT image over the whole volume
    from mirt_amd import quantize_views
    ref = img.clone()
    zs_all = np.broadcast_to(np.asarray(cg.source_zs, np.float64), (cg.na,))
    others = {}
    for key, fmt, mm in (("f32_fast", "f32", "fast"), ("f16", "f16", "fast"),
                         ("bf16", "bf16", "fast")):
        inp = (tsa,) if fmt == "f32" else quantize_views(tsa, fmt)
        p2 = FdkPlan(plan.geom, beta, np.asarray(ig.mask_or), fmt=fmt, math_mode=mm,
                     helix=(zs_all, na1_half, float(cg.pitch)))
        if not args.no_autotune:
            p2.autotune(*inp, out=img, reps=2)
        ms2 = time_steps(lambda: p2(*inp, out=img), args.steps, args.warmup)
        k2 = p2.profile(*inp, out=img, reps=max(2, args.steps))["back"]
        p2(*inp, out=img)
        torch.cuda.synchronize()
        err = float(torch.linalg.vector_norm(img - ref) / torch.linalg.vector_norm(ref))
        c2 = p2.config
        others[key] = {"ms_per_step": ms2, "kernel_ms": k2,
                       "value": U / (ms2 * 1e-3), "relL2_vs_exact": err,
                       "launch": f"waves={c2['waves']} zper={c2['zper']} vstage={c2['vstage']}"}
        log(f"helix {key}: {ms2:.3f} ms/step, back_kernel {k2:.3f} ms, relL2 vs EXACT {err:.2e}")
        del p2
    result["other_paths"] = others
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        if O.have_reference("fixed"):
            nthread = cpu_threads()
            tsa_host = tsa.cpu().numpy()
            g = oracle_geom(cg, ig)

            def slab(nzs):
                z0 = (ig.nz - nzs) // 2      # central slices, same z coordinates
                off = ig.offset_z + (ig.nz - 1) / 2 - z0 - (nzs - 1) / 2
                gs = O.geom((ig.nx, ig.ny, nzs), (ig.dx, ig.dy, ig.dz),
                            (ig.offset_x, ig.offset_y, off), cg.dso, cg.dsd, cg.dfs,
                            cg.ns, cg.nt, (cg.ds, cg.dt), (cg.offset_s, cg.offset_t))
                t = time.perf_counter()
                O.ref_back_ts_helix(gs, tsa_host, beta, zs, mask, na1_half=na1_half,
                                    source_dz_per_rad=dpr, pitch=float(cg.pitch),
                                    nthread=nthread)
                return time.perf_counter() - t, z0

            tp, _ = slab(8)                  # per-view set-up included: conservative
            nzs = int(max(8, min(ig.nz, 8 * 15.0 / max(tp, 1e-3))))
            t, z0 = slab(nzs)
            u = helix_update_count(cg, ig, beta, zs, izb, dev, z_range=(z0, z0 + nzs))
            result["cpu_baseline"] = {
                "value": u / t, "unit": "voxel-view updates/s (in-window)",
                "cores": nthread, "kind": "reference",
                "sample": f"reference C (fdk-helix fdk_ts_back_t, gcc -O3, {nthread} "
                          f"pthreads), all {cg.na} views, central {nzs} of {ig.nz} slices "
                          f"({u:.3e} in-window updates) in {t:.2f} s; host CPU {cpu_model()}",
                "seconds": t}
    return result


def slab_times(args, geom, beta, mask, ig, tsa, head, cfg, axis):
    """A G-way split of the volume for G GPUs, shown on this GPU: part g is
    the same scan with nz/G slices (axis "z": z-slabs) or ny/G image rows
    (axis "y": strips of columns), the image centre moved so that the part
    covers slices / rows g*n/G ... (g+1)*n/G - 1 of the full volume.  Each part
    is backprojected by its own plan with the full volume's launch shape; the
    slowest part is what G GPUs would take (no reduction: parts concatenate).
    """
    import torch
    from mirt_amd import FdkPlan
    from mirt_amd._lib import Geom
    G = args.slabs
    fmt, mm = PATHS[head]
    a = prepare(tsa, fmt)
    n = ig.nz if axis == "z" else ig.ny
    ns_ = n // G
    rows = []
    for gi in range(G):
        gs = Geom.from_buffer_copy(geom)
        msk = mask
        if axis == "z":
            gs.nz = ns_
            gs.offset_z = geom.offset_z + (ig.nz - ns_) / 2.0 - gi * ns_
        else:
            # y is the image's second axis; dy < 0 flips nothing here: the
            # part is rows gi*ns_ ... of the mask and of the volume
            gs.ny = ns_
            gs.offset_y = geom.offset_y + (ig.ny - ns_) / 2.0 - gi * ns_
            msk = np.asarray(mask)[:, gi * ns_:(gi + 1) * ns_]
        plan = FdkPlan(gs, beta, msk, fmt=fmt, math_mode=mm)
        plan.set_config(cfg["waves"], cfg["zper"], cfg["vstage"])
        if "occupancy" in cfg:
            plan.set_occupancy(cfg["occupancy"])
        out = torch.empty((gs.ny, gs.nx, gs.nz), dtype=torch.float32, device=tsa.device)
        ms = plan.profile(*a, out=out, reps=max(2, args.steps))["back"]
        rows.append({"part": gi, axis: [gi * ns_, (gi + 1) * ns_], "kernel_ms": ms,
                     "updates_effective": plan.count_updates(), "launch": plan.config})
        del plan, out
    slowest = max(r["kernel_ms"] for r in rows)
    log(f"{G} parts along {axis} of {ns_}: slowest {slowest:.3f} ms, "
        f"sum {sum(r['kernel_ms'] for r in rows):.3f} ms")
    return {"G": G, "axis": axis, "per_part": ns_, "slowest_kernel_ms": slowest,
            "parts": rows}


def batch_leg(args, dev):
    """The uneven_scans batch as an extra entry of the default run: all 32
    scans, FP16 storage, autotuned, graph-replayed; ms, updates, and every
    scan's error against the FP64 oracle on seeded columns."""
    import copy
    a = copy.copy(args)
    a.workload, a.path, a.nscan, a.launch = "uneven_scans", "f16", 32, None
    a.steps, a.warmup = min(args.steps, 5), min(args.warmup, 1)
    a.fmt_only, a.no_cpu_baseline = True, True
    a.errors_scans, a.error_cols = 32, 128
    t0 = time.perf_counter()
    r = run_batch(a, dev, errors_without_baseline=True)
    errs = r.get("errors", {})
    worst = max((v["relL2_f16_vs_fp64"] for v in errs.values()), default=None)
    leg = {"metric": r["metric"], "value": r["value"], "unit": r["unit"],
           "ms_per_step": r["ms_per_step"], "steps": a.steps, "warmup": a.warmup,
           "updates_nominal": r["updates_nominal"], "updates_effective": r["updates_effective"],
           "gups_effective": r["updates_effective"] / (r["ms_per_step"] * 1e-3) / 1e9,
           "kernel_ms": r["kernel_ms"], "launch": r["config"]["launch"],
           "roofline": r["roofline"], "scans_checked": len(errs),
           "worst_relL2_f16_vs_fp64": worst,
           "columns_per_scan": a.error_cols,
           "seconds": time.perf_counter() - t0}
    log(f"uneven_scans leg: {leg['ms_per_step']:.3f} ms/step, worst scan relL2 vs FP64 "
        f"{worst if worst is None else f'{worst:.2e}'}, {leg['seconds']:.1f} s")
    return leg


def run_batch(args, dev, errors_without_baseline=False):
    import torch
    from mirt_amd import FdkBatch, FdkPlan, geom_of, quantize_views
    from mirt_amd.cbct_back import _views
    from mirt_amd.workloads import filtered_views, uneven_scans

    head = args.path or HEADLINE["uneven_scans"]
    fmt, mm = PATHS[head]
    scans = uneven_scans(args.nscan)
    t0 = time.perf_counter()
    projs, scales, hosts = [], [], []
    for sc in scans:
        tsa = filtered_views(sc, dev)
        if args.errors_scans > len(hosts):
            hosts.append(tsa.to("cpu", non_blocking=False).numpy())
        a = prepare(tsa, fmt)
        projs.append(a[0])
        scales.append(a[1] if len(a) > 1 else None)
        del tsa
    torch.cuda.synchronize()
    log(f"{len(scans)} scans ready in {time.perf_counter() - t0:.1f} s")
    scales = None if fmt == "f32" else scales
    betas = [_views(sc.cg, sc.cg.na) for sc in scans]
    geoms = [geom_of(sc.cg, sc.ig) for sc in scans]
    masks = [sc.ig.mask_or for sc in scans]
    batch = FdkBatch(list(zip(geoms, betas, masks)), fmt=fmt, math_mode=mm)
    outs = [torch.empty((sc.ig.ny, sc.ig.nx, sc.ig.nz), dtype=torch.float32, device=dev)
            for sc in scans]
    if args.launch:
        batch.set_config(*args.launch[:3])
    elif not args.no_autotune:
        t0 = time.perf_counter()
        tuned = batch.autotune(projs, scales, outs, reps=2)
        log(f"autotune {time.perf_counter() - t0:.1f} s: {tuned:.3f} ms, "
            f"config {batch.config}")
    if args.batch_form == "per_scan" and not batch.per_scan:
        # each scan's plan takes the launch shape its autotune remembered
        batch.set_per_scan()
        log("one launch per scan (--batch-form per_scan)")
    launch = dict(batch.config)
    if batch.per_scan:
        launch = {"mode": "one launch per scan", "workgroups": launch["workgroups"],
                  "slices_per_lane": batch.scan_zper}
    else:
        launch["mode"] = "persistent"
        launch["half_height_scans"] = sum(z < launch["zper"] for z in batch.scan_zper)
    log(f"per-scan slices per lane: {batch.scan_zper}")
    ms = time_steps(lambda: batch(projs, scales, outs), args.steps, args.warmup)
    prof = batch.profile(projs, scales, outs, reps=max(2, args.steps))
    U = sum(sc.ig.nx * sc.ig.ny * sc.ig.nz * sc.cg.na for sc in scans)
    U_eff = batch.count_updates()
    log(f"batch: {ms:.3f} ms/step, kernel {prof['back']:.3f} ms, {U / ms * 1e3:.3e} upd/s")

    # the same scans one plan after another (each replayed from its graph)
    ms_seq = None
    if not args.fmt_only:
        plans = [FdkPlan(geoms[i], betas[i], masks[i], fmt=fmt, math_mode=mm)
                 for i in range(len(scans))]

        def sequential():
            for i, p in enumerate(plans):
                p(projs[i], scales[i] if scales else None, out=outs[i])
        ms_seq = time_steps(sequential, args.steps, args.warmup)
        log(f"sequential plans: {ms_seq:.3f} ms/step")
        del plans
    batch(projs, scales, outs)
    torch.cuda.synchronize()

    kms = prof["back"]
    achieved = FLOP_PER_UPDATE * U_eff / (kms * 1e-3) / 1e12
    bkernel = "back_kernel" if batch.per_scan else "batch_kernel"
    btraffic = profiled_traffic("uneven_scans", head, batch.config, bkernel)
    result = {
        "metric": "voxel-view updates/s (GUPS), uneven_scans batch",
        "value": U / (ms * 1e-3),
        "unit": "voxel-view updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: {len(scans)} seeded uneven scans, 10-30 ellipsoids each, "
                f"analytic flat cone-beam line integrals, FDK weight + ramp filter, {head} "
                "storage",
        "config": {"workload": f"uneven_scans: {len(scans)} scans in one batch execute (one "
                               "graph replay; edge {128,192,256,384}, nz=edge/2, 180-720 "
                               "views, flat 256-768 x 128-512 det)",
                   "storage": head, "accumulate": "f32", "launch": launch},
        "updates_nominal": U, "updates_effective": U_eff,
        "sequential_plans_ms_per_step": ms_seq,
        "kernel_ms": prof,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": btraffic,
                     "valu_issue": (profiled_valu_issue(f"{bkernel}_uneven_scans_{head}")
                                    if btraffic is not None else None),
                     "note": f"{FLOP_PER_UPDATE} FP32 flop x {U_eff} in-support updates / "
                             f"{kms:.3f} ms of {bkernel} "
                             f"({'one launch per scan' if batch.per_scan else 'persistent'}; "
                             "HIP events); traffic and VALU issue per batch execute from the "
                             "rocprofv3 summary in profiles/"},
        "cpu_baseline": None,
    }
    if (not args.no_cpu_baseline or errors_without_baseline) and hosts:
        errs = {}
        for i, h in enumerate(hosts):
            sc = scans[i]
            g = oracle_geom(sc.cg, sc.ig)
            m = np.ascontiguousarray(np.asarray(masks[i]).T).astype(np.uint8)
            e = column_errors(g, h, betas[i], m, {head: outs[i].cpu().numpy()},
                              ncol=args.error_cols)
            errs[sc.name] = {k: v for k, v in e.items() if "fp64" in k}
            if i == 0 and not args.no_cpu_baseline:
                result["cpu_baseline"] = cpu_reference_rate(
                    g, h, betas[i], m, (sc.ig.nx, sc.ig.ny, sc.ig.nz), target_s=10.0)
            hosts[i] = None
        result["errors"] = errs
    return result


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clinical_arc",
                    choices=["clinical_arc", "plain_flat", "large_flat", "cpu_small",
                             "uneven_scans", "clinical_helix"])
    ap.add_argument("--path", choices=list(PATHS) + ["mat", "helix"], default=None,
                    help="storage/math path of the headline (default per workload); "
                         "mat: cbct_back_mat, the FP64 use_mex=0 route; helix: the "
                         "helical fdk,ts,back (workload clinical_helix)")
    ap.add_argument("--nscan", type=int, default=32, help="uneven_scans: number of scans")
    ap.add_argument("--slabs", type=int, default=0, metavar="G",
                    help="also time the G z-slabs a G-GPU split would give each GPU")
    ap.add_argument("--batch-form", choices=["auto", "per_scan"], default="auto",
                    help="uneven_scans: the form the batch autotuner picks, or one launch "
                         "per scan whatever it picks (profiling runs with counters, which "
                         "serialise the two streams)")
    ap.add_argument("--errors-scans", type=int, default=4,
                    help="uneven_scans: scans checked against the FP64 oracle")
    ap.add_argument("--error-cols", type=int, default=256,
                    help="uneven_scans: seeded columns per scan checked against FP64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full-error", action="store_true",
                    help="skip the whole-volume FP64 comparison of the headline")
    ap.add_argument("--no-batch-leg", action="store_true",
                    help="clinical_arc: skip the uneven_scans batch entry")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB three times (torch copy kernel): the reference "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the plans' default launch shape instead of the tuned one")
    ap.add_argument("--launch", type=lambda v: tuple(int(x) for x in v.split(",")),
                    default=None, metavar="WAVES,ZPER,VSTAGE[,WAVES_PER_SIMD]",
                    help="fixed launch shape (no autotune), e.g. 4,32,2,8 (profiling runs); "
                         "the fourth value picks the 4- or 8-waves-per-SIMD kernel build")
    ap.add_argument("--fmt-only", action="store_true",
                    help="time only the headline path (profiling runs)")
    ap.add_argument("--filter-only", action="store_true",
                    help="time only the FDK weight + filter kernel on the workload's "
                         "views (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a G-GPU split into column strips is timed part by part on one GPU with "
              "--slabs G)", file=sys.stderr)
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        a = torch.ones(2 ** 28, device=dev)         # 1 GiB of float32
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": a.numel() * 4,
                          "bytes_written": a.numel() * 4, "repeats": 3}), flush=True)
        return
    if args.filter_only:
        from mirt_amd.workloads import WORKLOADS
        cg = WORKLOADS[args.workload]().cg
        print(json.dumps({"workload": args.workload,
                          "fdk_filter": filter_times(args, cg, dev)}), flush=True)
        return
    if args.path == "mat":
        result = run_mat(args, dev)
    elif args.path == "helix" or args.workload == "clinical_helix":
        result = run_helix(args, dev)
    elif args.workload == "uneven_scans":
        result = run_batch(args, dev)
    else:
        result = run_single(args, dev)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
