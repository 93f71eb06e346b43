"""GPU: the helical / cone-parallel 'fdk,ts,back' (fdkhip_fdk_ts_back_helix)
against the reference C and its restatement.

Bar: bit for bit -- the kernel does the reference's float operations in its
order (doubles where it uses double: atan2, asin, the square root of its
pow(., 0.5)):
* the fixtures the reference C wrote (tests/golden/helix);
* seeded native helical flat/arc cases, cone-parallel ones with and without
  w3d weights, helical and axial, masks and offsets, against the
  restatement (and the reference C when oracle/_ref is built);
* a larger helical scan (96 x 96 x 64, 3 turns of 120 views);
* the device-resident entry equals the host one; cbct_back_helix's default
  (GPU) route equals the same host steps with the restatement as back_call.
"""
from __future__ import annotations

import ctypes as C
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_helix import HELIX_GRID, _seeded, load_helix

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HELIX_GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "helix", "*.npz")))


def gpu_helix(g, proj, beta, zs, mask, *, na1_half, cone_par, w3d, source_dz_per_rad,
              pitch):
    """fdkhip_fdk_ts_back_helix through the gateway mirror; C-order arrays of
    the oracle's conventions in and out ((na, ns, nt) views, (ny, nx, nz))."""
    from mirt_amd import fdk_mex_helix
    na = proj.shape[0]
    ptsa = np.asfortranarray(np.asarray(proj, np.float32).transpose(2, 1, 0))   # [nt ns na]
    m = np.ones((g.nx, g.ny), np.uint8) if mask is None else \
        np.asfortranarray(np.asarray(mask, np.uint8).reshape(g.ny, g.nx).T)
    img = fdk_mex_helix(
        "fdk,ts,back", np.array([g.nx, g.ny, g.nz], np.int32),
        np.array([g.dx, g.dy, g.dz]), np.array([g.offset_x, g.offset_y, g.offset_z]), m,
        g.dso, g.dsd, g.dfs, np.array([g.ds, g.dt]), np.array([g.offset_s, g.offset_t]),
        ptsa, np.ascontiguousarray(beta, np.float64), np.ascontiguousarray(zs, np.float64),
        np.array([na1_half], np.int32), np.array([cone_par], np.int32),
        np.array([w3d], np.int32), np.array([source_dz_per_rad]),
        np.array([pitch], np.float32), np.array([1], np.int32))
    assert img.shape == (g.nz, g.nx, g.ny) and na == beta.size
    return np.ascontiguousarray(img.transpose(2, 1, 0))                         # (ny, nx, nz)


@pytest.mark.parametrize("path", HELIX_GOLDEN, ids=[os.path.basename(p) for p in HELIX_GOLDEN])
def test_helix_gpu_matches_golden(device, path):
    g, d, ops = load_helix(path)
    print(f"helix golden {os.path.basename(path)}", flush=True)
    got = gpu_helix(g, d["proj"], d["beta"], d["source_zs"], d["mask"], **ops)
    assert np.array_equal(got, d["out_fixed"]), float(np.abs(got - d["out_fixed"]).max())


@pytest.mark.parametrize("dfs,pitch,cone_par,w3d", HELIX_GRID)
def test_helix_gpu_matches_restatement(device, dfs, pitch, cone_par, w3d):
    g, proj, beta, zs, mask, ops = _seeded(dfs, pitch=pitch, cone_par=cone_par, w3d=w3d,
                                           seed=int(10 * pitch) + 3 * cone_par + w3d)
    print(f"helix dfs={dfs} pitch={pitch} cone_par={cone_par} w3d={w3d}", flush=True)
    got = gpu_helix(g, proj, beta, zs, mask, **ops)
    want = O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops)
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    if O.have_reference():
        assert np.array_equal(got, O.ref_back_ts_helix(g, proj, beta, zs, mask, nthread=4,
                                                       **ops))


def test_helix_gpu_parallel_beam(device):
    g, proj, beta, zs, mask, ops = _seeded(0.0, pitch=1.0, cone_par=0, w3d=0, seed=5,
                                           dso=np.inf, dsd=np.inf)
    got = gpu_helix(g, proj, beta, zs, mask, **ops)
    assert np.array_equal(got, O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops))


def test_helix_gpu_larger_scan(device):
    """96 x 96 x 64 at 1 mm, 3 turns of 120 views, pitch 1, arc detector:
    several slice chunks per column and view windows that cut them."""
    rng = np.random.default_rng(77)
    g = O.geom((96, 96, 64), (1.0, -1.0, 1.0), (0.0, 0.0, 0.0), 541.0, 949.075, 0.0,
               128, 32, (1.3, 1.2), (1.25, 0.0))
    na360, na = 120, 360
    zfov = 541.0 / 949.075 * 32 * 1.2
    dz_view = zfov / na360
    zs = (np.arange(na) - na / 2) * dz_view
    beta = 2 * np.pi * np.arange(na) / na360
    proj = rng.standard_normal((na, 128, 32)).astype(np.float32)
    ops = dict(na1_half=na360 // 2, cone_par=0, w3d=0,
               source_dz_per_rad=dz_view / (2 * np.pi / na360), pitch=1.0)
    print("helix 96x96x64, 360 views", flush=True)
    got = gpu_helix(g, proj, beta, zs, None, **ops)
    want = O.back_ts_helix_f32(g, proj, beta, zs, None, **ops)
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    assert (np.abs(want) > 0).mean() > 0.5


def test_helix_direct_equals_host(device):
    import torch
    from mirt_amd import _lib
    from mirt_amd._lib import Geom
    lib = _lib.load()
    g, proj, beta, zs, mask, ops = _seeded(np.inf, pitch=1.0, cone_par=1, w3d=1, seed=3)
    want = gpu_helix(g, proj, beta, zs, mask, **ops)
    pt = torch.from_numpy(np.ascontiguousarray(proj)).to(device)
    img = torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=device)
    gg = Geom(*[getattr(g, f[0]) for f in Geom._fields_])
    b = np.ascontiguousarray(beta, np.float64)
    z = np.ascontiguousarray(zs, np.float64)
    m = np.ascontiguousarray(mask, np.uint8)                           # (ny, nx) = [nx ny]
    ms = C.c_float(-1)
    s = torch.cuda.current_stream()
    P = lambda a: C.c_void_p(a.ctypes.data)
    _lib.check(lib.fdkhip_fdk_ts_back_helix_direct(
        C.byref(gg), C.c_void_p(pt.data_ptr()), proj.shape[0], P(b), P(z), ops["na1_half"],
        ops["cone_par"], ops["w3d"], ops["source_dz_per_rad"], ops["pitch"], P(m),
        C.c_void_p(img.data_ptr()), C.c_void_p(s.cuda_stream), C.byref(ms)))
    assert np.array_equal(img.cpu().numpy(), want)
    assert ms.value > 0


def test_cbct_back_helix_gpu_route(device):
    from mirt_amd import cbct_back_helix
    from mirt_amd.geom import ct_geom, image_geom
    from test_helix_host import helix_back_call
    cg = ct_geom("fan", ns=48, nt=16, na=96, ds=1.5, dt=1.6, offset_s=0.25, dsd=500.0,
                 dso=300.0, dfs=np.inf, orbit=720, pitch=0.9, source_z0="center")
    ig = image_geom(nx=20, ny=18, nz=24, dx=2.0, dz=1.0, mask="all-but-edge")
    proj = np.asfortranarray(np.random.default_rng(9).standard_normal(cg.dim)
                             .astype(np.float32))
    for kw in (dict(ia_skip=1), dict(ia_skip=2, cone_par=1, w3d=1)):
        with pytest.warns(UserWarning):
            got = cbct_back_helix(proj, cg, ig, **kw)
        with pytest.warns(UserWarning):
            want = cbct_back_helix(proj, cg, ig, back_call=helix_back_call, **kw)
        assert np.array_equal(got, want)


def _helix_plan(g, beta, zs, mask, ops, fmt="f32", math_mode="exact"):
    from mirt_amd import FdkPlan
    from mirt_amd._lib import Geom
    gg = Geom(*[getattr(g, f[0]) for f in Geom._fields_])
    m = None if mask is None else np.asarray(mask, np.uint8).reshape(g.ny, g.nx).T
    return FdkPlan(gg, beta, m, fmt=fmt, math_mode=math_mode,
                   helix=(zs, ops["na1_half"], ops["pitch"]))


def _window_binds(g, zs, ops):
    """True when a view's slice window (2 na1_half views of source travel) is
    shorter than the detector's height at the isocentre, so that it, not the
    cone, ends the slice range of the columns far from the source."""
    window = 2 * ops["na1_half"] * abs(zs[1] - zs[0]) / g.dz
    return window < g.nt * g.dt * g.dso / g.dsd / g.dz


@pytest.mark.parametrize("dfs,pitch", [(0.0, 1.0), (np.inf, 1.0), (0.0, 0.3)])
def test_helix_plan_every_launch_shape(device, dfs, pitch):
    """Native helical geometry on the axial pipeline (fdkhip_plan_create_helix):
    per-view source heights and slice windows, each z-group visiting only the
    views whose window meets it.  Every launch shape and stage depth is the
    reference bit for bit (restatement; reference C when built), and the
    direct entry's cone_par = 0 route equals it too.  Pitch 1 windows are
    wider than the cone; pitch 0.3 windows end the slice ranges."""
    import torch
    from mirt_amd import _lib
    g, proj, beta, zs, mask, ops = _seeded(dfs, pitch=pitch, cone_par=0, w3d=0, seed=41, nz=70)
    assert _window_binds(g, zs, ops) == (pitch < 1)
    want = O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops)
    if O.have_reference():
        assert np.array_equal(want, O.ref_back_ts_helix(g, proj, beta, zs, mask, nthread=4,
                                                        **ops))
    plan = _helix_plan(g, beta, zs, mask, ops)
    pt = torch.from_numpy(np.ascontiguousarray(proj)).to(device)
    for waves, zper in [(8, 16), (8, 8), (4, 16), (4, 32), (4, 8), (4, 48), (4, 64), (8, 32)]:
        for vstage in (0, 3, 8):
            print(f"helix plan dfs={dfs} waves={waves} zper={zper} vstage={vstage}", flush=True)
            plan.set_config(waves, zper, vstage)
            got = plan(pt).cpu().numpy()
            assert np.array_equal(got, want), (waves, zper, vstage,
                                               float(np.abs(got - want).max()))
    assert plan.count_updates() > 0
    # the direct entry (cone_par = 0) runs on such a plan
    from mirt_amd._lib import Geom
    gg = Geom(*[getattr(g, f[0]) for f in Geom._fields_])
    img = torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=device)
    P = lambda a: C.c_void_p(a.ctypes.data)
    ms = C.c_float(-1)
    m = np.ascontiguousarray(mask, np.uint8)
    b = np.ascontiguousarray(beta, np.float64)
    z = np.ascontiguousarray(zs, np.float64)
    _lib.check(_lib.load().fdkhip_fdk_ts_back_helix_direct(
        C.byref(gg), C.c_void_p(pt.data_ptr()), proj.shape[0], P(b), P(z), ops["na1_half"],
        0, 0, ops["source_dz_per_rad"], ops["pitch"], P(m), C.c_void_p(img.data_ptr()),
        C.c_void_p(torch.cuda.current_stream().cuda_stream), C.byref(ms)))
    assert np.array_equal(img.cpu().numpy(), want) and ms.value > 0


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("pitch", [1.0, 0.3])
def test_helix_plan_half_storage(device, fmt, pitch):
    """FP16/BF16-stored views on a helical plan (FAST arithmetic, closed-form
    slice coordinate): within the storage bar of the exact FP32 result
    (relative L2 1e-3 for FP16, 4e-3 for BF16, the formats' unit roundoffs
    2^-11 and 2^-8 with margin), also where the views' slice windows, not
    the cone, end the slice ranges (pitch 0.3): a slice outside a view's
    window must add nothing even where its row is on the detector.  The
    default launch shape and every shape with one or several chunks per
    consumer wave."""
    import torch
    from mirt_amd.cbct_back import quantize_views
    g, proj, beta, zs, mask, ops = _seeded(0.0, pitch=pitch, cone_par=0, w3d=0, seed=43, nz=70)
    assert _window_binds(g, zs, ops) == (pitch < 1)
    want = O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops)
    plan = _helix_plan(g, beta, zs, mask, ops, fmt=fmt)
    pt = torch.from_numpy(np.ascontiguousarray(proj)).to(device)
    q, sc = quantize_views(pt, fmt)
    for waves, zper in [(0, 0), (4, 16), (8, 8), (4, 48), (4, 64), (8, 32)]:
        if waves:
            plan.set_config(waves, zper, 0)
        got = plan(q, sc).cpu().numpy()
        err = O.rel_l2(got, want)
        print(f"helix plan {fmt} pitch {pitch} {plan.config}: relL2 {err:.2e}", flush=True)
        assert err <= {"f16": 1e-3, "bf16": 4e-3}[fmt], (waves, zper, err)
