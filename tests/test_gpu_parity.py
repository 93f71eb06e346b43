"""GPU parity: the HIP backprojector against the CPU oracle.

Bars (stated in each test):
* EXACT FP32 mode (the host entry fdkhip_fdk_ts_back): bit-for-bit equal to
  the reference C (oracle/_ref, fdk-helix axial path) and to the oracle's FP32
  restatement;
* FAST FP32 mode: relative L2 <= 1e-6 against the reference C;
* FP16 storage: relative L2 <= 1e-3 against the FP64 oracle;
* BF16 storage: relative L2 <= 1e-2 against the FP64 oracle (reported);
* FP16/BF16 storage against the FP64 oracle run on the stored views
  themselves (dequantised, which is exact): relative L2 <= 1e-5.  Storage
  rounding cancels out of that comparison, so what remains is the kernel's
  FP32 arithmetic (weights, positions, accumulation) -- a missed or extra
  slice, a wrong row or a weight rounded to 16 bits fails it.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import CASES, make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mex_args(g, mask):
    return (np.array([g.nx, g.ny, g.nz], np.int32),
            np.array([g.dx, g.dy, g.dz], np.float64),
            np.array([g.offset_x, g.offset_y, g.offset_z], np.float64),
            np.asfortranarray(mask.T), float(g.dso), float(g.dsd), float(g.dfs),
            np.array([g.ds, g.dt], np.float64),
            np.array([g.offset_s, g.offset_t], np.float64))


def gpu_ts_back(g, proj, beta, mask, nthread=4):
    """fdk_mex('fdk,ts,back', ...) on the GPU; returns C-order (ny, nx, nz)."""
    from mirt_amd import fdk_mex
    ptsa = np.asfortranarray(proj.transpose(2, 1, 0))   # (nt, ns, na) Fortran
    out = fdk_mex("fdk,ts,back", *_mex_args(g, mask), ptsa,
                  np.ascontiguousarray(beta, np.float64), np.array([nthread], np.int32))
    # out is (nz, nx, ny) Fortran == memory [nz nx ny]; view as C-order (ny, nx, nz)
    return np.ascontiguousarray(out.transpose(2, 1, 0))


def reference(which, g, proj, beta, mask):
    if O.have_reference(which):
        return O.ref_back_ts(which, g, proj, beta, mask, nthread=4)
    return O.back_ts_f32(g, proj, beta, mask, shipped_weights=(which == "shipped"))


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("data", ["noise", "smooth"])
def test_exact_mode_bit_for_bit(device, name, data):
    g, proj, beta, mask = make_case(name, seed=1, data=data)
    print(f"case={name} data={data} shape={g.nx}x{g.ny}x{g.nz} na={len(beta)}", flush=True)
    got = gpu_ts_back(g, proj, beta, mask)
    want = reference("fixed", g, proj, beta, mask)
    mine = O.back_ts_f32(g, proj, beta, mask)
    assert np.array_equal(mine, want)         # oracle restatement pinned
    diff = np.abs(got - want)
    assert np.array_equal(got, want), \
        f"max|diff|={diff.max():.3e} at {np.unravel_index(diff.argmax(), diff.shape)} " \
        f"relL2={O.rel_l2(got, want):.2e}"


def test_host_entry_cache(device):
    """fdkhip_fdk_ts_back keeps its last plan and device buffers between
    calls: repeated calls, new views in the same geometry, new angles, a new
    mask, a new geometry (larger and smaller buffers) and a call after
    fdkhip_release_host_cache each stay bit for bit equal to the oracle's
    FP32 restatement."""
    from mirt_amd import _lib
    lib = _lib.load()
    g, proj, beta, mask = make_case("arc", seed=21, data="noise")
    g2, proj2, beta2, mask2 = make_case("flat_full", seed=22, data="noise")
    rng = np.random.default_rng(23)
    proj_b = rng.standard_normal(proj.shape).astype(np.float32)
    beta_b = beta + 0.01
    mask_b = mask.copy()
    mask_b.reshape(-1)[::3] = 0
    calls = [("first", g, proj, beta, mask), ("again", g, proj, beta, mask),
             ("new views", g, proj_b, beta, mask), ("new angles", g, proj_b, beta_b, mask),
             ("new mask", g, proj_b, beta_b, mask_b), ("new geometry", g2, proj2, beta2, mask2),
             ("back", g, proj, beta, mask)]
    for k, (what, gg, pp, bb, mm) in enumerate(calls):
        if k == len(calls) - 1:
            _lib.check(lib.fdkhip_release_host_cache())
        print(f"host entry cache: {what}", flush=True)
        got = gpu_ts_back(gg, pp, bb, mm)
        assert np.array_equal(got, O.back_ts_f32(gg, pp, bb, mm)), what
    _lib.check(lib.fdkhip_release_host_cache())
    _lib.check(lib.fdkhip_release_host_cache())   # idempotent


def _plan(g, beta, mask, fmt="f32", math_mode="exact"):
    from mirt_amd import FdkPlan
    from mirt_amd._lib import Geom
    geom = Geom(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.offset_x, g.offset_y, g.offset_z,
                g.dso, g.dsd, g.dfs, g.ns, g.nt, g.ds, g.dt, g.offset_s, g.offset_t)
    return FdkPlan(geom, beta, mask.T, fmt=fmt, math_mode=math_mode)


@pytest.mark.parametrize("name", sorted(CASES))
def test_plan_fast_f32(device, name):
    import torch
    g, proj, beta, mask = make_case(name, seed=2, data="noise")
    print(f"case={name} plan fast f32", flush=True)
    plan = _plan(g, beta, mask, "f32", "fast")
    img = plan(torch.from_numpy(proj).to(device)).cpu().numpy()
    want = reference("fixed", g, proj, beta, mask)
    err = O.rel_l2(img, want)
    assert err <= 1e-6, err


@pytest.mark.parametrize("name", ["flat", "arc", "arc_tall"])
def test_plan_exact_graph_replay(device, name):
    """Plan path (graph replay, several replays with eager work between them)
    equals the reference bit for bit in EXACT mode."""
    import torch
    g, proj, beta, mask = make_case(name, seed=3, data="smooth")
    print(f"case={name} graph replay", flush=True)
    plan = _plan(g, beta, mask, "f32", "exact")
    want = reference("fixed", g, proj, beta, mask)
    p = torch.from_numpy(proj).to(device)
    out = torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=device)
    for rep in range(3):
        out.fill_(float("nan"))
        plan(p, out=out)
        torch.cuda.synchronize()
        _ = (p * 2).sum().item()               # eager work between replays
        assert np.array_equal(out.cpu().numpy(), want), f"replay {rep}"
    direct = plan(p, graph=False).cpu().numpy()
    assert np.array_equal(direct, want)


@pytest.mark.parametrize("fmt,bar", [("f16", 1e-3), ("bf16", 1e-2)])
@pytest.mark.parametrize("name", ["flat", "arc", "arc_tall"])
def test_low_precision_storage(device, name, fmt, bar):
    import torch
    from mirt_amd import quantize_views
    g, proj, beta, mask = make_case(name, seed=4, data="smooth")
    plan = _plan(g, beta, mask, fmt)
    q, sc = quantize_views(torch.from_numpy(proj).to(device), fmt)
    img = plan(q, sc).cpu().numpy()
    ref64 = O.back_ts_f64(g, proj, beta, mask)
    err = O.rel_l2(img, ref64)
    print(f"case={name} fmt={fmt} relL2 vs FP64 = {err:.3e}", flush=True)
    assert err <= bar, err


def test_quantize_views(device):
    import torch
    from mirt_amd import quantize_views
    rng = np.random.default_rng(5)
    proj = (rng.standard_normal((7, 33, 21)) * np.logspace(-3, 3, 7)[:, None, None]).astype(np.float32)
    proj[3] = 0.0
    p = torch.from_numpy(proj).to(device)
    for fmt, tdt in (("f16", np.float16), ("bf16", None)):
        q, sc = quantize_views(p, fmt)
        sc = sc.cpu().numpy()
        assert np.all(sc > 0)
        e = np.log2(sc)
        assert np.array_equal(e, np.round(e)), "scales must be powers of two"
        assert sc[3] == 1.0
        deq = q.float().cpu().numpy() * sc[:, None, None]
        mx = np.abs(proj).max(axis=(1, 2))
        nz = mx > 0
        assert np.all(np.abs(q.float().cpu().numpy()).max(axis=(1, 2))[nz] <= 32768)
        rel = np.abs(deq - proj).max(axis=(1, 2))[nz] / mx[nz]
        assert np.all(rel <= (2 ** -11 if fmt == "f16" else 2 ** -8)), rel
        if tdt is not None:
            want = (proj / sc[:, None, None]).astype(np.float16)
            assert np.array_equal(q.cpu().numpy(), want)


@pytest.mark.parametrize("name", ["flat", "arc", "par"])
def test_st_back(device, name):
    from mirt_amd import fdk_mex
    g, proj, beta, mask = make_case(name, seed=6, data="noise")
    psta = np.ascontiguousarray(proj.transpose(0, 2, 1))      # [na nt ns] C == [ns nt na]
    args = _mex_args(g, mask)
    out = fdk_mex("fdk,st,back", *args, np.asfortranarray(psta.transpose(2, 1, 0)),
                  np.ascontiguousarray(beta), np.array([1], np.int32))
    got = np.ascontiguousarray(out.transpose(2, 1, 0))       # C (nz, ny, nx)
    want = O.ref_back_st("fixed", g, psta, beta, mask) if O.have_reference() \
        else O.back_st_f32(g, psta, beta, mask)
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("name", sorted(CASES))
def test_count_updates(device, name):
    g, proj, beta, mask = make_case(name, seed=7)
    plan = _plan(g, beta, mask)
    _, n = O.back_ts_f32(g, proj, beta, mask, count=True)
    assert plan.count_updates() == n


def test_lds_overflow_path(device):
    """Views whose footprint exceeds the LDS slot gather from HBM; the result
    is unchanged (bit for bit in EXACT mode)."""
    import torch
    g, proj, beta, mask = make_case("flat_full", seed=8, data="noise")
    plan = _plan(g, beta, mask, "f32", "exact")
    plan.set_config(8, 16, 0, 128)  # slot of 128 floats: below every footprint here
    want = reference("fixed", g, proj, beta, mask)
    img = plan(torch.from_numpy(proj).to(device)).cpu().numpy()
    assert np.array_equal(img, want)


@pytest.mark.parametrize("waves,zper", [(4, 16), (8, 16), (4, 8), (8, 8), (4, 32), (4, 48),
                                        (4, 64), (8, 32)])
def test_all_configs(device, waves, zper):
    """EXACT at every built launch shape, including several interleaved
    16-slice chunks per consumer wave (4 x 48, 4 x 64, 8 x 32), FP32 EXACT
    and FAST: bit for bit with the reference C (EXACT), within FP32 rounding
    of it (FAST)."""
    import torch
    g, proj, beta, mask = make_case("arc_tall", seed=9, data="noise")
    want = reference("fixed", g, proj, beta, mask)
    plan = _plan(g, beta, mask, "f32", "exact")
    plan.set_config(waves, zper, 0, 0)
    assert plan.config["waves"] == waves and plan.config["zper"] == zper
    img = plan(torch.from_numpy(proj).to(device)).cpu().numpy()
    assert np.array_equal(img, want)
    fast = _plan(g, beta, mask, "f32", "fast")
    fast.set_config(waves, zper, 0, 0)
    img = fast(torch.from_numpy(proj).to(device)).cpu().numpy()
    assert O.rel_l2(img, want) <= 1e-5


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_occupancy_builds_same_bits(device, fmt):
    """The 8-waves-per-SIMD kernel build of one-chunk FP16/BF16 shapes gives
    the same bits as the 4-wave build (same instructions, other register
    budget); multi-chunk and FP32 shapes keep the 4-wave build."""
    import torch
    from mirt_amd import FdkhipError, quantize_views
    g, proj, beta, mask = make_case("wide_cone", seed=14, data="noise")
    q, sc = quantize_views(torch.from_numpy(proj).to(device), fmt)
    plan = _plan(g, beta, mask, fmt)
    for waves, zper, vstage in ((4, 32, 2), (4, 16, 3), (8, 8, 0)):
        plan.set_config(waves, zper, vstage)
        plan.set_occupancy(4)
        assert plan.occupancy == 4
        img4 = plan(q, sc).cpu().numpy()
        plan.set_occupancy(8)
        assert plan.occupancy == 8
        img8 = plan(q, sc).cpu().numpy()
        print(f"{fmt} {waves}x{zper} vstage {vstage}: 4 vs 8 waves/SIMD equal", flush=True)
        assert np.array_equal(img4, img8), (waves, zper, vstage)
    plan.set_config(4, 48, 0)
    assert plan.occupancy == 4          # no 8-wave build of multi-chunk shapes
    f32 = _plan(g, beta, mask, "f32")
    f32.set_occupancy(8)
    assert f32.occupancy == 4           # nor of FP32 ones
    with pytest.raises(FdkhipError):
        plan.set_occupancy(6)


@pytest.mark.parametrize("waves,zper", [(4, 16), (8, 16), (4, 8), (8, 8), (4, 32), (4, 48),
                                        (4, 64), (8, 32)])
def test_all_configs_f16(device, waves, zper):
    """FP16 storage at every built launch shape (including several interleaved
    16-slice chunks per consumer wave: 4 x 48, 4 x 64, 8 x 32) within 1e-3 of
    the FP64 restatement."""
    import torch
    from mirt_amd import quantize_views
    g, proj, beta, mask = make_case("arc_tall", seed=9, data="smooth")
    plan = _plan(g, beta, mask, "f16")
    plan.set_config(waves, zper, 0, 0)
    q, sc = quantize_views(torch.from_numpy(proj).to(device), "f16")
    img = plan(q, sc).cpu().numpy()
    err = O.rel_l2(img, O.back_ts_f64(g, proj, beta, mask))
    print(f"f16 waves={waves} zper={zper} relL2 vs FP64 = {err:.3e}", flush=True)
    assert err <= 1e-3, err


def _stored(q, sc):
    """The views a FP16/BF16 plan reads, as FP32 (exact: power-of-two scales)."""
    return (q.float() * sc[:, None, None]).cpu().numpy()


SHAPES = [(4, 16), (8, 16), (4, 8), (8, 8), (4, 32), (4, 48), (4, 64), (8, 32)]


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_direct_vs_fp64_of_stored_views(device, name, fmt):
    """FP16/BF16 storage, FP32 interpolation and accumulation: against the FP64
    restatement on the dequantised views within 1e-5 (every case, noise views)."""
    import torch
    from mirt_amd import quantize_views
    g, proj, beta, mask = make_case(name, seed=11, data="noise")
    print(f"case={name} fmt={fmt} vs FP64 of stored views", flush=True)
    q, sc = quantize_views(torch.from_numpy(proj).to(device), fmt)
    plan = _plan(g, beta, mask, fmt)
    img = plan(q, sc).cpu().numpy()
    want = O.back_ts_f64(g, _stored(q, sc), beta, mask)
    err = O.rel_l2(img, want)
    print(f"  relL2 = {err:.3e}", flush=True)
    assert err <= 1e-5, err


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("waves,zper", SHAPES)
def test_direct_all_configs_stored_views(device, waves, zper, fmt):
    """Every built launch shape, tall and wide-cone volumes (partial chunks,
    clamped slices, lanes without work), against FP64 of the stored views."""
    import torch
    from mirt_amd import quantize_views
    for name in ("arc_tall", "wide_cone"):
        g, proj, beta, mask = make_case(name, seed=12, data="noise")
        print(f"case={name} fmt={fmt} waves={waves} zper={zper}", flush=True)
        q, sc = quantize_views(torch.from_numpy(proj).to(device), fmt)
        plan = _plan(g, beta, mask, fmt)
        plan.set_config(waves, zper, 0, 0)
        img = plan(q, sc).cpu().numpy()
        err = O.rel_l2(img, O.back_ts_f64(g, _stored(q, sc), beta, mask))
        print(f"  relL2 = {err:.3e}", flush=True)
        assert err <= 1e-5, err


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_direct_hbm_gather_path(device, fmt):
    """Views whose footprint exceeds the LDS slot (a 128-word slot) gather from
    the caller's views in HBM; same bar as the staged path."""
    import torch
    from mirt_amd import quantize_views
    for name in ("flat_full", "wide_cone"):
        for waves, zper in ((8, 16), (4, 48)):
            g, proj, beta, mask = make_case(name, seed=13, data="noise")
            print(f"case={name} fmt={fmt} waves={waves} zper={zper} slot cap 128", flush=True)
            q, sc = quantize_views(torch.from_numpy(proj).to(device), fmt)
            plan = _plan(g, beta, mask, fmt)
            plan.set_config(waves, zper, 0, 128)
            img = plan(q, sc).cpu().numpy()
            err = O.rel_l2(img, O.back_ts_f64(g, _stored(q, sc), beta, mask))
            print(f"  relL2 = {err:.3e}", flush=True)
            assert err <= 1e-5, err
