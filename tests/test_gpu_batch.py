"""GPU: batches of scans in one persistent launch, and the autotuner.

Bars:
* EXACT FP32 batch: every scan bit-for-bit equal to the oracle's FP32
  restatement of the reference kernel (which the single-scan tests pin to the
  reference C) and to a single-scan plan;
* FAST FP32 / FP16 batch: equal to single-scan FAST plans of the same format
  within relative L2 1e-6 (same arithmetic, different launch);
* graph replays with new projection data match direct launches bit for bit;
* per-scan z-group heights (the launch's slices per lane or half of it, in
  one launch) leave EXACT results bit for bit unchanged;
* one launch per scan (each plan in its own launch shape) gives the same
  EXACT bits, FP16/BF16 within 1e-6 of the persistent kernel;
* the autotuner leaves results unchanged (EXACT: bit for bit).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import CASES, make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# uneven scans: different geometries, view counts, masks, and one scan whose
# image misses the field of view entirely
BATCH = [("flat", 3), ("arc", 4), ("par", 5), ("arc_tall", 6), ("flat_full", 7)]


def _geom(g):
    from mirt_amd._lib import Geom
    return Geom(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.offset_x, g.offset_y, g.offset_z,
                g.dso, g.dsd, g.dfs, g.ns, g.nt, g.ds, g.dt, g.offset_s, g.offset_t)


def _cases(data="noise"):
    return [make_case(name, seed=seed, data=data) for name, seed in BATCH]


def _off_fov_case():
    """A small arc scan whose image sits far outside the detector's reach."""
    g, proj, beta, mask = make_case("arc", seed=9)
    g = O.OrcGeom.from_buffer_copy(g)
    g.offset_z = 500.0
    return g, proj, beta, mask


def _batch(cases, fmt="f32", math_mode="exact"):
    from mirt_amd import FdkBatch
    return FdkBatch([(_geom(g), beta, mask.T) for g, _, beta, mask in cases],
                    fmt=fmt, math_mode=math_mode)


def _plan(g, beta, mask, fmt="f32", math_mode="exact"):
    from mirt_amd import FdkPlan
    return FdkPlan(_geom(g), beta, mask.T, fmt=fmt, math_mode=math_mode)


def test_batch_exact_bit_for_bit(device):
    import torch
    cases = _cases() + [_off_fov_case()]
    print(f"batch exact: {len(cases)} scans", flush=True)
    b = _batch(cases)
    projs = [torch.from_numpy(p).to(device) for _, p, _, _ in cases]
    outs = [o.cpu().numpy() for o in b(projs)]
    for i, (g, proj, beta, mask) in enumerate(cases):
        want = O.back_ts_f32(g, proj, beta, mask)
        d = np.abs(outs[i] - want)
        assert np.array_equal(outs[i], want), \
            f"scan {i}: max|diff|={d.max():.3e} relL2={O.rel_l2(outs[i], want):.2e}"
        plan_img = _plan(g, beta, mask)(projs[i]).cpu().numpy()
        assert np.array_equal(outs[i], plan_img), f"scan {i}: batch != plan"
    assert not outs[-1].any()   # out of the field of view: zeros


def test_batch_count_updates(device):
    cases = _cases()
    b = _batch(cases)
    total = 0
    for g, _, beta, mask in cases:
        total += _plan(g, beta, mask).count_updates()
    assert b.count_updates() == total > 0
    cfg = b.config
    assert cfg["items"] > 0 and cfg["workgroups"] > 0


@pytest.mark.parametrize("fmt", ["f32", "f16"])
def test_batch_fast_matches_plans(device, fmt):
    import torch
    from mirt_amd import quantize_views
    cases = _cases()
    print(f"batch fast {fmt}", flush=True)
    b = _batch(cases, fmt=fmt, math_mode="fast")
    projs, scales = [], []
    for _, p, _, _ in cases:
        t = torch.from_numpy(p).to(device)
        if fmt == "f16":
            q, s = quantize_views(t, "f16")
            projs.append(q)
            scales.append(s)
        else:
            projs.append(t)
    outs = b(projs, scales if scales else None)
    for i, (g, _, beta, mask) in enumerate(cases):
        plan = _plan(g, beta, mask, fmt, "fast")
        want = plan(projs[i], scales[i] if scales else None)
        err = O.rel_l2(outs[i].cpu().numpy(), want.cpu().numpy())
        assert err <= 1e-6, (i, err)


def test_batch_graph_replay_new_data(device):
    import torch
    cases = _cases()
    b = _batch(cases)
    projs = [torch.from_numpy(p).to(device) for _, p, _, _ in cases]
    outs = [torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=device)
            for g, _, _, _ in cases]
    for rep in range(3):
        # new data in the same buffers: the replayed graph must see it
        for t, (g, _, _, _) in zip(projs, cases):
            t.copy_(torch.from_numpy(
                np.random.default_rng(100 + rep).standard_normal(t.shape, dtype=np.float32)))
        b(projs, outs=outs, graph=True)
        got = [o.cpu().numpy() for o in outs]
        direct = [o.cpu().numpy() for o in b(projs, graph=False)]
        for i in range(len(cases)):
            assert np.array_equal(got[i], direct[i]), (rep, i)


@pytest.mark.parametrize("waves,zper,vstage", [(8, 8, 4), (4, 16, 0), (4, 32, 8), (8, 16, 1)])
def test_batch_launch_shapes(device, waves, zper, vstage):
    import torch
    cases = _cases()
    print(f"batch shape waves={waves} zper={zper} vstage={vstage}", flush=True)
    b = _batch(cases)
    b.set_config(waves, zper, vstage)
    projs = [torch.from_numpy(p).to(device) for _, p, _, _ in cases]
    outs = [o.cpu().numpy() for o in b(projs)]
    for i, (g, proj, beta, mask) in enumerate(cases):
        assert np.array_equal(outs[i], O.back_ts_f32(g, proj, beta, mask)), i


@pytest.mark.parametrize("waves,zper,vstage", [(4, 16, 0), (4, 32, 0), (8, 16, 0), (4, 16, 8)])
def test_batch_per_scan_heights(device, waves, zper, vstage):
    """Scans at the launch's slices per lane and at half of it in one launch:
    EXACT results bit for bit equal to the oracle, for a fixed mix and for
    the automatic choice (a deep stage shrinks the slot, so more scans move
    to the half height)."""
    import torch
    cases = _cases()
    b = _batch(cases)
    b.set_config(waves, zper, vstage)
    auto = b.scan_zper
    print(f"batch heights waves={waves} zper={zper} vstage={vstage} auto={auto}", flush=True)
    assert set(auto) <= {zper, zper // 2}
    projs = [torch.from_numpy(p).to(device) for _, p, _, _ in cases]
    want = [O.back_ts_f32(g, proj, beta, mask) for g, proj, beta, mask in cases]
    mix = [zper if i % 2 else zper // 2 for i in range(len(cases))]
    for heights in (mix, [zper // 2] * len(cases), None):
        b.set_zper(heights)
        assert b.scan_zper == (auto if heights is None else heights)
        outs = [o.cpu().numpy() for o in b(projs)]
        for i in range(len(cases)):
            assert np.array_equal(outs[i], want[i]), (heights, i)


def test_batch_f16_heights_vs_fp64(device):
    """FP16 batch with scans at 16 and at 8 slices per lane in one launch
    (the independent-row slice update and the straight-line groups of
    16-slice consumers): every scan within 1e-3 of the FP64 restatement."""
    import torch
    from mirt_amd import quantize_views
    cases = _cases(data="smooth")
    b = _batch(cases, fmt="f16", math_mode="fast")
    b.set_config(4, 16, 0)
    b.set_zper([16, 8, 16, 8, 16])
    projs, scales = [], []
    for _, p, _, _ in cases:
        q, s = quantize_views(torch.from_numpy(p).to(device), "f16")
        projs.append(q)
        scales.append(s)
    outs = b(projs, scales)
    for i, (g, proj, beta, mask) in enumerate(cases):
        err = O.rel_l2(outs[i].cpu().numpy(), O.back_ts_f64(g, proj, beta, mask))
        print(f"batch f16 scan {i} zper={b.scan_zper[i]} relL2 vs FP64 = {err:.2e}", flush=True)
        assert err <= 1e-3, (i, err)


def test_batch_per_scan_mode(device):
    """One launch per scan: EXACT bit for bit equal to the oracle (and so to
    the persistent kernel), with each plan in its own launch shape, also
    after the plans' shapes change; graph replays and direct launches agree;
    set_config returns to the persistent kernel."""
    import torch
    cases = _cases() + [_off_fov_case()]
    b = _batch(cases)
    projs = [torch.from_numpy(p).to(device) for _, p, _, _ in cases]
    want = [O.back_ts_f32(g, proj, beta, mask) for g, proj, beta, mask in cases]
    b.set_per_scan()
    cfg = b.config
    print(f"batch per-scan: config {cfg}, slices per lane {b.scan_zper}", flush=True)
    assert b.per_scan and cfg["waves"] == 0 and cfg["workgroups"] > 0
    for graph in (True, False, True):
        outs = [o.cpu().numpy() for o in b(projs, graph=graph)]
        for i in range(len(cases)):
            assert np.array_equal(outs[i], want[i]), (graph, i)
    # the plans of these geometries tuned elsewhere in the process: a batch
    # set to per-scan launches afterwards takes their tuned shapes
    for (g, proj, beta, mask), t in zip(cases[:3], projs):
        _plan(g, beta, mask).autotune(t, reps=1)
    b2 = _batch(cases)
    b2.set_per_scan()
    outs = [o.cpu().numpy() for o in b2(projs)]
    for i in range(len(cases)):
        assert np.array_equal(outs[i], want[i]), ("tuned plans", i)
    b.set_config(4, 16, 0)
    assert not b.per_scan and b.config["zper"] == 16
    outs = [o.cpu().numpy() for o in b(projs)]
    for i in range(len(cases)):
        assert np.array_equal(outs[i], want[i]), ("persistent again", i)


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_batch_per_scan_mode_16bit(device, fmt):
    """One launch per scan with FP16/BF16 storage: equal to the persistent
    kernel within relative L2 1e-6 (same arithmetic; only the FP32 summation
    order of views gathered from HBM may differ with the launch shape)."""
    import torch
    from mirt_amd import quantize_views
    cases = _cases(data="smooth")
    b = _batch(cases, fmt=fmt, math_mode="fast")
    projs, scales = [], []
    for _, p, _, _ in cases:
        q, s = quantize_views(torch.from_numpy(p).to(device), fmt)
        projs.append(q)
        scales.append(s)
    ref = [o.cpu().numpy() for o in b(projs, scales)]
    b.set_per_scan()
    outs = [o.cpu().numpy() for o in b(projs, scales)]
    for i in range(len(cases)):
        err = O.rel_l2(outs[i], ref[i])
        assert err <= 1e-6, (i, err)


def test_plan_autotune_keeps_results(device):
    import torch
    g, proj, beta, mask = make_case("arc_tall", seed=11)
    plan = _plan(g, beta, mask)
    t = torch.from_numpy(proj).to(device)
    ms = plan.autotune(t, reps=2)
    assert ms > 0
    cfg = plan.config
    img = plan(t).cpu().numpy()
    assert np.array_equal(img, O.back_ts_f32(g, proj, beta, mask)), cfg
    # a new plan of the same geometry starts from the tuned choice
    again = _plan(g, beta, mask)
    assert again.config == cfg


def test_batch_autotune_keeps_results(device):
    """Whichever the tuner keeps (a persistent shape or one launch per scan),
    EXACT results stay bit for bit."""
    import torch
    cases = _cases()
    b = _batch(cases)
    projs = [torch.from_numpy(p).to(device) for _, p, _, _ in cases]
    assert b.autotune(projs, reps=1) > 0
    print(f"batch autotune: per_scan={b.per_scan} config {b.config}", flush=True)
    outs = [o.cpu().numpy() for o in b(projs)]
    for i, (g, proj, beta, mask) in enumerate(cases):
        assert np.array_equal(outs[i], O.back_ts_f32(g, proj, beta, mask)), i


def test_batch_argument_errors(device):
    import torch
    from mirt_amd import FdkhipError
    cases = _cases()[:2]
    b = _batch(cases)
    projs = [torch.from_numpy(p).to(device) for _, p, _, _ in cases]
    with pytest.raises(ValueError):
        b(projs[:1])
    with pytest.raises(FdkhipError):
        b.set_config(8, 32, 0)    # shape not built
    b.set_config(4, 16, 0)
    keep = b.scan_zper
    with pytest.raises(FdkhipError):
        b.set_zper([12, 16])      # neither the launch's height nor half of it
    with pytest.raises(ValueError):
        b.set_zper([16])
    assert b.scan_zper == keep and b.config["zper"] == 16
    b16 = _batch(cases, fmt="f16", math_mode="fast")
    with pytest.raises(ValueError):
        b16([p.half() for p in projs])   # no scales
