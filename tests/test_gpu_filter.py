"""GPU: the FDK weight and ramp filter on the device (fdkhip_fdk_filter_views:
one fused kernel -- weight, FFT of the zero-padded rows in LDS, response,
inverse FFT, transpose) and feldkamp's device route.

Bars (FP32 arithmetic against the FP64 torch restatement of
feldkamp_weight1 + fdk_filter, which tests/test_host.py pins to direct
convolution and to fbp_ramp's closed form):
* filtered views within relative L2 1e-6 and max |error| 1e-5 of the
  largest sample (FP32 weights, taps and accumulation over ns terms);
* feldkamp's device route within relative L2 1e-5 of its host route (FP64
  weight and filter) with FP32 storage, 1e-3 with FP16 storage, 4e-3 with
  BF16 storage.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fp64_reference(p, cg, window, w1cyl):
    import torch
    from mirt_amd.fdk import fdk_filter, feldkamp_weight1
    f = fdk_filter(feldkamp_weight1(torch.as_tensor(p, dtype=torch.float64), cg, w1cyl=w1cyl),
                   cg, window).numpy()                       # [ns nt na]
    return np.ascontiguousarray(f.transpose(2, 0, 1))        # (na, ns, nt) = [nt ns na]


@pytest.mark.parametrize("dfs", [0.0, np.inf])
@pytest.mark.parametrize("ns,nt,na,window,w1cyl", [(37, 5, 3, "ramp", True),
                                                    (64, 7, 4, "hann", False),
                                                    (888, 16, 6, "ramp", True)])
def test_weight_filter_views_vs_fp64(device, dfs, ns, nt, na, window, w1cyl):
    import torch
    from mirt_amd.fdk import weight_filter_views
    from mirt_amd.geom import ct_geom
    print(f"filter dfs={dfs} ns={ns} nt={nt} na={na} {window} w1cyl={w1cyl}", flush=True)
    cg = ct_geom("fan", ns=ns, nt=nt, na=na, ds=1.0239, dt=1.0964, dsd=949.0, dso=541.0,
                 dfs=dfs, offset_s=1.25)
    p = np.random.default_rng(ns + nt).standard_normal((ns, nt, na)).astype(np.float32)
    want = _fp64_reference(p, cg, window, w1cyl)
    sta = torch.from_numpy(np.ascontiguousarray(p.transpose(2, 1, 0))).to(device)  # s fastest
    got = weight_filter_views(sta, cg, window=window, w1cyl=w1cyl).cpu().numpy()
    assert got.shape == want.shape
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err <= 1e-6, err
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


def test_filter_views_argument_errors(device):
    import torch
    from mirt_amd import _lib
    lib = _lib.load()
    a = torch.zeros((2, 3, 4), device=device)
    w = torch.ones((3, 4), device=device)
    H = torch.ones(8, device=device)
    P = lambda t: C.c_void_p(t.data_ptr())
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.fdkhip_fdk_filter_views(P(a), 4, 3, 2, P(w), P(H), 8, P(a), s) != 0
    assert "alias" in lib.fdkhip_last_error().decode()
    assert lib.fdkhip_fdk_filter_views(P(a), -1, 3, 2, P(w), P(H), 8, P(a), s) != 0
    out = torch.empty((2, 4, 3), device=device)
    for bad in (6, 2, 8192):     # not a power of two, below ns, above 4096
        assert lib.fdkhip_fdk_filter_views(P(a), 4, 3, 2, P(w), P(H), bad, P(out), s) != 0
    # a flat response of 1 leaves the weighted views (transposed to [nt ns na])
    b = torch.randn((2, 3, 4), device=device)
    assert lib.fdkhip_fdk_filter_views(P(b), 4, 3, 2, P(w), P(H), 8, P(out), s) == 0
    torch.cuda.synchronize()
    assert torch.allclose(out, b.permute(0, 2, 1), rtol=1e-6, atol=1e-6)
    with pytest.raises(ValueError):
        from mirt_amd.fdk import weight_filter_views
        from mirt_amd.geom import ct_geom
        weight_filter_views(a.double(), ct_geom("fan", ns=4, nt=3, na=2, ds=1.0, dt=1.0,
                                                dsd=949.0, dso=541.0, dfs=np.inf))


@pytest.mark.parametrize("ns,nt", [(1, 3), (2, 9), (5, 8), (200, 17), (2048, 4)])
def test_filter_sizes(device, ns, nt):
    """Padded lengths from 2 to 4096 and row counts not a multiple of the
    kernel's eight rows per workgroup, against the FP64 restatement."""
    import torch
    from mirt_amd.fdk import weight_filter_views
    from mirt_amd.geom import ct_geom
    if ns == 1:
        pytest.skip("fdk_filter's ramp (fbp_ramp) needs an even padded length; ns = 1 has 1")
    cg = ct_geom("fan", ns=ns, nt=nt, na=3, ds=0.9, dt=1.1, dsd=949.0, dso=541.0, dfs=np.inf)
    p = np.random.default_rng(ns * nt).standard_normal((ns, nt, 3)).astype(np.float32)
    want = _fp64_reference(p, cg, "ramp", True)
    sta = torch.from_numpy(np.ascontiguousarray(p.transpose(2, 1, 0))).to(device)
    got = weight_filter_views(sta, cg).cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"filter ns={ns} nt={nt}: relL2 {err:.2e}", flush=True)
    assert err <= 1e-6, err


def _window(name, ns):
    """A window argument of feldkamp: a name, or (for "array") samples
    ii = -npad/2 .. npad/2-1 of a Gaussian taper of the padded length."""
    if name != "array":
        return name
    npad = 2 ** int(np.ceil(np.log2(2 * ns - 1)))
    ii = np.arange(npad) - npad / 2
    return np.exp(-(ii / (npad / 4)) ** 2)


@pytest.mark.parametrize("window", ["ramp", "hann", "hamming,0.6", "array"])
@pytest.mark.parametrize("fmt", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("ia_skip", [1, 3])
def test_feldkamp_device_route_vs_host_route(device, ia_skip, fmt, window):
    """feldkamp on device views (view skipping before the filter, the fused
    weight + filter kernel, then the plan; FP16/BF16: quantised filtered
    views) against the host route (FP64 weight and FFT filter, then the
    backprojector on FP32 views).  Bars (relative L2): FP32 1e-5; FP16
    storage 1e-3; BF16 storage 4e-3, its unit roundoff 2^-8 (eight times
    FP16's: 7 stored fraction bits against 10)."""
    import torch
    from mirt_amd.fdk import feldkamp
    from mirt_amd.geom import ct_geom, image_geom
    cg = ct_geom("fan", ns=96, nt=24, na=72, ds=1.0, dt=1.0, dsd=949.0, dso=541.0, dfs=0.0)
    ig = image_geom(nx=40, ny=36, nz=16, dx=1.2, dz=1.0)
    p = np.asfortranarray(np.random.default_rng(3).random(cg.dim)).astype(np.float32)
    w = _window(window, cg.ns)
    w1cyl = window != "hann"
    print(f"feldkamp device route ia_skip={ia_skip} {fmt} {window} w1cyl={w1cyl}", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        host = feldkamp(cg, ig, np.asarray(p, np.float64), ia_skip=ia_skip, window=w,
                        w1cyl=w1cyl)
        dev = feldkamp(cg, ig, torch.from_numpy(np.ascontiguousarray(p)).to(device),
                       ia_skip=ia_skip, window=w, w1cyl=w1cyl, fmt=fmt)
    dev = dev.cpu().numpy()
    assert dev.shape == host.shape
    err = np.linalg.norm(dev - host) / np.linalg.norm(host)
    print(f"  relL2 {err:.2e}", flush=True)
    assert err <= {"f32": 1e-5, "f16": 1e-3, "bf16": 4e-3}[fmt], err


@pytest.mark.parametrize("fmt", ["f32", "f16"])
def test_feldkamp_device_route_extrapolate_t(device, fmt):
    """extrapolate_t on the device route: the filtered first and last rows
    replicated (cbct_back.m:95-104) before the plan, against the host route;
    a volume taller than the detector so that the replicated rows are read;
    offset_t with extrapolation is rejected as the reference does."""
    import torch
    from mirt_amd.fdk import feldkamp
    from mirt_amd.geom import ct_geom, image_geom
    cg = ct_geom("fan", ns=80, nt=16, na=60, ds=1.0, dt=1.0, dsd=949.0, dso=541.0,
                 dfs=np.inf)
    ig = image_geom(nx=36, ny=36, nz=30, dx=1.2, dz=1.0)
    p = np.asfortranarray(np.random.default_rng(5).random(cg.dim)).astype(np.float32)
    pd = torch.from_numpy(np.ascontiguousarray(p)).to(device)
    print(f"feldkamp extrapolate_t {fmt}", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        host = feldkamp(cg, ig, np.asarray(p, np.float64), extrapolate_t=9)
        flat = feldkamp(cg, ig, np.asarray(p, np.float64))
        dev = feldkamp(cg, ig, pd, extrapolate_t=9, fmt=fmt).cpu().numpy()
    assert np.linalg.norm(host - flat) > 0.05 * np.linalg.norm(host)    # rows are read
    err = np.linalg.norm(dev - host) / np.linalg.norm(host)
    print(f"  relL2 {err:.2e}", flush=True)
    assert err <= (1e-5 if fmt == "f32" else 1e-3), err
    cg2 = ct_geom("fan", ns=80, nt=16, na=60, ds=1.0, dt=1.0, dsd=949.0, dso=541.0,
                  dfs=np.inf, offset_t=0.5)
    with pytest.raises(ValueError, match="offset_t not yet done"):
        feldkamp(cg2, ig, pd, extrapolate_t=3)
