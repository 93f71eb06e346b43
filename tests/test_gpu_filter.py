"""GPU: the FDK weight and ramp filter on the device (fdkhip_fdk_filter_views:
one fused kernel -- weight, FFT of the zero-padded rows in LDS, response,
inverse FFT, transpose) and feldkamp's device route.

Bars (FP32 arithmetic against the FP64 torch restatement of
feldkamp_weight1 + fdk_filter, which tests/test_host.py pins to direct
convolution and to fbp_ramp's closed form):
* filtered views within relative L2 1e-6 and max |error| 1e-5 of the
  largest sample (FP32 weights, taps and accumulation over ns terms);
* feldkamp's device route within relative L2 1e-5 of its host route.
"""
from __future__ import annotations

import ctypes as C

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


def test_feldkamp_device_route_vs_host_route(device):
    import torch
    from mirt_amd.fdk import feldkamp
    from mirt_amd.geom import ct_geom, image_geom
    cg = ct_geom("fan", ns=96, nt=24, na=72, ds=1.0, dt=1.0, dsd=949.0, dso=541.0, dfs=0.0)
    ig = image_geom(nx=40, ny=36, nz=16, dx=1.2, dz=1.0)
    p = np.asfortranarray(np.random.default_rng(3).random(cg.dim)).astype(np.float32)
    host = feldkamp(cg, ig, np.asarray(p, np.float64))
    dev = feldkamp(cg, ig, torch.from_numpy(np.ascontiguousarray(p)).to(device))
    dev = dev.cpu().numpy()
    assert dev.shape == host.shape
    err = np.linalg.norm(dev - host) / np.linalg.norm(host)
    print(f"feldkamp device vs host relL2 {err:.2e}", flush=True)
    assert err <= 1e-5, err
