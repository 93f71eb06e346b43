"""GPU: cbct_back's use_mex = 1 route through fdkhip_cbct_back (permutes,
edge-row replication and angle scale on the device) against the host route
(numpy permutes around fdk_mex, fbp/cbct_back.m:74-137 step by step).

Bar: bit-for-bit equal (both run the EXACT kernel on the same views; the
permutes move values and the scale is one single-precision multiply).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan(kind, seed, offset_t=0.0):
    from mirt_amd.geom import ct_geom, image_geom
    dfs = 0.0 if kind == "arc" else np.inf
    cg = ct_geom("fan", ns=72, nt=40, na=36, ds=1.3, dt=1.1, offset_s=1.25,
                 offset_t=offset_t, dsd=949.0, dso=541.0, dfs=dfs, orbit=360)
    ig = image_geom(nx=40, ny=36, nz=24, dx=1.1, dz=0.9, offset_z=0.5)
    rng = np.random.default_rng(seed)
    proj = rng.standard_normal((cg.ns, cg.nt, cg.na)).astype(np.float32)
    return cg, ig, np.asfortranarray(proj)


@pytest.mark.parametrize("kind", ["flat", "arc"])
@pytest.mark.parametrize("ia_skip,extrapolate_t,scale_dang",
                         [(1, 0, True), (2, 0, True), (1, 3, False), (3, 2, True)])
def test_device_route_equals_host_route(device, kind, ia_skip, extrapolate_t, scale_dang):
    from mirt_amd import cbct_back, fdk_mex
    cg, ig, proj = _scan(kind, seed=ia_skip * 10 + extrapolate_t)
    print(f"cbct_back {kind} ia_skip={ia_skip} extrapolate_t={extrapolate_t} "
          f"scale_dang={scale_dang}", flush=True)
    dev = cbct_back(proj, cg, ig, ia_skip=ia_skip, extrapolate_t=extrapolate_t,
                    scale_dang=scale_dang)
    host = cbct_back(proj, cg, ig, ia_skip=ia_skip, extrapolate_t=extrapolate_t,
                     scale_dang=scale_dang, back_call=fdk_mex)
    assert dev.shape == host.shape == (ig.nx, ig.ny, ig.nz)
    assert np.array_equal(dev, host), float(np.abs(dev - host).max())
    assert np.abs(dev).max() > 0


def test_device_route_refuses_offset_t_with_extrapolation(device):
    from mirt_amd import cbct_back
    cg, ig, proj = _scan("flat", seed=1, offset_t=0.5)
    with pytest.raises(ValueError, match="offset_t not yet done"):
        cbct_back(proj, cg, ig, extrapolate_t=2)


def test_device_route_cache(device):
    """fdkhip_cbct_back keeps its plan and buffers between calls (shared with
    fdkhip_fdk_ts_back's): the same call twice, new views in the same
    geometry, a new geometry, and the first again after
    fdkhip_release_host_cache, each bit for bit equal to the host route."""
    from mirt_amd import _lib, cbct_back, fdk_mex
    cg, ig, proj = _scan("arc", seed=31)
    cg2, ig2, proj2 = _scan("flat", seed=32)
    rng = np.random.default_rng(33)
    proj_b = np.asfortranarray(rng.standard_normal(proj.shape).astype(np.float32))
    calls = [("first", cg, ig, proj), ("again", cg, ig, proj), ("new views", cg, ig, proj_b),
             ("new geometry", cg2, ig2, proj2), ("after release", cg, ig, proj)]
    for what, c, i, p in calls:
        if what == "after release":
            _lib.check(_lib.load().fdkhip_release_host_cache())
        print(f"cbct_back cache: {what}", flush=True)
        dev = cbct_back(p, c, i, ia_skip=2, extrapolate_t=1)
        host = cbct_back(p, c, i, ia_skip=2, extrapolate_t=1, back_call=fdk_mex)
        assert np.array_equal(dev, host), what
