"""fdkhip_back_noalloc: the allocation-free backprojection entry point.

It must equal the plan path it is a variant of (fdkhip_plan_create with
FDKHIP_MATH_FAST + fdkhip_plan_execute_direct) bit for bit, on two calls in
a row on one GPU (the second with a smaller scan, so that the shared device
scratch holds the first call's tables when the second starts), and refuse
scans beyond its fixed scratch with a clear error.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import make_case


def _geom(g):
    from mirt_amd._lib import Geom
    return Geom(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.offset_x, g.offset_y, g.offset_z,
                g.dso, g.dsd, g.dfs, g.ns, g.nt, g.ds, g.dt, g.offset_s, g.offset_t)


def _noalloc(geom, beta, mask_xy, fmt, q, sc, out, stream):
    from mirt_amd import _lib
    beta = np.ascontiguousarray(beta, np.float64)
    # [nx ny], x fastest (MATLAB order), as FdkPlan hands it to the library
    mask = None if mask_xy is None else np.asfortranarray(np.asarray(mask_xy) != 0).astype(
        np.uint8, order="F")
    return _lib.load().fdkhip_back_noalloc(
        C.byref(geom), len(beta), beta.ctypes.data_as(C.c_void_p),
        None if mask is None else mask.ctypes.data_as(C.c_void_p), _lib.FMT_NAMES[fmt],
        C.c_void_p(q.data_ptr()), C.c_void_p(sc.data_ptr()), C.c_void_p(out.data_ptr()),
        C.c_void_p(stream))


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_noalloc_equals_plan_path(device, fmt):
    import torch
    from mirt_amd import FdkPlan, _lib, quantize_views
    s = torch.cuda.current_stream(device)
    for name in ("arc_tall", "flat"):          # the second scan is the smaller one
        g, proj, beta, mask = make_case(name, seed=41, data="noise")
        print(f"noalloc {fmt} case={name}", flush=True)
        q, sc = quantize_views(torch.from_numpy(proj).to(device), fmt)
        geom = _geom(g)
        want = FdkPlan(geom, beta, mask.T, fmt=fmt)(q, sc, graph=False)
        got = torch.full((g.ny, g.nx, g.nz), float("nan"), device=device)
        _lib.check(_noalloc(geom, beta, mask.T, fmt, q, sc, got, s.cuda_stream))
        torch.cuda.synchronize()
        assert torch.equal(got, want)


def test_noalloc_limits():
    """Argument checks run before any device call (no GPU needed)."""
    from mirt_amd import _lib
    import torch
    g, proj, beta, mask = make_case("flat", seed=42)
    t = torch.zeros(1)
    geom = _geom(g)
    rc = _noalloc(geom, beta, None, "f32", t, t, t, 0)
    assert rc == _lib.FDKHIP_EUNSUPPORTED
    assert "F16 or BF16" in _lib.load().fdkhip_last_error().decode()
    big = np.zeros(16385)
    rc = _noalloc(geom, big, None, "f16", t, t, t, 0)
    assert rc == _lib.FDKHIP_EUNSUPPORTED
    assert "at most 16384 views" in _lib.load().fdkhip_last_error().decode()
    geom.nx, geom.ny = 4096, 2048
    rc = _noalloc(geom, beta, None, "f16", t, t, t, 0)
    assert rc == _lib.FDKHIP_EUNSUPPORTED
    assert "columns without allocation" in _lib.load().fdkhip_last_error().decode()
