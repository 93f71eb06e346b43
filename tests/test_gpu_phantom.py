"""GPU parity of ``fdkhip_ellipsoid_proj`` (ellipsoid_proj, fbp/ellipsoid_proj.m:35-125).

Bar: relative error <= 1e-10 (max |difference| over max |value|, FP64)
against the numpy restatement ``oracle/ellipsoid.py``, which tests/test_host.py
pins by closed forms and ray marching; and the closed-form sphere chords of
the parallel beam directly.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    from mirt_amd.geom import ct_geom
    return {
        "arc": ct_geom("fan", ns=96, nt=24, na=20, ds=1.0239, dt=1.0964, dsd=949.0, dso=541.0,
                       dfs=0.0, offset_s=1.25),
        "flat": ct_geom("fan", ns=80, nt=40, na=18, ds=1.6, dt=1.6, dsd=949.0, dso=541.0,
                        dfs=np.inf, offset_t=0.5),
        "par": ct_geom("fan", ns=64, nt=16, na=12, ds=2.0, dt=2.0, dsd=np.inf),
        "helix": ct_geom("fan", ns=72, nt=16, na=30, ds=1.0239, dt=1.0964, dsd=949.0,
                         dso=541.0, dfs=0.0, orbit=720, pitch=1.0, source_z0="center"),
    }


def _phantom(seed, n=14):
    from mirt_amd.phantom import random_ellipsoids
    return random_ellipsoids(seed, n, 120.0, 40.0, inserts=3)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("oversample", [1, 2])
@pytest.mark.parametrize("name", ["arc", "flat", "par", "helix"])
def test_ellipsoid_kernel_vs_oracle(device, name, oversample):
    from mirt_amd.phantom import ellipsoid_proj
    from oracle.ellipsoid import ellipsoid_proj as oracle_proj
    cg = _cases()[name]
    ell = _phantom(31)
    print(f"ellipsoid_proj case={name} oversample={oversample}", flush=True)
    got = ellipsoid_proj(cg, ell, oversample=oversample, device=device).cpu().numpy()
    want = oracle_proj(cg.s, cg.t, cg.ar, cg.source_zs, cg.dso, cg.dod, cg.dfs, ell, oversample)
    assert got.shape == want.shape == (cg.ns, cg.nt, cg.na)
    err = _rel(got, want)
    print(f"  rel err = {err:.2e}", flush=True)
    assert err <= 1e-10, err


def test_ellipsoid_kernel_sphere_chords(device):
    from mirt_amd.geom import ct_geom
    from mirt_amd.phantom import ellipsoid_proj
    cg = ct_geom("fan", ns=41, nt=9, na=4, ds=1.0, dt=1.0, dsd=np.inf)
    ell = np.array([[0, 0, 0, 10, 10, 10, 0, 0, 0.5]])
    p = ellipsoid_proj(cg, ell, device=device).cpu().numpy()
    want = 2 * 0.5 * np.sqrt(np.clip(100 - cg.s[:, None] ** 2 - cg.t[None, :] ** 2, 0, None))
    for ia in range(cg.na):
        assert np.allclose(p[:, :, ia], want, rtol=0, atol=1e-12)


def test_ellipsoid_kernel_views_and_errors(device):
    from mirt_amd._lib import FdkhipError
    from mirt_amd.geom import ct_geom
    from mirt_amd.phantom import ellipsoid_proj
    from oracle.ellipsoid import ellipsoid_proj as oracle_proj
    cg = _cases()["arc"]
    ell = _phantom(32, n=5)
    got = ellipsoid_proj(cg, ell, device=device, views=[3, 7, 11]).cpu().numpy()
    want = oracle_proj(cg.s, cg.t, cg.ar[[3, 7, 11]], 0.0, cg.dso, cg.dod, cg.dfs, ell)
    assert _rel(got, want) <= 1e-10
    bad = ell.copy()
    bad[0, 7] = 10.0
    with pytest.raises(FdkhipError, match="z rotation not done"):
        ellipsoid_proj(cg, bad, device=device)
    odd = ct_geom("fan", ns=8, nt=4, na=2, ds=1.0, dt=1.0, dsd=949.0, dso=541.0, dfs=100.0)
    with pytest.raises(FdkhipError, match="not done"):
        ellipsoid_proj(odd, ell, device=device)
