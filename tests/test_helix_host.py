"""Host logic of the helical path, runnable without a GPU.

* cbct_back_helix (mex/src/fdk-helix/cbct_back_helix.m) with an injected
  back_call computed by the oracle -- the reference's own fake-backend
  strategy: the 19 gateway operands (na1_half, source_dz_per_rad, pitch as
  single, the view subset and its heights), permutes, row replication and the
  per-slice scale, against a direct call of the restatement;
* the per-slice scale against its closed forms (axial: pi / (na/ia_skip);
  a helical slice whose window holds a full turn: 0.5 * dbeta) and MATLAB's
  int32 rounding (half away from zero, saturating);
* fdk_mex_helix argument checks (the fdk-helix gateway's messages).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O


def helix_back_call(cmd, nxyz, dxyz, cxyz, mask, dso, dsd, dfs, dst, ost, proj, beta,
                    source_zs, na1_half, cone_par, w3d, source_dz_per_rad, pitch, nthread):
    """A back_call with the helical gateway's contract, computed by the
    oracle's restatement (tests only)."""
    assert cmd == "fdk,ts,back"
    assert na1_half.dtype == np.int32 and cone_par.dtype == np.int32
    assert w3d.dtype == np.int32 and pitch.dtype == np.float32
    assert source_dz_per_rad.dtype == np.float64 and nthread.dtype == np.int32
    proj = np.asarray(proj, np.float32)
    nt, ns = proj.shape[0], proj.shape[1]
    g = O.geom(nxyz, dxyz, cxyz, dso, dsd, dfs, ns, nt, dst, ost)
    pc = np.ascontiguousarray(np.asfortranarray(proj).reshape(-1, order="F"))
    m = np.asfortranarray(mask).reshape(-1, order="F")
    out = O.back_ts_helix_f32(g, pc, np.atleast_1d(beta), np.atleast_1d(source_zs), m,
                              na1_half=int(na1_half[0]), cone_par=int(cone_par[0]),
                              w3d=int(w3d[0]), source_dz_per_rad=float(source_dz_per_rad[0]),
                              pitch=float(pitch[0]))
    return np.asfortranarray(out.transpose(2, 1, 0))                 # MATLAB [nz nx ny]


@pytest.fixture(scope="module")
def helix_scan():
    from mirt_amd.geom import ct_geom, image_geom
    cg = ct_geom("fan", ns=40, nt=16, na=72, ds=1.5, dt=1.6, offset_s=0.25, dsd=500.0,
                 dso=300.0, dfs=0.0, orbit=540, pitch=1.0, source_z0="center")
    ig = image_geom(nx=14, ny=12, nz=20, dx=2.0, dz=1.0, mask="all-but-edge")
    rng = np.random.default_rng(3)
    proj = np.asfortranarray(rng.standard_normal(cg.dim).astype(np.float32))
    return cg, ig, proj


def _direct(cg, ig, proj, ia_skip, cone_par=0, w3d=0, nt=None):
    """The restatement called directly; nt: detector rows of proj when the
    edge rows were replicated (the heights stay the scan's)."""
    from mirt_amd.cbct_back_helix import helix_operands
    g = O.geom((ig.nx, ig.ny, ig.nz), (ig.dx, ig.dy, ig.dz),
               (ig.offset_x, ig.offset_y, ig.offset_z), cg.dso, cg.dsd, cg.dfs,
               cg.ns, nt or cg.nt, (cg.ds, cg.dt), (cg.offset_s, cg.offset_t))
    beta = np.deg2rad(cg.orbit_start + cg.orbit * np.arange(cg.na) / cg.na)[::ia_skip]
    zs = np.asarray(cg.source_zs)[::ia_skip]
    tsa = np.ascontiguousarray(proj.transpose(2, 0, 1)[::ia_skip])
    mask = np.ascontiguousarray(ig.mask_or.T.astype(np.uint8))
    na1_half, dpr = helix_operands(cg, ia_skip)
    out = O.back_ts_helix_f32(g, tsa, beta, zs, mask, na1_half=na1_half, cone_par=cone_par,
                              w3d=w3d, source_dz_per_rad=dpr, pitch=cg.pitch)
    return out.transpose(1, 0, 2)


@pytest.mark.parametrize("ia_skip", [1, 2])
def test_cbct_back_helix_host_steps(helix_scan, ia_skip):
    from mirt_amd.cbct_back_helix import cbct_back_helix
    cg, ig, proj = helix_scan
    got = cbct_back_helix(proj, cg, ig, ia_skip=ia_skip, back_call=helix_back_call,
                          scale_dang=False)
    assert got.shape == (ig.nx, ig.ny, ig.nz)
    assert np.array_equal(got, _direct(cg, ig, proj, ia_skip))
    assert np.abs(got).max() > 0


def test_cbct_back_helix_cone_par_and_scale(helix_scan):
    from mirt_amd.cbct_back_helix import cbct_back_helix, helix_operands, slice_scales
    cg, ig, proj = helix_scan
    with pytest.warns(UserWarning, match="end slices"):
        got = cbct_back_helix(proj, cg, ig, back_call=helix_back_call, cone_par=1, w3d=1)
    sc = slice_scales(cg, ig, 1, helix_operands(cg, 1)[0]).astype(np.float32)
    want = _direct(cg, ig, proj, 1, cone_par=1, w3d=1) * sc[None, None, :]
    assert np.array_equal(got, want)


def test_cbct_back_helix_extrapolate_t(helix_scan):
    from mirt_amd.cbct_back_helix import cbct_back_helix
    cg, ig, proj = helix_scan
    got = cbct_back_helix(proj, cg, ig, back_call=helix_back_call, scale_dang=False,
                          extrapolate_t=3)
    big = np.concatenate([np.repeat(proj[:, :1], 3, axis=1), proj,
                          np.repeat(proj[:, -1:], 3, axis=1)], axis=1)
    assert np.array_equal(got, _direct(cg, ig, big, 1, nt=cg.nt + 6))


def test_int32_rounding_like_matlab():
    from mirt_amd.cbct_back_helix import _int32
    assert [_int32(v) for v in (0.5, -0.5, 1.5, 2.5, -2.5, 2.4999)] == [1, -1, 2, 3, -3, 2]
    assert _int32(1e12) == 2 ** 31 - 1 and _int32(-1e12) == -2 ** 31


def test_helix_operands_closed_form():
    from mirt_amd.cbct_back_helix import helix_operands
    from mirt_amd.geom import ct_geom
    cg = ct_geom("fan", ns=16, nt=8, na=984 * 3, ds=1.0, dt=1.0, dsd=949.0, dso=541.0,
                 dfs=0.0, orbit=360 * 3, pitch=1.0)
    na1_half, dpr = helix_operands(cg)
    assert na1_half == 492
    assert math.isclose(dpr, cg.source_dz_per_view * 984 / (2 * math.pi), rel_tol=1e-14)
    assert helix_operands(cg, 3)[0] == 164


def test_slice_scales_closed_forms(helix_scan):
    from mirt_amd.cbct_back_helix import helix_operands, slice_scales
    from mirt_amd.geom import ct_geom
    cg, ig, _ = helix_scan
    axial = ct_geom("fan", ns=40, nt=16, na=72, ds=1.5, dt=1.6, dsd=500.0, dso=300.0,
                    dfs=0.0, orbit=360)
    s = slice_scales(axial, ig, 2, helix_operands(axial, 2)[0])
    assert np.allclose(s, math.pi / 36, rtol=1e-15, atol=0)
    na1_half = helix_operands(cg, 1)[0]           # 24 views per half turn
    s = slice_scales(cg, ig, 1, na1_half)
    dbeta = math.radians(cg.orbit) / cg.na
    # central slices: their window holds 2 * na1_half views = one full turn
    assert math.isclose(s[ig.nz // 2], 0.5 * dbeta, rel_tol=1e-14)
    assert s[0] > s[ig.nz // 2] and s[-1] > s[ig.nz // 2]   # end slices see less orbit


def test_fdk_mex_helix_argument_checks():
    from mirt_amd.cbct_back import FdkMexError, fdk_mex_helix
    args = ["fdk,ts,back", np.array([4, 4, 4], np.int32), np.ones(3), np.zeros(3),
            np.ones((4, 4), np.uint8), 541.0, 949.0, 0.0, np.ones(2), np.zeros(2),
            np.zeros((6, 8, 2), np.float32), np.zeros(2), np.zeros(2),
            np.array([1], np.int32), np.array([0], np.int32), np.array([0], np.int32),
            np.array([0.0]), np.array([0.0], np.float32), np.array([1], np.int32)]
    cases = [(0, "fdk,xx,back", "command 'fdk,xx,back' unknown"),
             (13, np.array([1], np.int64), "mxIsScalarInt32"),
             (17, np.array([0.0]), "mxIsRealSingle"),
             (16, np.array([0.0], np.float32), "mxIsRealDouble"),
             (12, np.zeros(2, np.float32), "mxIsRealDouble"),
             (11, np.zeros(3), "beta length=3 vs na=size(proj,3)=2"),
             (10, np.zeros((6, 8, 2), np.float64), "mxIsRealSingle")]
    for pos, val, msg in cases:
        a = list(args)
        a[pos] = val
        with pytest.raises(FdkMexError, match=msg.replace("(", r"\(").replace(")", r"\)")):
            fdk_mex_helix(*a)
    with pytest.raises(FdkMexError, match="usage"):
        fdk_mex_helix(*args, 1)
