"""Host-side logic, runnable without a GPU.

* libfdkhip.so loads and exports every function include/fdkhip.h declares;
  argument errors are reported before any device call (codes + messages).
* fdk_mex mirrors the gateway's checks and messages (fdk,mex.c:39-90, 175-214).
* cbct_back's host steps (fbp/cbct_back.m:74-137) with an injected back_call
  -- the reference's own fake-backend strategy (cbct_back.m:19,38): angles,
  permutes, ia_skip, extrapolate_t, the angular scale, use_mex 1/2/3.
* ct_geom / image_geom conventions; fbp_ramp closed form; fdk_filter against
  direct convolution with the ramp taps; ellipsoid_proj sphere chords and the
  parallel-beam closed form.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT
from oracle import oracle as O


# --------------------------------------------------------------------------
# C ABI

def test_library_exports_header_symbols():
    from mirt_amd import _lib
    lib = _lib.load()
    hdr = open(os.path.join(ROOT, "include", "fdkhip.h")).read()
    declared = set(re.findall(r"\b(fdkhip_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert declared <= set(_lib.exported_symbols()) | {"fdkhip_version"}
    assert "gfx950" in _lib.version()


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _call_ts(**over):
    from mirt_amd import _lib
    lib = _lib.load()
    args = dict(nxyz=np.array([4, 4, 4], np.int32), dxyz=np.ones(3), cxyz=np.zeros(3),
                mask=np.ones(16, np.uint8), dso=541.0, dsd=949.0, dfs=0.0,
                dst=np.ones(2), ost=np.zeros(2), proj=np.zeros(8 * 6 * 2, np.float32),
                ns=8, nt=6, na=2, beta=np.zeros(2), nthread=1,
                image=np.zeros(64, np.float32))
    args.update(over)
    rc = lib.fdkhip_fdk_ts_back(_ptr(args["nxyz"]), _ptr(args["dxyz"]), _ptr(args["cxyz"]),
                                _ptr(args["mask"]), args["dso"], args["dsd"], args["dfs"],
                                _ptr(args["dst"]), _ptr(args["ost"]), _ptr(args["proj"]),
                                args["ns"], args["nt"], args["na"], _ptr(args["beta"]),
                                args["nthread"], _ptr(args["image"]))
    return rc, lib.fdkhip_last_error().decode()


def test_cabi_argument_errors_before_device():
    from mirt_amd import _lib
    rc, msg = _call_ts(nthread=251)
    assert rc == _lib.FDKHIP_ETHREADS and msg == "too many threads"
    rc, msg = _call_ts(nthread=0)
    assert rc == _lib.FDKHIP_EARG and "nthread" in msg
    rc, msg = _call_ts(nxyz=None)
    assert rc == _lib.FDKHIP_EARG and "mx_nxyz" in msg
    rc, msg = _call_ts(beta=None)
    assert rc == _lib.FDKHIP_EARG and "beta length" in msg
    rc, msg = _call_ts(image=None)
    assert rc == _lib.FDKHIP_EARG and "image" in msg
    rc, msg = _call_ts(na=-1)
    assert rc == _lib.FDKHIP_EARG


def test_batch_argument_errors_before_device():
    import ctypes as C
    from mirt_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.fdkhip_batch_create(None, 1, None, None, None, None, 0, 0) == _lib.FDKHIP_EARG
    assert lib.fdkhip_batch_create(C.byref(h), 0, None, None, None, None, 0, 0) \
        == _lib.FDKHIP_EARG
    assert "nscan" in lib.fdkhip_last_error().decode()
    assert h.value is None
    assert lib.fdkhip_batch_execute(None, None, None, None, None) == _lib.FDKHIP_EARG
    assert lib.fdkhip_batch_set_config(None, 8, 16, 0) == _lib.FDKHIP_EARG
    assert lib.fdkhip_batch_count_updates(None, None) == _lib.FDKHIP_EARG
    assert lib.fdkhip_batch_destroy(None) == _lib.FDKHIP_OK
    assert lib.fdkhip_plan_autotune(None, None, None, None, None, 1, None) \
        == _lib.FDKHIP_EARG


def test_cbct_back_entry_argument_errors_before_device():
    import ctypes as C
    from mirt_amd import _lib
    from mirt_amd._lib import Geom
    lib = _lib.load()
    img = np.zeros(8, np.float32)
    g = Geom(2, 2, 2, 1.0, -1.0, 1.0, 0, 0, 0, 541.0, 949.0, 0.0, 8, 4, 1.0, 1.0, 0.0, 0.5)
    beta = np.zeros(2)
    proj = np.zeros(64, np.float32)
    P = lambda a: C.c_void_p(a.ctypes.data)
    assert lib.fdkhip_cbct_back(None, P(proj), 2, P(beta), 1, None, 0, 1.0, P(img)) \
        == _lib.FDKHIP_EARG
    rc = lib.fdkhip_cbct_back(C.byref(g), P(proj), 2, P(beta), 1, None, 2, 1.0, P(img))
    assert rc == _lib.FDKHIP_EARG
    assert lib.fdkhip_last_error().decode() == "offset_t not yet done for extrapolation"
    assert lib.fdkhip_cbct_back(C.byref(g), P(proj), 2, P(beta), 0, None, 0, 1.0, P(img)) \
        == _lib.FDKHIP_EARG
    assert lib.fdkhip_cbct_back(C.byref(g), P(proj), 2, None, 1, None, 0, 1.0, P(img)) \
        == _lib.FDKHIP_EARG


def test_cbct_back_mat_argument_errors_before_device():
    """cbct_back_mat's own failures (an undefined sprime for a dfs that is
    neither 0 nor inf, a row index 0 when nt = 1) and bad counts are refused
    before any device call, by both entry points."""
    import ctypes as C
    from mirt_amd import _lib
    from mirt_amd._lib import Geom
    lib = _lib.load()
    img = np.zeros(8, np.float64)
    beta = np.zeros(2)
    proj = np.zeros(64, np.float32)
    P = lambda a: C.c_void_p(a.ctypes.data)
    g = lambda **kw: Geom(*[kw.get(k, v) for k, v in zip(
        [f[0] for f in Geom._fields_],
        (2, 2, 2, 1.0, -1.0, 1.0, 0, 0, 0, 541.0, 949.0, 0.0, 8, 4, 1.0, 1.0, 0.0, 0.5))])
    host = lambda gg, na=2, skip=1, beta=beta: lib.fdkhip_cbct_back_mat(
        C.byref(gg), P(proj), 0, na, None if beta is None else P(beta), None, skip, 0.0,
        None, 1.0, P(img))
    direct = lambda gg, na=2, skip=1, beta=beta: lib.fdkhip_cbct_back_mat_direct(
        C.byref(gg), P(proj), 0, na, None if beta is None else P(beta), None, skip, 0.0,
        None, 1.0, P(img), None, None)
    for call in (host, direct):
        assert call(g(dfs=500.0)) == _lib.FDKHIP_EARG
        assert "dfs must be 0 (arc) or inf (flat)" in lib.fdkhip_last_error().decode()
        assert call(g(nt=1)) == _lib.FDKHIP_EARG
        assert "nt >= 2" in lib.fdkhip_last_error().decode()
        assert call(g(), skip=0) == _lib.FDKHIP_EARG
        assert call(g(), na=-1) == _lib.FDKHIP_EARG
        assert call(g(), beta=None) == _lib.FDKHIP_EARG
        # a parallel geometry takes any dfs (cbct_back.m:211: par before arc/flat)
        assert call(g(dfs=500.0, dso=np.inf, dsd=np.inf, nz=0)) == 0
    assert lib.fdkhip_cbct_back_mat(None, P(proj), 0, 2, P(beta), None, 1, 0.0, None, 1.0,
                                    P(img)) == _lib.FDKHIP_EARG


def test_cbct_back_helix_mat_argument_errors_before_device():
    """fdkhip_cbct_back_helix_mat rejects a slice's view range outside the
    views, cone_par other than 0/1 and a missing range table before it
    touches the device."""
    from mirt_amd import _lib
    lib = _lib.load()
    g = _lib.Geom(4, 4, 2, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 541.0, 949.0, np.inf, 8, 6, 1.0, 1.0,
                  0.0, 0.0)
    proj = np.zeros((8, 6, 3), np.float32, order="F")
    beta = np.zeros(3)
    zs = np.zeros(3)
    img = np.zeros((4, 4, 2))
    P = lambda a: None if a is None else C.c_void_p(a.ctypes.data)
    def call(views, cone_par=0):
        return lib.fdkhip_cbct_back_helix_mat(C.byref(g), P(proj), 0, 3, P(beta), P(zs), 1,
                                              0.0, P(views), None, cone_par, 0.0, None, P(img))
    assert call(np.array([[0, 3], [0, 2]], np.int32)) == _lib.FDKHIP_EARG
    assert "outside" in lib.fdkhip_last_error().decode()
    assert call(np.array([[-1, 2], [0, 2]], np.int32)) == _lib.FDKHIP_EARG
    assert call(np.array([[0, 2], [0, 2]], np.int32), cone_par=2) == _lib.FDKHIP_EARG
    assert "cone_par" in lib.fdkhip_last_error().decode()
    assert call(None) == _lib.FDKHIP_EARG


def test_cbct_back_helix_mat_python_checks():
    from mirt_amd import cbct_back_helix
    from mirt_amd.cbct_back import FdkMexError
    from mirt_amd.geom import ct_geom, image_geom
    ig = image_geom(nx=8, ny=8, nz=4, dx=1.0, dz=1.0)
    cg = ct_geom("fan", ns=16, nt=8, na=25, ds=1.0, dt=1.0, dsd=949.0, dso=541.0,
                 dfs=np.inf, orbit=361, pitch=0.5)
    proj = np.zeros((16, 8, 25), np.float32)
    with pytest.warns(UserWarning), pytest.raises(FdkMexError, match="bad na/orbit"):
        cbct_back_helix(proj, cg, ig, use_mex=0)
    cg = ct_geom("fan", ns=16, nt=8, na=24, ds=1.0, dt=1.0, dsd=949.0, dso=541.0,
                 dfs=np.inf, orbit=360, pitch=0.5)
    with pytest.raises(FdkMexError, match="Min/Max"):
        cbct_back_helix(np.zeros((16, 8, 24), np.float32), cg, ig, use_mex=0, cone_par=1,
                        w3d=1, extrapolate_t=1)


def test_cbct_back_mat_python_checks():
    from mirt_amd import cbct_back
    from mirt_amd.geom import ct_geom, image_geom
    cg = ct_geom("fan", ns=8, nt=4, na=3, ds=1.0, dt=1.0, dsd=949.0, dso=541.0, dfs=0.0)
    ig = image_geom(nx=4, ny=4, nz=2, dx=1.0, dz=1.0)
    with pytest.raises(ValueError, match="proj shape"):
        cbct_back(np.zeros((8, 4, 2), np.float32), cg, ig, use_mex=0)
    with pytest.raises(ValueError, match="ia_skip"):
        cbct_back(np.zeros((8, 4, 3), np.float32), cg, ig, use_mex=0, ia_skip=0)


def test_quantize_argument_errors():
    from mirt_amd import _lib
    lib = _lib.load()
    rc = lib.fdkhip_quantize_views(None, 4, 4, 2, 0, None, None, None)
    assert rc == _lib.FDKHIP_EARG
    assert "F16 or BF16" in lib.fdkhip_last_error().decode()
    # empty input is a no-op
    assert lib.fdkhip_quantize_views(None, 0, 4, 2, 1, None, None, None) == 0


# --------------------------------------------------------------------------
# fdk_mex gateway mirror (argument checks only: no device call happens)

def _mex_args():
    return ["fdk,ts,back", np.array([4, 4, 4], np.int32), np.ones(3), np.zeros(3),
            np.ones((4, 4), np.uint8), 541.0, 949.0, 0.0, np.ones(2), np.zeros(2),
            np.zeros((6, 8, 2), np.float32), np.zeros(2), np.array([1], np.int32)]


@pytest.mark.parametrize("pos,val,msg", [
    (0, "fdk,xx,back", "command 'fdk,xx,back' unknown"),
    (1, np.array([4, 4, 4], np.int64), "mxIsInt32"),
    (1, np.array([4, 4], np.int32), "mx_nxyz must have 3 elements"),
    (2, np.ones(3, np.float32), "mxIsRealDouble"),
    (3, np.zeros(2), "mx_cxyz must have 3 elements"),
    (4, np.ones((4, 4), np.int8), "mxIsUint8"),
    (4, np.ones((4, 4, 1), np.uint8), "mask must be 2d"),
    (4, np.ones((4, 5), np.uint8), "mask size not nx by ny?"),
    (8, np.ones(3), "mx_dst must have 2 elements"),
    (9, np.zeros(1), "mx_ost must have 2 elements"),
    (10, np.zeros((6, 8, 2), np.float64), "mxIsRealSingle"),
    (10, np.zeros((6, 8, 2, 1), np.float32), "mx_proj must be 2d or 3d"),
    (11, np.zeros(3), "beta length=3 vs na=size(proj,3)=2"),
    (12, np.array([1], np.int64), "mxIsScalarInt32"),
])
def test_fdk_mex_checks(pos, val, msg):
    from mirt_amd import FdkMexError, fdk_mex
    args = _mex_args()
    args[pos] = val
    with pytest.raises(FdkMexError) as e:
        fdk_mex(*args)
    assert str(e.value) == msg


def test_fdk_mex_usage():
    from mirt_amd import FdkMexError, fdk_mex
    args = _mex_args()
    with pytest.raises(FdkMexError, match="usage error"):
        fdk_mex(*args, np.zeros(1))                 # nrhs != 13
    args[0] = 3
    with pytest.raises(FdkMexError, match="usage error"):
        fdk_mex(*args)


# --------------------------------------------------------------------------
# cbct_back host steps with an injected back_call (the oracle, on the CPU)

def oracle_back_call(cmd, nxyz, dxyz, cxyz, mask, dso, dsd, dfs, dst, ost, proj, beta,
                     nthread):
    """A back_call with fdk_mex's contract, computed by the oracle (tests only)."""
    proj = np.asarray(proj, np.float32)
    if cmd == "fdk,ts,back":
        nt, ns = proj.shape[0], proj.shape[1]
    else:
        ns, nt = proj.shape[0], proj.shape[1]
    g = O.geom(nxyz, dxyz, cxyz, dso, dsd, dfs, ns, nt, dst, ost)
    beta = np.atleast_1d(np.asarray(beta, np.float64))
    pc = np.ascontiguousarray(np.asfortranarray(proj).reshape(-1, order="F"))
    m = np.asfortranarray(mask).reshape(-1, order="F")
    if cmd == "fdk,ts,back":
        out = O.back_ts_f32(g, pc, beta, m)                       # C (ny, nx, nz)
        return np.asfortranarray(out.transpose(2, 1, 0))          # MATLAB [nz nx ny]
    out = O.back_st_f32(g, pc, beta, m)                           # C (nz, ny, nx)
    return np.asfortranarray(out.transpose(2, 1, 0))              # MATLAB [nx ny nz]


@pytest.fixture(scope="module")
def small_scan():
    from mirt_amd.geom import ct_geom, image_geom
    cg = ct_geom("fan", ns=40, nt=16, na=12, ds=1.5, dt=1.6, offset_s=0.25, dsd=500.0,
                 dso=300.0, dfs=0.0, orbit=360)
    ig = image_geom(nx=14, ny=12, nz=10, dx=2.0, dz=1.8, mask="all-but-edge")
    rng = np.random.default_rng(3)
    proj = np.asfortranarray(rng.standard_normal(cg.dim).astype(np.float32))
    return cg, ig, proj


def direct_backprojection(cg, ig, proj, ia_skip=1):
    """[nx ny nz] by the oracle directly (no cbct_back logic)."""
    g = O.geom((ig.nx, ig.ny, ig.nz), (ig.dx, ig.dy, ig.dz),
               (ig.offset_x, ig.offset_y, ig.offset_z), cg.dso, cg.dsd, cg.dfs,
               cg.ns, cg.nt, (cg.ds, cg.dt), (cg.offset_s, cg.offset_t))
    beta = np.deg2rad(cg.orbit_start + cg.orbit * np.arange(cg.na) / cg.na)[::ia_skip]
    tsa = np.ascontiguousarray(proj.transpose(2, 0, 1)[::ia_skip])   # (na, ns, nt)
    mask = np.ascontiguousarray(ig.mask_or.T.astype(np.uint8))
    out = O.back_ts_f32(g, tsa, beta, mask)                              # (ny, nx, nz)
    return out.transpose(1, 0, 2)


def test_cbct_back_use_mex_1_2_3(small_scan):
    from mirt_amd import cbct_back
    from mirt_amd.cbct_back import angle_scale
    cg, ig, proj = small_scan
    want = direct_backprojection(cg, ig, proj)
    img1 = cbct_back(proj, cg, ig, back_call=oracle_back_call, scale_dang=False)
    assert img1.shape == (ig.nx, ig.ny, ig.nz)
    assert np.array_equal(img1, want)
    img2 = cbct_back(proj, cg, ig, use_mex=2, back_call=oracle_back_call, scale_dang=False)
    assert np.array_equal(img2, want)   # per-view sums in view order: same roundings
    img3 = cbct_back(proj, cg, ig, use_mex=3, back_call=oracle_back_call, scale_dang=False)
    assert img3.shape == want.shape and np.isfinite(img3).all()
    scaled = cbct_back(proj, cg, ig, back_call=oracle_back_call)
    s = np.float32(0.5 * np.deg2rad(360.0) / cg.na)
    assert angle_scale(cg) == s
    assert np.array_equal(scaled, s * want)


def test_cbct_back_ia_skip_and_extrapolate(small_scan):
    from mirt_amd import cbct_back
    cg, ig, proj = small_scan
    want = direct_backprojection(cg, ig, proj, ia_skip=3)
    img = cbct_back(proj, cg, ig, ia_skip=3, back_call=oracle_back_call, scale_dang=False)
    assert np.array_equal(img, want)
    # extrapolate_t: rows replicated outward (cbct_back.m:86-94)
    seen = {}

    def spy(cmd, *a):
        seen["proj"] = a[9]
        return oracle_back_call(cmd, *a)
    cbct_back(proj, cg, ig, extrapolate_t=3, back_call=spy, scale_dang=False)
    p = seen["proj"]                      # [nt+6 ns na]
    assert p.shape == (cg.nt + 6, cg.ns, cg.na)
    assert np.array_equal(p[0], proj[:, 0, :]) and np.array_equal(p[2], proj[:, 0, :])
    assert np.array_equal(p[-1], proj[:, -1, :])
    assert np.array_equal(p[3:-3], proj.transpose(1, 0, 2))


def test_cbct_back_rejects_helix_for_mex_routes(small_scan):
    """cbct_back.m:46-49 refuses a helix for use_mex >= 1 only; use_mex = 0
    (cbct_back_mat) takes source_zs (GPU test in test_gpu_mat.py)."""
    from mirt_amd import cbct_back
    from mirt_amd.geom import ct_geom
    cg, ig, proj = small_scan
    helix = ct_geom("fan", ns=40, nt=16, na=12, ds=1.5, dt=1.6, dsd=500.0, dso=300.0,
                    pitch=1.0)
    for use_mex in (1, 2, 3):
        with pytest.raises(ValueError, match="helix not yet supported"):
            cbct_back(proj, helix, ig, use_mex=use_mex, back_call=oracle_back_call)
    from mirt_amd.geom import ct_geom as cgf
    cg2 = cgf("fan", ns=40, nt=16, na=12, ds=1.5, dt=1.6, dsd=500.0, dso=300.0,
              offset_t=0.5)
    with pytest.raises(ValueError, match="offset_t not yet done"):
        cbct_back(proj, cg2, ig, extrapolate_t=2, back_call=oracle_back_call)


# --------------------------------------------------------------------------
# geometry conventions

def test_ct_geom_conventions():
    from mirt_amd.geom import ct_geom
    cg = ct_geom("ge1", nt=800, na=400, down=8)         # cbct_back.m:293
    assert (cg.ns, cg.nt, cg.na) == (112, 100, 50)       # 4*ceil(888/8/4), 2*ceil(.../2)
    assert cg.ds == pytest.approx(1.0239 * 8) and cg.dt == pytest.approx(1.0964 * 8)
    assert cg.ws == (112 - 1) / 2 + 1.25 and cg.dfs == 0 and cg.dsd == 949.075
    cg = ct_geom("fan", ns=888, nt=64, na=984, offset_s=1.25, dsd=949, dod=408)
    assert cg.dso == 541 and cg.dt == -1.0               # dt defaults to -ds
    assert np.allclose(cg.s[:2], (np.arange(2) - cg.ws) * 1.0)
    assert np.allclose(cg.ar, np.deg2rad(np.arange(984) / 984 * 360))
    cg = ct_geom("fan", ns=100, nt=8, dsd=np.inf)         # parallel handled gracefully
    assert math.isinf(cg.dso) and cg.na == 2 * math.floor(100 * math.pi / 2 / 2)
    cg = ct_geom("fan", ns=64, nt=8, na=10, dsd=1000, dso=500, dt=("dz", 0.5))
    assert cg.dt == pytest.approx(0.5 * 1000 / 500)
    with pytest.raises(ValueError):
        ct_geom("fan", ns=64, nt=8, dsd=1000)


def test_image_geom_conventions():
    from mirt_amd.geom import image_geom
    ig = image_geom(nx=512, ny=496, nz=480, fov=500, mask="all-but-edge", down=8)
    assert (ig.nx, ig.ny, ig.nz) == (64, 62, 60)
    assert ig.dx == pytest.approx(500 / 512 * 8) and ig.dy == -ig.dx and ig.dz == ig.dx
    mo = ig.mask_or
    assert mo.shape == (64, 62) and not mo[0].any() and not mo[:, -1].any()
    assert mo[1:-1, 1:-1].all()
    ig = image_geom(nx=10, ny=8, nz=6, dx=1.0, offsets="dsp")
    assert (ig.offset_x, ig.offset_y, ig.offset_z) == (0.5, 0.5, 0.5)
    assert np.allclose(ig.x, (np.arange(10) - ig.wx) * 1.0)
    m = np.zeros((8, 6, 4), bool)
    m[2, 3, 1] = True
    ig = image_geom(nx=8, ny=6, nz=4, dx=1.0, mask=m, down=2)
    assert ig.mask.shape == (4, 3, 2) and ig.mask[1, 1, 0] and ig.mask.sum() == 1


# --------------------------------------------------------------------------
# FDK weight / filter, phantom

def test_fbp_ramp_closed_form():
    from mirt_amd.fdk import fbp_ramp
    h = fbp_ramp("flat", 16, 2.0)
    nn = np.arange(-8, 8)
    assert h[8] == 1 / 4 / 4
    odd = nn % 2 == 1
    assert np.allclose(h[odd], -1 / (np.pi * nn[odd]) ** 2 / 4)
    assert np.all(h[(nn % 2 == 0) & (nn != 0)] == 0)
    ha = fbp_ramp("arc", 16, 2.0, 1e7)                     # arc -> flat as dsd -> inf
    assert np.allclose(ha, h, rtol=1e-6)


def test_fdk_filter_matches_direct_convolution():
    from mirt_amd.fdk import fbp_ramp, fdk_filter
    from mirt_amd.geom import ct_geom
    for dfs in (0.0, np.inf):
        cg = ct_geom("fan", ns=37, nt=3, na=2, ds=1.3, dt=1.0, dsd=800.0, dso=500.0,
                     dfs=dfs)
        rng = np.random.default_rng(1)
        p = rng.standard_normal((37, 3, 2))
        got = fdk_filter(torch.as_tensor(p), cg).numpy()
        # direct: zero-padded circular convolution of length npad with ds*h
        npad = 2 ** math.ceil(math.log2(2 * 37 - 1))
        h = cg.ds * fbp_ramp("flat" if math.isinf(dfs) else "arc", npad, cg.ds,
                             None if math.isinf(dfs) else cg.dsd)
        hs = np.fft.ifftshift(h)          # tap for lag k at index k mod npad
        want = np.zeros_like(p)
        for i in range(37):
            for j in range(37):
                want[i] += hs[(i - j) % npad] * p[j]
        assert np.allclose(got, want, atol=1e-10)


@pytest.mark.parametrize("ns", [37, 64])
@pytest.mark.parametrize("window", ["ramp", "hann"])
def test_toeplitz_filter_matches_fdk_filter(ns, window):
    """The Toeplitz matrix of the filter's taps (inverse DFT of fdk_filter's
    response: the direct-convolution form the GPU filter is checked against)
    filters rows exactly as fdk_filter's zero-padded FFT does, for arc and
    flat detectors."""
    from mirt_amd.fdk import fdk_filter, toeplitz_matrix
    from mirt_amd.geom import ct_geom
    for dfs in (0.0, np.inf):
        cg = ct_geom("fan", ns=ns, nt=3, na=2, ds=1.3, dt=1.0, dsd=800.0, dso=500.0, dfs=dfs)
        p = np.random.default_rng(ns).standard_normal((ns, 3, 2))
        want = fdk_filter(torch.as_tensor(p), cg, window).numpy()
        T = toeplitz_matrix(cg, ns, window)                 # T[s', s]
        got = np.einsum("ab,btv->atv", T, p)
        assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("ns", [2, 5, 37, 64, 888])
@pytest.mark.parametrize("window", ["ramp", "hann"])
def test_padded_response_is_fdk_filters_response(ns, window):
    """padded_response (the full-length H that fdkhip_fdk_filter_views takes)
    filters a zero-padded row by a complex FFT exactly as fdk_filter's
    real FFT with the non-negative half of the response does."""
    from mirt_amd.fdk import fdk_filter, padded_response
    from mirt_amd.geom import ct_geom
    for dfs in (0.0, np.inf):
        cg = ct_geom("fan", ns=ns, nt=2, na=1, ds=0.9, dt=1.0, dsd=949.0, dso=541.0, dfs=dfs)
        p = np.random.default_rng(ns).standard_normal((ns, 2, 1))
        npad, H = padded_response(cg, ns, window)
        assert npad == 2 ** math.ceil(math.log2(2 * ns - 1))
        assert np.array_equal(H[1:], H[1:][::-1])               # real and even
        got = np.real(np.fft.ifft(H[:, None, None] * np.fft.fft(p, n=npad, axis=0),
                                  axis=0))[:ns]
        want = fdk_filter(torch.as_tensor(p), cg, window).numpy()
        assert np.allclose(got, want, rtol=0, atol=1e-12 * max(np.abs(want).max(), 1e-300))


@pytest.fixture(scope="module")
def fft_core_host(tmp_path_factory):
    """tests/fft_core_host.cpp: the FFT passes of mirt_amd/csrc/fft_core.h
    (the filter kernel's own index and butterfly code) built for the host."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("fft") / "libfftcore.so")
    subprocess.run([cxx, "-O1", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "mirt_amd", "csrc"),
                    os.path.join(ROOT, "tests", "fft_core_host.cpp"), "-o", out], check=True)
    lib = C.CDLL(out)
    lib.fft_core_twiddles.restype = C.c_int
    lib.fft_core_twiddles.argtypes = [C.c_int, C.c_void_p]
    lib.fft_core_forward.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    return lib


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_fft_core_passes_match_numpy(fft_core_host, n):
    """The Stockham passes (radix 8, then 4 or 2), the pass-by-pass twiddle
    table and the half-input first pass (a row zero-padded to twice its
    length or more, the upper half never read) give the DFT: FP32 against
    numpy's FP64 FFT, relative L2 below 4e-7."""
    lib = fft_core_host
    rng = np.random.default_rng(n)
    tw = np.zeros(max(8 * n, 8), np.complex64)
    ne = lib.fft_core_twiddles(n, tw.ctypes.data)
    assert 0 <= ne <= tw.size
    for half in (0, 1):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        if half:
            x[n // 2:] = 0
        buf = x.copy()
        if half:
            buf[n // 2:] = np.nan          # must not be read
        lib.fft_core_forward(buf.ctypes.data, n, tw.ctypes.data, half)
        ref = np.fft.fft(x.astype(np.complex128))
        assert np.linalg.norm(buf - ref) <= 4e-7 * np.linalg.norm(ref)
    # the inverse as the filter kernel takes it: conj(fft(conj(X))) / n
    X = np.fft.fft(rng.standard_normal(n)).astype(np.complex64)
    buf = np.conj(X)
    lib.fft_core_forward(buf.ctypes.data, n, tw.ctypes.data, 0)
    got = np.conj(buf) / n
    ref = np.fft.ifft(X.astype(np.complex128))
    assert np.linalg.norm(got - ref) <= 4e-7 * np.linalg.norm(ref)


def test_fft_core_twiddle_layout(fft_core_host):
    """Pass p of span Ns and radix R owns Ns*R entries, entry r*R + m being
    exp(-2 pi i m r / (Ns R)); the passes follow each other (N = 2048: spans
    1, 8, 64, 512 with radices 8, 8, 8, 4)."""
    lib = fft_core_host
    tw = np.zeros(8 * 2048, np.complex64)
    ne = lib.fft_core_twiddles(2048, tw.ctypes.data)
    assert ne == 1 * 8 + 8 * 8 + 64 * 8 + 512 * 4
    off = 8                                  # after the span-1 pass (no twiddles used)
    for ns_, r in ((8, 8), (64, 8), (512, 4)):
        want = np.exp(-2j * np.pi * np.outer(np.arange(ns_), np.arange(r)) / (ns_ * r))
        assert np.allclose(tw[off:off + ns_ * r].reshape(ns_, r), want, atol=1e-7)
        off += ns_ * r


def test_fbp2_window_forms():
    """fbp2_window.m: boxcar/ramp/'' are ones; every window is fft-shifted
    (its centre ii = 0 first, value 1); 'hann' is 'hanning,1'; a cutoff c
    keeps the round(c n) - 1 samples with |ii| < round(c n)/2 (even widths)
    and zeroes the rest; the windows are symmetric about ii = 0; Hamming's
    edge tends to 0.08; arrays pass through shifted; bad names and lengths
    fail.  fdk_filter's response takes them (fdk_fan_filter :50-58)."""
    from mirt_amd.fdk import fbp2_window, fdk_filter_response
    from mirt_amd.geom import ct_geom
    n = 128
    for w in ("", "boxcar", "ramp"):
        assert np.array_equal(fbp2_window(n, w), np.ones(n))
    assert np.array_equal(fbp2_window(n, "hann"), fbp2_window(n, "hanning,1"))
    for name, c in [("hann50", 0.5), ("hann75", 0.75), ("hann80", 0.8), ("boxcar,0.6", 0.6),
                    ("hamming,0.6", 0.6), ("hanning,0.3", 0.3)]:
        win = fbp2_window(n, name)
        assert win[0] == 1.0                                   # ii = 0 first
        centred = np.fft.ifftshift(win)                        # ii = -n/2 .. n/2-1
        wdt = int(np.floor(c * n + 0.5))
        assert np.count_nonzero(centred) == wdt - 1 + (wdt % 2)
        k = np.arange(1, n // 2)
        assert np.array_equal(centred[n // 2 + k], centred[n // 2 - k])
        if name.startswith("hamming"):
            edge = centred[n // 2 + (wdt - 1) // 2]
            assert 0.08 < edge < 0.1
    ramp = np.linspace(0, 1, n)
    assert np.array_equal(fbp2_window(n, ramp), np.fft.fftshift(ramp))
    with pytest.raises(ValueError, match="unknown window"):
        fbp2_window(n, "triangle")
    with pytest.raises(ValueError, match="bad window length"):
        fbp2_window(n, np.ones(n - 1))
    cg = ct_geom("fan", ns=40, nt=2, na=1, ds=1.0, dt=1.0, dsd=949.0, dso=541.0, dfs=np.inf)
    H = fdk_filter_response(cg, 128, "ramp")
    Hb = fdk_filter_response(cg, 128, "boxcar,0.5")
    assert np.array_equal(Hb, H * np.fft.fftshift(np.fft.ifftshift(fbp2_window(128, "boxcar,0.5"))))
    assert Hb[0] == H[0] and np.count_nonzero(Hb) < np.count_nonzero(H)
    with pytest.raises(ValueError, match="bad window length"):
        fdk_filter_response(cg, 128, np.ones(64))


def test_feldkamp_weight1_center_value():
    from mirt_amd.fdk import feldkamp_weight1
    from mirt_amd.geom import ct_geom
    cg = ct_geom("fan", ns=5, nt=3, na=1, ds=1.0, dt=1.0, dsd=1000.0, dso=600.0, dfs=np.inf)
    w = feldkamp_weight1(torch.ones((5, 3, 1), dtype=torch.float64), cg).numpy()
    assert w[2, 1, 0] == pytest.approx(600.0 / 1000.0)      # dso/sqrt(dsd^2) at s=t=0
    assert w[0, 0, 0] == pytest.approx(600.0 / math.sqrt(1000.0 ** 2 + 4 + 1))


def _ell_oracle(cg, ell, oversample=1, zs=None):
    from oracle.ellipsoid import ellipsoid_proj
    return ellipsoid_proj(cg.s, cg.t, cg.ar, cg.source_zs if zs is None else zs,
                          cg.dso, cg.dod, cg.dfs, ell, oversample)


def test_ellipsoid_proj_parallel_closed_form():
    from mirt_amd.geom import ct_geom
    cg = ct_geom("fan", ns=41, nt=5, na=3, ds=1.0, dt=1.0, dsd=np.inf)
    ell = np.array([[0, 0, 0, 10, 10, 10, 0, 0, 0.5]])
    p = _ell_oracle(cg, ell)
    u = cg.s[:, None]
    v = cg.t[None, :]
    want = 2 * 0.5 * np.sqrt(np.clip(100 - u ** 2 - v ** 2, 0, None))
    for ia in range(3):
        assert np.allclose(p[:, :, ia], want, atol=1e-12)


def _ray(cg, b, s, t, zs=0.0):
    """Source position and unit direction of the ray to detector sample (s, t)
    at angle b: source on +y at b = 0, flat detector at y = -(dsd - dso),
    arc detector on the circle of radius dsd about the source."""
    dsd = cg.dsd
    src = np.array([-cg.dso * math.sin(b), cg.dso * math.cos(b), zs])
    if cg.dfs == 0:
        gam = s / dsd
        d_loc = np.array([dsd * math.sin(gam), -dsd * math.cos(gam), t])
    else:
        d_loc = np.array([s, -dsd, t])
    rot = np.array([[math.cos(b), -math.sin(b), 0], [math.sin(b), math.cos(b), 0], [0, 0, 1]])
    d = rot @ d_loc
    return src, d / np.linalg.norm(d)


def test_ellipsoid_proj_cone_chords():
    """Arc and flat cone beams: sphere chord from the ray's distance to the centre."""
    from mirt_amd.geom import ct_geom
    for dfs in (0.0, np.inf):
        cg = ct_geom("fan", ns=33, nt=7, na=2, ds=2.0, dt=2.0, dsd=949.0, dso=541.0,
                     dfs=dfs, offset_s=0.25)
        c, r, val = np.array([10.0, -5.0, 3.0]), 40.0, 0.7
        ell = np.array([[*c, r, r, r, 0, 0, val]])
        p = _ell_oracle(cg, ell)
        for ia, b in enumerate(cg.ar):
            for i in range(cg.ns):
                for j in range(cg.nt):
                    src, d = _ray(cg, b, cg.s[i], cg.t[j])
                    w = c - src
                    dist2 = w @ w - (w @ d) ** 2
                    chord = 2 * val * math.sqrt(max(r * r - dist2, 0))
                    assert p[i, j, ia] == pytest.approx(chord, abs=1e-9)


def test_ellipsoid_proj_rotated_ellipsoid_by_ray_marching():
    """A rotated, unequal-axis ellipsoid off centre, with a source height:
    every sample equals the length (times value) of the ray inside the
    ellipsoid, counted by marching along the ray -- pins the rotation sign,
    the axis order and the source-height shift of ellipsoid_proj.m:97-118."""
    from mirt_amd.geom import ct_geom
    rng = np.random.default_rng(5)
    cx, cy, cz, rx, ry, rz, ang, val = 12.0, -7.0, 4.0, 35.0, 15.0, 20.0, 30.0, 0.9
    ell = np.array([[cx, cy, cz, rx, ry, rz, ang, 0, val]])
    th = math.radians(ang)
    rot = np.array([[math.cos(th), math.sin(th), 0], [-math.sin(th), math.cos(th), 0], [0, 0, 1]])
    for dfs in (0.0, np.inf):
        cg = ct_geom("fan", ns=24, nt=6, na=3, ds=3.0, dt=3.0, dsd=949.0, dso=541.0, dfs=dfs)
        zs = np.array([0.0, 5.0, -8.0])
        p = _ell_oracle(cg, ell, zs=zs)
        step = 0.01
        lam = np.arange(400.0, 700.0, step)
        for _ in range(12):
            ia, i, j = rng.integers(cg.na), rng.integers(cg.ns), rng.integers(cg.nt)
            src, d = _ray(cg, cg.ar[ia], cg.s[i], cg.t[j], zs[ia])
            pts = src[None, :] + lam[:, None] * d[None, :] - np.array([cx, cy, cz])
            q = pts @ rot.T                      # ellipsoid frame
            inside = (q[:, 0] / rx) ** 2 + (q[:, 1] / ry) ** 2 + (q[:, 2] / rz) ** 2 <= 1
            assert p[i, j, ia] == pytest.approx(val * inside.sum() * step, abs=3 * step * val)


def test_ellipsoid_proj_oversample_is_the_subsample_mean():
    from mirt_amd.geom import ct_geom
    from oracle.ellipsoid import ellipsoid_proj
    cg = ct_geom("fan", ns=9, nt=5, na=2, ds=1.5, dt=1.2, dsd=949.0, dso=541.0, dfs=0.0)
    ell = np.array([[3, 2, 1, 20, 12, 9, 40, 0, 1.0], [-5, 0, -2, 8, 8, 8, 0, 0, 0.5]])
    p2 = _ell_oracle(cg, ell, oversample=2)
    off = np.array([-0.25, 0.25])
    fine = ellipsoid_proj((off[:, None] * 1.5 + cg.s[None, :]).T.reshape(-1),
                          (off[:, None] * 1.2 + cg.t[None, :]).T.reshape(-1),
                          cg.ar, 0.0, cg.dso, cg.dod, cg.dfs, ell)
    want = fine.reshape(9, 2, 5, 2, 2).mean(axis=(1, 3))
    assert np.allclose(p2, want, rtol=1e-14, atol=1e-14)


# --------------------------------------------------------------------------
# workload definitions (the benchmark configurations)

def test_workloads_match_the_benchmark_configurations():
    from mirt_amd.workloads import WORKLOADS, uneven_scans
    want = {   # ns, nt, na, ds, dt, arc, (nx, ny, nz), dx, dz, ellipsoids, inserts, I0
        "cpu_small": (128, 96, 64, 1.0, 1.0, False, (64, 64, 32), 1.0, 1.0, 10, 0, None),
        "plain_flat": (512, 384, 360, 0.8, 0.8, False, (256, 256, 128), 0.5, 0.5, 20, 0, 1e5),
        "clinical_arc": (888, 64, 984, 1.0239, 1.0964, True, (512, 512, 256), 0.7, 0.625, 30,
                         0, 2e5),
        "large_flat": (1024, 768, 1440, 0.4, 0.4, False, (1024, 1024, 512), 0.25, 0.25, 40,
                       12, 1e5),
    }
    for name, (ns, nt, na, ds, dt, arc, nxyz, dx, dz, nell, nins, i0) in want.items():
        sc = WORKLOADS[name]()
        cg, ig = sc.cg, sc.ig
        assert (cg.ns, cg.nt, cg.na) == (ns, nt, na), name
        assert (cg.ds, cg.dt, cg.dso, cg.dsd) == (ds, dt, 541.0, 949.0), name
        assert (cg.dfs == 0) == arc and (arc or np.isinf(cg.dfs)), name
        assert cg.orbit == 360 and (ig.nx, ig.ny, ig.nz) == nxyz, name
        assert (ig.dx, ig.dy, ig.dz) == (dx, -dx, dz), name
        assert len(sc.ellipsoids) == nell + nins and sc.i0 == i0, name
        mu = sc.ellipsoids[:nell, 8]
        assert np.all((mu >= 0.005) & (mu <= 0.02)), name
        assert np.all(sc.ellipsoids[nell:, 8] == 0.05), name
    assert WORKLOADS["clinical_arc"]().cg.offset_s == 1.25
    hx = WORKLOADS["clinical_helix"]()
    cg, ig = hx.cg, hx.ig
    assert (cg.ns, cg.nt, cg.na, cg.ds, cg.dt, cg.offset_s) == (888, 64, 2952, 1.0239,
                                                                 1.0964, 1.25)
    assert cg.dfs == 0 and cg.orbit == 1080 and cg.pitch == 1.0
    assert (ig.nx, ig.ny, ig.nz, ig.dx, ig.dz) == (512, 512, 192, 0.7, 0.625)
    # the source travels 3 turns x (pitch * zfov) = the volume's 120 mm, centred
    zs = cg.source_zs
    assert zs[-1] - zs[0] == pytest.approx(3 * cg.zfov * (1 - 1 / cg.na), rel=1e-12)
    assert abs(zs[0] + zs[-1]) < cg.source_dz_per_view + 1e-12
    assert ig.nz * ig.dz == pytest.approx(3 * cg.zfov, rel=1e-4)
    scans = uneven_scans(32)
    assert len(scans) == 32
    for sc in scans:
        e = sc.ig.nx
        assert e in (128, 192, 256, 384) and sc.ig.ny == e and sc.ig.nz == e // 2
        assert 180 <= sc.cg.na <= 720 and 256 <= sc.cg.ns <= 768 and 128 <= sc.cg.nt <= 512
        assert 10 <= len(sc.ellipsoids) <= 30 and np.isinf(sc.cg.dfs)
    # seeded: the same scans every time
    again = uneven_scans(32)
    assert all(np.array_equal(a.ellipsoids, b.ellipsoids) for a, b in zip(scans, again))
