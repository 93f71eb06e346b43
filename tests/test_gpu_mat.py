"""GPU: cbct_back_mat (fbp/cbct_back.m:142-270, the use_mex = 0 route) on
the device against the oracle's restatement (oracle.back_mat_f64).

Bar:
* flat and parallel detectors: bit for bit.  Both sides do the same IEEE
  double operations in the same order (no FMA contraction on the device, a
  correctly rounded double quotient on both sides) and take cos/sin of each
  angle from the same host libm;
* arc detector: relative L2 <= 1e-13 and every voxel within 1e-12 of the
  image's largest magnitude -- the device's double atan2 (ROCm device libm)
  and the host's may differ in the last place.
At the benchmark size (plain_flat, 256 x 256 x 128, 360 views) the route is
held to what the reference's own cbct_back_test checks (cbct_back.m:335-364):
use_mex = 0 against the mex route with extrapolate_t rows, on views with
zero edge columns where the two edge rules coincide, within FP32 rounding
of the mex route.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_mat import MAT_CASES, mat_case

pytestmark = pytest.mark.gpu


def _geom(g):
    from mirt_amd._lib import Geom
    return Geom(*[getattr(g, f[0]) for f in Geom._fields_])


def _host_call(g, proj, beta, *, zs=None, ia_skip=1, offset_source=0.0, mask=None,
               scale=1.0):
    """fdkhip_cbct_back_mat on host arrays in MATLAB order."""
    from mirt_amd import _lib
    lib = _lib.load()
    fp64 = proj.dtype == np.float64
    p = np.asfortranarray(proj)
    b = np.ascontiguousarray(beta, np.float64)
    z = None if zs is None else np.ascontiguousarray(zs, np.float64)
    m = None if mask is None else np.asfortranarray(mask).astype(np.uint8, order="F")
    img = np.empty((g.nx, g.ny, g.nz), np.float64, order="F")
    P = lambda a: None if a is None else C.c_void_p(a.ctypes.data)
    _lib.check(lib.fdkhip_cbct_back_mat(C.byref(_geom(g)), P(p), int(fp64), p.shape[2],
                                        P(b), P(z), ia_skip, offset_source, P(m), scale,
                                        P(img)))
    return img


def _compare(kind, got, want):
    if kind == "arc":
        assert O.rel_l2(got, want) <= 1e-13
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    else:
        assert np.array_equal(got, want), float(np.abs(got - want).max())


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("name", sorted(MAT_CASES))
def test_mat_kernel_matches_oracle(device, name, dtype):
    g, proj, beta, mask = mat_case(name, seed=21, zero_edges=False)
    if dtype == "f32":
        proj = proj.astype(np.float32)
    else:
        proj = proj + 1e-3 * np.random.default_rng(1).standard_normal(proj.shape)
    na = proj.shape[2]
    zs = None if name == "par" else np.random.default_rng(2).uniform(-6, 6, na)
    osrc = 5.5 if name == "arc" else 0.0
    print(f"mat {name} {dtype}", flush=True)
    for ia_skip, m, scale in [(1, None, 1.0), (3, mask, 0.37)]:
        got = _host_call(g, proj, beta, zs=zs, ia_skip=ia_skip, offset_source=osrc, mask=m,
                         scale=scale)
        want = O.back_mat_f64(g, proj, beta, source_zs=zs, ia_skip=ia_skip,
                              offset_source=osrc, mask=m, scale=scale)
        _compare(name, got, want)
        if m is not None:
            assert np.all(got[~m] == 0)


def test_mat_edge_cases(device):
    """A sample exactly on the last column and on the last row, views of ones
    and of a ramp, a view count not divisible by ia_skip, na = 0, nz not a
    multiple of the slices per lane, and non-finite samples in column ns
    (weight 0 there: MATLAB's 0 * inf is NaN, and so is the device's)."""
    # ws = (12+1)/2 + 1.5 = 8: at beta = 0 column ix = 8 lands exactly on bs = 12 = ns;
    # slices beyond the 5 rows clamp to bt = 1 and bt = nt = 5
    ge = O.geom((9, 7, 37), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), np.inf, np.inf, 0.0,
                12, 5, (1.0, 1.0), (1.5, 0.0))
    ones = np.ones((12, 5, 5))
    beta = np.array([0.0, 0.4, 1.3, 2.0, 3.0])
    for ia_skip in (1, 2, 4, 7):
        got = _host_call(ge, ones, beta, ia_skip=ia_skip)
        _compare("par", got, O.back_mat_f64(ge, ones, beta, ia_skip=ia_skip))
    ramp = np.broadcast_to(np.arange(1.0, 13.0)[:, None, None], (12, 5, 5)).copy()
    _compare("par", _host_call(ge, ramp, beta), O.back_mat_f64(ge, ramp, beta))
    bad = ramp.copy()
    bad[-1, 2, 0] = np.inf
    got = _host_call(ge, bad, beta)
    want = O.back_mat_f64(ge, bad, beta)
    assert np.isnan(want).any() and np.isinf(want).any()
    assert np.array_equal(got, want, equal_nan=True)
    empty = _host_call(ge, np.zeros((12, 5, 0)), np.zeros(0))
    assert np.array_equal(empty, np.zeros((9, 7, 37)))


def test_mat_direct_equals_host_entry(device):
    """Device-resident entry on all views with ia_skip striding on the device
    == the host entry; kernel time from HIP events is reported."""
    import torch
    from mirt_amd import _lib
    lib = _lib.load()
    g, proj, beta, mask = mat_case("arc", seed=4, zero_edges=False)
    zs = np.linspace(-3, 3, proj.shape[2])
    want = _host_call(g, proj, beta, zs=zs, ia_skip=2, offset_source=-2.0, mask=mask,
                      scale=0.5)
    pt = torch.from_numpy(np.ascontiguousarray(proj.transpose(2, 1, 0))).to(device)
    img = torch.empty((g.nz, g.ny, g.nx), dtype=torch.float64, device=device)
    b = np.ascontiguousarray(beta)
    m = np.asfortranarray(mask).astype(np.uint8, order="F")
    ms = C.c_float(-1.0)
    P = lambda a: C.c_void_p(a.ctypes.data)
    s = torch.cuda.current_stream()
    _lib.check(lib.fdkhip_cbct_back_mat_direct(
        C.byref(_geom(g)), C.c_void_p(pt.data_ptr()), 1, proj.shape[2], P(b), P(zs), 2,
        -2.0, P(m), 0.5, C.c_void_p(img.data_ptr()), C.c_void_p(s.cuda_stream),
        C.byref(ms)))
    got = img.cpu().numpy().transpose(2, 1, 0)
    assert np.array_equal(got, want)
    assert ms.value > 0


def test_cbct_back_use_mex_0_route(device):
    """cbct_back(..., use_mex=0) with ct_geom/image_geom: the angles of
    cbct_back.m:157, a helix (source_zs from the pitch), the mask,
    offset_source and the double angle scale of cbct_back.m:264-266."""
    from mirt_amd import cbct_back
    from mirt_amd.cbct_back import _views, angle_scale64
    from mirt_amd.geom import ct_geom, image_geom
    for dfs, pitch, osrc in [(0.0, 0.0, 3.0), (np.inf, 0.5, 0.0), (0.0, 0.7, -1.5)]:
        cg = ct_geom("fan", ns=70, nt=32, na=40, ds=1.3, dt=1.1, offset_s=1.25,
                     offset_t=0.3, dsd=949.0, dso=541.0, dfs=dfs, orbit=360, pitch=pitch)
        ig = image_geom(nx=36, ny=34, nz=21, dx=1.4, dz=1.2, offset_z=0.5,
                        mask="all-but-edge")
        proj = np.asfortranarray(np.random.default_rng(7).standard_normal(
            (cg.ns, cg.nt, cg.na)).astype(np.float32))
        with pytest.warns(UserWarning):
            got = cbct_back(proj, cg, ig, use_mex=0, ia_skip=3, offset_source=osrc)
        g = O.geom((ig.nx, ig.ny, ig.nz), (ig.dx, ig.dy, ig.dz),
                   (ig.offset_x, ig.offset_y, ig.offset_z), cg.dso, cg.dsd, cg.dfs,
                   cg.ns, cg.nt, (cg.ds, cg.dt), (cg.offset_s, cg.offset_t))
        want = O.back_mat_f64(g, proj, _views(cg, cg.na), source_zs=cg.source_zs,
                              ia_skip=3, offset_source=osrc, mask=ig.mask_or,
                              scale=angle_scale64(cg, 3))
        assert got.dtype == np.float64 and got.shape == (ig.nx, ig.ny, ig.nz)
        _compare("arc" if dfs == 0 else "flat", got, want)
        if pitch:
            assert np.any(cg.source_zs != 0)


def test_mat_route_vs_mex_route_full_size(device):
    """plain_flat (256 x 256 x 128, 360 views of 512 x 384), ia_skip = 3, views
    with zero edge columns: use_mex = 0 against the FP64 restatement of the
    reference C on the views with replicated edge rows (3000 sampled columns,
    relative L2 <= 1e-11), and against the mex route with extrapolate_t (the
    pair the reference's cbct_back_test compares) within the FP32 rounding of
    the mex route (the FP32 reference departs from FP64 by 1.1e-3 on these
    noise views; a shifted bin or a wrong weight departs by > 1e-1)."""
    import math
    from mirt_amd import cbct_back
    from mirt_amd.cbct_back import _views
    from mirt_amd.workloads import plain_flat
    sc = plain_flat()
    cg, ig = sc.cg, sc.ig
    rng = np.random.default_rng(11)
    proj = np.asfortranarray(rng.standard_normal((cg.ns, cg.nt, cg.na)).astype(np.float32))
    proj[0] = 0
    proj[-1] = 0
    ext = int(math.ceil(1.3 * cg.nt / 2))
    print("mat vs mex plain_flat", flush=True)
    with pytest.warns(UserWarning):
        mat = cbct_back(proj, cg, ig, use_mex=0, ia_skip=3, scale_dang=False)
    g = O.geom((ig.nx, ig.ny, ig.nz), (ig.dx, ig.dy, ig.dz),
               (ig.offset_x, ig.offset_y, ig.offset_z), cg.dso, cg.dsd, cg.dfs,
               cg.ns, cg.nt + 2 * ext, (cg.ds, cg.dt), (cg.offset_s, cg.offset_t))
    p = proj[:, :, ::3]
    p = np.concatenate([np.repeat(p[:, :1], ext, axis=1), p,
                        np.repeat(p[:, -1:], ext, axis=1)], axis=1)
    tsa = np.ascontiguousarray(p.transpose(2, 0, 1))
    cols = np.sort(rng.choice(ig.nx * ig.ny, 3000, replace=False)).astype(np.int32)
    f64 = O.back_ts_f64(g, tsa, _views(cg, cg.na)[::3], cols=cols)
    got = mat.reshape(ig.nx * ig.ny, ig.nz, order="F")[cols]
    err64 = O.rel_l2(got, f64)
    mex = cbct_back(proj, cg, ig, use_mex=1, ia_skip=3, extrapolate_t=ext, scale_dang=False)
    err32 = O.rel_l2(mex, mat)
    print(f"rel_l2(mat, f64 restatement) = {err64:.3e}, rel_l2(mex, mat) = {err32:.3e}",
          flush=True)
    assert err64 <= 1e-11
    assert err32 <= 3e-3


@pytest.mark.parametrize("case", ["arc", "flat", "flat_cone_par", "flat_pitch0"])
def test_cbct_back_helix_use_mex_0(device, case):
    """cbct_back_helix(..., use_mex=0): the helical MATLAB backprojector
    (mex/src/fdk-helix/cbct_back_helix.m:185-350) on the device against the
    oracle's restatement (oracle.back_helix_mat_f64 with the slice ranges and
    scales of oracle.helix_mat_slices): per-slice view ranges on a helix,
    ia_skip 1 and 2, scale_dang on and off, cone-parallel views, a mask,
    float32 and float64 views.  Same bar as the axial route (bit for bit
    for flat; arc and cone_par within 1e-13, the device's atan2 / asin may
    differ from the host's in the last place)."""
    from mirt_amd import cbct_back_helix
    from mirt_amd.cbct_back import _views
    from mirt_amd.geom import ct_geom, image_geom
    dfs = 0.0 if case == "arc" else np.inf
    pitch = 0.0 if case == "flat_pitch0" else 0.8
    cone_par = int(case == "flat_cone_par")
    cg = ct_geom("fan", ns=90, nt=40, na=48, ds=1.3, dt=1.1, offset_s=0.75, offset_t=0.2,
                 dsd=949.0, dso=541.0, dfs=dfs, orbit=360, pitch=pitch)
    ig = image_geom(nx=34, ny=32, nz=23, dx=1.4, dz=1.6, offset_z=0.25, mask="all-but-edge")
    g = O.geom((ig.nx, ig.ny, ig.nz), (ig.dx, ig.dy, ig.dz),
               (ig.offset_x, ig.offset_y, ig.offset_z), cg.dso, cg.dsd, cg.dfs,
               cg.ns, cg.nt, (cg.ds, cg.dt), (cg.offset_s, cg.offset_t))
    rng = np.random.default_rng(11)
    for ia_skip, dtype, scale_dang in [(1, np.float32, True), (2, np.float64, False)]:
        print(f"helix mat {case} ia_skip={ia_skip} {dtype.__name__} scale={scale_dang}",
              flush=True)
        proj = np.asfortranarray(rng.standard_normal((cg.ns, cg.nt, cg.na)).astype(dtype))
        with pytest.warns(UserWarning, match="always extrapolates"):
            got = cbct_back_helix(proj, cg, ig, use_mex=0, ia_skip=ia_skip,
                                  offset_source=1.5 if case == "arc" else 0.0,
                                  scale_dang=scale_dang, cone_par=cone_par)
        first, last, sc = O.helix_mat_slices(g, cg.na, cg.orbit, cg.pitch, cg.source_zs,
                                             ia_skip)
        if pitch:
            assert len(set(zip(first, last))) > 1
        dz_per_rad = cg.source_dz_per_view / ((np.pi / 180 * abs(cg.orbit)) / (cg.na / ia_skip))
        want = O.back_helix_mat_f64(g, proj, _views(cg, cg.na), cg.source_zs, first, last,
                                    ia_skip=ia_skip, offset_source=1.5 if case == "arc" else 0.0,
                                    mask=ig.mask_or, slice_scale=sc if scale_dang else None,
                                    cone_par=cone_par, source_dz_per_rad=dz_per_rad)
        assert got.dtype == np.float64 and got.shape == (ig.nx, ig.ny, ig.nz)
        assert np.abs(want).max() > 0
        _compare("arc" if (case == "arc" or cone_par) else "flat", got, want)
