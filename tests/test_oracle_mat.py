"""Pinning the cbct_back_mat restatement (oracle.back_mat_f64), CPU only.

cbct_back_mat (fbp/cbct_back.m:142-270) is MATLAB code that cannot run here,
so its restatement is pinned by what fixes it from outside:

* the reference C (oracle/_ref): where every view has zero edge columns and
  its edge rows are replicated far enough, the MATLAB edge rules (zero beyond
  [1, ns], rows clamped to [1, nt]) and the mex's (one-bin zero fall-off)
  give the same image -- the comparison cbct_back_test makes between
  use_mex=0 and the mex with extrapolate_t (cbct_back.m:335-364).  Flat, arc
  and parallel geometries, a mask, ia_skip = 3; a geometry wider than the
  detector in both s and t, so that both edge rules act;
* the reference C's one-bin fall-off at the s edges of a view of ones: there
  cbct_back_mat must add exactly 0, and the reference's full weight inside;
* offset_source (arc) and per-view source_zs (helix) through shifts of the
  image offsets in the reference C: at beta = 0, r_loop = x_beta -
  offset_source is the x of an image moved by offset_source; zc - source_zs
  is the z of an image moved by source_zs;
* closed forms on the parallel geometry: a view linear in s backprojects to
  the 1-based bin position bs = x/ds + ws inside [1, ns] and to 0 outside; a
  view linear in t gives (views whose s is in range) x clamp(z/dt + wt, 1,
  nt).  These fix ws, wt, the t clamp and the hard s cut independently of
  the reference.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

need_ref = pytest.mark.skipif(not O.have_reference(), reason="reference C not built")

# geometries wider than the detector in s and taller than it in t
MAT_CASES = {
    "flat": dict(nxyz=(48, 44, 40), dxyz=(6.0, -6.0, 2.0), cxyz=(0.5, -0.25, 0.75),
                 dso=541.0, dsd=949.0, dfs=np.inf, ns=96, nt=48, dst=(4.0, 1.5),
                 ost=(0.75, 0.4), na=24),
    "arc": dict(nxyz=(48, 44, 40), dxyz=(6.0, -6.0, 2.0), cxyz=(0.5, -0.25, 0.75),
                dso=541.0, dsd=949.075, dfs=0.0, ns=96, nt=48, dst=(4.0, 1.5),
                ost=(0.75, 0.4), na=24),
    "par": dict(nxyz=(40, 38, 30), dxyz=(1.0, -1.0, 1.0), cxyz=(0.25, 0.0, -0.5),
                dso=np.inf, dsd=np.inf, dfs=0.0, ns=44, nt=20, dst=(1.0, 1.0),
                ost=(0.5, -0.3), na=18),
}


def mat_case(name, seed=0, zero_edges=True):
    c = MAT_CASES[name]
    g = O.geom(c["nxyz"], c["dxyz"], c["cxyz"], c["dso"], c["dsd"], c["dfs"],
               c["ns"], c["nt"], c["dst"], c["ost"])
    rng = np.random.default_rng(seed)
    # FP32-representable samples: the reference reads single views
    proj = rng.standard_normal((c["ns"], c["nt"], c["na"])).astype(np.float32).astype(np.float64)
    if zero_edges:
        proj[0] = 0
        proj[-1] = 0
    beta = np.linspace(0, 2 * np.pi, c["na"], endpoint=False) + 0.2
    mask = rng.random((g.nx, g.ny)) > 0.1
    return g, proj, beta, mask


def rows_needed(g) -> int:
    """Edge rows to replicate so that every voxel's t lies on the extended
    detector: |t| <= mag_max * max|z| / |dt| + |offset_t| + (nt+1)/2."""
    rmax = math.hypot(g.nx * abs(g.dx), g.ny * abs(g.dy))
    zmax = (g.nz / 2 + abs(g.offset_z) + 1) * abs(g.dz)
    mag = 1.0 if math.isinf(g.dsd) or math.isinf(g.dso) else g.dsd / (g.dso - rmax)
    return int(math.ceil(mag * zmax / abs(g.dt) + abs(g.offset_t))) + 2


# The reference C computes in FP32 (float geometry, a float t_bin walk over
# the slices): on these 6 mm voxels its images differ from exact arithmetic
# by ~1e-4 relative L2.  The FP64 restatement of the same C (back_ts_f64,
# pinned to the reference and to closed forms in test_oracle.py) carries the
# comparison down to double rounding.
TOL_REF_F32 = 3e-4
TOL_F64 = 1e-11


def ref_mex(g, proj_sta, beta, mask, ext, *, offset_x=None, offset_z=None, f64=False):
    """Reference C 'fdk,ts,back' (f64: its FP64 restatement) on views with
    `ext` edge rows replicated at both ends (cbct_back.m:86-94); returns the
    float64 [nx ny nz] image."""
    ge = O.geom((g.nx, g.ny, g.nz), (g.dx, g.dy, g.dz),
                (g.offset_x if offset_x is None else offset_x, g.offset_y,
                 g.offset_z if offset_z is None else offset_z),
                g.dso, g.dsd, g.dfs, g.ns, g.nt + 2 * ext, (g.ds, g.dt),
                (g.offset_s, g.offset_t))
    p = np.asarray(proj_sta, np.float32)
    if p.ndim == 2:
        p = p[:, :, None]
    p = np.concatenate([np.repeat(p[:, :1], ext, axis=1), p,
                        np.repeat(p[:, -1:], ext, axis=1)], axis=1)
    tsa = np.ascontiguousarray(p.transpose(2, 0, 1))          # [nt ns na]
    m = None if mask is None else np.ascontiguousarray(np.asarray(mask, np.uint8).T)
    if f64:
        out = O.back_ts_f64(ge, tsa, np.atleast_1d(beta), m)
    else:
        out = O.ref_back_ts("fixed", ge, tsa, np.atleast_1d(beta), m, nthread=4)
    return out.transpose(1, 0, 2).astype(np.float64)          # [nz nx ny] -> [nx ny nz]


def check_both(got, make):
    """got against the reference C (FP32 tolerance) and its FP64 restatement."""
    assert O.rel_l2(got, make(False)) < TOL_REF_F32
    assert O.rel_l2(got, make(True)) < TOL_F64


# --------------------------------------------------------------------------
# against the reference C

@need_ref
@pytest.mark.parametrize("name", sorted(MAT_CASES))
def test_mat_matches_reference_c_with_zero_edges(name):
    g, proj, beta, mask = mat_case(name, seed=5)
    ia_skip = 3
    ext = rows_needed(g)
    got = O.back_mat_f64(g, proj, beta, ia_skip=ia_skip, mask=mask)
    check_both(got, lambda f64: ref_mex(g, proj[:, :, ::ia_skip], beta[::ia_skip], mask,
                                        ext, f64=f64))
    assert np.all(got[~mask] == 0)
    # both edge rules act in this geometry: some voxel projects outside the
    # detector's columns in some view, and some outside its rows
    inside = O.back_mat_f64(g, np.ones((g.ns, g.nt)), beta[:1])
    assert np.any(inside[mask] == 0)
    tall = ref_mex(g, np.ones((g.ns, g.nt)), beta[:1], None, ext)
    short = ref_mex(g, np.ones((g.ns, g.nt)), beta[:1], None, 0)
    assert np.any((tall > 0) & (short == 0))


@need_ref
@pytest.mark.parametrize("name", ["flat", "arc", "par"])
def test_mat_zeroes_the_reference_falloff_band(name):
    """A view of ones: the reference C ramps from full weight to 0 over the
    bin beyond each s edge (its zero padding); cbct_back_mat adds 0 there
    (bs outside [1, ns] reads the zero column) and full weight inside."""
    g, _, beta, _ = mat_case(name)
    ext = rows_needed(g)
    b = beta[3:4] + 0.05
    ones = np.ones((g.ns, g.nt))
    narrow = ref_mex(g, ones, b, None, ext)
    gw = O.geom((g.nx, g.ny, g.nz), (g.dx, g.dy, g.dz), (g.offset_x, g.offset_y, g.offset_z),
                g.dso, g.dsd, g.dfs, g.ns + 4, g.nt, (g.ds, g.dt), (g.offset_s, g.offset_t))
    full = ref_mex(gw, np.ones((g.ns + 4, g.nt)), b, None, ext)   # same bins, 2 more each side
    got = O.back_mat_f64(g, ones, b)
    frac = np.divide(narrow, full, out=np.zeros_like(narrow), where=full > 0)
    band = (frac > 1e-3) & (frac < 0.999)
    inner = frac > 1 - 1e-5
    assert band.sum() > 0 and inner.sum() > 0
    assert np.all(got[band] == 0)
    assert np.all(got[full == 0] == 0)
    np.testing.assert_allclose(got[inner], full[inner], rtol=2e-6)


@need_ref
def test_mat_offset_source_is_an_image_shift():
    """Arc detector, beta = 0: x_beta = xc, so r_loop = xc - offset_source is
    the x of the image moved by offset_source (offset_x + offset_source/dx),
    which the reference C (no offset_source) computes."""
    g, proj, _, mask = mat_case("arc", seed=9)
    ext = rows_needed(g)
    beta = np.array([0.0])
    for off in (7.5, -12.25):
        got = O.back_mat_f64(g, proj[:, :, :1], beta, offset_source=off, mask=mask)
        check_both(got, lambda f64: ref_mex(g, proj[:, :, :1], beta, mask, ext,
                                            offset_x=g.offset_x + off / g.dx, f64=f64))
        same = O.back_mat_f64(g, proj[:, :, :1], beta, offset_source=0.0, mask=mask)
        assert O.rel_l2(got, same) > 1e-2      # the offset matters


@need_ref
@pytest.mark.parametrize("name", ["flat", "arc"])
def test_mat_source_zs_is_a_per_view_image_shift(name):
    """tprime = mag * (zc - source_zs(ia)): view ia sees the image moved by
    source_zs(ia) in z; with ia_skip = 2 the heights of views 1, 3, 5, ...
    (1-based) are the ones used."""
    g, proj, beta, mask = mat_case(name, seed=3)
    na = 8
    proj, beta = proj[:, :, :na], beta[:na]
    zs = np.random.default_rng(4).uniform(-9, 9, na)
    ext = rows_needed(g) + int(10 / abs(g.dt)) + 2
    got = O.back_mat_f64(g, proj, beta, source_zs=zs, ia_skip=2, mask=mask)
    check_both(got, lambda f64: sum(
        ref_mex(g, proj[:, :, ia], beta[ia:ia + 1], mask, ext,
                offset_z=g.offset_z + zs[ia] / g.dz, f64=f64) for ia in range(0, na, 2)))
    wrong = sum(ref_mex(g, proj[:, :, ia], beta[ia:ia + 1], mask, ext,
                        offset_z=g.offset_z + zs[ia + 1] / g.dz) for ia in range(0, na, 2))
    assert O.rel_l2(got, wrong) > 1e-2


# --------------------------------------------------------------------------
# closed forms (parallel beam)

def _par_bins(g, beta):
    """1-based s bin of every column for each view: s = x cos b + y sin b."""
    xc = (np.arange(g.nx) - ((g.nx - 1) / 2 + g.offset_x)) * g.dx
    yc = (np.arange(g.ny) - ((g.ny - 1) / 2 + g.offset_y)) * g.dy
    s = xc[:, None, None] * np.cos(beta) + yc[None, :, None] * np.sin(beta)
    return s / g.ds + (g.ns + 1) / 2 + g.offset_s                    # [nx ny nv]


def test_mat_s_ramp_closed_form():
    g, _, _, _ = mat_case("par")
    srow = np.arange(1, g.ns + 1, dtype=np.float64)
    proj = np.repeat(srow[:, None], g.nt, axis=1)                   # p(s, t) = s
    b = np.array([0.6])
    got = O.back_mat_f64(g, proj, b)
    bs = _par_bins(g, b)[:, :, 0]
    want = np.where((bs >= 1) & (bs <= g.ns), bs, 0.0)
    assert np.any(want == 0) and np.any(want > 0)
    np.testing.assert_allclose(got, np.repeat(want[:, :, None], g.nz, axis=2),
                               rtol=0, atol=1e-12)


def test_mat_t_ramp_closed_form():
    g, _, beta, _ = mat_case("par")
    trow = np.arange(1, g.nt + 1, dtype=np.float64)
    proj = np.repeat(trow[None, :], g.ns, axis=0)                   # p(s, t) = t
    got = O.back_mat_f64(g, np.repeat(proj[:, :, None], beta.size, axis=2), beta)
    bs = _par_bins(g, beta)
    count = ((bs >= 1) & (bs <= g.ns)).sum(axis=2)                  # views in range
    zc = (np.arange(g.nz) - ((g.nz - 1) / 2 + g.offset_z)) * g.dz
    bt = np.clip(zc / g.dt + (g.nt + 1) / 2 + g.offset_t, 1, g.nt)
    assert bt.min() == 1 and bt.max() == g.nt                      # both clamps act
    want = count[:, :, None] * bt[None, None, :]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-12)


def test_mat_scale_and_linearity():
    g, proj, beta, mask = mat_case("flat", seed=2, zero_edges=False)
    a = O.back_mat_f64(g, proj, beta, mask=mask)
    b = O.back_mat_f64(g, 2 * proj, beta, mask=mask, scale=0.25)
    np.testing.assert_array_equal(b, 0.5 * a)
    c = O.back_mat_f64(g, proj[:, :, :5], beta[:5], mask=mask) + \
        O.back_mat_f64(g, proj[:, :, 5:], beta[5:], mask=mask)
    assert O.rel_l2(c, a) < 1e-13


def test_angle_scale64_closed_form():
    from mirt_amd.geom import ct_geom
    from mirt_amd.cbct_back import angle_scale64
    cg = ct_geom("fan", ns=16, nt=8, na=984, ds=1.0, dt=1.0, dsd=949.0, dod=408.0,
                 dfs=0.0)
    assert abs(angle_scale64(cg) - math.pi / 984) < 1e-17
    assert abs(angle_scale64(cg, 3) - 3 * math.pi / 984) < 1e-16


# --------------------------------------------------------------------------
# the helical cbct_back_mat (mex/src/fdk-helix/cbct_back_helix.m:185-350)

def _helix_case(name, na=24, seed=5, span=30.0):
    g, proj, beta, mask = mat_case(name, seed=seed, zero_edges=False)
    proj, beta = proj[:, :, :na], np.linspace(0, 2 * np.pi, na, endpoint=False) + 0.1
    zs = np.linspace(-span / 2, span / 2, na)        # one turn of a helix
    return g, proj, beta, mask, zs


@pytest.mark.parametrize("name", ["flat", "arc"])
@pytest.mark.parametrize("ia_skip", [1, 3])
def test_helix_mat_pitch0_is_the_axial_mat(name, ia_skip):
    """pitch = 0: every slice takes views 1:ia_skip:na and the per-slice
    scale is the axial scale_dang factor, so the helical backprojector is
    cbct_back_mat with its 0.5 * deg2rad(|orbit|) / (na / ia_skip)."""
    g, proj, beta, mask, zs = _helix_case(name)
    na = beta.size
    first, last, sc = O.helix_mat_slices(g, na, 360.0, 0.0, zs, ia_skip)
    assert np.all(first == 0) and np.all(last == na - 1)
    axial = 0.5 * ((math.pi / 180) * 360.0) / (na / ia_skip)
    np.testing.assert_allclose(sc, axial, rtol=4e-16)
    got = O.back_helix_mat_f64(g, proj, beta, zs, first, last, ia_skip=ia_skip, mask=mask,
                               slice_scale=sc)
    want = O.back_mat_f64(g, proj, beta, source_zs=zs, ia_skip=ia_skip, mask=mask,
                          scale=axial)
    assert O.rel_l2(got, want) < 1e-15


def test_helix_mat_slices_closed_form():
    """na = 24 over 360 degrees: na1_half = 12; a slice whose nearest source
    is view 5 (1-based) takes views 1..16, a share 16/24 of the orbit, and
    its scale is pi / deg2rad(240) * deg2rad(360) / 24 = pi / 16."""
    g = O.geom((4, 4, 3), (1.0, 1.0, 10.0), (0.0, 0.0, 0.0), 541.0, 949.0, np.inf,
               8, 6, (1.0, 1.0), (0.0, 0.0))
    zs = np.arange(24) * 10.0 - 40.0                  # slice zc = -10, 0, 10 -> views 4, 5, 6
    first, last, sc = O.helix_mat_slices(g, 24, 360.0, 1.0, zs)
    np.testing.assert_array_equal(first, [0, 0, 0])
    np.testing.assert_array_equal(last, [14, 15, 16])
    np.testing.assert_allclose(sc[1], math.pi / 16, rtol=1e-15)
    # the ranges are the views within [-12, +11] of the nearest one (brute force)
    zc = (np.arange(3) - 1.0) * 10.0
    for iz in range(3):
        mid = int(np.argmin(np.abs(zc[iz] - zs)))
        want = [ia for ia in range(24) if -12 <= ia - mid <= 11]
        assert (first[iz], last[iz]) == (want[0], want[-1])
    with pytest.raises(ValueError, match="bad na/orbit"):
        O.helix_mat_slices(g, 25, 361.0, 1.0, np.zeros(25))


@pytest.mark.parametrize("name", ["flat", "arc"])
@pytest.mark.parametrize("ia_skip", [1, 2])
def test_helix_mat_slice_is_the_axial_mat_over_its_views(name, ia_skip):
    """pitch != 0: slice iz is cbct_back_mat over its own views
    first:ia_skip:last only (same sums in the same order), times its scale."""
    g, proj, beta, mask, zs = _helix_case(name, span=float(g_span(name)))
    na = beta.size
    first, last, sc = O.helix_mat_slices(g, na, 360.0, 1.0, zs, ia_skip)
    assert len(set(zip(first, last))) > 1                     # the ranges differ
    got = O.back_helix_mat_f64(g, proj, beta, zs, first, last, ia_skip=ia_skip, mask=mask,
                               slice_scale=sc)
    for iz in range(g.nz):
        sel = np.arange(first[iz], last[iz] + 1, ia_skip)
        one = O.back_mat_f64(g, proj[:, :, sel], beta[sel], source_zs=zs[sel], mask=mask)
        np.testing.assert_array_equal(got[:, :, iz], sc[iz] * one[:, :, iz])


def g_span(name):
    """z extent of a case's volume (so that the helix spans its slices)."""
    c = MAT_CASES[name]
    return c["nxyz"][2] * c["dxyz"][2]


def test_helix_mat_cone_parallel_ray_geometry():
    """Cone-parallel views (rebinned fan rays, one source per column): a view
    linear in the 1-based bins, p = a + b bs + c bt, backprojects (bilinear
    interpolation being exact on it, no weight) to a + b bs + c bt at the
    sample each voxel's ray hits.  That sample is computed here from the ray
    itself: the ray through voxel V is parallel to the view's central ray
    u = (sin beta, -cos beta), at lateral offset s = V . (cos beta, sin beta);
    its source S lies on the source circle (radius dso) at that offset; S's
    angle on the circle (source at angle phi = (-dso sin phi, dso cos phi))
    gives its height zs + (phi_S - beta) * dz_per_rad; the detector lies dsd
    from S along u, so t = dsd / ((V - S) . u) * (z - z_S)."""
    g = O.geom((12, 10, 6), (4.0, 4.0, 3.0), (0.3, -0.2, 0.1), 541.0, 949.0, np.inf,
               120, 60, (2.0, 1.5), (0.4, -0.3))
    beta = np.array([0.3, 1.9, 4.1])
    zs = np.array([-2.0, 5.0, 11.0])
    dz_per_rad = 7.5
    a, b, c = 0.5, 0.01, -0.02
    sb = np.arange(1, g.ns + 1, dtype=np.float64)
    tb = np.arange(1, g.nt + 1, dtype=np.float64)
    view = a + b * sb[:, None] + c * tb[None, :]
    proj = np.repeat(view[:, :, None], beta.size, axis=2)
    first = np.zeros(g.nz, np.int64)
    last = np.full(g.nz, beta.size - 1)
    ws = (g.ns + 1) / 2 + g.offset_s
    wt = (g.nt + 1) / 2 + g.offset_t
    xc = (np.arange(g.nx) - ((g.nx - 1) / 2 + g.offset_x)) * g.dx
    yc = (np.arange(g.ny) - ((g.ny - 1) / 2 + g.offset_y)) * g.dy
    zc = (np.arange(g.nz) - ((g.nz - 1) / 2 + g.offset_z)) * g.dz
    want = np.zeros((g.nx, g.ny, g.nz))
    for ia, bb in enumerate(beta):
        u = np.array([math.sin(bb), -math.cos(bb)])
        e = np.array([math.cos(bb), math.sin(bb)])
        for ix in range(g.nx):
            for iy in range(g.ny):
                V = np.array([xc[ix], yc[iy]])
                s = V @ e
                S = s * e - math.sqrt(g.dso ** 2 - s ** 2) * u      # on the circle, behind V
                phi = math.atan2(-S[0], S[1])
                dphi = (phi - bb + math.pi) % (2 * math.pi) - math.pi
                z_src = zs[ia] + dphi * dz_per_rad
                D = (V - S) @ u
                for iz in range(g.nz):
                    t = g.dsd / D * (zc[iz] - z_src)
                    bs, bt = s / g.ds + ws, t / g.dt + wt
                    assert 1 < bs < g.ns and 1 < bt < g.nt          # inside: no clamp
                    want[ix, iy, iz] += a + b * bs + c * bt
    got = O.back_helix_mat_f64(g, proj, beta, zs, first, last, cone_par=1,
                               source_dz_per_rad=dz_per_rad)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
