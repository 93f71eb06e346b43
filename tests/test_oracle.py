"""Pinning the CPU oracle (runs without a GPU).

* FP32 restatement == reference C compiled from its own sources (oracle/_ref),
  bit for bit, on the golden fixtures (tests/golden, written by the reference
  C) and on fresh seeded cases; likewise the 'fdk,st,back' order and the
  as-shipped (swapped-weight) variant.
* FP64 restatement against an independent per-voxel evaluation: the
  backprojection written as a sum of tent (triangle) kernels over detector
  samples, evaluated directly at every voxel (no slice loop, no t_bin
  accumulation, no padding), plus closed forms: a constant view on the
  parallel geometry gives na*c, on the flat geometry sum(mag^2)*c; linearity;
  masked columns stay zero.  The as-shipped weights fail the tent check, which
  shows the check can see a swapped weight.
"""
from __future__ import annotations

import glob
import math
import os

import numpy as np
import pytest

from conftest import CASES, make_case
from oracle import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load_golden(path):
    d = np.load(path, allow_pickle=False)
    g = O.geom(d["nxyz"], d["dxyz"], d["cxyz"], float(d["dso"]), float(d["dsd"]),
               float(d["dfs"]), int(d["ns"]), int(d["nt"]), d["dst"], d["ost"])
    return g, d


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_f32_matches_golden(path):
    g, d = load_golden(path)
    got = O.back_ts_f32(g, d["proj"], d["beta"], d["mask"])
    assert np.array_equal(got, d["out_fixed"])
    got_s = O.back_ts_f32(g, d["proj"], d["beta"], d["mask"], shipped_weights=True)
    assert np.array_equal(got_s, d["out_shipped"])
    psta = np.ascontiguousarray(d["proj"].transpose(0, 2, 1))
    got_st = O.back_st_f32(g, psta, d["beta"], d["mask"])
    assert np.array_equal(got_st, d["out_st"])


@pytest.mark.skipif(not O.have_reference(), reason="reference C not built")
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("seed", [11, 12])
def test_oracle_f32_matches_reference_c(name, seed):
    g, proj, beta, mask = make_case(name, seed=seed)
    want = O.ref_back_ts("fixed", g, proj, beta, mask, nthread=5)
    assert np.array_equal(O.back_ts_f32(g, proj, beta, mask), want)
    want_s = O.ref_back_ts("shipped", g, proj, beta, mask, nthread=1)
    assert np.array_equal(O.back_ts_f32(g, proj, beta, mask, shipped_weights=True), want_s)


# --------------------------------------------------------------------------
# independent per-voxel evaluation (tent kernels), float64

def tent_backproject(g, proj, beta, mask, cols=None):
    """sum over views of w2 * sum_{i,j} V[i,j] tri(sb - i) tri(t - j), evaluated
    directly at each voxel centre; returns (ny, nx, nz) or (ncol, nz)."""
    nx, ny, nz, ns, nt = g.nx, g.ny, g.nz, g.ns, g.nt
    wx = (nx - 1) / 2 + g.offset_x
    wy = (ny - 1) / 2 + g.offset_y
    wz = (nz - 1) / 2 + g.offset_z
    ws = (ns - 1) / 2 + g.offset_s
    wt = (nt - 1) / 2 + g.offset_t
    jj = np.arange(ny * nx) if cols is None else np.asarray(cols)
    ix, iy = jj % nx, jj // nx
    keep = mask.reshape(-1)[jj] != 0
    x = g.dx * (ix - wx)
    y = g.dy * (iy - wy)
    zrel = np.arange(nz) - wz
    out = np.zeros((jj.size, nz))
    par = math.isinf(g.dso) or math.isinf(g.dsd)
    arc = (not par) and g.dfs == 0
    si = np.arange(ns)
    ti = np.arange(nt)
    for ia, b in enumerate(beta):
        V = proj[ia].astype(np.float64)                    # [ns][nt]
        u = x * math.cos(b) + y * math.sin(b)             # along the detector
        L = g.dso - (-x * math.sin(b) + y * math.cos(b))  # source-to-voxel depth
        if par:
            s, mag, w2 = u, np.ones_like(u), np.ones_like(u)
        elif arc:
            s, mag = g.dsd * np.arctan2(u, L), g.dsd / L
            w2 = g.dsd ** 2 / (L ** 2 + u ** 2)
        else:
            s, mag = g.dsd * u / L, g.dsd / L
            w2 = mag ** 2
        sb = s / g.ds + ws                                # [ncol]
        t = (mag * g.dz / g.dt)[:, None] * zrel[None, :] + wt   # [ncol, nz]
        ks = np.maximum(0, 1 - np.abs(sb[:, None] - si[None, :]))   # [ncol, ns]
        kt = np.maximum(0, 1 - np.abs(t[:, :, None] - ti[None, None, :]))  # [ncol,nz,nt]
        val = np.einsum("cs,st,czt->cz", ks, V, kt)
        out += w2[:, None] * val
    out[~keep] = 0
    return out.reshape(ny, nx, nz) if cols is None else out


TINY = {
    "tiny_flat": O.geom((4, 4, 4), (1.0, -1.0, 1.0), (0.0, 0.0, 0.0), 100.0, 180.0,
                        np.inf, 8, 6, (1.0, 1.0), (0.0, 0.0)),
    "tiny_arc": O.geom((4, 4, 4), (1.0, -1.0, 1.0), (0.25, 0.0, -0.5), 100.0, 180.0,
                       0.0, 8, 6, (1.0, 1.0), (0.5, 0.25)),
    "tiny_par": O.geom((4, 4, 4), (1.0, -1.0, 1.0), (0.0, 0.0, 0.0), np.inf, np.inf,
                       0.0, 8, 6, (1.0, 1.0), (0.0, 0.0)),
}


@pytest.mark.parametrize("name", sorted(TINY))
def test_f64_brute_force_tiny(name):
    g = TINY[name]
    rng = np.random.default_rng(21)
    proj = rng.standard_normal((2, g.ns, g.nt)).astype(np.float32)
    beta = np.array([0.3, 2.1])
    mask = np.ones((g.ny, g.nx), np.uint8)
    got = O.back_ts_f64(g, proj, beta, mask)
    want = tent_backproject(g, proj, beta, mask)
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_f64_matches_tent_on_golden_inputs(path):
    g, d = load_golden(path)
    got = O.back_ts_f64(g, d["proj"], d["beta"], d["mask"])
    want = tent_backproject(g, d["proj"], d["beta"], d["mask"])
    assert O.rel_l2(got, want) < 1e-12
    # the FP32 reference sits within FP32 rounding of the FP64 answer
    assert O.rel_l2(d["out_fixed"], want) < 2e-5
    # the as-shipped vertical weights are visibly different (checker power);
    # not on the parallel fixture, where t_inc = 1 and wt = 5.5 make every
    # vertical weight exactly 0.5, so the swap changes nothing there
    if math.isfinite(g.dso):
        assert O.rel_l2(d["out_shipped"], want) > 1e-2


def test_f64_impulse_footprint():
    """A unit impulse (ct_geom.unitv) backprojects to w2 * tent(s) * tent(t)."""
    g = O.geom((24, 24, 16), (1.0, -1.0, 1.0), (0.0, 0.0, 0.0), 300.0, 500.0,
               np.inf, 40, 20, (1.0, 1.0), (0.0, 0.0))
    proj = np.zeros((3, g.ns, g.nt), np.float32)
    proj[1, 21, 9] = 1.0
    beta = np.array([0.0, 0.7, 2.0])
    mask = np.ones((g.ny, g.nx), np.uint8)
    got = O.back_ts_f64(g, proj, beta, mask)
    want = tent_backproject(g, proj, beta, mask)
    assert np.count_nonzero(want) > 10
    assert np.allclose(got, want, atol=1e-13)
    shipped = O.back_ts_f32(g, proj, beta, mask, shipped_weights=True)
    assert not np.allclose(shipped, want, atol=1e-3)


def test_constant_view_closed_forms():
    c = 2.5
    # parallel beam: w2 = 1, every interior voxel collects na * c
    g = O.geom((12, 12, 6), (1.0, -1.0, 1.0), (0.0, 0.0, 0.0), np.inf, np.inf, 0.0,
               40, 20, (1.0, 1.0), (0.0, 0.0))
    na = 7
    proj = np.full((na, g.ns, g.nt), c, np.float32)
    beta = np.linspace(0, np.pi, na, endpoint=False)
    img = O.back_ts_f64(g, proj, beta, None)
    assert np.allclose(img, na * c, rtol=1e-14)
    # flat detector: voxel = c * sum_views (dsd / (dso - v))^2 for voxels whose
    # rays hit the detector interior
    g = O.geom((10, 10, 4), (1.0, -1.0, 1.0), (0.0, 0.0, 0.0), 400.0, 700.0, np.inf,
               64, 48, (1.0, 1.0), (0.0, 0.0))
    proj = np.full((na, g.ns, g.nt), c, np.float32)
    img = O.back_ts_f64(g, proj, beta, None)
    wx = wy = (10 - 1) / 2
    iy, ix = np.mgrid[0:10, 0:10]
    x, y = 1.0 * (ix - wx), -1.0 * (iy - wy)
    want = np.zeros((10, 10))
    for b in beta:
        v = -x * np.sin(b) + y * np.cos(b)
        want += (700.0 / (400.0 - v)) ** 2
    assert np.allclose(img, c * want[:, :, None], rtol=1e-13)


def test_linearity_and_mask():
    g, p1, beta, mask = make_case("arc", seed=31)
    _, p2, _, _ = make_case("arc", seed=32)
    a, b = 0.75, -1.5
    lhs = O.back_ts_f64(g, (a * p1.astype(np.float64) + b * p2).astype(np.float32), beta, mask)
    rhs = a * O.back_ts_f64(g, p1, beta, mask) + b * O.back_ts_f64(g, p2, beta, mask)
    assert O.rel_l2(lhs, rhs) < 1e-7     # the combination is rounded to float32 once
    img = O.back_ts_f64(g, p1, beta, mask)
    assert np.all(img[mask == 0] == 0)
    img32 = O.back_ts_f32(g, p1, beta, mask)
    assert np.all(img32[mask == 0] == 0)
    assert np.count_nonzero(img32[mask != 0]) > 0


def test_column_subset_matches_full():
    g, proj, beta, mask = make_case("flat", seed=41)
    full32 = O.back_ts_f32(g, proj, beta, mask)
    full64 = O.back_ts_f64(g, proj, beta, mask)
    cols = np.array([0, 5, 77, g.nx * g.ny - 1, 300], np.int32)
    sub32 = O.back_ts_f32(g, proj, beta, mask, cols=cols)
    sub64 = O.back_ts_f64(g, proj, beta, mask, cols=cols)
    f32 = full32.reshape(-1, g.nz)
    f64 = full64.reshape(-1, g.nz)
    assert np.array_equal(sub32, f32[cols])
    assert np.array_equal(sub64, f64[cols])


def test_dz_over_dt_negative_fails():
    g = O.geom((8, 8, 4), (1.0, -1.0, 1.0), (0, 0, 0), 541.0, 949.0, 0.0, 16, 8,
               (1.0, -1.0), (0.0, 0.0))   # ct_geom's default dt = -ds
    proj = np.ones((2, 16, 8), np.float32)
    with pytest.raises(RuntimeError, match="dz/dt"):
        O.back_ts_f32(g, proj, np.array([0.0, 1.0]), None)
    if O.have_reference():
        with pytest.raises(RuntimeError):
            O.ref_back_ts("fixed", g, proj, np.array([0.0, 1.0]), None)
