"""Pinning the helical / cone-parallel restatement (oracle.back_ts_helix_f32),
CPU only.

The fdk-helix kernel (mex/src/fdk-helix/mex/fdk,ts.c:55-251 with
fdk,ts,t.c:39-167 and fdk,w3d,pow,tab.c) is compiled from the reference
sources into oracle/_ref, so the restatement is held to it bit for bit:
* on the fixtures the reference C wrote (tests/golden/helix, always run);
* on fresh seeded cases: native helical flat/arc/parallel, cone-parallel
  with and without w3d weights, with and without a pitch, masks, offsets.
Its axial special case (source_zs = 0, pitch = 0, cone_par = 0) is the axial
restatement bit for bit, and a constant source height is the axial
backprojection of an image moved by that height (to FP32 rounding).
The w3d weights and the cone-parallel geometry change the result (the
checks see them).
"""
from __future__ import annotations

import glob
import importlib.util
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
HELIX_GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "helix", "*.npz")))
need_ref = pytest.mark.skipif(not O.have_reference(), reason="reference C not built")


def _make_golden_module():
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def load_helix(path):
    d = np.load(path, allow_pickle=False)
    g = O.geom(d["nxyz"], d["dxyz"], d["cxyz"], float(d["dso"]), float(d["dsd"]),
               float(d["dfs"]), int(d["ns"]), int(d["nt"]), d["dst"], d["ost"])
    ops = dict(na1_half=int(d["na1_half"]), cone_par=int(d["cone_par"]), w3d=int(d["w3d"]),
               source_dz_per_rad=float(d["source_dz_per_rad"]), pitch=float(d["pitch"]))
    return g, d, ops


@pytest.mark.parametrize("path", HELIX_GOLDEN, ids=[os.path.basename(p) for p in HELIX_GOLDEN])
def test_helix_restatement_matches_golden(path):
    g, d, ops = load_helix(path)
    got = O.back_ts_helix_f32(g, d["proj"], d["beta"], d["source_zs"], d["mask"], **ops)
    assert np.array_equal(got, d["out_fixed"])
    assert np.abs(d["out_fixed"]).max() > 0


def test_helix_golden_set_complete():
    names = {os.path.basename(p)[:-4] for p in HELIX_GOLDEN}
    assert names == set(_make_golden_module().HELIX_CASES)


def _seeded(dfs, *, pitch, cone_par, w3d, seed, nt=32, dso=541.0, dsd=949.0, offz=0.0,
            ost=(0.75, 0.3), nz=30):
    rng = np.random.default_rng(seed)
    g = O.geom((20, 18, nz), (1.5, -1.5, 1.0), (0.3, -0.2, offz), dso, dsd, dfs, 44, nt,
               (2.0, 1.6), ost if not w3d else (ost[0], 0.0))
    na360, na = 36, 54
    proj = rng.standard_normal((na, 44, nt)).astype(np.float32)
    if w3d:
        proj[:, :, 0] = 0
        proj[:, :, -1] = 0
    beta = 0.1 + 2 * np.pi * np.arange(na) / na360
    dz_view = pitch * nz * 1.0 / na360
    zs = (np.arange(na) - na / 2) * dz_view
    mask = (rng.random((18, 20)) > 0.1).astype(np.uint8)
    ops = dict(na1_half=na360 // 2, cone_par=cone_par, w3d=w3d,
               source_dz_per_rad=dz_view / (2 * np.pi / na360), pitch=pitch)
    return g, proj, beta, zs, mask, ops


HELIX_GRID = [
    # (dfs, pitch, cone_par, w3d)
    (np.inf, 1.0, 0, 0), (0.0, 1.0, 0, 0), (np.inf, 0.0, 0, 0), (0.0, 0.6, 0, 0),
    (np.inf, 1.0, 1, 0), (np.inf, 0.0, 1, 0), (np.inf, 1.0, 1, 1), (0.0, 1.0, 1, 3),
    (np.inf, 0.0, 1, 1), (np.inf, 0.0, 1, 10),
]


@need_ref
@pytest.mark.parametrize("dfs,pitch,cone_par,w3d", HELIX_GRID)
def test_helix_restatement_matches_reference_c(dfs, pitch, cone_par, w3d):
    g, proj, beta, zs, mask, ops = _seeded(dfs, pitch=pitch, cone_par=cone_par, w3d=w3d,
                                           seed=int(10 * pitch) + 3 * cone_par + w3d)
    want = O.ref_back_ts_helix(g, proj, beta, zs, mask, nthread=4, **ops)
    got = O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops)
    assert np.array_equal(got, want)
    assert np.abs(want).max() > 0


@need_ref
def test_helix_parallel_beam_matches_reference_c():
    g, proj, beta, zs, mask, ops = _seeded(0.0, pitch=1.0, cone_par=0, w3d=0, seed=5,
                                           dso=np.inf, dsd=np.inf)
    want = O.ref_back_ts_helix(g, proj, beta, zs, mask, nthread=2, **ops)
    assert np.array_equal(O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops), want)


def test_helix_axial_case_is_the_axial_restatement():
    g, proj, beta, _, mask, ops = _seeded(0.0, pitch=0.0, cone_par=0, w3d=0, seed=8)
    ops.update(source_dz_per_rad=0.0)
    got = O.back_ts_helix_f32(g, proj, beta, np.zeros(beta.size), mask, **ops)
    assert np.array_equal(got, O.back_ts_f32(g, proj, beta, mask))


def test_helix_constant_height_is_a_z_shift():
    """source_zs = c for every view, no pitch: the axial image of a volume
    moved by c (offset_z + c/dz), to FP32 rounding of the shifted centre."""
    g, proj, beta, _, mask, ops = _seeded(np.inf, pitch=0.0, cone_par=0, w3d=0, seed=9)
    c = 3.7
    got = O.back_ts_helix_f32(g, proj, beta, np.full(beta.size, c), mask, **ops)
    gs = O.geom((g.nx, g.ny, g.nz), (g.dx, g.dy, g.dz), (g.offset_x, g.offset_y,
                g.offset_z + c / g.dz), g.dso, g.dsd, g.dfs, g.ns, g.nt, (g.ds, g.dt),
                (g.offset_s, g.offset_t))
    want = O.back_ts_f64(gs, proj, beta, mask)
    assert O.rel_l2(got, want) < 1e-5
    assert O.rel_l2(got, O.back_ts_f32(g, proj, beta, mask)) > 1e-2


def test_helix_weights_and_cone_parallel_change_the_result():
    g, proj, beta, zs, mask, ops = _seeded(np.inf, pitch=1.0, cone_par=1, w3d=1, seed=2)
    w = O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops)
    ops0 = dict(ops, w3d=0)
    w0 = O.back_ts_helix_f32(g, proj, beta, zs, mask, **ops0)
    native = O.back_ts_helix_f32(g, proj, beta, zs, mask, **dict(ops0, cone_par=0))
    assert O.rel_l2(w, w0) > 1e-3
    assert O.rel_l2(w0, native) > 1e-2
