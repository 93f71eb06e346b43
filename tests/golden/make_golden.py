"""Write the golden fixtures of the FDK backprojector (tests/golden/*.npz).

The reference ships no stored vectors for this path, so the fixtures are
produced here by the reference C itself, compiled from its own sources by
oracle/Makefile (oracle/_ref):
  * ``fixed``   -- mex/src/fdk-helix/mex/fdk_ts_back_t on its axial path;
  * ``shipped`` -- mex/src/fdk/fdk-orig/fdk_ts_back_t (swapped vertical weights);
  * ``st``      -- fdk_st_back1 ('fdk,st,back' order) from fdk-helix/mex.
Each fixture holds the seeded inputs (projections [nt ns na] as a C-order
(na, ns, nt) float32 array, beta, mask [nx ny] as C-order (ny, nx)), the
scalar geometry and the three outputs.  ``helix/*.npz`` hold the helical and
cone-parallel cases of the fdk-helix kernel (``fixed`` output only).

Run:  make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

# name: (nxyz, dxyz, cxyz, dso, dsd, dfs, ns, nt, dst, ost, na, beta0, orbit)
CASES = {
    "flat_small": ((16, 14, 10), (1.0, -1.1, 0.9), (0.3, -0.2, 0.5), 541.0, 949.0,
                   np.inf, 24, 16, (1.3, 1.1), (1.25, 0.3), 12, 0.1, 360.0),
    "arc_small": ((16, 14, 10), (1.0, -1.1, 0.9), (0.3, -0.2, 0.5), 541.0, 949.075,
                  0.0, 24, 16, (1.3, 1.1), (1.25, 0.3), 12, 0.1, 360.0),
    "par_small": ((15, 13, 9), (1.0, -1.0, 1.0), (0.0, 0.0, 0.0), np.inf, np.inf,
                  0.0, 22, 12, (1.0, 1.0), (0.0, 0.0), 10, 0.0, 180.0),
    # ge1-like arc, coarse: quarter-detector offset, short detector
    "ge1_coarse": ((20, 20, 24), (2.8, -2.8, 2.5), (0.0, 0.0, 0.0), 541.0, 949.075,
                   0.0, 28, 8, (4.0956, 4.3856), (1.25, 0.0), 16, 0.0, 360.0),
    # flat detector, partial orbit, odd sizes, centred image offsets
    "flat_short_scan": ((11, 17, 13), (1.5, -1.5, 1.2), (0.5, 0.5, 0.5), 500.0,
                        900.0, np.inf, 30, 20, (1.6, 1.4), (0.0, -0.5), 9, -0.3, 220.0),
}


def make(name, spec, seed):
    nxyz, dxyz, cxyz, dso, dsd, dfs, ns, nt, dst, ost, na, b0, orbit = spec
    g = O.geom(nxyz, dxyz, cxyz, dso, dsd, dfs, ns, nt, dst, ost)
    rng = np.random.default_rng(seed)
    proj = rng.standard_normal((na, ns, nt)).astype(np.float32)
    beta = b0 + np.deg2rad(orbit) * np.arange(na) / na
    mask = (rng.random((nxyz[1], nxyz[0])) > 0.15).astype(np.uint8)
    out_fixed = O.ref_back_ts("fixed", g, proj, beta, mask, nthread=3)
    out_shipped = O.ref_back_ts("shipped", g, proj, beta, mask, nthread=2)
    psta = np.ascontiguousarray(proj.transpose(0, 2, 1))
    out_st = O.ref_back_st("fixed", g, psta, beta, mask)
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"),
        nxyz=np.array(nxyz, np.int32), dxyz=np.array(dxyz), cxyz=np.array(cxyz),
        dso=dso, dsd=dsd, dfs=dfs, ns=ns, nt=nt, dst=np.array(dst), ost=np.array(ost),
        proj=proj, beta=beta, mask=mask,
        out_fixed=out_fixed, out_shipped=out_shipped, out_st=out_st)


# Helical / cone-parallel cases of the fdk-helix kernel (the 19-operand
# 'fdk,ts,back'), written to tests/golden/helix/:
# name: (nxyz, dxyz, cxyz, dfs, ns, nt, dst, ost, na, views per turn, pitch,
#        cone_par, w3d).  dso 541, dsd 949; source_zs as ct_geom.m builds it
# from the pitch (source_z0 = -na/2 * dz_per_view), na1_half and
# source_dz_per_rad as cbct_back_helix.m:110-111 computes them.  w3d tables
# of 16 entries (64 bytes), so that the reference's read one past its table
# lands in calloc's zeroed slack (fdk,w3d,pow,tab.c:45-46).
HELIX_CASES = {
    "helix_flat": ((18, 16, 30), (1.5, -1.5, 1.0), (0.3, -0.2, 0.0), np.inf, 40, 24,
                   (2.0, 1.6), (0.75, 0.3), 60, 40, 1.0, 0, 0),
    "helix_arc": ((18, 16, 30), (1.5, -1.5, 1.0), (0.3, -0.2, 0.4), 0.0, 40, 24,
                  (2.0, 1.6), (1.25, 0.0), 60, 40, 0.8, 0, 0),
    "helix_conepar_w3d": ((18, 16, 30), (1.5, -1.5, 1.0), (0.3, -0.2, 0.0), np.inf, 40,
                          32, (2.0, 1.6), (0.75, 0.0), 60, 40, 1.0, 1, 1),
    "axial_conepar_w3d": ((18, 16, 20), (1.5, -1.5, 1.0), (0.3, -0.2, 0.0), np.inf, 40,
                          32, (2.0, 1.6), (0.75, 0.0), 30, 30, 0.0, 1, 2),
}


def helix_inputs(spec, seed):
    """Seeded inputs and helical operands of one HELIX_CASES entry."""
    (nxyz, dxyz, cxyz, dfs, ns, nt, dst, ost, na, na360, pitch, cone_par, w3d) = spec
    g = O.geom(nxyz, dxyz, cxyz, 541.0, 949.0, dfs, ns, nt, dst, ost)
    rng = np.random.default_rng(seed)
    proj = rng.standard_normal((na, ns, nt)).astype(np.float32)
    if w3d:   # zero edge rows: samples whose w3d index leaves the table add 0
        proj[:, :, 0] = 0
        proj[:, :, -1] = 0
    beta = 0.2 + 2 * np.pi * np.arange(na) / na360
    zfov = nxyz[2] * dxyz[2]
    dz_view = pitch * zfov / na360
    source_zs = (np.arange(na) - na / 2) * dz_view
    mask = (rng.random((nxyz[1], nxyz[0])) > 0.15).astype(np.uint8)
    ops = dict(na1_half=int(np.ceil(na360 / 2)), cone_par=cone_par, w3d=w3d,
               source_dz_per_rad=dz_view / (2 * np.pi / na360), pitch=pitch)
    return g, proj, beta, source_zs, mask, ops


def make_helix(name, spec, seed):
    g, proj, beta, source_zs, mask, ops = helix_inputs(spec, seed)
    out = O.ref_back_ts_helix(g, proj, beta, source_zs, mask, nthread=3, **ops)
    os.makedirs(os.path.join(HERE, "helix"), exist_ok=True)
    nxyz, dxyz, cxyz, dfs, ns, nt, dst, ost = spec[:8]
    np.savez_compressed(
        os.path.join(HERE, "helix", f"{name}.npz"),
        nxyz=np.array(nxyz, np.int32), dxyz=np.array(dxyz), cxyz=np.array(cxyz),
        dso=541.0, dsd=949.0, dfs=dfs, ns=ns, nt=nt, dst=np.array(dst), ost=np.array(ost),
        proj=proj, beta=beta, source_zs=source_zs, mask=mask,
        na1_half=ops["na1_half"], cone_par=ops["cone_par"], w3d=ops["w3d"],
        source_dz_per_rad=ops["source_dz_per_rad"], pitch=np.float32(ops["pitch"]),
        out_fixed=out)


if __name__ == "__main__":
    if not (O.have_reference("fixed") and O.have_reference("shipped")):
        sys.exit("reference C not built: make -C oracle ref")
    for i, (name, spec) in enumerate(CASES.items()):
        make(name, spec, seed=100 + i)
        print("wrote", name)
    for i, (name, spec) in enumerate(HELIX_CASES.items()):
        make_helix(name, spec, seed=200 + i)
        print("wrote helix/" + name)
