"""Synthetic scan workloads for benchmarks and full-size tests.

Each workload is a circular cone-beam scan in MIRT's conventions (ct_geom /
image_geom, dso = 541 mm, dsd = 949 mm, views evenly spaced over 360 deg)
with a seeded random-ellipsoid phantom; :func:`filtered_views` turns one into
the FDK-filtered views cbct_back receives, following feldkamp.m's steps
(analytic line integrals, optional Poisson noise, cosine weight, ramp
filter).

* ``cpu_small``     flat 128 x 96 at 1.0 mm, 64 views, 64 x 64 x 32 at 1.0 mm,
                     10 ellipsoids, noise-free;
* ``plain_flat``    flat 512 x 384 at 0.8 mm, 360 views, 256 x 256 x 128 at
                     0.5 mm, 20 ellipsoids, Poisson I0 = 1e5;
* ``clinical_arc``  arc 888 x 64 (ds 1.0239, dt 1.0964, offset_s 1.25),
                     984 views, 512 x 512 x 256 at 0.7 / 0.625 mm,
                     30 ellipsoids, Poisson I0 = 2e5;
* ``large_flat``    flat 1024 x 768 at 0.4 mm, 1440 views, 1024 x 1024 x 512 at
                     0.25 mm, 40 ellipsoids plus small high-contrast inserts,
                     Poisson I0 = 1e5;
* ``clinical_helix`` the ``clinical_arc`` detector on a helix: 3 turns of
                     984 views at pitch 1 (source_z0 centred), 512 x 512 x 192
                     at 0.7 / 0.625 mm (the 120 mm the source travels),
                     30 ellipsoids, Poisson I0 = 2e5;
* ``uneven_scans``  seeded scans of uneven size: volume edge in
                     {128, 192, 256, 384} with nz = edge/2, views in [180, 720],
                     flat detector columns in [256, 768] and rows in
                     [128, 512], 10-30 ellipsoids each, noise-free.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .geom import CTGeom, ImageGeom, ct_geom, image_geom


@dataclass
class Scan:
    name: str
    cg: CTGeom
    ig: ImageGeom
    ellipsoids: np.ndarray      # rows of random_ellipsoids
    i0: float | None            # Poisson incident count, None = noise-free
    seed: int


def _scan(name, seed, *, ns, nt, na, ds, dt, dfs, nxyz, dx, dz, n_ell, i0,
          offset_s=0.0, inserts=0, turns=1, pitch=0.0):
    from .phantom import random_ellipsoids
    helix = dict(pitch=pitch, source_z0="center") if pitch else {}
    cg = ct_geom("fan", ns=ns, nt=nt, na=na, ds=ds, dt=dt, offset_s=offset_s,
                 dsd=949.0, dso=541.0, dfs=dfs, orbit=360 * turns, **helix)
    ig = image_geom(nx=nxyz[0], ny=nxyz[1], nz=nxyz[2], dx=dx, dz=dz)
    ell = random_ellipsoids(seed, n_ell, ig.nx * ig.dx, ig.nz * ig.dz, inserts=inserts)
    return Scan(name, cg, ig, ell, i0, seed)


def cpu_small(seed: int = 1000) -> Scan:
    return _scan("cpu_small", seed, ns=128, nt=96, na=64, ds=1.0, dt=1.0, dfs=np.inf,
                 nxyz=(64, 64, 32), dx=1.0, dz=1.0, n_ell=10, i0=None)


def plain_flat(seed: int = 1500) -> Scan:
    return _scan("plain_flat", seed, ns=512, nt=384, na=360, ds=0.8, dt=0.8, dfs=np.inf,
                 nxyz=(256, 256, 128), dx=0.5, dz=0.5, n_ell=20, i0=1e5)


def clinical_arc(seed: int = 2000) -> Scan:
    return _scan("clinical_arc", seed, ns=888, nt=64, na=984, ds=1.0239, dt=1.0964,
                 dfs=0.0, offset_s=1.25, nxyz=(512, 512, 256), dx=0.7, dz=0.625,
                 n_ell=30, i0=2e5)


def clinical_helix(seed: int = 2500) -> Scan:
    return _scan("clinical_helix", seed, ns=888, nt=64, na=3 * 984, ds=1.0239, dt=1.0964,
                 dfs=0.0, offset_s=1.25, nxyz=(512, 512, 192), dx=0.7, dz=0.625,
                 n_ell=30, i0=2e5, turns=3, pitch=1.0)


def large_flat(seed: int = 3000) -> Scan:
    return _scan("large_flat", seed, ns=1024, nt=768, na=1440, ds=0.4, dt=0.4, dfs=np.inf,
                 nxyz=(1024, 1024, 512), dx=0.25, dz=0.25, n_ell=40, i0=1e5, inserts=12)


# uneven scans: every volume spans 200 mm in-plane; the flat detector spans
# 640 x 320 mm (enough for the cone through the whole volume) at the drawn
# pixel counts
UNEVEN_FOV_MM = 200.0
UNEVEN_DET_MM = (640.0, 320.0)


def uneven_scans(n: int = 32, seed: int = 4000) -> list[Scan]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        edge = int(rng.choice([128, 192, 256, 384]))
        na = int(rng.integers(180, 721))
        ns = int(rng.integers(256, 769))
        nt = int(rng.integers(128, 513))
        n_ell = int(rng.integers(10, 31))
        dx = UNEVEN_FOV_MM / edge
        out.append(_scan(f"uneven{i:02d}", seed + 1 + i, ns=ns, nt=nt, na=na,
                         ds=UNEVEN_DET_MM[0] / ns, dt=UNEVEN_DET_MM[1] / nt, dfs=np.inf,
                         nxyz=(edge, edge, edge // 2), dx=dx, dz=dx, n_ell=n_ell, i0=None))
    return out


WORKLOADS = {"cpu_small": cpu_small, "plain_flat": plain_flat,
             "clinical_arc": clinical_arc, "large_flat": large_flat,
             "clinical_helix": clinical_helix}


def filtered_views(scan: Scan, device):
    """FDK-filtered views of ``scan`` on ``device``: float32 tensor
    (na, ns, nt), i.e. [nt ns na] with t fastest, as the plans take them."""
    import torch
    from .fdk import fdk_filter, feldkamp_weight1
    from .phantom import ellipsoid_proj, transmission_noise
    p = ellipsoid_proj(scan.cg, scan.ellipsoids, device=device)   # [ns nt na] float64
    if scan.i0 is not None:
        p = transmission_noise(p, scan.i0, seed=scan.seed)
    p = fdk_filter(feldkamp_weight1(p, scan.cg), scan.cg)
    return p.permute(2, 0, 1).contiguous().to(torch.float32)
