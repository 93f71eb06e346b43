"""Analytic cone-beam projections of ellipsoids (synthetic scan inputs).

``ellipsoid_proj`` (fbp/ellipsoid_proj.m:35-125) runs on the GPU through the
library's ``fdkhip_ellipsoid_proj`` kernel (mirt_amd/csrc/phantom_kernel.hip,
FP64, the detector-to-parallel coordinate maps of
fbp/ir_coord_cb_{flat,arc}_to_par.m).  Also the seeded random phantoms and the
Poisson transmission noise used by the benchmark configurations.

Ellipsoid parameter rows are MIRT's ``[cx cy cz rx ry rz xy_angle_deg
z_angle_deg amplitude]``; z rotation is unsupported, as in the reference
(ellipsoid_proj.m:94).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .geom import CTGeom


def ellipsoid_proj(cg: CTGeom, params, *, oversample: int = 1, device="cuda",
                   views=None, stream=None):
    """Line integrals ``[ns nt na]`` (MATLAB index order) as a float64 torch
    tensor on ``device`` (a HIP device: the kernel is the only path), whose
    memory is MATLAB's (s fastest).  ``views`` optionally restricts to a list
    of view indices (0-based)."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("ellipsoid_proj runs on the GPU (fdkhip_ellipsoid_proj); "
                           f"got device {dev}")
    params = np.ascontiguousarray(np.atleast_2d(np.asarray(params, dtype=np.float64)))
    if params.shape[1] != 9:
        raise ValueError("9 parameters per ellipsoid")
    beta = np.asarray(cg.ar, dtype=np.float64)
    zs = np.asarray(cg.source_zs, dtype=np.float64)
    if views is not None:
        beta = beta[np.asarray(views)]
        zs = zs[np.asarray(views)]
    beta = np.ascontiguousarray(beta)
    zs = np.ascontiguousarray(np.broadcast_to(zs, beta.shape))
    s = np.ascontiguousarray(cg.s, dtype=np.float64)
    t = np.ascontiguousarray(cg.t, dtype=np.float64)
    out = torch.empty((beta.size, t.size, s.size), dtype=torch.float64, device=dev)
    if stream is None:
        stream = torch.cuda.current_stream(dev).cuda_stream
    p = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    _lib.check(_lib.load().fdkhip_ellipsoid_proj(
        p(s), s.size, p(t), t.size, p(beta), p(zs), beta.size, float(cg.dso), float(cg.dod),
        float(cg.dfs), p(params), params.shape[0], int(oversample),
        C.c_void_p(out.data_ptr()), C.c_void_p(stream)))
    return out.permute(2, 1, 0)


def random_ellipsoids(seed: int, n: int, fov_xy: float, fov_z: float, *,
                      mu=(0.005, 0.02), inserts: int = 0, insert_mu=0.05,
                      insert_r=(1.0, 3.0)):
    """Seeded phantom rows: centres uniform in the inner 60 % of the FOV,
    semi-axes 5-40 % of the FOV, rotation about z uniform in [0, 180) deg,
    attenuation uniform in ``mu`` per mm; optional small high-contrast
    inserts (semi-axes in ``insert_r`` mm, attenuation ``insert_mu``)."""
    rng = np.random.default_rng(seed)
    rows = []
    half = np.array([fov_xy, fov_xy, fov_z]) / 2
    for _ in range(n):
        c = rng.uniform(-0.6, 0.6, 3) * half
        r = rng.uniform(0.05, 0.40, 3) * np.array([fov_xy, fov_xy, fov_z])
        rows.append([*c, *r, rng.uniform(0, 180), 0.0, rng.uniform(*mu)])
    for _ in range(inserts):
        c = rng.uniform(-0.6, 0.6, 3) * half
        r = rng.uniform(*insert_r, 3)
        rows.append([*c, *r, rng.uniform(0, 180), 0.0, insert_mu])
    return np.asarray(rows, dtype=np.float64)


def transmission_noise(proj, I0: float, seed: int):
    """Poisson counts n ~ Poisson(I0 exp(-proj)), returned as -log(max(n,1)/I0)."""
    g = torch.Generator(device=proj.device)
    g.manual_seed(int(seed))
    lam = I0 * torch.exp(-proj.to(torch.float64))
    n = torch.poisson(lam, generator=g)
    return -torch.log(torch.clamp(n, min=1.0) / I0)
