"""Analytic cone-beam projections of ellipsoids (synthetic scan inputs).

Restatement of ``ellipsoid_proj`` (fbp/ellipsoid_proj.m:35-125) with the
detector-to-parallel coordinate maps ``ir_coord_cb_flat_to_par`` and
``ir_coord_cb_arc_to_par`` (fbp/ir_coord_cb_{flat,arc}_to_par.m), written with
torch so the same code runs on the CPU (small cases) and on the GPU (full-size
scans).  Also the seeded random phantoms and the Poisson transmission noise
used by the benchmark configurations.

Ellipsoid parameter rows are MIRT's ``[cx cy cz rx ry rz xy_angle_deg
z_angle_deg amplitude]``; z rotation is unsupported, as in the reference
(ellipsoid_proj.m:94).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .geom import CTGeom


def _coords(cg: CTGeom, ss, tt):
    """(uu, vv, azim, polar) of each detector sample at beta = 0."""
    dso, dod = cg.dso, cg.dod
    if math.isinf(dso):                       # parallel beam
        return ss, tt, torch.zeros_like(ss), torch.zeros_like(ss)
    Ds, Dd = dso, dod
    Dc = Ds + Dd
    if math.isinf(cg.dfs):                    # flat: ir_coord_cb_flat_to_par.m
        r2 = Dc * Dc + ss * ss
        uu = Ds * ss / torch.sqrt(r2)
        vv = Ds * tt / torch.sqrt(r2 + tt * tt) * Dc / torch.sqrt(r2)
        polar = -torch.atan(tt / torch.sqrt(r2))
        azim = torch.atan(ss / Dc)
        return uu, vv, azim, polar
    if cg.dfs == 0:                           # arc: ir_coord_cb_arc_to_par.m
        Rf = Dc
        p1 = Rf * torch.sin(ss / Rf)
        p2 = Ds - Rf * torch.cos(ss / Rf)
        e1 = p1 - 0.0
        e2 = p2 - dso
        e3 = tt - 0.0
        en = torch.sqrt(e1 * e1 + e2 * e2 + e3 * e3)
        e1, e2, e3 = e1 / en, e2 / en, e3 / en
        polar = -torch.asin(e3)
        azim = -torch.atan(e1 / e2)
        uu = torch.cos(azim) * 0.0 + torch.sin(azim) * dso
        vv = (torch.sin(azim) * 0.0 - torch.cos(azim) * dso) * torch.sin(polar) \
            + 0.0 * torch.cos(polar)
        return uu, vv, azim, polar
    raise ValueError("dfs not done")


def ellipsoid_proj(cg: CTGeom, params, *, oversample: int = 1, device="cpu",
                   dtype=torch.float64, views=None):
    """Line integrals ``[ns nt na]`` (torch tensor, MATLAB index order).

    ``views`` optionally restricts to a list of view indices (0-based).
    """
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    if params.shape[1] != 9:
        raise ValueError("9 parameters per ellipsoid")
    ss = torch.as_tensor(cg.s, dtype=dtype, device=device)
    tt = torch.as_tensor(cg.t, dtype=dtype, device=device)
    beta = cg.ar
    szs = cg.source_zs
    if views is not None:
        beta = beta[np.asarray(views)]
        szs = szs[np.asarray(views)]
    if oversample > 1:
        No = oversample
        ds = float(cg.s[1] - cg.s[0]) if cg.ns > 1 else cg.ds
        dt = float(cg.t[1] - cg.t[0]) if cg.nt > 1 else cg.dt
        off = torch.arange(-(No - 1), No, 2, dtype=dtype, device=device) / (2 * No)
        ss = (off[:, None] * ds + ss[None, :]).T.reshape(-1)
        tt = (off[:, None] * dt + tt[None, :]).T.reshape(-1)
    sss, ttt = torch.meshgrid(ss, tt, indexing="ij")
    uu, vv, azim0, polar = _coords(cg, sss, ttt)
    cpol, spol = torch.cos(polar)[..., None], torch.sin(polar)[..., None]
    uu, vv, azim0 = uu[..., None], vv[..., None], azim0[..., None]
    na = len(beta)
    out = torch.zeros((sss.shape[0], sss.shape[1], na), dtype=dtype, device=device)
    # views in chunks: every elementwise step is over [ns, nt, views]
    chunk = max(1, int(2 ** 24 // max(1, sss.numel())))
    for v0 in range(0, na, chunk):
        v1 = min(na, v0 + chunk)
        bt = torch.as_tensor(np.asarray(beta[v0:v1], dtype=np.float64), dtype=dtype,
                             device=device)
        zs = torch.as_tensor(np.asarray(szs[v0:v1], dtype=np.float64), dtype=dtype,
                             device=device)
        acc = out[:, :, v0:v1]
        for p in params:
            cx, cy, cz, rx, ry, rz = (float(v) for v in p[:6])
            xang = math.radians(p[6])
            if p[7] != 0:
                raise ValueError("z rotation not done")
            val = float(p[8])
            az = bt + azim0                      # [ns, nt, nv]
            cz_eff = cz - zs
            ushift = cx * torch.cos(az) + cy * torch.sin(az)
            vshift = (cx * torch.sin(az) - cy * torch.cos(az)) * spol + cz_eff * cpol
            az = az - xang
            caz, saz = torch.cos(az), torch.sin(az)
            du, dv = uu - ushift, vv - vshift
            p1 = du * caz + dv * saz * spol
            p2 = du * saz - dv * caz * spol
            p3 = dv * cpol
            e1 = -saz * cpol
            e2 = caz * cpol
            e3 = spol
            A = e1 * e1 / rx ** 2 + e2 * e2 / ry ** 2 + e3 * e3 / rz ** 2
            B = p1 * e1 / rx ** 2 + p2 * e2 / ry ** 2 + p3 * e3 / rz ** 2
            Cc = p1 * p1 / rx ** 2 + p2 * p2 / ry ** 2 + p3 * p3 / rz ** 2 - 1
            # real part of 2*val*sqrt(B^2 - A C)/A: zero where the ray misses
            acc += 2 * val * torch.sqrt(torch.clamp(B * B - A * Cc, min=0)) / A
    if oversample > 1:
        No = oversample
        ns, nt = cg.ns, cg.nt
        out = out.reshape(ns, No, nt, No, na).mean(dim=(1, 3))
    return out


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
