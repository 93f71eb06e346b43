"""TEST INFRASTRUCTURE ONLY: numpy FP64 restatement of ``ellipsoid_proj``.

Checker for the device kernel ``fdkhip_ellipsoid_proj``
(mirt_amd/csrc/phantom_kernel.hip); never imported by the product.  Follows
fbp/ellipsoid_proj.m:35-125 operation by operation (the coordinate maps of
fbp/ir_coord_cb_flat_to_par.m and fbp/ir_coord_cb_arc_to_par.m, the chord
formula of :110-118, the real part of :123, oversampling :44-58 with
downsample3's averaging order, utilities/downsample3.m: along s, then t).

MATLAB is absent, so this restatement is pinned by closed forms
(tests/test_host.py): parallel-beam chords of a sphere and of a rotated
ellipsoid, cone-beam central-ray chords, shift invariance in z with the source
height, and averaging of sub-samples.
"""
from __future__ import annotations

import numpy as np


def par_coords(ss, tt, dso, dod, dfs):
    """(uu, vv, azim, polar) at beta = 0 (ellipsoid_proj.m:67-82)."""
    if np.isinf(dso):                                  # parallel beam
        return ss, tt, np.zeros_like(ss), np.zeros_like(ss)
    Ds, Dd = dso, dod
    Dc = Ds + Dd
    if np.isinf(dfs):                                  # ir_coord_cb_flat_to_par.m
        uu = Ds * ss / np.sqrt(Dc ** 2 + ss ** 2)
        vv = Ds * tt / np.sqrt(Dc ** 2 + ss ** 2 + tt ** 2) * Dc / np.sqrt(Dc ** 2 + ss ** 2)
        polar = -np.arctan(tt / np.sqrt(Dc ** 2 + ss ** 2))
        azim = np.arctan(ss / Dc)
        return uu, vv, azim, polar
    if dfs == 0:                                       # ir_coord_cb_arc_to_par.m
        Rf = 0 + Dc
        p1 = Rf * np.sin(ss / Rf)
        p2 = Ds + 0 - Rf * np.cos(ss / Rf)
        p3 = tt
        e1, e2, e3 = p1 - 0, p2 - dso, p3 - 0
        en = np.sqrt(e1 ** 2 + e2 ** 2 + e3 ** 2)
        e1, e2, e3 = e1 / en, e2 / en, e3 / en
        polar = -np.arcsin(e3)
        azim = -np.arctan(e1 / e2)
        uu = np.cos(azim) * 0 + np.sin(azim) * dso
        vv = (np.sin(azim) * 0 - np.cos(azim) * dso) * np.sin(polar) + 0 * np.cos(polar)
        return uu, vv, azim, polar
    raise ValueError("not done")


def ellipsoid_proj(s, t, beta, source_zs, dso, dod, dfs, params, oversample=1):
    """Line integrals ``[ns nt na]`` (numpy float64, MATLAB index order)."""
    params = np.atleast_2d(np.asarray(params, np.float64))
    if params.shape[1] != 9:
        raise ValueError("9 parameters per ellipsoid")
    s = np.asarray(s, np.float64)
    t = np.asarray(t, np.float64)
    beta = np.atleast_1d(np.asarray(beta, np.float64))
    zs = np.zeros_like(beta) if source_zs is None else \
        np.broadcast_to(np.asarray(source_zs, np.float64), beta.shape)
    if oversample > 1:
        No = oversample
        ds, dt = s[1] - s[0], t[1] - t[0]
        off = np.arange(-(No - 1), No, 2) / (2 * No)
        sf = (off[:, None] * ds + s[None, :]).T.reshape(-1)     # [No ns] -> s(:)
        tf = (off[:, None] * dt + t[None, :]).T.reshape(-1)
        fine = ellipsoid_proj(sf, tf, beta, zs, dso, dod, dfs, params, 1)
        ns, nt = s.size, t.size
        y = fine.reshape(ns, No, nt * No, -1).sum(axis=1) / No      # along s
        y = y.reshape(ns, nt, No, -1).sum(axis=2) / No              # along t
        return y
    sss, ttt = np.meshgrid(s, t, indexing="ij")
    uu, vv, azim0, polar = par_coords(sss, ttt, dso, dod, dfs)
    cpol, spol = np.cos(polar), np.sin(polar)
    out = np.zeros((s.size, t.size, beta.size))
    for p in params:
        cx, cy, cz, rx, ry, rz = p[:6]
        xang = np.deg2rad(p[6])
        if np.deg2rad(p[7]):
            raise ValueError("z rotation not done")
        val = p[8]
        for ib, b in enumerate(beta):
            az = b + azim0
            cz_eff = cz - zs[ib]
            ushift = cx * np.cos(az) + cy * np.sin(az)
            vshift = (cx * np.sin(az) - cy * np.cos(az)) * spol + cz_eff * cpol
            az = az - xang
            p1 = (uu - ushift) * np.cos(az) + (vv - vshift) * np.sin(az) * spol
            p2 = (uu - ushift) * np.sin(az) - (vv - vshift) * np.cos(az) * spol
            p3 = (vv - vshift) * cpol
            e1 = -np.sin(az) * cpol
            e2 = np.cos(az) * cpol
            e3 = spol
            A = e1 ** 2 / rx ** 2 + e2 ** 2 / ry ** 2 + e3 ** 2 / rz ** 2
            B = p1 * e1 / rx ** 2 + p2 * e2 / ry ** 2 + p3 * e3 / rz ** 2
            C = p1 ** 2 / rx ** 2 + p2 ** 2 / ry ** 2 + p3 ** 2 / rz ** 2 - 1
            d = B ** 2 - A * C
            # real part of 2 val sqrt(d) / A: zero where the ray misses
            out[:, :, ib] += np.where(d > 0, 2 * val * np.sqrt(np.maximum(d, 0)) / A, 0.0)
    return out
