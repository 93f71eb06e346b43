"""Helical cone-beam backprojection: ``cbct_back_helix`` and its gateway.

Host-side mirror of ``mex/src/fdk-helix/cbct_back_helix.m`` (the helical
version of ``cbct_back`` with the view weighting of Tang et al., by D. Kim):
the view angles, the helical operands of the 19-operand ``fdk,ts,back``
gateway (``na1_half``, ``source_dz_per_rad``, ``pitch``, ``cone_par``,
``w3d``), the row replication, the permutes, and its per-slice angular scale.
The backprojection runs on the GPU through ``fdkhip_fdk_ts_back_helix``
(:func:`mirt_amd.cbct_back.fdk_mex_helix`).  There is no CPU fallback.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from .cbct_back import FdkMexError, _views, fdk_mex, fdk_mex_helix
from .geom import CTGeom, ImageGeom, deg2rad

_I32 = (-2 ** 31, 2 ** 31 - 1)


def _int32(x) -> int:
    """MATLAB int32() of a double: round half away from zero, saturate."""
    x = float(x)
    r = math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)
    return int(min(max(r, _I32[0]), _I32[1]))


def helix_operands(cg: CTGeom, ia_skip: int = 1):
    """na1_half and source_dz_per_rad as cbct_back_helix.m:110-111 computes
    them: int32(ceil((na/ia_skip) / |orbit| * 180)) and the source height
    step per radian of the views used."""
    na1_half = _int32(math.ceil((cg.na / ia_skip) / abs(cg.orbit) * 180))
    dz_per_rad = float(cg.source_dz_per_view / (float(deg2rad(abs(cg.orbit)))
                                                / (cg.na / ia_skip)))
    return na1_half, dz_per_rad


def slice_scales(cg: CTGeom, ig: ImageGeom, ia_skip: int, na1_half: int) -> np.ndarray:
    """The per-slice "\\der angle" scale of cbct_back_helix.m:138-160, in
    double: a helical slice is scaled by the share of the orbit whose views
    reach it.  MATLAB evaluates the view counts in int32 arithmetic (each
    step rounded), which is kept."""
    na = cg.na
    orbit = abs(cg.orbit)
    wz = (ig.nz - 1) / 2 + ig.offset_z
    zc = (np.arange(ig.nz) - wz) * ig.dz
    zs = np.asarray(cg.source_zs, np.float64)
    out = np.empty(ig.nz)
    for iz in range(ig.nz):
        if cg.pitch == 0:
            orbit_iz = ((na - 1 + 1) / ia_skip) * orbit / (na / ia_skip)
        else:
            ia1_middle = int(np.argmin(np.abs(zc[iz] - zs))) + 1          # 1-based imin
            ia_min = max(_int32(ia1_middle - na1_half), 1)
            ia_max = min(_int32(ia1_middle + (na1_half - 1)), na)
            t = _int32(ia_max - ia_min + 1)
            t = _int32(t / ia_skip)
            t = _int32(t * orbit)
            orbit_iz = float(_int32(t / (na / ia_skip)))
        scale = math.pi / float(deg2rad(abs(orbit_iz)))
        out[iz] = scale * float(deg2rad(orbit)) / (na / ia_skip)
    return out


def cbct_back_helix(proj, cg: CTGeom, ig: ImageGeom, *, use_mex=None, ia_skip: int = 1,
                    offset_source: float = 0.0, nthread=None, back_call=None,
                    scale_dang: bool = True, extrapolate_t: int = 0, cone_par: int = 0,
                    w3d: int = 0):
    """Mirror of ``cbct_back_helix`` (mex/src/fdk-helix/cbct_back_helix.m:1-161).

    ``proj`` is ``[ns nt na]`` (numpy, MATLAB index order); returns the
    ``[nx ny nz]`` float32 backprojection.  ``use_mex`` 1 (default) calls
    ``back_call`` (default :func:`fdk_mex_helix`, the GPU) with the 19
    operands of the helical gateway; 3 calls ``back_call('fdk,st,back', ...)``
    with the 13 axial operands (default :func:`fdk_mex`), as the MATLAB does.
    ``cone_par`` = 1 takes views rebinned to cone-parallel geometry; ``w3d``
    is the exponent of the w3d view weights (cone-parallel only).
    ``use_mex=0`` runs the file's MATLAB backprojector (double precision,
    per-slice view ranges) on the GPU: :func:`cbct_back_helix_mat`.
    """
    if use_mex is None:
        use_mex = 1
    if nthread is None:
        nthread = 1
    if not use_mex:
        return cbct_back_helix_mat(proj, cg, ig, ia_skip=ia_skip, offset_source=offset_source,
                                   scale_dang=scale_dang, extrapolate_t=extrapolate_t,
                                   cone_par=cone_par, w3d=w3d)
    del offset_source   # accepted as in cbct_back_helix.m; the mex path ignores it
    ia_skip = int(ia_skip)
    mask = ig.mask_or
    nx, ny = mask.shape
    nz = ig.nz
    ns, nt, na = cg.ns, cg.nt, cg.na
    ia_list = np.arange(0, na, ia_skip)
    betas = _views(cg, na)
    proj = np.asarray(proj, dtype=np.float32)
    if proj.ndim == 2:
        proj = proj[:, :, None]
    if proj.shape != (ns, nt, na):
        raise ValueError(f"proj shape {proj.shape} != [ns nt na] = {(ns, nt, na)}")
    nthread_a = np.array([nthread], dtype=np.int32)
    na1_half, dz_per_rad = helix_operands(cg, ia_skip)
    offset_t = cg.offset_t
    if extrapolate_t > 0:   # cbct_back_helix.m:114-122
        e = int(extrapolate_t)
        proj = np.concatenate([np.repeat(proj[:, :1, :], e, axis=1), proj,
                               np.repeat(proj[:, -1:, :], e, axis=1)], axis=1)
        nt = nt + 2 * e
        if offset_t != 0:
            raise ValueError("offset_t not yet done for extrapolation")
    args = (np.array([nx, ny, nz], dtype=np.int32),
            np.array([ig.dx, ig.dy, ig.dz], dtype=np.float64),
            np.array([ig.offset_x, ig.offset_y, ig.offset_z], dtype=np.float64),
            mask.astype(np.uint8), float(cg.dso), float(cg.dsd), float(cg.dfs),
            np.array([cg.ds, cg.dt], dtype=np.float64),
            np.array([cg.offset_s, offset_t], dtype=np.float64))
    if use_mex == 1:
        if back_call is None:
            back_call = fdk_mex_helix
        ptsa = np.asfortranarray(np.transpose(proj, (1, 0, 2)))      # sta -> tsa
        zs = np.asarray(cg.source_zs, np.float64)
        img = back_call("fdk,ts,back", *args, np.asfortranarray(ptsa[:, :, ia_list]),
                        np.ascontiguousarray(betas[ia_list]),
                        np.ascontiguousarray(zs[ia_list]),
                        np.array([na1_half], np.int32), np.array([int(cone_par)], np.int32),
                        np.array([int(w3d)], np.int32), np.array([dz_per_rad]),
                        np.array([cg.pitch], np.float32), nthread_a)
        img = np.asfortranarray(np.transpose(img, (1, 2, 0)))        # zxy -> xyz
    elif use_mex == 3:
        warnings.warn("fdk,st does not handle projection view edges completely",
                      stacklevel=2)
        warnings.warn("fdk,st does not handle helical geometry", stacklevel=2)
        if back_call is None:
            back_call = fdk_mex
        img = back_call("fdk,st,back", *args, np.asfortranarray(proj[:, :, ia_list]),
                        np.ascontiguousarray(betas[ia_list]), nthread_a)
    else:
        raise FdkMexError("bug")
    if scale_dang:
        warnings.warn("todo: scaling for end slices in mex version", stacklevel=2)
        sc = slice_scales(cg, ig, ia_skip, na1_half).astype(np.float32)
        img = np.asfortranarray(img * sc[None, None, :])
    return img


def helix_mat_slices(cg: CTGeom, ig: ImageGeom, ia_skip: int = 1):
    """Views and scale of each slice in the helical MATLAB backprojector
    (cbct_back_helix.m:212-236, 341-345), in double as MATLAB computes them:
    ``(views, scale)`` with ``views[iz] = (first, last)`` 0-based inclusive
    (slice iz sums views first:ia_skip:last) and its ``scale_dang`` factor.
    na1 = na / orbit * 360 must be an integer (to 1e-4), as the file checks."""
    na, orbit = cg.na, cg.orbit
    na1 = na / orbit * 360
    if abs(round(na1) - na1) > 1e-4:
        raise FdkMexError("bad na/orbit")
    na1_half = math.ceil(na1 / 2)
    wz = (ig.nz - 1) / 2 + ig.offset_z
    zc = (np.arange(ig.nz) - wz) * ig.dz
    zs = np.broadcast_to(np.asarray(cg.source_zs, np.float64), (na,))
    views = np.empty((ig.nz, 2), np.int32)
    scale = np.empty(ig.nz)
    for iz in range(ig.nz):
        if cg.pitch == 0:
            ia_min, ia_max = 1, na
        else:
            ia1_middle = int(np.argmin(np.abs(zc[iz] - zs))) + 1          # imin, 1-based
            ia_min = max(ia1_middle - na1_half, 1)
            ia_max = min(ia1_middle + (na1_half - 1), na)
        views[iz] = (ia_min - 1, ia_max - 1)
        orbit_iz = ((ia_max - ia_min + 1) / ia_skip) * abs(orbit) / (na / ia_skip)
        sc = math.pi / float(deg2rad(abs(orbit_iz)))
        scale[iz] = sc * float(deg2rad(abs(orbit))) / (na / ia_skip)
    return views, scale


def cbct_back_helix_mat(proj, cg: CTGeom, ig: ImageGeom, *, ia_skip: int = 1,
                        offset_source: float = 0.0, scale_dang: bool = True,
                        extrapolate_t=0, cone_par: int = 0, w3d: int = 0):
    """The helical MATLAB backprojector, ``cbct_back_mat`` of
    mex/src/fdk-helix/cbct_back_helix.m:185-350, on the GPU
    (``fdkhip_cbct_back_helix_mat``): double precision, slice iz sums its own
    views ia_min:ia_skip:ia_max (:helix_mat_slices), cone-parallel geometry
    when ``cone_par``, the MATLAB edge rules (zero beyond detector columns
    [1, ns], rows clamped), then the slice's ``scale_dang`` factor.  ``proj``
    is ``[ns nt na]``; float32 views stay float32 on the way in, anything
    else enters as float64.  Returns the float64 ``[nx ny nz]`` image
    (F-ordered), zero outside ``ig.mask_or``.  ``w3d`` with ``cone_par``
    is rejected: the reference's code for it calls ``Min`` and ``Max``,
    which it does not define."""
    import ctypes as C
    from . import _lib
    from .cbct_back import _ptr, geom_of
    if not extrapolate_t:
        warnings.warn("mat version always extrapolates in t by replication", stacklevel=2)
    if cone_par and w3d:
        raise FdkMexError("cone_par with w3d in the MATLAB backprojector calls Min/Max, "
                          "which the reference does not define")
    ia_skip = int(ia_skip)
    if ia_skip < 1:
        raise ValueError("ia_skip must be >= 1")
    ns, nt, na = cg.ns, cg.nt, cg.na
    proj = np.asarray(proj)
    if proj.ndim == 2:
        proj = proj[:, :, None]
    if proj.shape != (ns, nt, na):
        raise ValueError(f"proj shape {proj.shape} != [ns nt na] = {(ns, nt, na)}")
    views, scale = helix_mat_slices(cg, ig, ia_skip)
    dz_per_rad = float(cg.source_dz_per_view / (float(deg2rad(abs(cg.orbit))) / (na / ia_skip)))
    fp64 = proj.dtype != np.float32
    proj = np.asfortranarray(proj, dtype=np.float64 if fp64 else np.float32)
    betas = np.ascontiguousarray(_views(cg, na), dtype=np.float64)
    zs = np.ascontiguousarray(np.broadcast_to(np.asarray(cg.source_zs, np.float64), (na,)))
    mask = np.asfortranarray(ig.mask_or != 0).astype(np.uint8, order="F")
    img = np.empty((ig.nx, ig.ny, ig.nz), dtype=np.float64, order="F")
    views = np.ascontiguousarray(views)
    scale = np.ascontiguousarray(scale) if scale_dang else None
    lib = _lib.load()
    _lib.check(lib.fdkhip_cbct_back_helix_mat(
        C.byref(geom_of(cg, ig)), _ptr(proj), int(fp64), na, _ptr(betas), _ptr(zs), ia_skip,
        float(offset_source), _ptr(views), None if scale is None else _ptr(scale),
        int(bool(cone_par)), dz_per_rad, _ptr(mask), _ptr(img)))
    return img


__all__ = ["cbct_back_helix", "cbct_back_helix_mat", "helix_mat_slices", "helix_operands",
           "slice_scales"]
