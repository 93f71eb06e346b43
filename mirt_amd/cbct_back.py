"""Cone-beam FDK backprojection: ``cbct_back`` and the ``fdk_mex`` gateway.

Host-side mirror of the reference interface for this path:

* :func:`fdk_mex` -- the mex gateway ``fdk_mex('fdk,ts,back' | 'fdk,st,back',
  nxyz, dxyz, cxyz, mask, dso, dsd, dfs, dst, ost, proj, beta, nthread)``
  (``mex/src/fdk/fdk-orig/fdk,mex.c:22-227``): same operands, same checks and
  messages, arrays in MATLAB column-major index order; computes on the GPU
  through ``libfdkhip.so``.
* :func:`cbct_back` -- ``fbp/cbct_back.m:1-137``: builds the view angles,
  casts to single, replicates edge rows (``extrapolate_t``), permutes
  ``[ns nt na] -> [nt ns na]``, calls ``back_call`` and scales by the
  angular step.
* :class:`FdkPlan` -- device-resident backprojection of torch tensors for
  repeated use (one plan per geometry; projections FP32, FP16 or BF16 with a
  per-view scale).

There is no CPU fallback anywhere in this module.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings

import numpy as np

from . import _lib
from ._lib import FdkhipError, Geom
from .geom import CTGeom, ImageGeom, deg2rad


class FdkMexError(ValueError):
    """Argument error raised where the reference gateway calls Fail."""


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _is_int32(a):
    return isinstance(a, np.ndarray) and a.dtype == np.int32


def _is_real_double(a):
    return isinstance(a, np.ndarray) and a.dtype == np.float64


def _scalar_double(x, name):
    if isinstance(x, np.ndarray):
        if x.dtype != np.float64 or x.size != 1:
            raise FdkMexError(f"{name} must be a double scalar")
        return float(x.reshape(-1)[0])
    if isinstance(x, (float, int)) and not isinstance(x, bool):
        return float(x)
    raise FdkMexError(f"{name} must be a double scalar")


def _gateway_operands(nxyz, dxyz, cxyz, mask, dso, dsd, dfs, dst, ost, proj):
    """Checks of fdk_mex_back (fdk,mex.c:39-90) on the operands both gateways
    share, in the reference's order and with its messages."""
    if not _is_int32(nxyz):
        raise FdkMexError("mxIsInt32")
    if nxyz.size != 3:
        raise FdkMexError("mx_nxyz must have 3 elements")
    nx, ny, nz = (int(v) for v in nxyz.reshape(-1))
    if not _is_real_double(dxyz):
        raise FdkMexError("mxIsRealDouble")
    if dxyz.size != 3:
        raise FdkMexError("mx_dxyz must have 3 elements")
    if not _is_real_double(cxyz):
        raise FdkMexError("mxIsRealDouble")
    if cxyz.size != 3:
        raise FdkMexError("mx_cxyz must have 3 elements")
    if not (isinstance(mask, np.ndarray) and mask.dtype == np.uint8):
        raise FdkMexError("mxIsUint8")
    if mask.ndim != 2:
        raise FdkMexError("mask must be 2d")
    if mask.shape != (nx, ny):
        raise FdkMexError("mask size not nx by ny?")
    dso = _scalar_double(dso, "dso")
    dsd = _scalar_double(dsd, "dsd")
    dfs = _scalar_double(dfs, "dfs")
    if not _is_real_double(dst):
        raise FdkMexError("mxIsRealDouble")
    if dst.size != 2:
        raise FdkMexError("mx_dst must have 2 elements")
    if not _is_real_double(ost):
        raise FdkMexError("mxIsRealDouble")
    if ost.size != 2:
        raise FdkMexError("mx_ost must have 2 elements")
    if not (isinstance(proj, np.ndarray) and proj.dtype == np.float32):
        raise FdkMexError("mxIsRealSingle")
    return (nx, ny, nz), dso, dsd, dfs


def _per_view(x, proj, na, name):
    """beta / source_zs: a double scalar for a 2-D proj, else [na] doubles."""
    if proj.ndim == 2:
        return np.asarray([_scalar_double(x, name)], dtype=np.float64)
    if not _is_real_double(x):
        raise FdkMexError("mxIsRealDouble")
    if name == "beta" and x.size != na:
        raise FdkMexError(f"beta length={x.size} vs na=size(proj,3)={na}")
    return x


def _scalar_int32(x):
    if not (isinstance(x, np.ndarray) and x.dtype == np.int32 and x.size == 1):
        raise FdkMexError("mxIsScalarInt32")
    return int(x.reshape(-1)[0])


def _c_operands(nxyz, dxyz, cxyz, mask, dst, ost, proj, beta):
    return (np.ascontiguousarray(nxyz.reshape(-1), dtype=np.int32),
            np.ascontiguousarray(dxyz.reshape(-1)), np.ascontiguousarray(cxyz.reshape(-1)),
            np.asfortranarray(mask), np.ascontiguousarray(dst.reshape(-1)),
            np.ascontiguousarray(ost.reshape(-1)), np.asfortranarray(proj),
            np.ascontiguousarray(np.asarray(beta, np.float64).reshape(-1)))


def fdk_mex(cmd, nxyz, dxyz, cxyz, mask, dso, dsd, dfs, dst, ost, proj, beta,
            nthread, *extra):
    """Mirror of the mex gateway ``fdk_mex`` (fdk,mex.c:175-214, 22-171).

    ``cmd`` is ``'fdk,ts,back'`` (proj ``[nt ns na]``, result ``[nz nx ny]``)
    or ``'fdk,st,back'`` (proj ``[ns nt na]``, result ``[nx ny nz]``).
    ``nxyz`` int32[3], ``dxyz``/``cxyz`` float64[3], ``mask`` uint8 ``[nx ny]``,
    ``dst``/``ost`` float64[2], ``proj`` float32 2-D or 3-D, ``beta`` float64
    (scalar for 2-D proj), ``nthread`` int32 scalar.  Returns a Fortran-order
    float32 array.
    """
    if not isinstance(cmd, str):
        raise FdkMexError("usage error. see above")
    if extra or cmd not in ("fdk,ts,back", "fdk,st,back"):
        if cmd not in ("fdk,ts,back", "fdk,st,back"):
            raise FdkMexError(f"command '{cmd}' unknown")
        raise FdkMexError("usage error. see above")   # nrhs != 13
    is_st = cmd.startswith("fdk,st")
    (nx, ny, nz), dso, dsd, dfs = _gateway_operands(nxyz, dxyz, cxyz, mask, dso, dsd, dfs,
                                                    dst, ost, proj)
    if proj.ndim not in (2, 3):
        raise FdkMexError("mx_proj must be 2d or 3d")
    na = 1 if proj.ndim == 2 else proj.shape[2]
    beta = _per_view(beta, proj, na, "beta")
    nthread = _scalar_int32(nthread)

    if is_st:
        ns, nt = proj.shape[0], proj.shape[1]
        out = np.zeros((nx, ny, nz), dtype=np.float32, order="F")
    else:
        nt, ns = proj.shape[0], proj.shape[1]
        out = np.zeros((nz, nx, ny), dtype=np.float32, order="F")

    lib = _lib.load()
    nxyz_c, dxyz_c, cxyz_c, maskf, dst_c, ost_c, projf, beta_c = _c_operands(
        nxyz, dxyz, cxyz, mask, dst, ost, proj, beta)
    fn = lib.fdkhip_fdk_st_back if is_st else lib.fdkhip_fdk_ts_back
    rc = fn(_ptr(nxyz_c), _ptr(dxyz_c), _ptr(cxyz_c), _ptr(maskf), dso, dsd, dfs,
            _ptr(dst_c), _ptr(ost_c), _ptr(projf), ns, nt, na, _ptr(beta_c),
            nthread, _ptr(out))
    _lib.check(rc)
    return out


def fdk_mex_helix(cmd, nxyz, dxyz, cxyz, mask, dso, dsd, dfs, dst, ost, proj, beta,
                  source_zs, na1_half, cone_par, w3d, source_dz_per_rad, pitch, nthread,
                  *extra):
    """Mirror of the helical mex gateway (mex/src/fdk-helix/mex/fdk,mex.c:22-236):
    ``fdk_mex('fdk,ts,back', ..., proj, beta, source_zs, na1_half, cone_par,
    w3d, source_dz_per_rad, pitch, nthread)`` -- 19 operands.  ``source_zs``
    float64 like ``beta``; ``na1_half``, ``cone_par``, ``w3d`` int32 scalars,
    ``source_dz_per_rad`` float64, ``pitch`` float32 (single).  Returns the
    ``[nz nx ny]`` float32 image (Fortran order).  The ``'fdk,st,back'``
    order of this gateway ignores the helical operands (fdk,mex.c:150-155);
    use :func:`fdk_mex` for it.
    """
    if not isinstance(cmd, str):
        raise FdkMexError("usage error. see above")
    if cmd not in ("fdk,ts,back", "fdk,st,back"):
        raise FdkMexError(f"command '{cmd}' unknown")
    if extra:
        raise FdkMexError("usage error. see above")   # nrhs != 19
    if cmd == "fdk,st,back":
        raise FdkMexError("'fdk,st,back' does not handle helical geometry; use fdk_mex")
    (nx, ny, nz), dso, dsd, dfs = _gateway_operands(nxyz, dxyz, cxyz, mask, dso, dsd, dfs,
                                                    dst, ost, proj)
    if proj.ndim not in (2, 3):
        raise FdkMexError("mx_proj must be 2d or 3d")
    na = 1 if proj.ndim == 2 else proj.shape[2]
    beta = _per_view(beta, proj, na, "beta")
    source_zs = _per_view(source_zs, proj, na, "source_zs")
    if source_zs.size < na:
        raise FdkMexError(f"source_zs length={source_zs.size} vs na={na}")
    nthread = _scalar_int32(nthread)
    na1_half = _scalar_int32(na1_half)
    cone_par = _scalar_int32(cone_par)
    w3d = _scalar_int32(w3d)
    if not _is_real_double(source_dz_per_rad):
        raise FdkMexError("mxIsRealDouble")
    if not (isinstance(pitch, np.ndarray) and pitch.dtype == np.float32):
        raise FdkMexError("mxIsRealSingle")
    nt, ns = proj.shape[0], proj.shape[1]
    out = np.zeros((nz, nx, ny), dtype=np.float32, order="F")
    lib = _lib.load()
    nxyz_c, dxyz_c, cxyz_c, maskf, dst_c, ost_c, projf, beta_c = _c_operands(
        nxyz, dxyz, cxyz, mask, dst, ost, proj, beta)
    zs_c = np.ascontiguousarray(source_zs.reshape(-1)[:na], dtype=np.float64)
    rc = lib.fdkhip_fdk_ts_back_helix(
        _ptr(nxyz_c), _ptr(dxyz_c), _ptr(cxyz_c), _ptr(maskf), dso, dsd, dfs,
        _ptr(dst_c), _ptr(ost_c), _ptr(projf), ns, nt, na, _ptr(beta_c), _ptr(zs_c),
        na1_half, cone_par, w3d, float(np.asarray(source_dz_per_rad).reshape(-1)[0]),
        float(pitch.reshape(-1)[0]), nthread, _ptr(out))
    _lib.check(rc)
    return out


def _views(cg: CTGeom, na: int):
    """betas of cbct_back.m:82: deg2rad(orbit_start + orbit * [0:na-1] / na)."""
    k = np.arange(na, dtype=np.float64)
    return deg2rad(np.asarray(cg.orbit_start, dtype=np.float64) + cg.orbit * k / na)


def angle_scale64(cg: CTGeom, ia_skip: int = 1) -> float:
    """The "\\der angle" scale 0.5 * deg2rad(|orbit|) / (na/ia_skip) in double
    (cbct_back.m:133-135 and 264-266)."""
    return 0.5 * float(deg2rad(abs(cg.orbit))) / (cg.na / ia_skip)


def angle_scale(cg: CTGeom, ia_skip: int = 1) -> np.float32:
    """cbct_back.m:133-135: the scale as it multiplies the single image of
    the mex route, i.e. narrowed to single."""
    return np.float32(angle_scale64(cg, ia_skip))


def cbct_back(proj, cg: CTGeom, ig: ImageGeom, *, use_mex=None, ia_skip: int = 1,
              offset_source: float = 0.0, nthread=None, back_call=None,
              scale_dang: bool = True, extrapolate_t: int = 0):
    """Mirror of ``cbct_back`` (fbp/cbct_back.m:1-137).

    ``proj`` is ``[ns nt na]`` (numpy, MATLAB index order); returns the
    backprojection ``[nx ny nz]`` as a float32 numpy array.  ``use_mex`` 1
    (default) calls ``back_call`` once with all views, 2 once per view and
    sums, 3 uses the ``'fdk,st,back'`` order; ``back_call`` defaults to
    :func:`fdk_mex` (the GPU gateway).  ``use_mex=0`` is the reference's MATLAB
    backprojector ``cbct_back_mat`` (cbct_back.m:142-270): double precision,
    its own edge rules, helical ``source_zs`` accepted; it returns a float64
    ``[nx ny nz]`` array and runs on the GPU through ``fdkhip_cbct_back_mat``.
    """
    if use_mex is None:
        use_mex = 1
    if nthread is None:
        nthread = 1
    device_path = back_call is None and use_mex == 1 and 1 <= int(nthread) <= 250
    if back_call is None:
        back_call = fdk_mex
    if not use_mex:
        return cbct_back_mat(proj, cg, ig, ia_skip=ia_skip, offset_source=offset_source,
                             scale_dang=scale_dang, extrapolate_t=extrapolate_t)
    if cg.pitch != 0 or np.any(cg.source_zs != 0):
        raise ValueError("helix not yet supported")
    del offset_source   # accepted as in cbct_back.m; the mex path ignores it

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
    if device_path:
        # permutes, edge-row replication and the angle scale on the device
        # (fdkhip_cbct_back); same arithmetic as the host route below
        return _cbct_back_device(proj, cg, ig, betas, ia_skip, int(extrapolate_t),
                                 scale_dang)
    nthread = np.array([nthread], dtype=np.int32)
    offset_t = cg.offset_t
    if extrapolate_t > 0:
        e = int(extrapolate_t)
        top = np.repeat(proj[:, :1, :], e, axis=1)
        bot = np.repeat(proj[:, -1:, :], e, axis=1)
        proj = np.concatenate([top, proj, bot], axis=1)
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
        ptsa = np.asfortranarray(np.transpose(proj, (1, 0, 2)))   # sta -> tsa
        img = back_call("fdk,ts,back", *args, np.asfortranarray(ptsa[:, :, ia_list]),
                        np.ascontiguousarray(betas[ia_list]), nthread)
        img = np.asfortranarray(np.transpose(img, (1, 2, 0)))    # zxy -> xyz
    elif use_mex == 2:
        ptsa = np.asfortranarray(np.transpose(proj, (1, 0, 2)))
        img = np.float32(0)
        for ia in ia_list:
            tmp = back_call("fdk,ts,back", *args, np.asfortranarray(ptsa[:, :, ia]),
                            float(betas[ia]), nthread)
            img = img + tmp
        img = np.asfortranarray(np.transpose(img, (1, 2, 0)))
    elif use_mex == 3:
        img = back_call("fdk,st,back", *args, np.asfortranarray(proj[:, :, ia_list]),
                        np.ascontiguousarray(betas[ia_list]), nthread)
    else:
        raise ValueError("bug")
    if scale_dang:
        img = angle_scale(cg, ia_skip) * img
    return img


def _cbct_back_device(proj, cg, ig, betas, ia_skip, extrapolate_t, scale_dang):
    """cbct_back's use_mex = 1 route through fdkhip_cbct_back: ``proj``
    [ns nt na] float32 in, [nx ny nz] float32 out."""
    lib = _lib.load()
    if extrapolate_t > 0 and cg.offset_t != 0:
        raise ValueError("offset_t not yet done for extrapolation")
    geom = geom_of(cg, ig)                       # detector rows before replication
    proj = np.asfortranarray(proj, dtype=np.float32)
    betas = np.ascontiguousarray(betas, dtype=np.float64)
    mask = np.asfortranarray(ig.mask_or != 0).astype(np.uint8, order="F")
    img = np.empty((ig.nx, ig.ny, ig.nz), dtype=np.float32, order="F")
    scale = float(angle_scale(cg, ia_skip)) if scale_dang else 1.0
    _lib.check(lib.fdkhip_cbct_back(C.byref(geom), _ptr(proj), int(cg.na), _ptr(betas),
                                    int(ia_skip), _ptr(mask), int(extrapolate_t), scale,
                                    _ptr(img)))
    return img


def cbct_back_mat(proj, cg: CTGeom, ig: ImageGeom, *, ia_skip: int = 1,
                  offset_source: float = 0.0, scale_dang: bool = True,
                  extrapolate_t=0):
    """``cbct_back_mat`` (fbp/cbct_back.m:142-270) on the GPU.

    ``proj`` is ``[ns nt na]`` (or ``[ns nt]`` when na = 1); float32 views stay
    float32 on the way in, anything else enters as float64 (MATLAB's
    ``proj1(1:ns,:) = proj(:,:,ia)`` converts to double).  Views 1:ia_skip:na
    at the angles of cbct_back.m:157 and the heights ``cg.source_zs``; returns
    the float64 ``[nx ny nz]`` image (F-ordered), zero outside ``ig.mask_or``.
    """
    if not extrapolate_t:
        warnings.warn("mat version always extrapolates in t by replication", stacklevel=2)
    zs = np.ascontiguousarray(np.broadcast_to(np.asarray(cg.source_zs, np.float64),
                                              (cg.na,)))
    if np.any(zs != 0):
        warnings.warn("helix not yet tested", stacklevel=2)
    ia_skip = int(ia_skip)
    if ia_skip < 1:
        raise ValueError("ia_skip must be >= 1")
    ns, nt, na = cg.ns, cg.nt, cg.na
    proj = np.asarray(proj)
    if proj.ndim == 2:
        proj = proj[:, :, None]
    if proj.shape != (ns, nt, na):
        raise ValueError(f"proj shape {proj.shape} != [ns nt na] = {(ns, nt, na)}")
    fp64 = proj.dtype != np.float32
    proj = np.asfortranarray(proj, dtype=np.float64 if fp64 else np.float32)
    betas = np.ascontiguousarray(_views(cg, na), dtype=np.float64)
    mask = np.asfortranarray(ig.mask_or != 0).astype(np.uint8, order="F")
    img = np.empty((ig.nx, ig.ny, ig.nz), dtype=np.float64, order="F")
    scale = angle_scale64(cg, ia_skip) if scale_dang else 1.0
    lib = _lib.load()
    _lib.check(lib.fdkhip_cbct_back_mat(C.byref(geom_of(cg, ig)), _ptr(proj), int(fp64), na,
                                        _ptr(betas), _ptr(zs), ia_skip, float(offset_source),
                                        _ptr(mask), float(scale), _ptr(img)))
    return img


def geom_of(cg: CTGeom, ig: ImageGeom, extrapolate_t: int = 0) -> Geom:
    """fdkhip_geom of a (ct_geom, image_geom) pair as cbct_back passes it."""
    return Geom(ig.nx, ig.ny, ig.nz, ig.dx, ig.dy, ig.dz,
                ig.offset_x, ig.offset_y, ig.offset_z,
                float(cg.dso), float(cg.dsd), float(cg.dfs),
                cg.ns, cg.nt + 2 * int(extrapolate_t), float(cg.ds), float(cg.dt),
                float(cg.offset_s), float(cg.offset_t))


class FdkPlan:
    """Device-resident 'fdk,ts,back' for one geometry (libfdkhip plan).

    ``proj`` passed to :meth:`__call__` is a CUDA (HIP) torch tensor holding
    ``[nt ns na]`` with t fastest, i.e. a contiguous tensor of shape
    ``(na, ns, nt)``, in the plan's storage format (``float32``, ``float16``
    or ``bfloat16`` with a ``(na,)`` float32 ``scale``).  The result is the
    ``[nz nx ny]`` image as a contiguous ``(ny, nx, nz)`` float32 tensor;
    ``image_xyz`` views it as ``[nx ny nz]`` without copying.
    """

    def __init__(self, geom: Geom, beta, mask=None, *, fmt="f32", math_mode="exact",
                 helix=None):
        self.lib = _lib.load()
        self.geom = geom
        self.fmt = _lib.FMT_NAMES[fmt] if isinstance(fmt, str) else int(fmt)
        self.math = {"exact": _lib.MATH_EXACT, "fast": _lib.MATH_FAST}[math_mode] \
            if isinstance(math_mode, str) else int(math_mode)
        beta = np.ascontiguousarray(np.atleast_1d(beta), dtype=np.float64)
        self.na = int(beta.size)
        self._mask = None
        if mask is not None:
            m = np.asarray(mask)
            if m.shape != (geom.nx, geom.ny):
                raise ValueError("mask must be [nx ny]")
            self._mask = np.asfortranarray(m != 0).astype(np.uint8, order="F")
        h = C.c_void_p()
        if helix is None:
            rc = self.lib.fdkhip_plan_create(C.byref(h), C.byref(geom), self.na,
                                             _ptr(beta), _ptr(self._mask), self.fmt, self.math)
        else:
            zs, na1_half, pitch = helix
            self._zs = np.ascontiguousarray(np.broadcast_to(np.asarray(zs, np.float64),
                                                            (self.na,)))
            rc = self.lib.fdkhip_plan_create_helix(
                C.byref(h), C.byref(geom), self.na, _ptr(beta), _ptr(self._zs),
                int(na1_half), float(pitch), _ptr(self._mask), self.fmt, self.math)
        _lib.check(rc)
        self.handle = h

    @classmethod
    def from_geoms(cls, cg: CTGeom, ig: ImageGeom, *, ia_skip=1, extrapolate_t=0, **kw):
        """Plan for views 1:ia_skip:na of ``cg``; ``extrapolate_t`` rows added
        above and below (the views it takes then have nt + 2 extrapolate_t
        rows)."""
        beta = _views(cg, cg.na)[::ia_skip]
        return cls(geom_of(cg, ig, extrapolate_t), beta, ig.mask_or, **kw)

    @classmethod
    def helix_from_geoms(cls, cg: CTGeom, ig: ImageGeom, *, ia_skip=1, **kw):
        """Plan for the helical scan ``cg`` in native helical geometry (views
        1:ia_skip:na at their source heights, the slice windows of
        cbct_back_helix's na1_half; fdkhip_plan_create_helix)."""
        from .cbct_back_helix import helix_operands
        na1_half, _ = helix_operands(cg, ia_skip)
        beta = _views(cg, cg.na)[::ia_skip]
        zs = np.broadcast_to(np.asarray(cg.source_zs, np.float64), (cg.na,))[::ia_skip]
        return cls(geom_of(cg, ig), beta, ig.mask_or,
                   helix=(zs, na1_half, float(cg.pitch)), **kw)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.fdkhip_plan_destroy(h)
            self.handle = None

    @property
    def config(self):
        cfg = (C.c_int32 * 6)()
        _lib.check(self.lib.fdkhip_plan_config(self.handle, cfg))
        keys = ("waves", "zper", "vstage", "z0", "zgroups", "workgroups")
        return dict(zip(keys, list(cfg)))

    def set_config(self, waves, zper, vstage=0, slot_cap=0):
        _lib.check(self.lib.fdkhip_plan_set_config(self.handle, waves, zper, vstage,
                                                   slot_cap))

    @property
    def occupancy(self) -> int:
        """Minimum waves per SIMD of the kernel build the plan launches (4, or
        8 for one-chunk FP16/BF16 shapes when chosen)."""
        w = C.c_int32()
        _lib.check(self.lib.fdkhip_plan_occupancy(self.handle, C.byref(w)))
        return int(w.value)

    def set_occupancy(self, waves_per_simd):
        """Choose the 4- or 8-waves-per-SIMD kernel build
        (fdkhip_plan_set_occupancy)."""
        _lib.check(self.lib.fdkhip_plan_set_occupancy(self.handle, int(waves_per_simd)))

    def count_updates(self) -> int:
        n = C.c_int64()
        _lib.check(self.lib.fdkhip_plan_count_updates(self.handle, C.byref(n)))
        return int(n.value)

    def _check_proj(self, proj, scale):
        import torch
        g = self.geom
        want = {_lib.F32: torch.float32, _lib.F16: torch.float16,
                _lib.BF16: torch.bfloat16}[self.fmt]
        if proj.dtype != want or not proj.is_cuda or not proj.is_contiguous():
            raise ValueError(f"proj must be a contiguous device tensor of dtype {want}")
        if proj.numel() != self.na * g.ns * g.nt:
            raise ValueError("proj must hold [nt ns na]")
        if self.fmt != _lib.F32:
            if scale is None or scale.dtype != torch.float32 or scale.numel() != self.na \
                    or not scale.is_cuda:
                raise ValueError("F16/BF16 projections need a (na,) float32 device scale")

    def __call__(self, proj, scale=None, out=None, *, stream=None, graph=True):
        import torch
        self._check_proj(proj, scale)
        g = self.geom
        if out is None:
            out = torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=proj.device)
        elif out.dtype != torch.float32 or not out.is_contiguous() \
                or out.numel() != g.nx * g.ny * g.nz:
            raise ValueError("out must be a contiguous float32 [nz nx ny] tensor")
        s = stream if stream is not None else torch.cuda.current_stream(proj.device)
        fn = self.lib.fdkhip_plan_execute if graph else self.lib.fdkhip_plan_execute_direct
        rc = fn(self.handle, C.c_void_p(proj.data_ptr()),
                C.c_void_p(scale.data_ptr()) if scale is not None else None,
                C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream))
        _lib.check(rc)
        return out

    def autotune(self, proj, scale=None, out=None, *, reps=3, stream=None):
        """Time the supported launch shapes, stage depths and kernel builds on
        these buffers (the backprojection kernel over the first 128 views,
        then the four fastest over all views) and keep the fastest
        (fdkhip_plan_autotune); returns its average milliseconds over the
        views of the last round.  ``out`` (or a scratch image) is
        overwritten."""
        import torch
        self._check_proj(proj, scale)
        g = self.geom
        if out is None:
            out = torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=proj.device)
        s = stream if stream is not None else torch.cuda.current_stream(proj.device)
        best = C.c_float()
        _lib.check(self.lib.fdkhip_plan_autotune(
            self.handle, C.c_void_p(proj.data_ptr()),
            C.c_void_p(scale.data_ptr()) if scale is not None else None,
            C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream), int(reps),
            C.byref(best)))
        return float(best.value)

    def profile(self, proj, scale=None, out=None, *, reps=5, stream=None):
        """Average device milliseconds of the launch steps over ``reps`` direct
        executions (fdkhip_plan_profile): ``{"fill", "pad", "back"}``."""
        import torch
        self._check_proj(proj, scale)
        g = self.geom
        if out is None:
            out = torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=proj.device)
        s = stream if stream is not None else torch.cuda.current_stream(proj.device)
        ms = (C.c_float * 3)()
        _lib.check(self.lib.fdkhip_plan_profile(
            self.handle, C.c_void_p(proj.data_ptr()),
            C.c_void_p(scale.data_ptr()) if scale is not None else None,
            C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream), int(reps), ms))
        return {"fill": ms[0], "pad": ms[1], "back": ms[2]}

    @staticmethod
    def image_xyz(img):
        """[nz nx ny] (tensor (ny, nx, nz)) viewed as [nx ny nz]."""
        return img.permute(1, 0, 2)


class FdkBatch:
    """Many scans backprojected by one persistent launch (libfdkhip batch).

    ``scans`` is a sequence of ``(geom, beta, mask)`` (mask may be None).
    :meth:`__call__` takes per-scan device projections (and scales for
    F16/BF16) laid out as for :class:`FdkPlan` and returns per-scan
    ``(ny, nx, nz)`` float32 images equal to what one :class:`FdkPlan` per
    scan gives.
    """

    def __init__(self, scans, *, fmt="f32", math_mode="exact"):
        self.lib = _lib.load()
        self.fmt = _lib.FMT_NAMES[fmt] if isinstance(fmt, str) else int(fmt)
        self.math = {"exact": _lib.MATH_EXACT, "fast": _lib.MATH_FAST}[math_mode] \
            if isinstance(math_mode, str) else int(math_mode)
        scans = list(scans)
        n = len(scans)
        if n == 0:
            raise ValueError("a batch needs at least one scan")
        self.geoms = [sc[0] for sc in scans]
        garr = (Geom * n)(*self.geoms)
        self._betas = [np.ascontiguousarray(np.atleast_1d(sc[1]), dtype=np.float64)
                       for sc in scans]
        self.na = [int(b.size) for b in self._betas]
        self._masks = []
        for (g, _, m) in scans:
            if m is None:
                self._masks.append(None)
                continue
            m = np.asarray(m)
            if m.shape != (g.nx, g.ny):
                raise ValueError("mask must be [nx ny]")
            self._masks.append(np.asfortranarray(m != 0).astype(np.uint8, order="F"))
        na_arr = (C.c_int32 * n)(*self.na)
        beta_arr = (C.c_void_p * n)(*[b.ctypes.data for b in self._betas])
        mask_arr = (C.c_void_p * n)(*[None if m is None else m.ctypes.data
                                      for m in self._masks])
        h = C.c_void_p()
        _lib.check(self.lib.fdkhip_batch_create(C.byref(h), n, garr, na_arr, beta_arr,
                                                mask_arr, self.fmt, self.math))
        self.handle = h
        self.nscan = n

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.fdkhip_batch_destroy(h)
            self.handle = None

    @property
    def config(self):
        """Launch shape of the persistent kernel; with one launch per scan,
        waves = zper = vstage = slot = 0 and workgroups = items = the
        workgroups of all the scans' launches."""
        cfg = (C.c_int32 * 6)()
        _lib.check(self.lib.fdkhip_batch_config(self.handle, cfg))
        keys = ("waves", "zper", "vstage", "slot", "workgroups", "items")
        return dict(zip(keys, list(cfg)))

    def set_config(self, waves, zper, vstage=0):
        """Persistent kernel in this launch shape (fdkhip_batch_set_config)."""
        _lib.check(self.lib.fdkhip_batch_set_config(self.handle, waves, zper, vstage))

    def set_per_scan(self):
        """One launch per scan in turn, each in its plan's own launch shape
        (fdkhip_batch_set_per_scan); ``set_config`` returns to the persistent
        kernel."""
        _lib.check(self.lib.fdkhip_batch_set_per_scan(self.handle))

    @property
    def per_scan(self) -> bool:
        """True when the batch runs one launch per scan."""
        return self.config["waves"] == 0

    @property
    def scan_zper(self):
        """Each scan's slices per lane (the launch's ``zper`` or half of it)."""
        z = (C.c_int32 * self.nscan)()
        _lib.check(self.lib.fdkhip_batch_scan_zper(self.handle, z))
        return list(z)

    def set_zper(self, zper=None):
        """Fix each scan's slices per lane (``None``: the automatic choice,
        half height for scans whose footprint overflows the LDS slot)."""
        if zper is not None and len(zper) != self.nscan:
            raise ValueError(f"expected {self.nscan} per-scan values")
        arr = None if zper is None else (C.c_int32 * self.nscan)(*[int(z) for z in zper])
        _lib.check(self.lib.fdkhip_batch_set_zper(self.handle, arr))

    def count_updates(self) -> int:
        n = C.c_int64()
        _lib.check(self.lib.fdkhip_batch_count_updates(self.handle, C.byref(n)))
        return int(n.value)

    def _arrays(self, projs, scales, outs):
        import torch
        if len(projs) != self.nscan:
            raise ValueError(f"expected {self.nscan} projection tensors")
        want = {_lib.F32: torch.float32, _lib.F16: torch.float16,
                _lib.BF16: torch.bfloat16}[self.fmt]
        if outs is None:
            outs = [torch.empty((g.ny, g.nx, g.nz), dtype=torch.float32, device=p.device)
                    for g, p in zip(self.geoms, projs)]
        for i, (g, p, o) in enumerate(zip(self.geoms, projs, outs)):
            if p.dtype != want or not p.is_cuda or not p.is_contiguous() \
                    or p.numel() != self.na[i] * g.ns * g.nt:
                raise ValueError(f"scan {i}: proj must be a contiguous {want} [nt ns na]")
            if o.dtype != torch.float32 or not o.is_contiguous() \
                    or o.numel() != g.nx * g.ny * g.nz:
                raise ValueError(f"scan {i}: out must be a contiguous float32 [nz nx ny]")
        n = self.nscan
        pa = (C.c_void_p * n)(*[p.data_ptr() for p in projs])
        oa = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
        sa = None
        if self.fmt != _lib.F32:
            if scales is None or len(scales) != n:
                raise ValueError("F16/BF16 batches need one (na,) float32 scale per scan")
            for i, sc in enumerate(scales):
                if sc.dtype != torch.float32 or not sc.is_cuda or sc.numel() != self.na[i]:
                    raise ValueError(f"scan {i}: scale must be a (na,) float32 device tensor")
            sa = (C.c_void_p * n)(*[sc.data_ptr() for sc in scales])
        return outs, pa, sa, oa

    def __call__(self, projs, scales=None, outs=None, *, stream=None, graph=True):
        import torch
        outs, pa, sa, oa = self._arrays(projs, scales, outs)
        s = stream if stream is not None else torch.cuda.current_stream(projs[0].device)
        fn = self.lib.fdkhip_batch_execute if graph else self.lib.fdkhip_batch_execute_direct
        _lib.check(fn(self.handle, pa, sa, oa, C.c_void_p(s.cuda_stream)))
        return outs

    def profile(self, projs, scales=None, outs=None, *, reps=5, stream=None):
        """Average device milliseconds (fdkhip_batch_profile):
        ``{"prep": fills + padded copies, "back": persistent kernel}``."""
        import torch
        outs, pa, sa, oa = self._arrays(projs, scales, outs)
        s = stream if stream is not None else torch.cuda.current_stream(projs[0].device)
        ms = (C.c_float * 2)()
        _lib.check(self.lib.fdkhip_batch_profile(self.handle, pa, sa, oa,
                                                 C.c_void_p(s.cuda_stream), int(reps), ms))
        return {"prep": ms[0], "back": ms[1]}

    def autotune(self, projs, scales=None, outs=None, *, reps=3, stream=None):
        """fdkhip_batch_autotune on these buffers (the persistent kernel's
        shapes, then one launch per scan with every plan tuned alone, each
        timed over the first 128 views of every scan); keeps the fastest and
        returns its average ms.  ``outs`` are overwritten."""
        import torch
        outs, pa, sa, oa = self._arrays(projs, scales, outs)
        s = stream if stream is not None else torch.cuda.current_stream(projs[0].device)
        best = C.c_float()
        _lib.check(self.lib.fdkhip_batch_autotune(self.handle, pa, sa, oa,
                                                  C.c_void_p(s.cuda_stream), int(reps),
                                                  C.byref(best)))
        return float(best.value)


def quantize_views(proj, fmt="f16", *, stream=None):
    """Per-view FP16/BF16 storage of ``[nt ns na]`` FP32 device projections.

    Returns ``(q, scale)``: ``q`` has ``proj``'s shape in the chosen format,
    ``scale`` is ``(na,)`` float32 with ``proj ~= q * scale`` per view
    (scale a power of two, see fdkhip_quantize_views)."""
    import torch
    lib = _lib.load()
    f = _lib.FMT_NAMES[fmt] if isinstance(fmt, str) else int(fmt)
    if proj.dtype != torch.float32 or not proj.is_cuda or not proj.is_contiguous():
        raise ValueError("proj must be a contiguous float32 device tensor (na, ns, nt)")
    na, ns, nt = proj.shape
    dt = torch.float16 if f == _lib.F16 else torch.bfloat16
    q = torch.empty(proj.shape, dtype=dt, device=proj.device)
    scale = torch.empty((na,), dtype=torch.float32, device=proj.device)
    s = stream if stream is not None else torch.cuda.current_stream(proj.device)
    rc = lib.fdkhip_quantize_views(C.c_void_p(proj.data_ptr()), ns, nt, na, f,
                                   C.c_void_p(q.data_ptr()), C.c_void_p(scale.data_ptr()),
                                   C.c_void_p(s.cuda_stream))
    _lib.check(rc)
    return q, scale


__all__ = ["fdk_mex", "fdk_mex_helix", "cbct_back", "cbct_back_mat", "FdkPlan", "FdkBatch", "FdkMexError",
           "FdkhipError", "quantize_views", "geom_of", "angle_scale", "angle_scale64"]
