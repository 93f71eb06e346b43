"""CPU checker for the FDK backprojector -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  It wraps two libraries built by
``oracle/Makefile``:

* ``oracle/lib/libfdk_oracle.so`` -- this directory's clean-room restatement
  (``fdk_oracle.c``): FP32 with the reference's operation order, FP64 error
  reference, the ``[ns nt]``-order variant, and the helical / cone-parallel
  path of the fdk-helix kernel (``back_ts_helix_f32``);
* ``oracle/_ref/libfdkref_{fixed,shipped}.so`` -- the reference C compiled from
  ``mex/src/fdk-helix/mex`` (debugged vertical weights, axial path) and
  ``mex/src/fdk/fdk-orig`` (as shipped), driven by ``ref_driver.c`` the way
  ``fdk_mex_back`` (``mex/src/fdk/fdk-orig/fdk,mex.c:22-171``) drives them.

and one numpy function of its own:

* :func:`back_helix_mat_f64` and :func:`helix_mat_slices` -- the helical
  ``cbct_back_mat`` of ``mex/src/fdk-helix/cbct_back_helix.m:185-350``
  (per-slice view ranges and scales, cone-parallel geometry; w3d excluded,
  it calls functions the reference does not define).  Pinned in
  ``tests/test_oracle_mat.py``: for pitch 0 it is :func:`back_mat_f64`
  with the axial scale; each slice equals :func:`back_mat_f64` over its own
  views; and linear views recover the sample positions that the
  cone-parallel ray geometry gives, computed from first principles.
* :func:`back_mat_f64` -- restatement of ``cbct_back_mat``
  (``fbp/cbct_back.m:142-270``), the reference's MATLAB backprojector, step
  by step and vectorised over the masked pixels as the MATLAB code is, in
  double precision with its edge rules.  Pinned in ``tests/test_oracle.py``
  against the reference C (views with zero edge columns and replicated edge
  rows, where the two backprojectors coincide -- the comparison the
  reference's own ``cbct_back_test`` makes, ``cbct_back.m:335-364``), against
  the reference C's one-bin fall-off band at the detector's s edges (which
  ``cbct_back_mat`` must zero), and against closed forms for linear views on
  the parallel geometry (the 1-based ``ws``/``wt``, the t clamp, the hard s
  cut).  ``offset_source`` and ``source_zs`` are pinned through shifts of the
  image offsets that the reference C computes.

Array conventions of the compiled functions follow the reference mex call
``fdk_mex('fdk,ts,back', nxyz, dxyz, cxyz, mask, dso, dsd, dfs, [ds dt],
[offset_s offset_t], proj, beta, nthread)``: ``proj`` is ``[nt ns na]`` with t
fastest (a C-order numpy array of shape ``(na, ns, nt)``), the result is
``[nz nx ny]`` with z fastest (C-order shape ``(ny, nx, nz)``), ``mask`` is
``[nx ny]`` with x fastest (C-order shape ``(ny, nx)``).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "lib", "libfdk_oracle.so")
LIB_REF = {
    "fixed": os.path.join(HERE, "_ref", "libfdkref_fixed.so"),
    "shipped": os.path.join(HERE, "_ref", "libfdkref_shipped.so"),
}


class OrcGeom(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("offset_x", C.c_double), ("offset_y", C.c_double), ("offset_z", C.c_double),
        ("dso", C.c_double), ("dsd", C.c_double), ("dfs", C.c_double),
        ("ns", C.c_int32), ("nt", C.c_int32),
        ("ds", C.c_double), ("dt", C.c_double),
        ("offset_s", C.c_double), ("offset_t", C.c_double),
    ]


def build(reference: bool = True) -> None:
    """Compile the restatement (and, if /root/reference exists, the reference)."""
    targets = ["oracle"]
    if reference and os.path.isdir("/root/reference/mex/src"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


_libs: dict = {}


def _load(path: str) -> C.CDLL:
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        _libs[path] = C.CDLL(path)
    return _libs[path]


def have_reference(which: str = "fixed") -> bool:
    return os.path.exists(LIB_REF[which])


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def geom(nxyz, dxyz, cxyz, dso, dsd, dfs, ns, nt, dst, ost) -> OrcGeom:
    return OrcGeom(int(nxyz[0]), int(nxyz[1]), int(nxyz[2]),
                   float(dxyz[0]), float(dxyz[1]), float(dxyz[2]),
                   float(cxyz[0]), float(cxyz[1]), float(cxyz[2]),
                   float(dso), float(dsd), float(dfs), int(ns), int(nt),
                   float(dst[0]), float(dst[1]), float(ost[0]), float(ost[1]))


def _prep(g: OrcGeom, proj, beta, mask):
    proj = _f32(proj)
    na = proj.size // (g.ns * g.nt)
    if proj.size != na * g.ns * g.nt:
        raise ValueError("proj size is not a multiple of ns*nt")
    beta = np.ascontiguousarray(np.atleast_1d(beta), dtype=np.float64)
    if beta.size != na:
        raise ValueError(f"beta length={beta.size} vs na={na}")
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        if mask.size != g.nx * g.ny:
            raise ValueError("mask size")
    return proj, na, beta, mask


def back_ts_f32(g: OrcGeom, proj, beta, mask=None, *, shipped_weights=False,
                cols=None, count=False):
    """FP32 restatement of fdk_ts_back1. Returns [nz nx ny] (or [ncol nz])."""
    lib = _load(LIB_ORACLE)
    proj, na, beta, mask = _prep(g, proj, beta, mask)
    if cols is not None:
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.empty((cols.size, g.nz), np.float32)
    else:
        out = np.empty((g.ny, g.nx, g.nz), np.float32)
    n = C.c_longlong(0)
    ok = lib.orc_back_ts_f32(C.byref(g), _ptr(proj, C.c_float), C.c_int32(na),
                             _ptr(beta, C.c_double), _ptr(mask, C.c_uint8),
                             C.c_int32(int(shipped_weights)),
                             _ptr(cols, C.c_int32),
                             C.c_int32(0 if cols is None else cols.size),
                             _ptr(out, C.c_float), C.byref(n))
    if not ok:
        raise RuntimeError("todo: dz/dt < 0 not done")
    return (out, n.value) if count else out


def back_ts_f64(g: OrcGeom, proj, beta, mask=None, *, cols=None):
    """FP64 restatement (error reference). Returns [nz nx ny] (or [ncol nz])."""
    lib = _load(LIB_ORACLE)
    proj, na, beta, mask = _prep(g, proj, beta, mask)
    if cols is not None:
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.empty((cols.size, g.nz), np.float64)
    else:
        out = np.empty((g.ny, g.nx, g.nz), np.float64)
    ok = lib.orc_back_ts_f64(C.byref(g), _ptr(proj, C.c_float), C.c_int32(na),
                             _ptr(beta, C.c_double), _ptr(mask, C.c_uint8),
                             _ptr(cols, C.c_int32),
                             C.c_int32(0 if cols is None else cols.size),
                             _ptr(out, C.c_double))
    if not ok:
        raise RuntimeError("todo: dz/dt < 0 not done")
    return out


def _helix_args(na, source_zs):
    zs = np.ascontiguousarray(np.broadcast_to(np.asarray(source_zs, np.float64), (na,)))
    return zs


def back_ts_helix_f32(g: OrcGeom, proj, beta, source_zs, mask=None, *, na1_half=0,
                      cone_par=0, w3d=0, source_dz_per_rad=0.0, pitch=0.0):
    """FP32 restatement of the fdk-helix fdk_ts_back_t with its helical
    arguments (fdk,ts.c:55-251, fdk,ts,t.c:39-167, fdk,w3d,pow,tab.c).
    Returns [nz nx ny]."""
    lib = _load(LIB_ORACLE)
    proj, na, beta, mask = _prep(g, proj, beta, mask)
    zs = _helix_args(na, source_zs)
    out = np.empty((g.ny, g.nx, g.nz), np.float32)
    ok = lib.orc_back_ts_helix_f32(C.byref(g), _ptr(proj, C.c_float), C.c_int32(na),
                                   _ptr(beta, C.c_double), _ptr(zs, C.c_double),
                                   C.c_int32(na1_half), C.c_int32(cone_par), C.c_int32(w3d),
                                   C.c_double(source_dz_per_rad), C.c_float(pitch),
                                   _ptr(mask, C.c_uint8), _ptr(out, C.c_float))
    if not ok:
        raise RuntimeError("todo: dz/dt < 0 not done")
    return out


def ref_back_ts_helix(g: OrcGeom, proj, beta, source_zs, mask=None, *, na1_half=0,
                      cone_par=0, w3d=0, source_dz_per_rad=0.0, pitch=0.0, nthread=1):
    """The reference C (oracle/_ref, fdk-helix build) 'fdk,ts,back' with its
    19 operands; returns [nz nx ny]."""
    lib = _load(LIB_REF["fixed"])
    proj, na, beta, mask = _prep(g, proj, beta, mask)
    zs = _helix_args(na, source_zs)
    if mask is None:
        mask = np.ones(g.nx * g.ny, np.uint8)
    out = np.empty((g.ny, g.nx, g.nz), np.float32)
    nxyz = np.array([g.nx, g.ny, g.nz], np.int32)
    dxyz = np.array([g.dx, g.dy, g.dz], np.float64)
    cxyz = np.array([g.offset_x, g.offset_y, g.offset_z], np.float64)
    dst = np.array([g.ds, g.dt], np.float64)
    ost = np.array([g.offset_s, g.offset_t], np.float64)
    ok = lib.ref_fdk_ts_back_helix(
        _ptr(out, C.c_float), _ptr(nxyz, C.c_int32), _ptr(dxyz, C.c_double),
        _ptr(cxyz, C.c_double), _ptr(mask, C.c_uint8), C.c_double(g.dso),
        C.c_double(g.dsd), C.c_double(g.dfs), _ptr(dst, C.c_double), _ptr(ost, C.c_double),
        _ptr(proj, C.c_float), C.c_int32(g.ns), C.c_int32(g.nt), C.c_int32(na),
        _ptr(beta, C.c_double), _ptr(zs, C.c_double), C.c_int32(na1_half),
        C.c_int32(cone_par), C.c_int32(w3d), C.c_double(source_dz_per_rad),
        C.c_float(pitch), C.c_int32(nthread))
    if not ok:
        raise RuntimeError("reference fdk_ts_back_t failed")
    return out


def back_st_f32(g: OrcGeom, proj_sta, beta, mask=None):
    """FP32 restatement of fdk_st_back1: proj [ns nt na] -> image [nx ny nz]."""
    lib = _load(LIB_ORACLE)
    proj, na, beta, mask = _prep(g, proj_sta, beta, mask)
    out = np.empty((g.nz, g.ny, g.nx), np.float32)
    lib.orc_back_st_f32(C.byref(g), _ptr(proj, C.c_float), C.c_int32(na),
                        _ptr(beta, C.c_double), _ptr(mask, C.c_uint8),
                        _ptr(out, C.c_float))
    return out


def ref_back_ts(which: str, g: OrcGeom, proj, beta, mask=None, nthread=1,
                timed=False):
    """Run the reference C (oracle/_ref) 'fdk,ts,back'; returns [nz nx ny].

    ``timed=True`` also returns the wall time of the fdk_ts_back_t call alone
    (CLOCK_MONOTONIC via time.perf_counter around the foreign call)."""
    lib = _load(LIB_REF[which])
    proj, na, beta, mask = _prep(g, proj, beta, mask)
    if mask is None:
        mask = np.ones(g.nx * g.ny, np.uint8)
    out = np.empty((g.ny, g.nx, g.nz), np.float32)
    nxyz = np.array([g.nx, g.ny, g.nz], np.int32)
    dxyz = np.array([g.dx, g.dy, g.dz], np.float64)
    cxyz = np.array([g.offset_x, g.offset_y, g.offset_z], np.float64)
    dst = np.array([g.ds, g.dt], np.float64)
    ost = np.array([g.offset_s, g.offset_t], np.float64)
    t0 = time.perf_counter()
    ok = lib.ref_fdk_ts_back(_ptr(out, C.c_float), _ptr(nxyz, C.c_int32),
                             _ptr(dxyz, C.c_double), _ptr(cxyz, C.c_double),
                             _ptr(mask, C.c_uint8), C.c_double(g.dso),
                             C.c_double(g.dsd), C.c_double(g.dfs),
                             _ptr(dst, C.c_double), _ptr(ost, C.c_double),
                             _ptr(proj, C.c_float), C.c_int32(g.ns),
                             C.c_int32(g.nt), C.c_int32(na),
                             _ptr(beta, C.c_double), C.c_int32(nthread))
    dt = time.perf_counter() - t0
    if not ok:
        raise RuntimeError("reference fdk_ts_back_t failed")
    return (out, dt) if timed else out


def ref_back_st(which: str, g: OrcGeom, proj_sta, beta, mask=None):
    """Reference 'fdk,st,back': proj [ns nt na] -> image [nx ny nz]."""
    lib = _load(LIB_REF[which])
    proj, na, beta, mask = _prep(g, proj_sta, beta, mask)
    if mask is None:
        mask = np.ones(g.nx * g.ny, np.uint8)
    out = np.empty((g.nz, g.ny, g.nx), np.float32)
    nxyz = np.array([g.nx, g.ny, g.nz], np.int32)
    dxyz = np.array([g.dx, g.dy, g.dz], np.float64)
    cxyz = np.array([g.offset_x, g.offset_y, g.offset_z], np.float64)
    dst = np.array([g.ds, g.dt], np.float64)
    ost = np.array([g.offset_s, g.offset_t], np.float64)
    lib.ref_fdk_st_back(_ptr(out, C.c_float), _ptr(nxyz, C.c_int32),
                        _ptr(dxyz, C.c_double), _ptr(cxyz, C.c_double),
                        _ptr(mask, C.c_uint8), C.c_double(g.dso),
                        C.c_double(g.dsd), C.c_double(g.dfs),
                        _ptr(dst, C.c_double), _ptr(ost, C.c_double),
                        _ptr(proj, C.c_float), C.c_int32(g.ns), C.c_int32(g.nt),
                        C.c_int32(na), _ptr(beta, C.c_double))
    return out


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def back_mat_f64(g: OrcGeom, proj_sta, beta, *, source_zs=None, ia_skip=1,
                 offset_source=0.0, mask=None, scale=1.0):
    """numpy restatement of cbct_back_mat (fbp/cbct_back.m:142-270).

    ``proj_sta`` has shape ``(ns, nt, na)`` and ``mask`` ``(nx, ny)`` (MATLAB
    index order, any memory order); returns the float64 ``(nx, ny, nz)`` image
    ``scale * embed(img2, mask)``.  Every step is the MATLAB expression's own
    double operation in its order (x.^2 as x*x); cos/sin of each angle come
    from libm (``math.cos``), as the product's host computes them.  One
    deviation, shared with the product: a NaN sample position adds 0 where
    MATLAB would fail on the index.
    """
    nx, ny, nz, ns, nt = g.nx, g.ny, g.nz, g.ns, g.nt
    proj = np.asarray(proj_sta, dtype=np.float64)
    if proj.ndim == 2:
        proj = proj[:, :, None]
    na = proj.shape[2]
    if proj.shape[:2] != (ns, nt):
        raise ValueError("proj must be (ns, nt, na)")
    beta = np.asarray(beta, dtype=np.float64).reshape(-1)
    zs = np.zeros(na) if source_zs is None else \
        np.broadcast_to(np.asarray(source_zs, np.float64), (na,))
    par = math.isinf(g.dsd) or math.isinf(g.dso)
    flat = par or math.isinf(g.dfs)                  # cbct_back.m:211-216
    if not flat and g.dfs != 0:
        raise ValueError("dfs must be 0 (arc) or inf (flat)")
    m = np.ones((nx, ny), bool) if mask is None else (np.asarray(mask) != 0)
    # cbct_back.m:160-165
    wx = (nx - 1) / 2 + g.offset_x
    wy = (ny - 1) / 2 + g.offset_y
    wz = (nz - 1) / 2 + g.offset_z
    xc, yc = np.meshgrid((np.arange(nx) - wx) * g.dx, (np.arange(ny) - wy) * g.dy,
                         indexing="ij")
    zc = (np.arange(nz) - wz) * g.dz
    xc = xc[m]                                       # cbct_back.m:182-183
    yc = yc[m]
    ws = (ns + 1) / 2 + g.offset_s                   # cbct_back.m:185-186
    wt = (nt + 1) / 2 + g.offset_t
    proj1 = np.zeros((ns + 1, nt))                   # cbct_back.m:190-191
    img2 = np.zeros((xc.size, nz))
    for ia in range(0, na, int(ia_skip)):           # cbct_back.m:201
        cb, sb = math.cos(beta[ia]), math.sin(beta[ia])
        x_beta = xc * cb + yc * sb                   # cbct_back.m:204-205
        y_betas = g.dso - (-xc * sb + yc * cb)
        mag = np.ones_like(y_betas) if par else g.dsd / y_betas
        if flat:
            sprime = mag * x_beta
        else:
            r_loop = x_beta - offset_source
            sprime = g.dsd * np.arctan2(r_loop, y_betas)
        tprime = mag[:, None] * (zc[None, :] - zs[ia])   # cbct_back.m:219
        bs = sprime / g.ds + ws
        bt = tprime / g.dt + wt
        bs = np.where((bs >= 1) & (bs <= ns), bs, ns + 1.0)
        bt = np.fmax(bt, 1.0)                        # max/min ignore NaN, as MATLAB's
        bt = np.fmin(bt, float(nt))
        is_ = np.floor(bs)
        it = np.floor(bt)
        is_[is_ == ns + 1] = ns
        it[it == nt] = nt - 1
        wr = (bs - is_)[:, None]
        wl = 1 - wr
        wu = bt - it
        wd = 1 - wu
        proj1[:ns, :] = proj[:, :, ia]
        i0 = (is_.astype(np.int64) - 1)[:, None]     # 0-based left column
        t0 = it.astype(np.int64) - 1                 # 0-based lower row
        p1 = wl * proj1[i0, t0] + wr * proj1[i0 + 1, t0]
        p2 = wl * proj1[i0, t0 + 1] + wr * proj1[i0 + 1, t0 + 1]
        p0 = wd * p1 + wu * p2
        if flat:
            p0 = p0 * (mag * mag)[:, None]
        else:
            p0 = p0 * (g.dsd * g.dsd) / (r_loop * r_loop + y_betas * y_betas)[:, None]
        img2 = img2 + p0
    img = np.zeros((nx, ny, nz))
    img[m] = img2                                    # embed (cbct_back.m:261)
    return scale * img


def helix_mat_slices(g: OrcGeom, na: int, orbit: float, pitch: float, source_zs,
                     ia_skip: int = 1):
    """Per-slice view ranges and scales of the helical MATLAB backprojector
    (mex/src/fdk-helix/cbct_back_helix.m:212-236, 341-345), in double as
    MATLAB computes them: returns ``(first, last, scale)`` with 0-based
    inclusive view ranges (slice iz uses views first:ia_skip:last) and the
    ``scale_dang`` factor of each slice."""
    na1 = na / orbit * 360                           # :212-216
    if abs(np.round(na1) - na1) > 1e-4:
        raise ValueError("bad na/orbit")
    na1_half = math.ceil(na1 / 2)
    wz = (g.nz - 1) / 2 + g.offset_z
    zc = (np.arange(g.nz) - wz) * g.dz
    zs = np.broadcast_to(np.asarray(source_zs, np.float64), (na,))
    first = np.empty(g.nz, np.int64)
    last = np.empty(g.nz, np.int64)
    scale = np.empty(g.nz)
    for iz in range(g.nz):
        if pitch == 0:                               # :226-228
            ia_min, ia_max = 1, na
        else:                                        # :229-235, 1-based
            ia1_middle = int(np.argmin(np.abs(zc[iz] - zs))) + 1
            ia_min = max(ia1_middle - na1_half, 1)
            ia_max = min(ia1_middle + (na1_half - 1), na)
        first[iz], last[iz] = ia_min - 1, ia_max - 1
        orbit_iz = ((ia_max - ia_min + 1) / ia_skip) * abs(orbit) / (na / ia_skip)
        sc = math.pi / ((math.pi / 180) * abs(orbit_iz))
        scale[iz] = sc * ((math.pi / 180) * abs(orbit)) / (na / ia_skip)
    return first, last, scale


def back_helix_mat_f64(g: OrcGeom, proj_sta, beta, source_zs, first, last, *, ia_skip=1,
                       offset_source=0.0, mask=None, slice_scale=None, cone_par=0,
                       source_dz_per_rad=0.0):
    """numpy restatement of the helical ``cbct_back_mat``
    (mex/src/fdk-helix/cbct_back_helix.m:185-350) without w3d: slice iz
    sums views first[iz]:ia_skip:last[iz] (0-based, inclusive) as
    ``img2 = img2 + p0`` in view order, with the cone-parallel geometry when
    ``cone_par`` (:245-265: mag = dsd / (sqrt(dso^2 - x_beta^2) - (dso -
    y_betas)), sprime = x_beta, source height offset asin(-x_beta / dso) *
    source_dz_per_rad, no distance weight), then times ``slice_scale[iz]``
    (or 1) and embedded in the mask.  Same edge rules, operation order and
    NaN deviation as :func:`back_mat_f64`."""
    nx, ny, nz, ns, nt = g.nx, g.ny, g.nz, g.ns, g.nt
    proj = np.asarray(proj_sta, dtype=np.float64)
    if proj.ndim == 2:
        proj = proj[:, :, None]
    na = proj.shape[2]
    beta = np.asarray(beta, dtype=np.float64).reshape(-1)
    zs = np.broadcast_to(np.asarray(source_zs, np.float64), (na,))
    par = math.isinf(g.dsd) or math.isinf(g.dso)
    flat = par or math.isinf(g.dfs)
    if not cone_par and not flat and g.dfs != 0:
        raise ValueError("dfs must be 0 (arc) or inf (flat)")
    m = np.ones((nx, ny), bool) if mask is None else (np.asarray(mask) != 0)
    wx = (nx - 1) / 2 + g.offset_x
    wy = (ny - 1) / 2 + g.offset_y
    wz = (nz - 1) / 2 + g.offset_z
    xc, yc = np.meshgrid((np.arange(nx) - wx) * g.dx, (np.arange(ny) - wy) * g.dy,
                         indexing="ij")
    zc = (np.arange(nz) - wz) * g.dz
    xc = xc[m]
    yc = yc[m]
    ws = (ns + 1) / 2 + g.offset_s
    wt = (nt + 1) / 2 + g.offset_t
    proj1 = np.zeros((ns + 1, nt))
    img = np.zeros((nx, ny, nz))
    with np.errstate(invalid="ignore", divide="ignore"):
        for iz in range(nz):
            img2 = np.zeros(xc.size)
            for ia in range(int(first[iz]), int(last[iz]) + 1, int(ia_skip)):
                cb, sb = math.cos(beta[ia]), math.sin(beta[ia])
                x_beta = xc * cb + yc * sb
                y_betas = g.dso - (-xc * sb + yc * cb)
                if cone_par:
                    mag = g.dsd / (np.sqrt(g.dso * g.dso - x_beta * x_beta) - (g.dso - y_betas))
                    sprime = x_beta
                    zoff = np.arcsin(-x_beta / g.dso) * source_dz_per_rad
                    tprime = mag * (zc[iz] - (zs[ia] + zoff))
                else:
                    mag = np.ones_like(y_betas) if par else g.dsd / y_betas
                    if flat:
                        sprime = mag * x_beta
                    else:
                        r_loop = x_beta - offset_source
                        sprime = g.dsd * np.arctan2(r_loop, y_betas)
                    tprime = mag * (zc[iz] - zs[ia])
                bs = sprime / g.ds + ws
                bt = tprime / g.dt + wt
                bs = np.where((bs >= 1) & (bs <= ns), bs, ns + 1.0)
                bt = np.fmin(np.fmax(bt, 1.0), float(nt))
                is_ = np.floor(bs)
                it = np.floor(bt)
                is_[is_ == ns + 1] = ns
                it[it == nt] = nt - 1
                wr = bs - is_
                wl = 1 - wr
                wu = bt - it
                wd = 1 - wu
                proj1[:ns, :] = proj[:, :, ia]
                i0 = is_.astype(np.int64) - 1
                t0 = it.astype(np.int64) - 1
                p1 = wl * proj1[i0, t0] + wr * proj1[i0 + 1, t0]
                p2 = wl * proj1[i0, t0 + 1] + wr * proj1[i0 + 1, t0 + 1]
                p0 = wd * p1 + wu * p2
                if not cone_par:
                    if flat:
                        p0 = p0 * (mag * mag)
                    else:
                        p0 = p0 * (g.dsd * g.dsd) / (r_loop * r_loop + y_betas * y_betas)
                img2 = img2 + p0
            if slice_scale is not None:
                img2 = slice_scale[iz] * img2
            img[:, :, iz][m] = img2
    return img
