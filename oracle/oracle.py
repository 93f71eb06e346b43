"""CPU checker for the FDK backprojector -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  It wraps two libraries built by
``oracle/Makefile``:

* ``oracle/lib/libfdk_oracle.so`` -- this directory's clean-room restatement
  (``fdk_oracle.c``): FP32 with the reference's operation order, FP64 error
  reference, and the ``[ns nt]``-order variant;
* ``oracle/_ref/libfdkref_{fixed,shipped}.so`` -- the reference C compiled from
  ``mex/src/fdk-helix/mex`` (debugged vertical weights, axial path) and
  ``mex/src/fdk/fdk-orig`` (as shipped), driven by ``ref_driver.c`` the way
  ``fdk_mex_back`` (``mex/src/fdk/fdk-orig/fdk,mex.c:22-171``) drives them.

Array conventions follow the reference mex call
``fdk_mex('fdk,ts,back', nxyz, dxyz, cxyz, mask, dso, dsd, dfs, [ds dt],
[offset_s offset_t], proj, beta, nthread)``: ``proj`` is ``[nt ns na]`` with t
fastest (a C-order numpy array of shape ``(na, ns, nt)``), the result is
``[nz nx ny]`` with z fastest (C-order shape ``(ny, nx, nz)``), ``mask`` is
``[nx ny]`` with x fastest (C-order shape ``(ny, nx)``).
"""
from __future__ import annotations

import ctypes as C
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
