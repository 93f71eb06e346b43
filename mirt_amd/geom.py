"""Cone-beam CT and image geometry, as ``ct_geom`` / ``image_geom`` define them.

Host-side restatement of the conventions ``cbct_back`` consumes
(``fbp/ct_geom.m``, ``fbp/image_geom.m``).  Arrays follow MATLAB index order:
a projection stack is indexed ``[is, it, ia]`` (shape ``(ns, nt, na)``), an
image ``[ix, iy, iz]``; ``dims`` are given in that order.

Only what the cone-beam backprojection path needs is restated: the fan-beam
('fan', 'ge1', 'ge2') CT geometry with its sample positions, angles, source
positions and down-sampling, and the 3-D image geometry with offsets,
voxel sizes, masks and down-sampling.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np


def _mround(x):
    """MATLAB round(): halves away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def deg2rad(d):
    """MATLAB deg2rad: (pi/180) * d."""
    return (math.pi / 180) * np.asarray(d, dtype=np.float64)


# ---------------------------------------------------------------------------
# ct_geom

@dataclass
class CTGeom:
    """Restatement of the ``ct_geom`` strum (fbp/ct_geom.m).

    Field meaning as in ``ct_geom.m:1-90``; distances in mm, angles in
    degrees (``orbit``, ``orbit_start``), offsets in detector pixels.
    """
    type: str = "fan"
    ns: int = 0
    nt: int = 0
    na: int = 0
    ds: float = 1.0
    dt: float = -1.0
    offset_s: float = 0.0
    offset_t: float = 0.0
    dsd: float = math.inf
    dso: float = math.inf
    dod: float = 0.0
    dfs: float = 0.0
    orbit: float = 360.0
    orbit_start: object = 0.0
    pitch: float = 0.0
    source_z0: float = 0.0
    user_source_zs: object = None
    down: object = 1
    units: str = "mm"

    # ct_geom.m:282-289
    @property
    def ws(self) -> float:
        return (self.ns - 1) / 2 + self.offset_s

    @property
    def wt(self) -> float:
        return (self.nt - 1) / 2 + self.offset_t

    # ct_geom.m:293-305
    @property
    def s(self) -> np.ndarray:
        return self.ds * (np.arange(self.ns, dtype=np.float64) - self.ws)

    @property
    def t(self) -> np.ndarray:
        return self.dt * (np.arange(self.nt, dtype=np.float64) - self.wt)

    # ct_geom.m:309-320
    @property
    def ad(self) -> np.ndarray:
        return (np.arange(self.na, dtype=np.float64) / self.na * self.orbit
                + np.asarray(self.orbit_start, dtype=np.float64))

    @property
    def ar(self) -> np.ndarray:
        return deg2rad(self.ad)

    @property
    def dim(self):
        return (self.ns, self.nt, self.na)

    def zeros(self):
        return np.zeros(self.dim, dtype=np.float32, order="F")

    def ones(self):
        return np.ones(self.dim, dtype=np.float32, order="F")

    def unitv(self, is_=None, it=None, ia=None):
        """ct_geom.m:207-217; 1-based indices like MATLAB."""
        out = self.zeros()
        if is_ is None:
            is_ = math.floor(self.ns / 2 + 1)
            it = math.floor(self.nt / 2 + 1)
            ia = 1
        out[is_ - 1, it - 1, ia - 1] = 1
        return out

    # ct_geom.m:468-477
    def gamma_s(self, ss):
        if self.dfs == 0:
            return ss / self.dsd
        if math.isinf(self.dfs):
            return np.arctan(ss / self.dsd)
        raise ValueError("dfs not done")

    @property
    def gamma(self):
        return self.gamma_s(self.s)

    @property
    def gamma_max_abs(self):
        return float(np.max(np.abs(self.gamma)))

    @property
    def rmax(self) -> float:
        """ct_geom.m:238-257"""
        smax = float(np.max(np.abs(self.s)))
        if math.isinf(self.dso):
            return smax
        return self.dso * math.sin(self.gamma_max_abs)

    @property
    def cone_angle(self) -> float:
        if math.isinf(self.dso) or math.isinf(self.dsd):
            return 0.0
        return math.atan((self.nt * self.dt) / 2 / self.dsd)

    @property
    def zfov(self) -> float:
        if math.isinf(self.dso) or math.isinf(self.dsd):
            return self.nt * self.dt
        return self.dso / self.dsd * self.nt * self.dt

    # ct_geom.m:437-453
    @property
    def source_dz_per_view(self) -> float:
        if self.user_source_zs is not None:
            raise ValueError("undefined")
        if self.na == 1 or not self.pitch:
            return 0.0
        if np.ndim(self.orbit) != 0 or self.orbit == 0:
            raise ValueError("dz_per_view undefined for user angles")
        na_per_360 = self.na * (360 / self.orbit)
        return self.pitch * self.zfov / na_per_360

    @property
    def source_zs(self) -> np.ndarray:
        """ct_geom.m:458-466"""
        if self.user_source_zs is not None:
            return np.asarray(self.user_source_zs, dtype=np.float64)
        return self.source_z0 + np.arange(self.na) * self.source_dz_per_view

    def downsample(self, down) -> "CTGeom":
        """ct_geom.m:330-400 (axial and user-specified cases)."""
        st = copy.deepcopy(self)
        down = np.atleast_1d(down)
        st.down = np.asarray(st.down) * down if np.ndim(st.down) else st.down * down
        if down.size == 1:
            down_s = down_t = down_a = int(down[0])
        elif down.size == 3:
            down_s, down_t, down_a = (int(d) for d in down)
        else:
            raise ValueError(f"bad down {down}")
        st.ns = int(4 * math.ceil(self.ns / down_s / 4))
        st.nt = int(2 * math.ceil(self.nt / down_t / 2))
        user = False
        if self.user_source_zs is not None:
            st.user_source_zs = np.asarray(self.user_source_zs)[::down_a]
            user = True
        if np.ndim(self.orbit_start) and np.size(self.orbit_start) > 1:
            st.orbit_start = np.asarray(self.orbit_start)[::down_a]
            user = True
        st.ds = self.ds * down_s
        st.dt = self.dt * down_t
        if user:
            st.na = len(range(0, self.na, down_a))
        elif np.all(np.diff(self.source_zs) == 0):
            st.na = int(max(1, _mround(self.na / down_a)))
        else:
            nturn = self.orbit / 360
            na1 = self.na / nturn
            tol = 1e-4
            if abs(_mround(na1) - na1) < tol:
                na2 = _mround(na1 / down_a)
                tmp = nturn * na2
                if abs(_mround(tmp) - tmp) < tol:
                    st.na = int(_mround(tmp))
                else:
                    st.na = int(math.floor(tmp))
                    st.orbit = 360 / na2 * st.na
            else:
                st.na = int(_mround(self.na / down_a))
        return st


def ct_geom(type: str, **kw) -> CTGeom:
    """Mirror of ``ct_geom(type, ...)`` for 'fan', 'ge1' and 'ge2'.

    Option names as in ct_geom.m (aliases dis_src_det -> dsd, dis_src_iso ->
    dso, dis_iso_det -> dod, dis_foc_src -> dfs, nbeta -> na, source_zs ->
    user_source_zs).  ``dt`` may be ``('dz', dz)`` for dz*dsd/dso.
    """
    if type == "ge1":
        return _ct_geom_ge1(**kw)
    if type == "ge2":
        return _ct_geom_ge1(**{"nt": 64, "dt": 949.0750 / 541 * 0.625, **kw})
    if type != "fan":
        raise ValueError(f"unknown sinotype {type}")
    subs = {"src_det_dis": "dsd", "dis_src_det": "dsd", "dis_src_iso": "dso",
            "dis_iso_det": "dod", "dis_foc_src": "dfs", "nbeta": "na",
            "source_zs": "user_source_zs"}
    kw = {subs.get(k, k): v for k, v in kw.items()}
    down = kw.pop("down", 1)
    st = CTGeom(type="fan", dt=None, dsd=None, dso=None, dod=None, na=None)
    for k, v in kw.items():
        if not hasattr(st, k):
            raise TypeError(f"unknown ct_geom option {k}")
        setattr(st, k, v)
    # ct_geom.m:605-631 (ct_geom_fan)
    if (st.dsd is not None and math.isinf(st.dsd)) or (st.dso is not None and math.isinf(st.dso)):
        st.dsd, st.dso, st.dod = math.inf, math.inf, 1.0
    if (st.dsd is None) + (st.dso is None) + (st.dod is None) > 1:
        raise ValueError("must provide at least two of dsd, dso, dod")
    if st.dsd is None:
        st.dsd = st.dso + st.dod
    if st.dso is None:
        st.dso = st.dsd - st.dod
    if st.dod is None:
        st.dod = st.dsd - st.dso
    if not (math.isinf(st.dsd) and math.isinf(st.dso)) and st.dso + st.dod != st.dsd:
        raise ValueError("bad fan-beam distances")
    if st.dt is None:
        st.dt = -st.ds
    if isinstance(st.dt, (tuple, list)):
        if len(st.dt) == 2 and st.dt[0] == "dz":
            st.dt = st.dt[1] * st.dsd / st.dso
        else:
            raise ValueError("bad dt cell")
    if st.na is None:
        st.na = 2 * math.floor(st.ns * math.pi / 2 / 2)
    for name in ("ns", "nt", "na"):
        if round(getattr(st, name)) != getattr(st, name):
            raise ValueError(f"{name} must be integer")
        setattr(st, name, int(getattr(st, name)))
    if st.source_z0 == "center":
        st.source_z0 = -st.na / 2 * st.source_dz_per_view
    if np.any(np.asarray(down) != 1):
        st = st.downsample(down)
        st.down = down
    if isinstance(st.orbit, str) and st.orbit == "short":
        st.orbit = 180 + 2 * math.degrees(max(st.gamma))
    return st


def _ct_geom_ge1(**kw) -> CTGeom:
    """ct_geom.m:812-839: GE LightSpeed, arc detector (published geometry)."""
    na = 984
    if kw.get("orbit") == "short":
        na = 642
        kw["orbit"] = na / 984 * 360
    base = dict(ns=888, nt=1, na=na, orbit=360, offset_s=1.25, dsd=949.075,
                dso=541, dfs=0, ds=1.0239, dt=1.0964)
    base.update(kw)
    return ct_geom("fan", **base)


# ---------------------------------------------------------------------------
# image_geom

@dataclass
class ImageGeom:
    """Restatement of the 3-D ``image_geom`` strum (fbp/image_geom.m)."""
    nx: int = 0
    ny: int = 0
    nz: int = 0
    dx: float = 1.0
    dy: float = -1.0
    dz: float = 1.0
    offset_x: float = 0.0
    offset_y: float = 0.0
    offset_z: float = 0.0
    fov: float = 0.0
    zfov: float = 0.0
    down: object = 1
    mask: np.ndarray = field(default=None, repr=False)

    @property
    def dim(self):
        return (self.nx, self.ny, self.nz)

    # image_geom.m:439-465
    @property
    def wx(self):
        return (self.nx - 1) / 2 + self.offset_x

    @property
    def wy(self):
        return (self.ny - 1) / 2 + self.offset_y

    @property
    def wz(self):
        return (self.nz - 1) / 2 + self.offset_z

    @property
    def x(self):
        return (np.arange(self.nx) - self.wx) * self.dx

    @property
    def y(self):
        return (np.arange(self.ny) - self.wy) * self.dy

    @property
    def z(self):
        return (np.arange(self.nz) - self.wz) * self.dz

    @property
    def mask_or(self) -> np.ndarray:
        """image_geom.m:348-349: [nx ny] logical or over z."""
        return np.any(self.mask, axis=2)

    @property
    def np(self) -> int:
        return int(np.sum(self.mask))

    def zeros(self):
        return np.zeros(self.dim, dtype=np.float32, order="F")

    def ones(self):
        return np.ones(self.dim, dtype=np.float32, order="F")

    def downsample(self, down) -> "ImageGeom":
        """image_geom.m:288-337."""
        st = copy.deepcopy(self)
        downv = np.broadcast_to(np.atleast_1d(down), (3,)).astype(int)
        st.down = self.down * int(downv[0]) if np.size(down) == 1 else np.asarray(self.down) * downv

        def down_round(val, dd, d):
            out = val / d
            if out != _mround(out):
                out = 2 * _mround(out / 2)
            return int(out), dd * d

        st.nx, st.dx = down_round(self.nx, self.dx, downv[0])
        st.ny, st.dy = down_round(self.ny, self.dy, downv[1])
        st.nz, st.dz = down_round(self.nz, self.dz, downv[2])
        st.offset_x = self.offset_x / downv[0]
        st.offset_y = self.offset_y / downv[1]
        st.offset_z = self.offset_z / downv[2]
        mdim = self.mask.shape
        if tuple(np.array(st.dim) * downv) == tuple(mdim):
            # downsample3(mask, down) > 0: inclusive or over each block
            m = self.mask.reshape(downv[0], st.nx, downv[1], st.ny, downv[2], st.nz,
                                  order="F")
            st.mask = np.asfortranarray(m.any(axis=(0, 2, 4)))
        else:
            raise ValueError("bug: bad mask size.  need to address mask downsampling")
        return st


def image_geom(**kw) -> ImageGeom:
    """Mirror of ``image_geom(...)`` for 3-D images (image_geom.m:63-260).

    Options: nx, ny, nz, dx, dy, dz, offset_x/y/z, offsets ('dsp' or 3
    numbers), fov, zfov, down, mask (logical [nx ny nz] array or
    'all-but-edge').
    """
    nx = kw.pop("nx")
    ny = kw.pop("ny", None)
    nz = kw.pop("nz", None)
    dx = kw.pop("dx", None)
    dy = kw.pop("dy", None)
    dz = kw.pop("dz", None)
    ox = kw.pop("offset_x", 0.0)
    oy = kw.pop("offset_y", 0.0)
    oz = kw.pop("offset_z", 0.0)
    offsets = kw.pop("offsets", "")
    fov = kw.pop("fov", None)
    zfov = kw.pop("zfov", None)
    down = kw.pop("down", 1)
    mask = kw.pop("mask", None)
    if kw:
        raise TypeError(f"unknown image_geom options {sorted(kw)}")
    if ny is None:
        ny = nx
    if nz is None and dz is None and zfov is None:
        raise ValueError("only 3-D image geometry is restated (nz, dz or zfov required)")
    if offsets == "dsp":
        ox = 0.5 * (nx % 2 == 0)
        oy = 0.5 * (ny % 2 == 0)
        oz = 0.5 * ((nz or 0) % 2 == 0)
    elif not isinstance(offsets, str):
        ox, oy, oz = (float(v) for v in offsets)
    if dx is None:
        if fov is None:
            raise ValueError("dx or fov required")
        fovv = np.atleast_1d(fov)
        if fovv.size == 1:
            dx = float(fovv[0]) / nx
        elif fovv.size == 3:
            dx, dy, dz = fovv[0] / nx, fovv[1] / ny, fovv[2] / nz
            zfov = float(fovv[2])
            fov = float(fovv[0])
        else:
            raise ValueError("bad fov")
    elif fov is None:
        fov = nx * dx
    if isinstance(dy, str):
        dy = eval(dy, {"dx": dx})  # 'dx', '-dx', '2*dx' as in image_geom.m:158-160
    if dy is None:
        dy = -dx
    if nz is None:
        nz = zfov / dz
        if round(nz) != nz:
            raise ValueError("bad zfov/dz")
        nz = int(round(nz))
    if dz is None:
        dz = dx if zfov is None else zfov / nz
    if zfov is None:
        zfov = nz * dz
    all_but_edge = isinstance(mask, str) and mask == "all-but-edge"
    if all_but_edge or mask is None:
        mask_arr = np.ones((nx, ny, nz), dtype=bool, order="F")
        if all_but_edge:
            mask_arr[[0, -1], :, :] = False
            mask_arr[:, [0, -1], :] = False
    else:
        mask_arr = np.asfortranarray(np.asarray(mask, dtype=bool))
        if mask_arr.shape != (nx, ny, nz):
            raise ValueError(f"bad input mask size, nx={nx} ny={ny}")
    st = ImageGeom(nx=int(nx), ny=int(ny), nz=int(nz), dx=float(dx), dy=float(dy),
                   dz=float(dz), offset_x=float(ox), offset_y=float(oy),
                   offset_z=float(oz), fov=float(fov), zfov=float(zfov), down=1,
                   mask=mask_arr)
    if np.any(np.asarray(down) != 1):
        st = st.downsample(down)
        if all_but_edge:
            m = st.mask.copy(order="F")
            m[[0, -1], :, :] = False
            m[:, [0, -1], :] = False
            st.mask = m
    return st
