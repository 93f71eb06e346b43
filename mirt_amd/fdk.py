"""FDK steps around the backprojector: cosine pre-weight, ramp filter, driver.

* :func:`feldkamp_weight1` -- fbp/feldkamp.m:103-132 (``w1cyl`` weighting);
* :func:`fbp_ramp` -- fbp/fbp_ramp.m:38-60 (band-limited ramp taps, 'flat'
  and 'arc');
* :func:`fdk_filter` -- fbp/fdk_filter.m:1-63 (zero-padded FFT filtering
  along s, ``ifft(...,'symmetric')`` taken as the real inverse transform of
  the non-negative frequencies);
* :func:`feldkamp` -- fbp/feldkamp.m:1-98: weight, filter, backproject.

The weight and the filter are written with torch so they run where the
projections live (the GPU for full-size scans); the backprojection is the
HIP kernel in libfdkhip.so.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .cbct_back import FdkPlan, angle_scale, cbct_back
from .geom import CTGeom, ImageGeom


def _t(x, like):
    return torch.as_tensor(x, dtype=like.dtype, device=like.device)


def feldkamp_weight1(proj, cg: CTGeom, *, w1cyl: bool = True):
    """feldkamp.m:103-132: multiply every view ``[ns nt]`` by ww1."""
    ns, nt = proj.shape[0], proj.shape[1]
    ds, dt, dsd, dso = cg.ds, cg.dt, cg.dsd, cg.dso
    ss = (np.arange(ns) - (ns - 1) / 2 - cg.offset_s) * ds
    tt = (np.arange(nt) - (nt - 1) / 2 - cg.offset_t) * dt
    ss, tt = np.meshgrid(ss, tt, indexing="ij")
    if math.isinf(cg.dfs):
        if w1cyl:
            ww1 = dso / np.sqrt(dsd ** 2 + ss ** 2 + tt ** 2)
        else:
            ww1 = dso / np.sqrt(dsd ** 2 + ss ** 2 + tt ** 2) * np.sqrt(1 + (tt / dsd) ** 2)
    elif cg.dfs == 0:
        if w1cyl:
            ww1 = (dso / dsd) * np.cos(ss / dsd) / np.sqrt(1 + (tt / dsd) ** 2)
        else:
            ww1 = (dso / dsd) * np.cos(ss / (dsd * np.sqrt(1 + (tt / dsd) ** 2)))
    else:
        raise ValueError("other configurations not implemented")
    return proj * _t(ww1, proj)[:, :, None]


def fbp_ramp(type: str, n: int, ds: float, dsd: float = None) -> np.ndarray:
    """fbp_ramp.m:38-60: samples of the band-limited ramp, nn = -n/2..n/2-1."""
    if n % 2:
        raise ValueError("n must be even")
    nn = np.arange(-(n // 2), n // 2)
    h = np.zeros(n)
    odd = (nn % 2) == 1
    if type == "arc":
        if n / 2 * ds / dsd > 0.9 * math.pi / 2:
            raise ValueError("physically impossible arc geometry")
        h[nn == 0] = 1 / (4 * ds ** 2)
        h[odd] = -1 / (math.pi * dsd * np.sin(nn[odd] * ds / dsd)) ** 2
    elif type == "flat":
        h[n // 2] = 1 / 4
        h[odd] = -1 / (math.pi * nn[odd]) ** 2
        h = h / ds ** 2
    else:
        raise ValueError("bad fan type")
    return h


def fdk_filter_response(cg: CTGeom, npadh: int, window: str = "ramp") -> np.ndarray:
    """H of fdk_filter.m:24-33 / fdk_fan_filter :41-63, times ds."""
    if math.isinf(cg.dsd) or math.isinf(cg.dfs):
        h = fbp_ramp("flat", npadh, cg.ds)
    elif cg.dfs == 0:
        h = fbp_ramp("arc", npadh, cg.ds, cg.dsd)
    else:
        raise ValueError("dfs not done")
    H = np.real(np.fft.fft(np.fft.fftshift(h)))
    if window == "ramp":
        win = np.ones(npadh)
    elif window == "hann":
        win = 0.5 * (1 - np.cos(2 * math.pi * np.arange(npadh) / npadh))   # periodic
    else:
        raise ValueError(f"window {window} not provided")
    H = H * np.fft.fftshift(win)
    return cg.ds * H


def fdk_filter(proj, cg: CTGeom, window: str = "ramp"):
    """fdk_filter.m: filter each view along s with zero padding to
    2^ceil(log2(2 ns - 1)).  ``proj`` is ``[ns nt na]`` (torch)."""
    ns = proj.shape[0]
    npadh = 2 ** math.ceil(math.log2(2 * ns - 1))
    H = fdk_filter_response(cg, npadh, window)
    Hh = _t(H[: npadh // 2 + 1], proj)
    P = torch.fft.rfft(proj, n=npadh, dim=0)
    out = torch.fft.irfft(P * Hh[:, None, None], n=npadh, dim=0)
    return out[:ns]


def feldkamp(cg: CTGeom, ig: ImageGeom, proj, *, window="ramp", w1cyl=True,
             ia_skip=1, extrapolate_t=0, fmt="f32", math_mode="exact"):
    """feldkamp.m:1-98.  ``proj`` ``[ns nt na]``: a numpy array (host path,
    returns numpy ``[nx ny nz]``) or a device torch tensor (returns a device
    ``[nx ny nz]`` view; FP16/BF16 storage with ``fmt``)."""
    if cg.pitch != 0 or np.any(cg.source_zs != 0):
        raise ValueError("sorry, helical CT unsupported")
    if isinstance(proj, np.ndarray):
        p = torch.as_tensor(proj, dtype=torch.float64)
        p = fdk_filter(feldkamp_weight1(p, cg, w1cyl=w1cyl), cg, window)
        return cbct_back(p.numpy().astype(np.float32), cg, ig, ia_skip=ia_skip,
                         extrapolate_t=extrapolate_t)
    p = fdk_filter(feldkamp_weight1(proj, cg, w1cyl=w1cyl), cg, window)
    if extrapolate_t:
        raise NotImplementedError("extrapolate_t on the device path")
    tsa = p[:, :, ::ia_skip].permute(2, 0, 1).contiguous().to(torch.float32)  # [nt ns na]
    plan = FdkPlan.from_geoms(cg, ig, ia_skip=ia_skip, fmt=fmt, math_mode=math_mode)
    if fmt == "f32":
        img = plan(tsa)
    else:
        from .cbct_back import quantize_views
        q, sc = quantize_views(tsa, fmt)
        img = plan(q, sc)
    img = img * float(angle_scale(cg, ia_skip))
    return FdkPlan.image_xyz(img)
