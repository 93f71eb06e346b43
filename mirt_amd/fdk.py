"""FDK steps around the backprojector: cosine pre-weight, ramp filter, driver.

* :func:`feldkamp_weight1` -- fbp/feldkamp.m:103-132 (``w1cyl`` weighting);
* :func:`fbp_ramp` -- fbp/fbp_ramp.m:38-60 (band-limited ramp taps, 'flat'
  and 'arc');
* :func:`fdk_filter` -- fbp/fdk_filter.m:1-63 (zero-padded FFT filtering
  along s, ``ifft(...,'symmetric')`` taken as the real inverse transform of
  the non-negative frequencies);
* :func:`weight_filter_views` -- the weight and the filter on the device
  through ``fdkhip_fdk_filter_views`` (one fused HIP kernel: weight, FFT of
  the zero-padded rows in LDS, the response, inverse FFT, transpose), FP32,
  producing the ``[nt ns na]`` views the plans take;
* :func:`feldkamp` -- fbp/feldkamp.m:1-98: weight, filter, backproject.

:func:`feldkamp_weight1` and :func:`fdk_filter` are written with torch in
any precision, on any device (FP64 on the host for the host route and for
synthesising benchmark views); :func:`feldkamp`'s device route uses
:func:`weight_filter_views`.  The backprojection is the HIP kernel in
libfdkhip.so.
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
    ww1 = weight_table(cg, proj.shape[0], proj.shape[1], w1cyl=w1cyl)
    return proj * _t(ww1, proj)[:, :, None]


def weight_table(cg: CTGeom, ns: int, nt: int, *, w1cyl: bool = True) -> np.ndarray:
    """ww1 of feldkamp.m:103-132 as an FP64 ``[ns nt]`` array."""
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
    return ww1


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


def _matlab_round(x: float) -> float:
    """MATLAB round: half away from zero."""
    return math.copysign(math.floor(abs(x) + 0.5), x)


def fbp2_window(n: int, window) -> np.ndarray:
    """fbp2_window.m: the apodizing window of length n, fft-shifted (its
    centre sample ii = 0 first).  ``window``: '' / 'boxcar' / 'ramp' (ones),
    'boxcar,c', 'hamming,c', 'hanning,c' (cutoff c), 'hann' (c = 1),
    'hann50', 'hann75', 'hann80', or an array of length n (samples
    ii = -n/2 .. n/2-1)."""
    def support(cut):
        w = _matlab_round(cut * n)
        ii = np.arange(n) - n / 2
        return w, ii, np.abs(ii) < w / 2

    def boxcar(cut):
        return support(cut)[2].astype(np.float64)

    def hann(cut):
        w, ii, inside = support(cut)
        return 0.5 * (1 + np.cos(2 * math.pi * ii / w)) * inside

    def hamming(cut):
        w, ii, inside = support(cut)
        return (0.54 + 0.46 * np.cos(2 * math.pi * ii / w)) * inside

    def cutoff(prefix):
        try:
            return float(window[len(prefix):])
        except ValueError:
            raise ValueError(f"bad cutoff in window {window!r}") from None

    if isinstance(window, str):
        if window in ("", "boxcar", "ramp"):
            win = np.ones(n)
        elif window.startswith("boxcar,"):
            win = boxcar(cutoff("boxcar,"))
        elif window.startswith("hamming,"):
            win = hamming(cutoff("hamming,"))
        elif window.startswith("hanning,"):
            win = hann(cutoff("hanning,"))
        elif window in ("hann", "hann50", "hann75", "hann80"):
            win = hann({"hann": 1.0, "hann50": 0.5, "hann75": 0.75, "hann80": 0.80}[window])
        else:
            raise ValueError(f"unknown window {window}")
    else:
        win = np.asarray(window, dtype=np.float64).ravel()
        if win.size != n:
            raise ValueError("bad window length")
    return np.fft.fftshift(win)


def fdk_filter_response(cg: CTGeom, npadh: int, window="ramp") -> np.ndarray:
    """H of fdk_filter.m:24-33 / fdk_fan_filter :41-63, times ds.  ``window``:
    'ramp', 'hann' (periodic Hann, ir_hann_periodic), any fbp2_window name,
    or an array of npadh samples ii = -npadh/2 .. npadh/2-1."""
    if math.isinf(cg.dsd) or math.isinf(cg.dfs):
        h = fbp_ramp("flat", npadh, cg.ds)
    elif cg.dfs == 0:
        h = fbp_ramp("arc", npadh, cg.ds, cg.dsd)
    else:
        raise ValueError("dfs not done")
    H = np.real(np.fft.fft(np.fft.fftshift(h)))
    if isinstance(window, str):
        if window == "ramp":
            win = np.ones(npadh)
        elif window == "hann":
            win = 0.5 * (1 - np.cos(2 * math.pi * np.arange(npadh) / npadh))   # periodic
        else:
            win = np.fft.fftshift(fbp2_window(npadh, window))
    else:
        win = np.asarray(window, dtype=np.float64).ravel()
        if win.size != npadh:
            raise ValueError("bad window length")
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


def filter_taps(cg: CTGeom, ns: int, window: str = "ramp") -> np.ndarray:
    """Spatial taps of fdk_filter's zero-padded filter, FP64, for the lags
    -(ns-1)..ns-1: the inverse DFT of the response (fdk_filter.m:24-33), tap
    for lag k at index k mod npad.  With npad = 2^ceil(log2(2ns-1)) the padded
    circular convolution of a row of ns samples is the product with the
    Toeplitz matrix of these taps."""
    npadh = 2 ** math.ceil(math.log2(2 * ns - 1))
    h = np.real(np.fft.ifft(fdk_filter_response(cg, npadh, window)))
    return h[np.arange(-(ns - 1), ns) % npadh]


def toeplitz_matrix(cg: CTGeom, ns: int, window: str = "ramp") -> np.ndarray:
    """The filter as the FP64 C-order ``(ns, ns)`` array ``T[s', s] = tap(s' - s)``:
    row filtering as a matrix product (a direct-convolution form of
    fdk_filter, used to check it)."""
    taps = filter_taps(cg, ns, window)
    k = np.arange(ns)
    return taps[(k[:, None] - k[None, :]) + ns - 1]


def padded_response(cg: CTGeom, ns: int, window: str = "ramp"):
    """(npad, H): fdk_filter's padded length 2^ceil(log2(2 ns - 1)) and its
    response times ds on all npad frequencies, H[k] = H[npad - k] as
    ``irfft`` extends the non-negative half (fdk_filter.m:21-35)."""
    npadh = 2 ** math.ceil(math.log2(2 * ns - 1))
    H = fdk_filter_response(cg, npadh, window)
    k = np.arange(npadh)
    return npadh, H[np.minimum(k, npadh - k) % npadh]


_FILTER_TABLES = {}


def _filter_tables(cg: CTGeom, ns, nt, window, w1cyl, device):
    import torch
    wkey = window if isinstance(window, str) else \
        ("array", np.asarray(window, dtype=np.float64).tobytes())
    key = (ns, nt, cg.ds, cg.dt, cg.dsd, cg.dso, cg.dfs, cg.offset_s, cg.offset_t, wkey,
           w1cyl, str(device))
    if key not in _FILTER_TABLES:
        ww1 = torch.as_tensor(weight_table(cg, ns, nt, w1cyl=w1cyl).T.copy(),  # [ns nt], s fastest
                              dtype=torch.float32, device=device)
        npad, H = padded_response(cg, ns, window)
        Hd = torch.as_tensor(H, dtype=torch.float32, device=device)
        _FILTER_TABLES.clear()
        _FILTER_TABLES[key] = (ww1, npad, Hd)
    return _FILTER_TABLES[key]


def weight_filter_views(proj, cg: CTGeom, *, window: str = "ramp", w1cyl: bool = True,
                        stream=None):
    """feldkamp_weight1 + fdk_filter on the device (fdkhip_fdk_filter_views):
    ``proj`` is a float32 device tensor ``[ns nt na]`` (MATLAB order, s
    fastest: torch shape ``(na, nt, ns)``); returns the filtered views as a
    float32 ``(na, ns, nt)`` tensor, i.e. ``[nt ns na]`` with t fastest, as
    the plans take them.  FP32 throughout (the weight and the taps are
    computed in FP64 on the host, then rounded)."""
    import ctypes as C
    import torch
    from . import _lib
    if not (isinstance(proj, torch.Tensor) and proj.is_cuda and proj.dtype == torch.float32
            and proj.dim() == 3 and proj.is_contiguous()):
        raise ValueError("proj must be a contiguous float32 device tensor (na, nt, ns)")
    na, nt, ns = proj.shape
    lib = _lib.load()
    ww1, npad, H = _filter_tables(cg, ns, nt, window, w1cyl, proj.device)
    s = stream if stream is not None else torch.cuda.current_stream(proj.device)
    with torch.cuda.stream(s):      # the output is allocated on the stream it is written on
        tsa = torch.empty((na, ns, nt), dtype=torch.float32, device=proj.device)
    P = lambda t: C.c_void_p(t.data_ptr())
    _lib.check(lib.fdkhip_fdk_filter_views(P(proj), ns, nt, na, P(ww1), P(H), npad,
                                           P(tsa), C.c_void_p(s.cuda_stream)))
    return tsa


def feldkamp(cg: CTGeom, ig: ImageGeom, proj, *, window="ramp", w1cyl=True,
             ia_skip=1, extrapolate_t=0, fmt="f32", math_mode="exact"):
    """feldkamp.m:1-98.  ``proj`` ``[ns nt na]``: a numpy array (host path,
    returns numpy ``[nx ny nz]``) or a device torch tensor (returns a device
    ``[nx ny nz]`` view; FP16/BF16 storage with ``fmt``).  ``window`` as
    fdk_filter takes it (fdk_filter_response); ``extrapolate_t`` replicates
    the first and last filtered rows that many times (cbct_back.m:95-104)."""
    if cg.pitch != 0 or np.any(cg.source_zs != 0):
        raise ValueError("sorry, helical CT unsupported")
    if isinstance(proj, np.ndarray):
        p = torch.as_tensor(proj, dtype=torch.float64)
        p = fdk_filter(feldkamp_weight1(p, cg, w1cyl=w1cyl), cg, window)
        return cbct_back(p.numpy().astype(np.float32), cg, ig, ia_skip=ia_skip,
                         extrapolate_t=extrapolate_t)
    e = int(extrapolate_t)
    if e > 0 and cg.offset_t != 0:
        raise ValueError("offset_t not yet done for extrapolation")
    # [ns nt na] device views (torch shape (ns, nt, na)) -> MATLAB memory order
    sta = proj[:, :, ::ia_skip].permute(2, 1, 0).contiguous().to(torch.float32)
    tsa = weight_filter_views(sta, cg, window=window, w1cyl=w1cyl)        # [nt ns na]
    if e > 0:   # cbct_back.m:95-104: e copies of the first and last filtered rows
        tsa = torch.cat([tsa[:, :, :1].expand(-1, -1, e), tsa,
                         tsa[:, :, -1:].expand(-1, -1, e)], dim=2).contiguous()
    plan = FdkPlan.from_geoms(cg, ig, ia_skip=ia_skip, extrapolate_t=e, fmt=fmt,
                              math_mode=math_mode)
    if fmt == "f32":
        img = plan(tsa)
    else:
        from .cbct_back import quantize_views
        q, sc = quantize_views(tsa, fmt)
        img = plan(q, sc)
    img = img * float(angle_scale(cg, ia_skip))
    return FdkPlan.image_xyz(img)
