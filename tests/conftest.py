"""Shared test setup: repo root on sys.path, the `gpu` marker, seeded cases."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    if not gpu_available():
        pytest.skip("no HIP device")
    import torch
    return torch.device("cuda:0")


# A small geometry set covering the detector models, offsets, a flipped y
# axis, odd sizes and a mask (seeded).  Each entry: kwargs for oracle.geom plus
# the number of views.
CASES = {
    "flat": dict(nxyz=(40, 36, 24), dxyz=(1.0, -1.1, 0.9), cxyz=(0.3, -0.2, 0.5),
                 dso=541.0, dsd=949.0, dfs=np.inf, ns=64, nt=40, dst=(1.3, 1.1),
                 ost=(1.25, 0.3), na=30),
    "arc": dict(nxyz=(40, 36, 24), dxyz=(1.0, -1.1, 0.9), cxyz=(0.3, -0.2, 0.5),
                dso=541.0, dsd=949.075, dfs=0.0, ns=64, nt=40, dst=(1.3, 1.1),
                ost=(1.25, 0.3), na=30),
    "par": dict(nxyz=(33, 31, 20), dxyz=(1.0, -1.0, 1.0), cxyz=(0.0, 0.0, 0.0),
                dso=np.inf, dsd=np.inf, dfs=0.0, ns=50, nt=24, dst=(1.0, 1.0),
                ost=(0.0, 0.0), na=24),
    # a tall volume: several z-groups, many slices out of the cone
    "arc_tall": dict(nxyz=(24, 24, 200), dxyz=(0.7, -0.7, 0.625), cxyz=(0, 0, 0),
                     dso=541.0, dsd=949.075, dfs=0.0, ns=96, nt=64,
                     dst=(1.0239, 1.0964), ost=(1.25, 0.0), na=40),
    # a volume wider than the field of view and a steep cone: within one 8x8
    # tile some columns miss a view while the slice ranges of the others
    # differ by more than a z-chunk (ranges that miss a chunk must add
    # nothing)
    "wide_cone": dict(nxyz=(48, 48, 160), dxyz=(6.0, -6.0, 0.5), cxyz=(0.5, 0, 0),
                      dso=541.0, dsd=949.0, dfs=np.inf, ns=96, nt=48, dst=(4.0, 1.0),
                      ost=(0.0, 0.0), na=24),
    # a volume inside the cone everywhere (all slices in support)
    "flat_full": dict(nxyz=(32, 32, 48), dxyz=(0.5, -0.5, 0.5), cxyz=(0, 0, 0),
                      dso=541.0, dsd=949.0, dfs=np.inf, ns=128, nt=96,
                      dst=(0.8, 0.8), ost=(0.0, 0.0), na=36),
}


def make_case(name: str, seed: int = 0, data: str = "noise", mask_frac: float = 0.1):
    """(geom, proj [na ns nt] float32, beta, mask [ny nx] uint8) for a case."""
    from oracle import oracle as O
    c = CASES[name]
    g = O.geom(c["nxyz"], c["dxyz"], c["cxyz"], c["dso"], c["dsd"], c["dfs"],
               c["ns"], c["nt"], c["dst"], c["ost"])
    rng = np.random.default_rng(seed)
    na, ns, nt = c["na"], c["ns"], c["nt"]
    if data == "noise":
        proj = rng.standard_normal((na, ns, nt)).astype(np.float32)
    else:
        s = np.linspace(-1, 1, ns)[None, :, None]
        t = np.linspace(-1, 1, nt)[None, None, :]
        a = np.arange(na)[:, None, None]
        proj = (np.cos(3 * s + 0.05 * a) * np.exp(-t ** 2)
                + 0.1 * np.sin(5 * t + 0.1 * a)).astype(np.float32)
    beta = np.linspace(0, 2 * np.pi, na, endpoint=False) + 0.1
    mask = (rng.random((g.ny, g.nx)) > mask_frac).astype(np.uint8)
    return g, proj, beta, mask
