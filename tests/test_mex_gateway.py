"""The MATLAB mex gateway ``mirt_amd/mex/fdk_mex.c``.

MATLAB is not available here, so the gateway is compiled (gcc, as a MATLAB
``mex`` build compiles it) against the test stand-in of the mx API in
``tests/mex_stub`` and linked to ``libfdkhip.so``; the tests call its
``mexFunction`` with MATLAB-shaped operands (first index fastest).

* CPU: every operand check of the reference gateway, with its message
  (mex/src/fdk/fdk-orig/fdk,mex.c:39-91; command and operand counts
  :181-209), a failure the library reports (nthread > 250, cbct,mask2.c:738),
  and a well-formed call that passes every check and reaches
  ``fdkhip_fdk_ts_back`` (which, without a GPU, reports the HIP failure).
* GPU: well-formed 'fdk,ts,back' and 'fdk,st,back' calls through the
  gateway equal the reference C (oracle/_ref) bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, make_case

GATEWAY = os.path.join(ROOT, "mirt_amd", "mex", "fdk_mex.c")
STUB = os.path.join(ROOT, "tests", "mex_stub")
CLASS = {np.dtype(np.float64): 6, np.dtype(np.float32): 7, np.dtype(np.int8): 8,
         np.dtype(np.uint8): 9, np.dtype(np.int32): 12, np.dtype(np.uint32): 13}


class Gateway:
    """ctypes handle on the gateway + fake mx runtime."""

    def __init__(self, path):
        lib = C.CDLL(path)
        P = C.c_void_p
        lib.fmx_wrap.restype = P
        lib.fmx_wrap.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t), P]
        lib.fmx_string.restype = P
        lib.fmx_string.argtypes = [C.c_char_p]
        lib.fmx_free.argtypes = [P]
        lib.fmx_call.restype = C.c_int
        lib.fmx_call.argtypes = [C.c_int, C.POINTER(P), C.c_int, C.POINTER(P), C.c_char_p, C.c_int]
        lib.fmx_printed.restype = C.c_char_p
        lib.fmx_data.restype = P
        lib.fmx_data.argtypes = [P]
        lib.fmx_ndim.restype = C.c_int
        lib.fmx_ndim.argtypes = [P]
        lib.fmx_dim.restype = C.c_size_t
        lib.fmx_dim.argtypes = [P, C.c_int]
        lib.fmx_class.restype = C.c_int
        lib.fmx_class.argtypes = [P]
        self.lib = lib

    def _mx(self, x, keep):
        if isinstance(x, str):
            a = self.lib.fmx_string(x.encode())
            keep.append((a, None))
            return a
        cplx = np.iscomplexobj(x)
        arr = np.asfortranarray(x.real if cplx else x)
        if arr.ndim == 0:
            arr = arr.reshape(1, 1)
        if arr.ndim == 1:
            arr = arr.reshape(1, -1, order="F")   # a MATLAB row vector
        dims = (C.c_size_t * arr.ndim)(*arr.shape)
        a = self.lib.fmx_wrap(CLASS[arr.dtype], int(cplx), arr.ndim, dims,
                              arr.ctypes.data_as(C.c_void_p))
        keep.append((a, arr))
        return a

    def __call__(self, *args, nlhs=1):
        """Returns (error message or None, outputs as Fortran-order arrays)."""
        keep = []
        prhs = (C.c_void_p * max(len(args), 1))(*[self._mx(x, keep) for x in args])
        plhs = (C.c_void_p * max(nlhs, 1))()
        err = C.create_string_buffer(1024)
        rc = self.lib.fmx_call(nlhs, plhs, len(args), prhs, err, len(err))
        outs = []
        for i in range(nlhs if rc == 0 else 0):
            a = plhs[i]
            shape = tuple(self.lib.fmx_dim(a, k) for k in range(self.lib.fmx_ndim(a)))
            assert self.lib.fmx_class(a) == 7   # single
            n = int(np.prod(shape))
            buf = (C.c_float * n).from_address(self.lib.fmx_data(a))
            outs.append(np.array(np.frombuffer(buf, np.float32, n)).reshape(shape, order="F"))
            self.lib.fmx_free(a)
        for a, _ in keep:
            self.lib.fmx_free(a)
        return (err.value.decode() if rc else None), outs

    def printed(self) -> str:
        return self.lib.fmx_printed().decode()


@pytest.fixture(scope="module")
def gw(tmp_path_factory):
    from mirt_amd import build as B
    if not os.path.exists(B.LIB):
        pytest.skip("libfdkhip.so is not built")
    out = str(tmp_path_factory.mktemp("mex") / "libfdk_mex_gateway.so")
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-shared", "-fPIC",
           "-I", os.path.join(ROOT, "include"), "-I", STUB, GATEWAY,
           os.path.join(STUB, "fake_mx.c"), "-L", B.LIBDIR, "-lfdkhip",
           "-Wl,-rpath," + B.LIBDIR, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return Gateway(out)


def operands(g, proj, beta, mask, nthread=1, st=False):
    """The 12 operands of fdk_mex for a case (proj: [na ns nt] float32)."""
    p = proj.transpose(1, 2, 0) if st else proj.transpose(2, 1, 0)   # [ns nt na] / [nt ns na]
    return [np.array([g.nx, g.ny, g.nz], np.int32),
            np.array([g.dx, g.dy, g.dz]), np.array([g.offset_x, g.offset_y, g.offset_z]),
            np.ascontiguousarray(mask.T), np.float64(g.dso), np.float64(g.dsd),
            np.float64(g.dfs), np.array([g.ds, g.dt]), np.array([g.offset_s, g.offset_t]),
            np.asfortranarray(p), np.asarray(beta, np.float64), np.int32(nthread)]


@pytest.fixture(scope="module")
def case():
    return make_case("arc", seed=21, data="noise")


def test_help_and_usage(gw):
    assert gw(nlhs=0) == (None, [])                       # fdk_mex with no arguments: help
    assert "fdk,ts,back" in gw.printed()
    err, _ = gw("help")
    assert err == "fdk_mex: usage error. see above"       # fdk,mex.c:181-186
    err, _ = gw(np.float64(1), np.float64(2))
    assert err == "fdk_mex: usage error. see above"


def test_command_and_counts(gw, case):
    ops = operands(*case)
    assert gw("fdk,xx,back", *ops)[0] == "fdk_mex: command 'fdk,xx,back' unknown"
    assert gw("fdk,ts,back", *ops[:-1])[0] == "fdk_mex: usage error. see above"   # nrhs != 13
    assert gw("fdk,ts,back", *ops, nlhs=2)[0] == "fdk_mex: usage error. see above"
    assert gw("fdk,st,back", *ops, np.int32(1))[0] == "fdk_mex: usage error. see above"


# (operand index, replacement, message): each reference check of fdk,mex.c:39-91
def _bad_operands(g):
    return [
        (0, np.array([g.nx, g.ny, g.nz], np.float64), "mxIsInt32"),
        (0, np.array([g.nx, g.ny], np.int32), "mx_nxyz must have 3 elements"),
        (1, np.array([1, -1, 1], np.float32), "mxIsRealDouble"),
        (1, np.array([1, -1, 1], np.complex128), "mxIsRealDouble"),
        (1, np.array([1.0, -1.0]), "mx_dxyz must have 3 elements"),
        (2, np.array([0, 0, 0], np.int32), "mxIsRealDouble"),
        (2, np.zeros(4), "mx_cxyz must have 3 elements"),
        (3, np.ones((g.nx, g.ny)), "mxIsUint8"),
        (3, np.ones((g.nx, g.ny, 2), np.uint8), "mask must be 2d"),
        (3, np.ones((g.ny, g.nx), np.uint8), "mask size not nx by ny?"),
        (4, np.array([541.0, 541.0]), "mxIsScalarDouble"),
        (5, np.float32(949), "mxIsScalarDouble"),
        (6, np.int32(0), "mxIsScalarDouble"),
        (7, np.array([1.0, 1.0, 1.0]), "mx_dst must have 2 elements"),
        (7, np.array([1, 1], np.float32), "mxIsRealDouble"),
        (8, np.array([0.0]), "mx_ost must have 2 elements"),
        (9, np.zeros((4, 5, 3)), "mxIsRealSingle"),
        (9, np.zeros((4, 5, 3), np.complex64), "mxIsRealSingle"),
        (9, np.zeros((4, 5, 3, 2), np.float32), "mx_proj must be 2d or 3d"),
        (11, np.float64(1), "mxIsScalarInt32"),
        (11, np.array([1, 2], np.int32), "mxIsScalarInt32"),
    ]


@pytest.mark.parametrize("cmd", ["fdk,ts,back", "fdk,st,back"])
def test_operand_checks(gw, case, cmd):
    g = case[0]
    for i, bad, msg in _bad_operands(g):
        ops = operands(*case, st=cmd == "fdk,st,back")
        ops[i] = bad
        err, outs = gw(cmd, *ops)
        assert err == "fdk_mex: " + msg, (i, err)
        assert outs == []


def test_view_count_checks(gw, case):
    g, proj, beta, mask = case
    ops = operands(g, proj, beta, mask)
    na = proj.shape[0]
    ops[10] = np.asarray(beta[:-1], np.float64)
    assert gw("fdk,ts,back", *ops)[0] == \
        f"fdk_mex: beta length={na - 1} vs na=size(proj,3)={na}"
    ops[10] = np.asarray(beta, np.float32)
    assert gw("fdk,ts,back", *ops)[0] == "fdk_mex: mxIsRealDouble"
    # one [nt ns] view wants a scalar beta
    ops = operands(g, proj[:1], beta[:1], mask)
    ops[9] = np.asfortranarray(ops[9][:, :, 0])
    ops[10] = np.asarray(beta[:2], np.float64)
    assert gw("fdk,ts,back", *ops)[0] == "fdk_mex: mxIsScalarDouble"


def test_library_checks_reach_the_caller(gw, case):
    """Past the gateway's own checks, the library's message comes back
    (nthread > 250: "too many threads", cbct,mask2.c:738)."""
    err, _ = gw("fdk,ts,back", *operands(*case, nthread=300))
    assert err == "fdk_mex: too many threads"


def test_wellformed_call_reaches_the_library(gw, case):
    """Without a GPU the well-formed call passes every check and fails in the
    library's HIP set-up; with one it returns the [nz nx ny] image."""
    from conftest import gpu_available
    g = case[0]
    err, outs = gw("fdk,ts,back", *operands(*case))
    if gpu_available():
        assert err is None and outs[0].shape == (g.nz, g.nx, g.ny)
    else:
        assert err is not None and err.startswith("fdk_mex: ")
        checks = {m for _, _, m in _bad_operands(g)} | {"usage error. see above"}
        assert err[len("fdk_mex: "):] not in checks, err


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["arc", "flat", "par"])
def test_gateway_ts_back_equals_reference(gw, device, name):
    from oracle import oracle as O
    g, proj, beta, mask = make_case(name, seed=22, data="noise")
    print(f"mex gateway 'fdk,ts,back' case={name}", flush=True)
    err, outs = gw("fdk,ts,back", *operands(g, proj, beta, mask, nthread=4))
    assert err is None, err
    got = np.ascontiguousarray(outs[0].transpose(2, 1, 0))      # (ny, nx, nz)
    want = O.ref_back_ts("fixed", g, proj, beta, mask, nthread=4) if O.have_reference() \
        else O.back_ts_f32(g, proj, beta, mask)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_gateway_st_back_equals_reference(gw, device):
    from oracle import oracle as O
    g, proj, beta, mask = make_case("arc", seed=23, data="noise")
    print("mex gateway 'fdk,st,back'", flush=True)
    err, outs = gw("fdk,st,back", *operands(g, proj, beta, mask, st=True))
    assert err is None, err
    got = np.ascontiguousarray(outs[0].transpose(2, 1, 0))      # (nz, ny, nx)
    psta = np.ascontiguousarray(proj.transpose(0, 2, 1))       # [na nt ns] C
    want = O.ref_back_st("fixed", g, psta, beta, mask) if O.have_reference() \
        else O.back_st_f32(g, psta, beta, mask)
    assert np.array_equal(got, want)
