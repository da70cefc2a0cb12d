"""The kernels' per-pair arithmetic (2d4_amd/csrc/d4_field.h), compiled for the
host by tests/csrc/hostmath.hip, against the oracle -- on the CPU, before any
GPU run: bound (A) of helpers.py with the same E, c."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import TOL, scale_dsigma, scale_sigma, wrap

HM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "libhostmath.so")


@pytest.fixture(scope="module")
def hm():
    if not os.path.exists(HM):
        pytest.skip("tests/csrc/libhostmath.so not built (__graft_entry__.build())")
    L = C.CDLL(HM)
    vp = C.c_void_p
    L.hm_field.argtypes = [C.c_long, vp, vp, vp, vp, vp, vp, C.c_int]
    L.hm_field_dy.argtypes = [C.c_long, vp, vp, vp]
    L.hm_wrap.restype = C.c_double
    L.hm_wrap.argtypes = [C.c_double]
    return L


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _points(golden, n=100000, seed=21):
    g = golden["field"]
    rng = np.random.default_rng(seed)
    dx = np.concatenate([g["dx"], rng.uniform(-0.5, 0.5, n)])
    dy = np.concatenate([g["dy"], rng.uniform(-0.5, 0.5, n)])
    ok = np.hypot(dx, dy) > 1e-200  # r = 0 is undefined (NaN) in the reference
    return dx[ok], dy[ok]


@pytest.mark.parametrize("exact", [False, True])
def test_kernel_math_matches_oracle(hm, orc, golden, exact):
    dx, dy = _points(golden)
    e, c = orc.inputs(dx, dy)
    s0, d0 = orc.field(dx, dy, e, c, exact=exact)
    n = dx.size
    s, d = np.empty(n), np.empty(n)
    hm.hm_field(n, _p(dx), _p(dy), _p(e), _p(c), _p(s), _p(d), 1 if exact else 0)
    S = scale_sigma(dx, dy, e, c)
    Sd = scale_dsigma(dx, dy, e, c)
    es = np.abs(s - s0) / np.maximum(np.abs(s0), S)
    ed = np.abs(d - d0) / np.maximum(np.abs(d0), Sd)
    assert es.max() <= TOL, es.max()
    assert ed.max() <= TOL, ed.max()
    # pure relative error: the bulk is at rounding level
    rel = np.abs(s - s0) / np.abs(s0)
    assert np.quantile(rel[np.isfinite(rel)], 0.999) < 1e-13


def test_kernel_dy_derivative_matches_oracle(hm, orc):
    rng = np.random.default_rng(22)
    dx = rng.uniform(-0.5, 0.5, 5000)
    dy = rng.uniform(-0.5, 0.5, 5000)
    keep = np.hypot(dx, dy) > 1e-3
    dx, dy = dx[keep], dy[keep]
    out = np.empty(dx.size)
    hm.hm_field_dy(dx.size, _p(dx), _p(dy), _p(out))
    e, c = orc.inputs(dx, dy)
    ref = np.array([orc.L.orc_dyxy4(a, b, ee, cc) for a, b, ee, cc in zip(dx, dy, e, c)])
    assert np.all(np.abs(out - ref) <= 1e-11 * (np.abs(ref) + scale_sigma(dx, dy, e, c) * 10))


def test_kernel_wrap_matches_reference_normalize(hm, golden):
    g = golden["utils"]
    out = np.array([hm.hm_wrap(v) for v in g["norm_in"]])
    assert np.array_equal(out, g["norm_out"])
    rng = np.random.default_rng(23)
    v = np.concatenate([rng.uniform(-3, 3, 20000), np.nextafter(np.arange(-4, 5) + 0.5, 0),
                        np.nextafter(np.arange(-4, 5) + 0.5, 1), np.arange(-4, 5) + 0.5])
    got = np.array([hm.hm_wrap(x) for x in v])
    want = np.array([wrap(x) for x in v])
    assert np.array_equal(got, want)
    assert np.all((got >= -0.5) & (got < 0.5))
