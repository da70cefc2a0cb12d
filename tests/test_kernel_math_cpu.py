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
    L.hm_div_by_reciprocal.argtypes = [C.c_long, vp, vp, vp]
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


def test_kernel_transcendentals_within_two_ulp(hm, orc):
    """exp2_pair / cos2pi_reduced (d4_field.h) against the libm values the reference uses:
    E within 2 ulp of exp2(PI2LN2INV dx), R the correctly rounded 1/E, and c within 2 ulp of
    cos(2 pi dy) in the units that matter to the field (c enters only as C - c and C c - 1):
    relative for |c| >= 1/2, absolute below (3 ulp(1/2): 1 - 2 sin^2 near c = 0)."""
    rng = np.random.default_rng(24)
    e_ = np.nextafter(0.5, 0)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, 200000), [-0.5, e_, 0.0, -1e-300, 0.25, -0.25]])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, 200000), [-0.5, e_, 0.0, 1e-300, 0.25, -0.25]])
    n = dx.size
    E, R, c = np.empty(n), np.empty(n), np.empty(n)
    hm.hm_inputs.argtypes = [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    hm.hm_inputs(n, _p(dx), _p(dy), _p(E), _p(R), _p(c))
    e0, c0 = orc.inputs(dx, dy)
    assert np.all(np.abs(E - e0) <= 2 * np.spacing(e0))
    assert np.array_equal(R, 1.0 / E)
    assert np.all(np.abs(c - c0) <= 3 * np.spacing(np.maximum(np.abs(c0), 0.5)))
    # the median error is at rounding level
    assert np.median(np.abs(E - e0) / np.spacing(e0)) <= 0.5


def test_kernel_own_inputs_within_bound_b(hm, orc):
    """The kernels' own exp2/cos/1-over-E feeding their field stay inside bound (B) of helpers.py,
    near pairs (where C0 - c cancels) included."""
    from helpers import pair_allowance
    rng = np.random.default_rng(31)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, 100000), rng.uniform(-2e-3, 2e-3, 4000)])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, 100000), rng.uniform(-2e-3, 2e-3, 4000)])
    n = dx.size
    s, d = np.empty(n), np.empty(n)
    hm.hm_field(n, _p(dx), _p(dy), None, None, _p(s), _p(d), 0)
    e, c = orc.inputs(dx, dy)
    a, s0 = pair_allowance(orc, dx, dy, e, c)
    ad, d0 = pair_allowance(orc, dx, dy, e, c, deriv=True)
    assert np.all(np.abs(s - s0) <= a)
    assert np.all(np.abs(d - d0) <= ad)


def test_division_by_correctly_rounded_reciprocal_with_one_correction_is_ieee_division(hm):
    """The triangular solves' U step, q = v r (r = 1/u), then q + (v - q u) r: the same quotient as v / u,
    bit for bit, over wide exponent ranges, divisors near 1 (the step matrices' diagonal) and mantissas of
    every pattern."""
    rng = np.random.default_rng(77)
    n = 2_000_000

    def spread(m, emin, emax):
        mant = rng.uniform(1.0, 2.0, m) * rng.choice([-1.0, 1.0], m)
        return np.ldexp(mant, rng.integers(emin, emax + 1, m))

    v = np.concatenate([spread(n // 2, -60, 60), rng.uniform(-4, 4, n // 2)])
    u = np.concatenate([spread(n // 4, -30, 30), 1.0 + spread(n // 4, -50, -1),
                        rng.uniform(0.5, 2.0, n // 4), spread(n // 4, -3, 3)])
    rng.shuffle(u)
    out = np.empty(n)
    hm.hm_div_by_reciprocal(n, _p(v), _p(u), _p(out))
    assert np.array_equal(out, v / u)
