"""The reference's d sigma / dx with its derivatives in dx and dy by dual numbers
(2d4_amd/csrc/d4_dual.h, compiled for the host by tests/csrc/hostmath.hip), against the
oracle: the value against the oracle's restatement of Field::xy_diff_x (Field.cpp:314-331),
the derivatives against central differences of that same oracle function -- on both sides of
the close-pair branch (dx^2 + dy^2 = 1e-6) and of dx = 0, where the wall sum changes form.
The storage weight's derivative against central differences of its closed form."""
import ctypes as C
import os

import numpy as np
import pytest

HM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "libhostmath.so")


@pytest.fixture(scope="module")
def hm():
    if not os.path.exists(HM):
        pytest.skip("tests/csrc/libhostmath.so not built (__graft_entry__.build())")
    L = C.CDLL(HM)
    vp = C.c_void_p
    L.hm_dsigma_dual.argtypes = [C.c_long, vp, vp, vp, vp, vp]
    L.hm_storage_weight.argtypes = [C.c_long, vp, vp, C.c_double, C.c_int, vp, vp, vp]
    return L


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _dual(hm, dx, dy):
    dx, dy = np.ascontiguousarray(dx), np.ascontiguousarray(dy)
    v, gx, gy = (np.empty_like(dx) for _ in range(3))
    hm.hm_dsigma_dual(dx.size, _p(dx), _p(dy), _p(v), _p(gx), _p(gy))
    return v, gx, gy


def _orc(orc, dx, dy):
    return np.array([orc.L.orc_dxy(a, b) for a, b in zip(dx, dy)])


def _fd(orc, dx, dy, h):
    fx = (_orc(orc, dx + h, dy) - _orc(orc, dx - h, dy)) / (2 * h)
    fy = (_orc(orc, dx, dy + h) - _orc(orc, dx, dy - h)) / (2 * h)
    return fx, fy


def test_value_is_the_reference_dsigma(hm, orc):
    rng = np.random.default_rng(5)
    dx, dy = rng.uniform(-0.5, 0.5, 4000), rng.uniform(-0.5, 0.5, 4000)
    v, _, _ = _dual(hm, dx, dy)
    ref = _orc(orc, dx, dy)
    assert np.all(np.abs(v - ref) <= 1e-11 * np.abs(ref) + 1e-9), np.max(np.abs(v - ref) / np.abs(ref))


def test_derivatives_match_central_differences_far(hm, orc):
    # dx, dy away from the branch lines (r^2 = 1e-6, dx = 0) by more than the difference step
    rng = np.random.default_rng(6)
    dx, dy = rng.uniform(-0.49, 0.49, 3000), rng.uniform(-0.49, 0.49, 3000)
    keep = (np.abs(dx) > 1e-3) & (dx * dx + dy * dy > 4e-6)
    dx, dy = dx[keep], dy[keep]
    h = 1e-6 * np.maximum(np.hypot(dx, dy), 1e-2)
    _, gx, gy = _dual(hm, dx, dy)
    fx, fy = _fd(orc, dx, dy, h)
    # central differences: truncation O(h^2 f'''), cancellation ~1e-16 |f| / h
    scale = np.abs(gx) + np.abs(gy) + np.abs(_orc(orc, dx, dy)) / np.hypot(dx, dy)
    assert np.max(np.abs(gx - fx) / scale) < 1e-6
    assert np.max(np.abs(gy - fy) / scale) < 1e-6


def test_derivatives_match_central_differences_close_pairs(hm, orc):
    # the Taylor branch of the central wall (r^2 <= 1e-6), differentiated as the reference writes it
    rng = np.random.default_rng(7)
    r = rng.uniform(1e-4, 9e-4, 400)
    a = rng.uniform(0, 2 * np.pi, 400)
    dx, dy = r * np.cos(a), r * np.sin(a)
    keep = np.abs(dx) > 2e-5
    dx, dy, r = dx[keep], dy[keep], r[keep]
    h = 1e-4 * r
    _, gx, gy = _dual(hm, dx, dy)
    fx, fy = _fd(orc, dx, dy, h)
    # d sigma/dx is anisotropic (it vanishes along lines through the origin): the scale is the gradient's size
    scale = np.abs(gx) + np.abs(gy) + np.abs(_orc(orc, dx, dy)) / r
    assert np.max(np.abs(gx - fx) / scale) < 1e-6
    assert np.max(np.abs(gy - fy) / scale) < 1e-6


def test_reference_dsigma_is_not_exactly_even(hm):
    # the reference's d sigma / dx is even under (dx, dy) -> (-dx, -dy) only up to its truncated, cosh-for-sinh
    # outer walls (median 3e-6 relative, up to 2e-3 where the value is small): the step's adjoint therefore
    # differentiates each pair where the reference evaluates it (lower index first) instead of relying on
    # the symmetry
    rng = np.random.default_rng(8)
    dx, dy = rng.uniform(-0.49, 0.49, 1000), rng.uniform(-0.49, 0.49, 1000)
    v1, gx1, gy1 = _dual(hm, dx, dy)
    v2, gx2, gy2 = _dual(hm, -dx, -dy)
    rel = np.abs(v1 - v2) / np.abs(v1)
    assert rel.max() > 1e-4 and rel.max() < 1e-2


@pytest.mark.parametrize("heaviside", [0, 1])
def test_storage_weight_derivative(hm, heaviside):
    rc = 0.05
    rng = np.random.default_rng(9)
    r = rng.uniform(0.2 * rc, 6.9 * rc, 500)
    a = rng.uniform(0, 2 * np.pi, 500)
    dx, dy = np.ascontiguousarray(r * np.cos(a)), np.ascontiguousarray(r * np.sin(a))
    v, gx, gy = (np.empty_like(dx) for _ in range(3))
    hm.hm_storage_weight(dx.size, _p(dx), _p(dy), rc, heaviside, _p(v), _p(gx), _p(gy))
    closed = np.where((r <= rc) | bool(heaviside), 1.0, np.exp(-((r - rc) ** 2) / rc ** 2))
    assert np.allclose(v, closed, rtol=1e-12, atol=0)  # exp of an exponent up to 36: its rounding x 36
    h = 1e-7

    def w(x, y):
        rr = np.hypot(x, y)
        return np.where((rr <= rc) | bool(heaviside), 1.0, np.exp(-((rr - rc) ** 2) / rc ** 2))

    keep = np.abs(r - rc) > 1e-5
    fx = (w(dx + h, dy) - w(dx - h, dy)) / (2 * h)
    fy = (w(dx, dy + h) - w(dx, dy - h)) / (2 * h)
    assert np.max(np.abs(gx - fx)[keep]) < 1e-6
    assert np.max(np.abs(gy - fy)[keep]) < 1e-6
