"""The oracle's tabulated field (the build's own format) pinned by what the
mathematics fixes: node values returned exactly, bilinear functions reproduced,
periodic wrap, nodes equal to the analytic field."""
import ctypes as C

import numpy as np
import pytest


def _p(a):
    return C.c_void_p(a.ctypes.data)


def test_nodes_returned_exactly(orc):
    n = 32
    t = np.random.default_rng(1).standard_normal(n * n)
    for kx, ky in [(0, 0), (3, 5), (31, 31), (16, 0), (0, 17)]:
        assert orc.L.orc_table_bilinear(_p(t), n, -0.5 + kx / n, -0.5 + ky / n) == t[ky * n + kx]


def test_bilinear_functions_reproduced(orc):
    # f = a + b x + c y + d x y at the nodes; away from the wrapping cells the
    # interpolant equals f (to rounding)
    n = 64
    a, b, c, d = 0.3, -1.7, 2.2, 0.9
    k = np.arange(n)
    X, Y = np.meshgrid(-0.5 + k / n, -0.5 + k / n)
    t = np.ascontiguousarray((a + b * X + c * Y + d * X * Y).ravel())
    rng = np.random.default_rng(2)
    for _ in range(200):
        x, y = rng.uniform(-0.5, 0.5 - 1.0 / n, 2)
        got = orc.L.orc_table_bilinear(_p(t), n, x, y)
        assert got == pytest.approx(a + b * x + c * y + d * x * y, abs=1e-14)


def test_periodic_wrap(orc):
    n = 16
    t = np.random.default_rng(3).standard_normal(n * n)
    # between the last column (x = 1/2 - 1/n) and the first (x = -1/2 = 1/2)
    x = 0.5 - 0.5 / n
    left, right = t[2 * n + n - 1], t[2 * n + 0]
    assert orc.L.orc_table_bilinear(_p(t), n, x, -0.5 + 2 / n) == pytest.approx(0.5 * (left + right), abs=1e-15)


def test_build_nodes_are_the_field(orc):
    n = 16
    ts, td = np.empty(n * n), np.empty(n * n)
    orc.L.orc_table_build(n, _p(ts), _p(td))
    assert ts[(n // 2) * n + n // 2] == 0.0 and td[(n // 2) * n + n // 2] == 0.0  # the origin
    for kx, ky in [(0, 0), (3, 9), (15, 2), (8, 1)]:
        x, y = -0.5 + kx / n, -0.5 + ky / n
        assert ts[ky * n + kx] == orc.L.orc_xy(x, y)
        assert td[ky * n + kx] == orc.L.orc_dxy(x, y)


def test_table_speed_pass_sums(orc):
    # pair forces cancel (each added once and subtracted once), tau enters as tau * b
    n, tn = 40, 64
    ts = np.empty(tn * tn)
    orc.L.orc_table_build(tn, _p(ts), None)
    xy = np.random.default_rng(4).uniform(-0.5, 0.5, (n, 2))
    v0, v1 = np.empty(n), np.empty(n)
    orc.L.orc_speed_pass_table(n, _p(xy), _p(ts), tn, 0.0, _p(v0))
    orc.L.orc_speed_pass_table(n, _p(xy), _p(ts), tn, 0.25, _p(v1))
    assert abs(v0.sum()) <= 1e-13 * np.abs(v0).sum()
    b = np.where(np.arange(n) < n // 2, 1.0, -1.0)
    np.testing.assert_allclose(v1 - v0, 0.25 * b, atol=1e-14)
