"""Tabulated-field passes (1024 x 1024 FP64 tables, bilinear) against the oracle's
bilinear on the same table."""
import numpy as np
import pytest

from helpers import all_pairs, pair_allowance, random_config

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


@pytest.fixture(scope="module")
def tables(ctx):
    ts, td = ctx.table_build_t(1024)
    return ts, td, ts.cpu().numpy().ravel().copy(), td.cpu().numpy().ravel().copy()


def test_table_nodes_match_the_field(ctx, orc, tables):
    # GPU-built nodes against the oracle's field at the same points, per node within bound (B)
    _, _, ts, td = tables
    n = 1024
    rng = np.random.default_rng(5)
    k = rng.integers(0, n * n, 4000)
    x, y = -0.5 + (k % n) / n, -0.5 + (k // n) / n
    keep = (x != 0) | (y != 0)
    k, x, y = k[keep], x[keep], y[keep]
    e, c = orc.inputs(x, y)
    s_ref, d_ref = orc.field(x, y, e, c)
    a_s, _ = pair_allowance(orc, x, y, e, c)
    a_d, _ = pair_allowance(orc, x, y, e, c, deriv=True)
    assert np.all(np.abs(ts[k] - s_ref) <= a_s)
    assert np.all(np.abs(td[k] - d_ref) <= a_d)


def test_table_eval_bit_exact(ctx, orc, tables):
    tsd, tdd, ts, td = tables
    rng = np.random.default_rng(6)
    dx = rng.uniform(-1.2, 1.2, 5000)
    dy = rng.uniform(-1.2, 1.2, 5000)
    s, d = ctx.table_eval_t(_dev(dx), _dev(dy), tsd, tdd)
    wx = np.array([orc.L.orc_normalize(v) for v in dx])
    wy = np.array([orc.L.orc_normalize(v) for v in dy])
    ref_s = np.array([orc.L.orc_table_bilinear(ts.ctypes.data, 1024, a, b) for a, b in zip(wx, wy)])
    ref_d = np.array([orc.L.orc_table_bilinear(td.ctypes.data, 1024, a, b) for a, b in zip(wx, wy)])
    np.testing.assert_array_equal(s.cpu().numpy(), ref_s)
    np.testing.assert_array_equal(d.cpu().numpy(), ref_d)


def test_table_speed_pass(ctx, orc, tables):
    tsd, _, ts, _ = tables
    n = 500
    xy = random_config(n, 7)
    v = ctx.speed_pass_table_t(_dev(xy), tsd, 0.1).cpu().numpy()
    ref = np.empty(n)
    orc.L.orc_speed_pass_table(n, xy.ctypes.data, ts.ctypes.data, 1024, 0.1, ref.ctypes.data)
    i, j, dx, dy = all_pairs(xy)
    f = np.array([orc.L.orc_table_bilinear(ts.ctypes.data, 1024, a, b) for a, b in zip(dx, dy)])
    scale = np.zeros(n)
    np.add.at(scale, i, np.abs(f))
    np.add.at(scale, j, np.abs(f))
    assert np.all(np.abs(v - ref) <= 1e-13 * scale + 1e-15)


def test_table_dense_jacobian(ctx, orc, pkg, tables):
    tsd, tdd, ts, td = tables
    n = 300
    xy = random_config(n, 8)
    p = pkg.jac_params(h=1e-3, tau_a=0.0, tau_b=0.1, n=n)
    vA, vB, M, d = ctx.jacobian_dense_table_t(_dev(xy), p, tsd, tdd)
    A = M.t().cpu().numpy()  # (row, col) = V + diag(s)
    v0 = ctx.speed_pass_table_t(_dev(xy), tsd, 0.0).cpu().numpy()
    np.testing.assert_allclose(vA.cpu().numpy(), v0, rtol=0, atol=1e-12)
    b = np.where(np.arange(n) < n // 2, 1.0, -1.0)
    i, j, dx, dy = all_pairs(xy)
    dd = np.array([orc.L.orc_table_bilinear(td.ctypes.data, 1024, a, c) for a, c in zip(dx, dy)])
    V = np.zeros((n, n))
    V[i, j] = 1e-3 * b[i] * b[j] * dd
    V[j, i] = V[i, j]
    np.testing.assert_array_equal(A[i, j], V[i, j])
    np.testing.assert_array_equal(A[j, i], V[j, i])
    s_ref = -V.sum(axis=1)
    assert np.all(np.abs(np.diag(A) - s_ref) <= 1e-13 * np.abs(V).sum(axis=1) + 1e-300)
