"""Adjoint of the pair-speed map (VJP kernel) and gradients through backward-Euler
steps (2d4_amd/adjoint.py), checked against the transpose of the Jacobian the
reference assembles, against the oracle's closed form of d sigma/dy, and
against central finite differences of the library's own forward map.  The
reference's own step is covered by test_gpu_theta_adjoint.py."""
from importlib import import_module

import numpy as np
import pytest

from helpers import all_pairs, random_config

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
adj = import_module("2d4_amd.adjoint")


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def _signs(n):
    return np.where(np.arange(n) < n // 2, 1.0, -1.0)


def test_vjp_is_transpose_of_reference_jacobian(ctx, orc):
    # the VJP with the reference's d sigma against -M^T w, M = the reference's
    # V + diag(s) at h = 1 (the oracle's assembly), per entry within 1e-12 of
    # the sum of |w_i M_ij| it adds up
    n = 300
    xy = random_config(n, 41)
    w = np.random.default_rng(2).standard_normal(n)
    gx, gy, gt = ctx.vjp_t(_dev(xy), _dev(w), exact=False)
    _, _, M, _ = orc.jacobian_dense(xy, 1.0, 0.0, 0.0)
    ref = -(M.T @ w)
    scale = np.abs(M).T @ np.abs(w)
    err = np.abs(gx.cpu().numpy() - ref) / scale
    assert err.max() <= 1e-12, err.max()
    assert gt == pytest.approx(float(_signs(n) @ w), rel=1e-14, abs=1e-14)


def test_vjp_exact_matches_exact_dense_jacobian(ctx, pkg):
    n = 256
    xy = random_config(n, 43)
    w = np.random.default_rng(3).standard_normal(n)
    gx, _, _ = ctx.vjp_t(_dev(xy), _dev(w), exact=True)
    p = pkg.jac_params(h=1.0, n=n, exact=True)
    _, _, M, _ = ctx.jacobian_dense_t(_dev(xy), p)
    A = M.t().cpu().numpy()   # (row, col)
    ref = -(A.T @ w)
    scale = np.abs(A).T @ np.abs(w)
    assert (np.abs(gx.cpu().numpy() - ref) / scale).max() <= 1e-12


def test_vjp_y_part_matches_closed_form(ctx, orc):
    # dL/dy_j = sum_{i != j} (w_j - w_i) b_i b_j d sigma/dy (x_i - x_j, y_i - y_j), with
    # d sigma/dy from the oracle's closed form (differentiation through c = cos 2 pi y)
    n = 64
    xy = random_config(n, 47)
    w = np.random.default_rng(4).standard_normal(n)
    _, gy, _ = ctx.vjp_t(_dev(xy), _dev(w), exact=True)
    b = _signs(n)
    i, j, dx, dy = all_pairs(xy)
    e, c = orc.inputs(dx, dy)
    d = np.array([orc.L.orc_dyxy4(a, bb, ee, cc) for a, bb, ee, cc in zip(dx, dy, e, c)])
    term = (w[j] - w[i]) * b[i] * b[j] * d
    ref = np.zeros(n)
    np.add.at(ref, j, term)
    np.add.at(ref, i, -term)
    scale = np.zeros(n)
    np.add.at(scale, j, np.abs(term))
    np.add.at(scale, i, np.abs(term))
    assert (np.abs(gy.cpu().numpy() - ref) / scale).max() <= 1e-12


def test_pair_speed_gradcheck(ctx):
    # torch's finite-difference check of the autograd Function (both positions and the stress)
    n = 24
    x, y = adj.separated_config(n, 53, 0.05)
    x = _dev(x).requires_grad_()
    y = _dev(y).requires_grad_()
    tau = torch.tensor(0.1, dtype=torch.float64, device="cuda:0", requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b_, t: adj.PairSpeed.apply(a, b_, t, ctx, True), (x, y, tau),
                                    eps=1e-6, atol=1e-5, rtol=1e-4)


def test_implicit_step_solves_backward_euler(ctx):
    n = 128
    x, y = adj.separated_config(n, 59, 0.05)
    x, y = _dev(x), _dev(y)
    xn = adj.ImplicitStep.apply(x, y, torch.tensor(0.2, dtype=torch.float64, device="cuda:0"), 1e-4, ctx)
    v = ctx.speed_pass_t(torch.stack([xn, y], 1).contiguous(), 0.2)
    res = (xn - x - 1e-4 * v).abs().max().item()
    assert res <= adj.ImplicitStep.RESIDUAL_FLOOR * 1e-3


def _grad_and_fd(ctx, n, T, coords_x, coords_y, seed, checkpoint_every=16, dt=2e-5):
    xy = np.stack(adj.separated_config(n, seed, 0.5 / np.sqrt(n)), axis=1)
    tau0 = 0.2
    x0 = _dev(xy[:, 0]).requires_grad_()
    y = _dev(xy[:, 1]).requires_grad_()
    tau = torch.tensor(tau0, dtype=torch.float64, device="cuda:0", requires_grad=True)
    L = adj.accumulated_strain(x0, y, tau, T, dt, ctx, checkpoint_every=checkpoint_every, method="backward_euler")
    L.backward()

    def loss(xv, yv, tv):
        with torch.no_grad():
            return adj.accumulated_strain(_dev(xv), _dev(yv), tv, T, dt, ctx, checkpoint_every=0,
                                          method="backward_euler").item()

    h = 1e-6
    fd = []
    for k in coords_x:
        xp, xm = xy[:, 0].copy(), xy[:, 0].copy()
        xp[k] += h
        xm[k] -= h
        fd.append(("x", k, (loss(xp, xy[:, 1], tau0) - loss(xm, xy[:, 1], tau0)) / (2 * h), x0.grad[k].item()))
    for k in coords_y:
        yp, ym = xy[:, 1].copy(), xy[:, 1].copy()
        yp[k] += h
        ym[k] -= h
        fd.append(("y", k, (loss(xy[:, 0], yp, tau0) - loss(xy[:, 0], ym, tau0)) / (2 * h), y.grad[k].item()))
    fd.append(("tau", 0, (loss(xy[:, 0], xy[:, 1], tau0 + h) - loss(xy[:, 0], xy[:, 1], tau0 - h)) / (2 * h),
               tau.grad.item()))
    return fd, (x0.grad.clone(), y.grad.clone(), tau.grad.clone())


def test_strain_gradient_matches_finite_differences(ctx):
    # a trajectory on which every implicit step converges directly, so the map
    # is smooth where the differences are taken (no sub-step changes, no stops
    # at the field's residual floor)
    n, T = 64, 16
    rng = np.random.default_rng(7)
    cx = rng.choice(n, 8, replace=False)
    cy = rng.choice(n, 4, replace=False)
    adj.reset_stats()
    fd, _ = _grad_and_fd(ctx, n, T, cx, cy, 61)
    assert adj.stats["failures"] == 0 and adj.stats["floor_stops"] == 0, adj.stats
    scale = max(abs(g) for _, _, _, g in fd)
    for kind, k, f, g in fd:
        print(f"{kind}[{k}] fd {f:.9e} adjoint {g:.9e}", flush=True)
        assert abs(f - g) <= 1e-5 * max(abs(g), 1e-3 * scale), (kind, k, f, g)


def test_split_steps_where_newton_fails(ctx):
    # at dt = 1e-3 a step of these positions moves dislocations farther than they
    # are apart and Newton fails on the whole step (checked); the sub-stepped
    # step ends at a backward-Euler solution of its last sub-step
    n = 128
    x, y = adj.separated_config(n, 59, 0.05)
    x, y = _dev(x), _dev(y)
    tau = torch.tensor(0.2, dtype=torch.float64, device="cuda:0")
    with pytest.raises(adj.NewtonFailure):
        adj.ImplicitStep.apply(x, y, tau, 1e-3, ctx)
    xn = adj.implicit_step(x, y, tau, 1e-3, ctx)
    assert torch.isfinite(xn).all()
    # the same end point from 10 sub-stepped steps of 1e-4 agrees to first order in dt
    xr = x
    for _ in range(10):
        xr = adj.implicit_step(xr, y, tau, 1e-4, ctx)
    assert (xn - xr).abs().max().item() <= 0.25 * (xr - x).abs().max().item()


def test_checkpointing_does_not_change_gradients(ctx):
    n, T = 96, 40
    _, g16 = _grad_and_fd(ctx, n, T, [], [], 67, checkpoint_every=16)
    _, g0 = _grad_and_fd(ctx, n, T, [], [], 67, checkpoint_every=0)
    for a, b in zip(g16, g0):
        torch.testing.assert_close(a, b, rtol=1e-12, atol=0.0)
