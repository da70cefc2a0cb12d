"""Gradients through the reference's implicit step (2d4_amd/adjoint.py: ThetaHalfStep, theta_step,
accumulated_strain(method="theta")), the map the simulation runs (simulation.cpp:98-171).

- the pair term of its backward (d4_theta_pair_vjp_dev) against a host restatement from the same
  dual-number header (tests/csrc/hostmath.hip; the header itself is pinned to the oracle on the CPU,
  tests/test_dual_cpu.py);
- the forward map against the oracle's accept-all loop (positions and strain);
- the full vector-Jacobian product of one half step by torch.autograd.gradcheck (every input, every
  output, finite differences of the library's own forward map);
- the strain gradient through several steps against central differences of the library's forward map
  and of the oracle's own loop, with and without the cutoff;
- checkpointing leaves the gradients unchanged."""
import ctypes as C
import os
from importlib import import_module

import numpy as np
import pytest

from helpers import all_pairs, random_config

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
adj = import_module("2d4_amd.adjoint")

HM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "libhostmath.so")
TAU = 0.2


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def _signs(n):
    return np.where(np.arange(n) < n // 2, 1.0, -1.0)


@pytest.fixture(scope="module")
def hm():
    if not os.path.exists(HM):
        pytest.skip("tests/csrc/libhostmath.so not built")
    L = C.CDLL(HM)
    vp = C.c_void_p
    L.hm_dsigma_dual.argtypes = [C.c_long, vp, vp, vp, vp, vp]
    L.hm_storage_weight.argtypes = [C.c_long, vp, vp, C.c_double, C.c_int, vp, vp, vp]
    return L


def _p(a):
    return C.c_void_p(a.ctypes.data)


OPTS = {
    "all_pairs": adj.ThetaOptions(),
    "taper": adj.ThetaOptions(cutoff_mult=0.8),  # rc = 0.8 / sqrt(n), Gaussian taper to 7 rc
    "heaviside": adj.ThetaOptions(cutoff_mult=1.6, heaviside=True),
}


@pytest.mark.parametrize("name", list(OPTS))
def test_pair_vjp_kernel_matches_host_restatement(ctx, hm, name):
    n = 150
    xy = random_config(n, 71)
    rng = np.random.default_rng(72)
    a, c, u, w, t = (rng.standard_normal(n) for _ in range(5))
    k = 1e-3
    p = OPTS[name].params(n, k, TAU, TAU)
    gx, gy = ctx.theta_pair_vjp_t(_dev(xy), p, *(_dev(v) for v in (a, c, u, w, t)))
    # host: each pair i < j once, at (x_i - x_j, y_i - y_j)
    i, j, dx, dy = all_pairs(xy)
    d2 = dx * dx + dy * dy
    rng2 = (1.0 if p.heaviside else 49.0) * p.cut_sq
    keep = d2 < rng2
    i, j, dx, dy = i[keep], j[keep], np.ascontiguousarray(dx[keep]), np.ascontiguousarray(dy[keep])
    v, mx, my = (np.empty_like(dx) for _ in range(3))
    hm.hm_dsigma_dual(dx.size, _p(dx), _p(dy), _p(v), _p(mx), _p(my))
    wv, wx, wy = (np.empty_like(dx) for _ in range(3))
    hm.hm_storage_weight(dx.size, _p(dx), _p(dy), p.cutoff, p.heaviside, _p(wv), _p(wx), _p(wy))
    b = _signs(n)
    P = -(a[i] * u[j] + a[j] * u[i] + c[i] * w[j] + c[j] * w[i]) - t[i] - t[j]
    kbb = k * b[i] * b[j]
    tx = kbb * (mx * wv + v * wx) * P
    ty = kbb * (my * wv + v * wy) * P
    rx, ry, sx, sy = (np.zeros(n) for _ in range(4))
    np.add.at(rx, i, tx)
    np.add.at(rx, j, -tx)
    np.add.at(ry, i, ty)
    np.add.at(ry, j, -ty)
    np.add.at(sx, i, np.abs(tx))
    np.add.at(sx, j, np.abs(tx))
    np.add.at(sy, i, np.abs(ty))
    np.add.at(sy, j, np.abs(ty))
    assert keep.sum() > 0
    assert np.all(np.abs(gx.cpu().numpy() - rx) <= 1e-12 * sx + 1e-300)
    assert np.all(np.abs(gy.cpu().numpy() - ry) <= 1e-12 * sy + 1e-300)


def _run_gpu(ctx, xy, T, dt, opts):
    x, y = _dev(xy[:, 0]), _dev(xy[:, 1])
    tau = torch.tensor(TAU, dtype=torch.float64, device="cuda:0")
    with torch.no_grad():
        xt = x
        for _ in range(T):
            xt = adj.theta_step(xt, y, tau, dt, ctx, opts)
    return xt.cpu().numpy(), float(_signs(xy.shape[0]) @ (xt - x).cpu().numpy())


def _run_oracle(orc, xy, T, dt, opts):
    kw = dict(dt0=dt, max_dt=dt, prec=1e10, dipole_prec=0.0, step_limit=T - 1, stress_kind=1, stress_init=TAU,
              stress_rate=0.0, cutoff_mult=opts.cutoff_mult if opts.cutoff_mult is not None else 1e20,
              heaviside=1 if opts.heaviside else 0, wfunc=opts.wfunc)
    out, st = orc.run(xy.copy(), **kw)
    assert st.ok_steps == T and st.failed_steps == 0
    return out, st.strain


@pytest.mark.parametrize("name", list(OPTS))
def test_theta_steps_match_oracle_loop(ctx, orc, name):
    n, T, dt = 64, 5, 1e-3
    xy = np.stack(adj.separated_config(n, 73, 0.5 / np.sqrt(n)), axis=1)
    xg, sg = _run_gpu(ctx, xy, T, dt, OPTS[name])
    xo, so = _run_oracle(orc, xy, T, dt, OPTS[name])
    d = xg - xo[:, 0]
    d -= np.rint(d)
    moved = np.abs(xg - xy[:, 0]).max()
    print(f"{name}: max |dx| {np.abs(d).max():.3e}, moved {moved:.3e}, strain {sg:.12e} vs {so:.12e}")
    assert moved > 1e-4
    # a different LU (threshold pivoting, blocked) and pair order: rounding, amplified by stiff close pairs
    assert np.abs(d).max() <= 1e-10
    assert abs(sg - so) <= 1e-10 * max(abs(so), 1e-3)


@pytest.mark.parametrize("name", ["all_pairs", "taper"])
def test_half_step_gradcheck(ctx, name):
    n = 12
    x, y = adj.separated_config(n, 79, 0.12)
    x = _dev(x).requires_grad_()
    y = _dev(y).requires_grad_()
    t0 = torch.tensor(TAU, dtype=torch.float64, device="cuda:0", requires_grad=True)
    t1 = torch.tensor(TAU + 0.05, dtype=torch.float64, device="cuda:0", requires_grad=True)
    opts = OPTS[name] if name == "all_pairs" else adj.ThetaOptions(cutoff_mult=0.9)

    def f(a, b_, s0, s1):
        return adj.ThetaHalfStep.apply(a, b_, s0, s1, 5e-3, ctx, opts)

    assert torch.autograd.gradcheck(f, (x, y, t0, t1), eps=1e-6, atol=1e-7, rtol=1e-5)


def _strain_grad(ctx, xy, T, dt, opts, checkpoint_every=16):
    x0 = _dev(xy[:, 0]).requires_grad_()
    y = _dev(xy[:, 1]).requires_grad_()
    tau = torch.tensor(TAU, dtype=torch.float64, device="cuda:0", requires_grad=True)
    L = adj.accumulated_strain(x0, y, tau, T, dt, ctx, checkpoint_every=checkpoint_every, method="theta", opts=opts)
    L.backward()
    return x0.grad.cpu().numpy(), y.grad.cpu().numpy(), tau.grad.item()


@pytest.mark.parametrize("name", ["all_pairs", "taper", "heaviside"])
def test_strain_gradient_matches_finite_differences(ctx, orc, name):
    n, T, dt = 48, 6, 1e-3
    opts = OPTS[name]
    xy = np.stack(adj.separated_config(n, 83, 0.5 / np.sqrt(n)), axis=1)
    gx, gy, gt = _strain_grad(ctx, xy, T, dt, opts, checkpoint_every=4)
    rng = np.random.default_rng(84)
    h = 1e-6

    def loss_gpu(a):
        with torch.no_grad():
            return _run_gpu(ctx, a, T, dt, opts)[1]

    def loss_orc(a, tau=TAU):
        kw = dict(dt0=dt, max_dt=dt, prec=1e10, dipole_prec=0.0, step_limit=T - 1, stress_kind=1, stress_init=tau,
                  stress_rate=0.0, cutoff_mult=opts.cutoff_mult if opts.cutoff_mult is not None else 1e20,
                  heaviside=1 if opts.heaviside else 0, wfunc=opts.wfunc)
        return orc.run(a.copy(), **kw)[1].strain

    # central differences at h and h/2, extrapolated (Richardson): a close pair gives this map a large third
    # derivative, and the plain h = 1e-6 difference is off by ~1e-5 relative at 6 steps (it converges to
    # the adjoint as h^2 when h shrinks: 1e-5, 1e-6, 1e-7 give errors 5e-3, 5e-5, 5e-7 on y[22])
    def rich(f, a, col, k, h):
        def d(hh):
            p_, m_ = a.copy(), a.copy()
            p_[k, col] += hh
            m_[k, col] -= hh
            return (f(p_) - f(m_)) / (2 * hh)
        return (4 * d(h / 2) - d(h)) / 3

    scale = max(np.abs(gx).max(), np.abs(gy).max())
    checks = [("x", k) for k in rng.choice(n, 6, replace=False)] + [("y", k) for k in rng.choice(n, 3, replace=False)]
    for kind, k in checks:
        col = 0 if kind == "x" else 1
        g = (gx if kind == "x" else gy)[k]
        f_gpu = rich(loss_gpu, xy, col, k, h)
        f_orc = rich(loss_orc, xy, col, k, h)
        print(f"{name} {kind}[{k}] adjoint {g:.9e} fd(gpu) {f_gpu:.9e} fd(oracle) {f_orc:.9e}", flush=True)
        tol = 1e-6 * max(abs(g), 1e-3 * scale)
        assert abs(f_gpu - g) <= tol, (kind, k)
        assert abs(f_orc - g) <= tol, (kind, k)
    ft = (4 * (loss_orc(xy, TAU + h / 2) - loss_orc(xy, TAU - h / 2)) / h
          - (loss_orc(xy, TAU + h) - loss_orc(xy, TAU - h)) / (2 * h)) / 3
    print(f"{name} tau adjoint {gt:.9e} fd(oracle) {ft:.9e}", flush=True)
    assert abs(ft - gt) <= 1e-6 * max(abs(gt), 1e-3)


def test_checkpointing_does_not_change_gradients(ctx):
    n, T, dt = 64, 20, 1e-3
    xy = np.stack(adj.separated_config(n, 89, 0.5 / np.sqrt(n)), axis=1)
    a = _strain_grad(ctx, xy, T, dt, OPTS["taper"], checkpoint_every=8)
    b = _strain_grad(ctx, xy, T, dt, OPTS["taper"], checkpoint_every=0)
    for u, v in zip(a, b):
        np.testing.assert_allclose(u, v, rtol=1e-12, atol=0.0)


def test_reference_schedule_replay(ctx):
    # the step sizes the reference's loop accepts (its error control, through the library's host
    # Simulation), replayed through the differentiable map: the same end point, and gradients of the
    # replayed map (schedule held fixed) against central differences of it
    n = 32
    xy = np.stack(adj.separated_config(n, 97, 0.5 / np.sqrt(n)), axis=1)
    t_end = 2e-3
    sizes, end = adj.reference_schedule(xy[:, 0], xy[:, 1], TAU, t_end, ctx, init_step_size=1e-5,
                                        max_step_size=1e-3)
    print(f"{len(sizes)} accepted steps, from {min(sizes):.3e} to {max(sizes):.3e}, total {sum(sizes):.6e}")
    assert len(sizes) >= 3 and sum(sizes) >= t_end
    x0 = _dev(xy[:, 0]).requires_grad_()
    y = _dev(xy[:, 1]).requires_grad_()
    tau = torch.tensor(TAU, dtype=torch.float64, device="cuda:0", requires_grad=True)
    L = adj.accumulated_strain(x0, y, tau, 0, 0.0, ctx, checkpoint_every=4, steps=sizes)
    L.backward()
    with torch.no_grad():
        xt = x0.detach()
        for h in sizes:
            xt = adj.theta_step(xt, y.detach(), tau.detach(), h, ctx)
    d = xt.cpu().numpy() - end[:, 0]
    d -= np.rint(d)
    print(f"replay vs the run: max |dx| {np.abs(d).max():.3e}")
    assert np.abs(d).max() <= 1e-9

    def loss(a):
        with torch.no_grad():
            return adj.accumulated_strain(_dev(a[:, 0]), _dev(a[:, 1]), TAU, 0, 0.0, ctx, checkpoint_every=0,
                                          steps=sizes).item()

    h = 1e-6
    g = x0.grad.cpu().numpy()
    for k in (3, 17, 29):
        def dd(hh):
            p_, m_ = xy.copy(), xy.copy()
            p_[k, 0] += hh
            m_[k, 0] -= hh
            return (loss(p_) - loss(m_)) / (2 * hh)
        f = (4 * dd(h / 2) - dd(h)) / 3
        print(f"x[{k}] adjoint {g[k]:.9e} fd {f:.9e}")
        assert abs(f - g[k]) <= 1e-6 * max(abs(g[k]), 1e-3 * np.abs(g).max())
