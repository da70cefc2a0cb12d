"""Gradients through the pair-speed map and through implicit steps.

No counterpart in the reference (which has no derivatives beyond its own
Jacobian); the forward map differentiated here is defined by this module:

  * PairSpeed: v_i(x, y, tau) = sum_{j != i} b_i b_j sigma(x_i - x_j, y_i - y_j)
    + tau b_i, the speed pass of simulation.cpp:287-333 (b_i = +1 for the first
    half of the sorted order, simulation_data.cpp:265-270).  Its backward is the
    hand-written VJP kernel (d4_vjp_dev): for L = sum_i w_i v_i,
    dL/dx_j = sum_{i != j} (w_j - w_i) b_i b_j d sigma/dx (x_i - x_j, y_i - y_j),
    the same with d sigma/dy for dL/dy_j, and dL/dtau = sum_i b_i w_i.
  * ImplicitStep: one backward-Euler step x' = x + dt v(x', y, tau), solved by
    Newton's method with the dense Jacobian from the pair kernel
    (M = -dt dv/dx, so the Newton matrix is I + M), each Newton system solved
    by the library's GPU dense LU (d4_lu_factor_dense_dev).  Its backward follows from
    the implicit function theorem: mu = (I + M)^{-T} lambda at the converged x',
    dL/dx = mu, dL/dy = dt (dv/dy)^T mu, dL/dtau = dt b^T mu -- the last two
    from the VJP kernel.
  * implicit_step: ImplicitStep over dt, or -- where Newton does not converge,
    as near a pair so close that dt |dv/dx| >> 1 and a step moves dislocations
    farther than they are apart -- a run of shorter implicit sub-steps (halved
    on failure, doubled after success) that ends exactly at dt; each sub-step
    is its own ImplicitStep in the autograd graph, and the same sub-steps are
    taken again when a checkpointed segment is recomputed.
  * accumulated_strain: sum_i b_i (x_T,i - x_0,i) after T steps (the plastic
    strain the reference accumulates, simulation.cpp:713-720, summed over steps),
    with checkpoint/recompute every `checkpoint_every` steps.

Every derivative here is the true derivative of sigma (exact = True), not the
reference's Jacobian entry, which replaces sinh by cosh for the |shift| >= 3
images (Field.cpp:269-283); with exact = False the VJP kernel is the transpose
of the reference's own Jacobian instead.
"""
from importlib import import_module

import torch
from torch.utils.checkpoint import checkpoint


def _pkg():
    return import_module("2d4_amd")


def _signs(n, device):
    return torch.where(torch.arange(n, device=device) < n // 2, 1.0, -1.0).to(torch.float64)


def _xy(x, y):
    return torch.stack([x, y], dim=1).contiguous()


class PairSpeed(torch.autograd.Function):
    """v = speeds of (x, y) at external stress tau; gradients by the VJP kernel."""

    @staticmethod
    def forward(ctx, x, y, tau, d4, exact=True):
        v = d4.speed_pass_t(_xy(x, y), float(tau))
        ctx.save_for_backward(x, y)
        ctx.d4, ctx.exact = d4, exact
        return v

    @staticmethod
    def backward(ctx, gv):
        x, y = ctx.saved_tensors
        gx, gy, gt = ctx.d4.vjp_t(_xy(x, y), gv.contiguous(), exact=ctx.exact, want_dy=True)
        return gx, gy, torch.tensor(gt, dtype=torch.float64, device=x.device), None, None


def _newton_matrix(d4, x, y, tau, dt):
    """(v(x), A) with A = I + M, M = -dt dv/dx (true derivative), stored column-major as the
    kernel writes it: the tensor's [col, row] entry is A's (row, col)."""
    n = x.numel()
    p = _pkg().jac_params(h=dt, tau_a=tau, tau_b=tau, n=n, exact=True)
    vA, _, A, _ = d4.jacobian_dense_t(_xy(x, y), p)
    A.diagonal().add_(1.0)
    return vA, A


def _solve(d4, A, rhs, transpose=False):
    """A^-1 rhs (or A^-T rhs) with the library's GPU dense LU; A column-major as _newton_matrix
    returns it.  A singular matrix is a Newton failure."""
    singular, _ = d4.lu_factor_dense_t(A, transpose=not transpose)  # the tensor read row-major is A^T
    if singular:
        raise NewtonFailure("singular Newton matrix")
    return d4.lu_solve_dense_t(rhs)


class NewtonFailure(RuntimeError):
    pass


# counts since the last reset_stats(): implicit solves, Newton failures (each
# one shortens a sub-step), and solves that ended at the residual floor
stats = {"solves": 0, "failures": 0, "floor_stops": 0}


def reset_stats():
    for k in stats:
        stats[k] = 0


class ImplicitStep(torch.autograd.Function):
    """x' solving x' = x + dt v(x', y, tau) (backward Euler), by Newton's method."""

    # The reference's field is a truncated image sum: sigma jumps by ~2e-8 where
    # a pair's wrapped dx passes +-1/2 (d sigma/dx by ~2e-4), so the step
    # equation has a residual floor of about dt * 2e-8 per pair near that line.
    # Newton stops there once no damped step lowers the residual any more.
    RESIDUAL_FLOOR = 1e-7

    @staticmethod
    def forward(ctx, x, y, tau, dt, d4, tol=1e-12, max_iter=30):
        t = float(tau)
        # ... and below that, the rounding of the positions themselves
        floor = ImplicitStep.RESIDUAL_FLOOR * dt + 64 * 2.0 ** -52 * (1.0 + float(x.abs().max()))
        stats["solves"] += 1
        xn = x.clone()
        v, A = _newton_matrix(d4, xn, y, t, dt)
        G = xn - x - dt * v
        for it in range(max_iter):
            delta = _solve(d4, A, G)
            if not bool(torch.isfinite(delta).all()):
                raise NewtonFailure("non-finite Newton correction")
            if float(delta.abs().max()) <= tol:  # converged: the remaining error is O(|delta|^2)
                xn = xn - delta
                break
            # step halving on the residual keeps Newton from overshooting near
            # close pairs, where the speed field is steep
            g0 = float(torch.linalg.vector_norm(G))
            lam = 1.0
            stalled = False
            while True:
                trial = xn - lam * delta
                vt = d4.speed_pass_t(_xy(trial, y), t)  # speeds only; the matrix once a trial is taken
                Gt = trial - x - dt * vt
                if float(torch.linalg.vector_norm(Gt)) < g0:
                    break
                if lam < 1e-3:
                    if float(G.abs().max()) <= floor:
                        stalled = True  # at the field's own residual floor: converged
                        break
                    raise NewtonFailure("no Newton step reduces the residual")
                lam *= 0.5
            if stalled:
                stats["floor_stops"] += 1
                break
            xn, G = trial, Gt
            _, A = _newton_matrix(d4, xn, y, t, dt)
        else:
            raise NewtonFailure(f"Newton did not converge in {max_iter} iterations "
                                f"(last correction {float(delta.abs().max()):.3e})")
        ctx.save_for_backward(xn, y)
        ctx.d4, ctx.dt, ctx.tau = d4, dt, t
        return xn

    @staticmethod
    def backward(ctx, lam):
        xn, y = ctx.saved_tensors
        _, A = _newton_matrix(ctx.d4, xn, y, ctx.tau, ctx.dt)
        mu = _solve(ctx.d4, A, lam.contiguous(), transpose=True)
        _, gy, gt = ctx.d4.vjp_t(_xy(xn, y), mu.contiguous(), exact=True, want_dy=True)
        return mu, ctx.dt * gy, torch.tensor(ctx.dt * gt, dtype=torch.float64, device=xn.device), \
            None, None, None, None


def implicit_step(x, y, tau, dt, d4, tol=1e-12, max_iter=30, max_halvings=24):
    """One backward-Euler step of dt, taken as sub-steps where Newton fails.

    The sub-step starts at dt; a failed one is halved, a successful one lets the
    next be twice as long (never beyond what is left of dt).  Sub-steps are whole
    multiples of dt / 2^max_halvings, so they add up to dt exactly.  Where a
    perturbation of the input changes which sub-steps succeed, the map jumps;
    gradients are those of the map with its sub-steps held fixed."""
    units_total = 1 << max_halvings
    done, sub = 0, units_total  # progress and sub-step, in units of dt / 2^max_halvings
    while done < units_total:
        span = min(sub, units_total - done)
        try:
            x = ImplicitStep.apply(x, y, tau, dt * span / units_total, d4, tol, max_iter)
        except NewtonFailure:
            stats["failures"] += 1
            if span <= 1:
                raise
            sub = span // 2
            continue
        done += span
        sub = min(2 * span, units_total)
    return x


def _segment(x, y, tau, steps, dt, d4, tol, max_iter):
    for _ in range(steps):
        x = implicit_step(x, y, tau, dt, d4, tol, max_iter)
    return x


def accumulated_strain(x0, y, tau, T, dt, d4=None, checkpoint_every=16, tol=1e-12, max_iter=30):
    """sum_i b_i (x_T,i - x_0,i) after T backward-Euler steps of dt from sorted
    positions (x0, y) at stress tau (a 0-d tensor to differentiate by it).
    Only every `checkpoint_every`-th state is kept for the backward pass; the
    steps in between are recomputed."""
    d4 = d4 if d4 is not None else _pkg().default_context()
    if not torch.is_tensor(tau):
        tau = torch.tensor(float(tau), dtype=torch.float64, device=x0.device)
    x = x0
    done = 0
    seg = checkpoint_every if checkpoint_every > 0 else T
    while done < T:
        k = min(seg, T - done)
        if checkpoint_every > 0 and torch.is_grad_enabled():
            x = checkpoint(_segment, x, y, tau, k, dt, d4, tol, max_iter, use_reentrant=False)
        else:
            x = _segment(x, y, tau, k, dt, d4, tol, max_iter)
        done += k
    return (_signs(x0.numel(), x0.device) * (x - x0)).sum()


def separated_config(n, seed, min_dist, device=None):
    """n sorted positions, uniform in the periodic unit cell but no two closer
    than min_dist (rejection sampling; the reference's generator likewise
    redraws y values closer than its shortest distance, init_config_gen.cpp:262-295).
    Returns (x, y) as float64 tensors on `device` (numpy arrays if None)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    pts = np.empty((0, 2))
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 200 * n:
            raise ValueError(f"cannot place {n} points {min_dist} apart")
        p = rng.uniform(-0.5, 0.5, 2)
        if len(pts):
            d = pts - p
            d -= np.rint(d)
            if np.min(np.einsum("ij,ij->i", d, d)) < min_dist * min_dist:
                continue
        pts = np.vstack([pts, p])
    b = np.where(np.arange(n) % 2 == 1, 1.0, -1.0)
    pts = pts[np.argsort(-(pts[:, 1] + b), kind="stable")]
    if device is None:
        return pts[:, 0].copy(), pts[:, 1].copy()
    return (torch.tensor(pts[:, 0], dtype=torch.float64, device=device),
            torch.tensor(pts[:, 1], dtype=torch.float64, device=device))
