"""Gradients through the pair-speed map and through implicit steps.

The reference computes no derivatives beyond its own Jacobian; the maps differentiated here are:

  * PairSpeed: v_i(x, y, tau) = sum_{j != i} b_i b_j sigma(x_i - x_j, y_i - y_j)
    + tau b_i, the speed pass of simulation.cpp:287-333 (b_i = +1 for the first
    half of the sorted order, simulation_data.cpp:265-270).  Its backward is the
    hand-written VJP kernel (d4_vjp_dev): for L = sum_i w_i v_i,
    dL/dx_j = sum_{i != j} (w_j - w_i) b_i b_j d sigma/dx (x_i - x_j, y_i - y_j),
    the same with d sigma/dy for dL/dy_j, and dL/dtau = sum_i b_i w_i.
  * ThetaHalfStep: one half of the reference's implicit step, the map the simulation runs
    (simulation.cpp:98-154).  At z, with k the half step: the pass gives the speeds v0, v1 at the
    half step's two times, V (k b_i b_j times the reference's own d sigma, Field.cpp:314-331, inside
    the storage radius), s = -V 1 and d = weight(s) (sim_utils.cpp:36-53);
    A = I + diag((1 + d)/2)(V + diag(s)) (simulation.cpp:516-522);
        w1 = z - A^-1 g1,  g1 = -k ((1 + d) v1 + (1 - d) v0) / 2,
        w2 = w1 - A^-1 g2, g2 = w1 - z - k ((1 + d) v(w1) + (1 - d) v0) / 2
    (calcGSolveAndUpdate, simulation.cpp:610-621).  A step of h is two half steps of h/2: stage II's
    matrix, rescaled from stage I's (simulation.cpp:546-608), is the pass at h/2, and stage I's large
    step only feeds the error estimate.  The backward pass is this map's exact vector-Jacobian
    product: two transposed solves with A, the VJP kernel at z and at w1, the derivative of weight(),
    and how A's entries move with the positions (d4_theta_pair_vjp_dev, the reference's d sigma
    differentiated by dual numbers).
  * theta_step: one accepted step of h.  The step sizes are taken as given: the reference chooses
    them from an error estimate (precision_handler.cpp); reference_schedule runs the reference's loop
    (the library's host Simulation) to record them, and gradients are those of the map with its step
    sequence held fixed.
  * ImplicitStep: a backward-Euler step x' = x + dt v(x', y, tau), solved by Newton's method with
    the dense Jacobian from the pair kernel, each Newton system by the library's GPU dense LU
    (d4_lu_factor_dense_dev); no reference counterpart (kept as an alternative map).  Its backward
    follows from the implicit function theorem: mu = (I + M)^{-T} lambda at the converged x',
    dL/dx = mu, dL/dy = dt (dv/dy)^T mu, dL/dtau = dt b^T mu.
  * implicit_step: ImplicitStep over dt, or -- where Newton does not converge, as near a pair so
    close that dt |dv/dx| >> 1 -- a run of shorter implicit sub-steps (halved on failure, doubled
    after success) that ends exactly at dt.
  * accumulated_strain: sum_i b_i (x_T,i - x_0,i) after T steps (the plastic strain the reference
    accumulates, simulation.cpp:167-171), with checkpoint/recompute every `checkpoint_every` steps;
    method "theta" (the reference's step, the default) or "backward_euler".
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


_WEIGHT_C = 0.412073  # sim_utils.cpp:47


def _weight_prime(s, wfunc):
    """d weight(s) / ds for sim_utils.cpp:36-53's weight functions (0 where s <= 0, where weight is 0)."""
    pos = s > 0
    sp = torch.where(pos, s, torch.zeros_like(s))
    if wfunc == "c":  # (s / (s + 1))^2
        g = 2.0 * sp / (sp + 1.0) ** 3
    elif wfunc == "p":  # s / (s + 1)
        g = 1.0 / (sp + 1.0) ** 2
    elif wfunc == "m":  # q / (q + 1.5), q = s + c s^2
        q = sp + _WEIGHT_C * sp * sp
        g = 1.5 * (1.0 + 2.0 * _WEIGHT_C * sp) / (q + 1.5) ** 2
    else:
        raise ValueError(f"unknown weight function {wfunc!r}")
    return torch.where(pos, g, torch.zeros_like(s))


class ThetaOptions:
    """The reference's pass settings of an implicit step: cutoff (from -u as SimulationData::updateCutOff,
    simulation_data.cpp:253-262; None: every pair), heaviside (-h) and the weight function (-w)."""

    def __init__(self, cutoff_mult=None, heaviside=False, wfunc="c"):
        self.cutoff_mult, self.heaviside, self.wfunc = cutoff_mult, bool(heaviside), wfunc

    def params(self, n, k, tau0, tau1):
        return _pkg().jac_params(h=k, tau_a=tau0, tau_b=tau1, n=n, cutoff_mult=self.cutoff_mult,
                                 heaviside=self.heaviside, wfunc=self.wfunc, exact=False)


class ThetaHalfStep(torch.autograd.Function):
    """w2 = H(z): one half step of the reference's implicit step (see the module docstring)."""

    @staticmethod
    def _pass(d4, z, y, p):
        """(v0, v1, Mdot, d) at z: Mdot[col, row] = V + diag(s), V as the reference stores it."""
        return d4.jacobian_dense_t(_xy(z, y), p, cutoff=True)

    @staticmethod
    def forward(ctx, z, y, tau0, tau1, k, d4, opts):
        n = z.numel()
        t0, t1 = float(tau0), float(tau1)
        p = opts.params(n, k, t0, t1)
        v0, v1, A, d = ThetaHalfStep._pass(d4, z, y, p)
        d4.dense_scale_t(A, d)  # A[col, row]: the tensor read row-major is A^T
        singular, _ = d4.lu_factor_dense_t(A, transpose=True)
        if singular:
            raise RuntimeError("singular implicit-step matrix")
        g1 = -k * ((1 + d) * v1 + (1 - d) * v0) / 2
        u = d4.lu_solve_dense_t(g1)
        w1 = z - u
        v2 = d4.speed_pass_t(_xy(w1, y), t1)
        g2 = w1 - z - k * ((1 + d) * v2 + (1 - d) * v0) / 2
        w = d4.lu_solve_dense_t(g2)
        ctx.save_for_backward(z, y, w1, u, w, v0, v1, v2, d)
        ctx.d4, ctx.k, ctx.opts, ctx.p = d4, k, opts, p
        return w1 - w

    @staticmethod
    def backward(ctx, wb2):
        z, y, w1, u, w, v0, v1, v2, d = ctx.saved_tensors
        d4, k, p = ctx.d4, ctx.k, ctx.p
        n = z.numel()
        _, _, Md, d_again = ThetaHalfStep._pass(d4, z, y, p)
        s = Md.diagonal().clone()  # s_i = -sum_j V_ij
        Mdot = Md.t()  # (row, col) view of V + diag(s)
        Mu, Mw = torch.mv(Mdot, u), torch.mv(Mdot, w)
        A = d4.dense_scale_t(Md, d_again)
        singular, _ = d4.lu_factor_dense_t(A, transpose=False)  # A^T: its solves give A^-T
        if singular:
            raise RuntimeError("singular implicit-step matrix")
        Dh = (1 + d) / 2
        r2 = -d4.lu_solve_dense_t(wb2.contiguous())  # adjoint of g2
        vb2 = -k * Dh * r2
        gx1, gy1, _ = d4.vjp_t(_xy(w1, y), vb2.contiguous(), exact=True, want_dy=True)
        wb1 = wb2 + r2 + gx1
        r1 = -d4.lu_solve_dense_t(wb1.contiguous())  # adjoint of g1
        vb1 = -k * Dh * r1
        vb0 = -k * (1 - d) / 2 * (r1 + r2)
        gx0, gy0, _ = d4.vjp_t(_xy(z, y), (vb0 + vb1).contiguous(), exact=True, want_dy=True)
        # the matrix: Abar = -(r1 u^T + r2 w^T); through diag((1 + d)/2), weight(s) and s = -V 1
        db = -k / 2 * ((v2 - v0) * r2 + (v1 - v0) * r1) - (r1 * Mu + r2 * Mw) / 2
        t = _weight_prime(s, ctx.opts.wfunc) * db - Dh * (r1 * u + r2 * w)
        gpx, gpy = d4.theta_pair_vjp_t(_xy(z, y), p, Dh * r1, Dh * r2, u, w, t)
        b = _signs(n, z.device)
        xb = wb1 - r2 + gx0 + gpx
        yb = gy1 + gy0 + gpy
        tb0 = (b * vb0).sum()
        tb1 = (b * (vb1 + vb2)).sum()
        return xb, yb, tb0, tb1, None, None, None


def theta_step(x, y, tau, h, d4, opts=None):
    """One accepted step of h of the reference's loop (simulation.cpp:88-171) at constant stress tau:
    two half steps of h/2."""
    opts = opts if opts is not None else ThetaOptions()
    x = ThetaHalfStep.apply(x, y, tau, tau, h / 2, d4, opts)
    return ThetaHalfStep.apply(x, y, tau, tau, h / 2, d4, opts)


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


def _segment(x, y, tau, steps, dt, d4, tol, max_iter, method, opts):
    """`steps` steps of dt, or the step sizes listed in `steps`."""
    sizes = steps if isinstance(steps, (list, tuple)) else [dt] * steps
    for h in sizes:
        if method == "theta":
            x = theta_step(x, y, tau, h, d4, opts)
        else:
            x = implicit_step(x, y, tau, h, d4, tol, max_iter)
    return x


def accumulated_strain(x0, y, tau, T, dt, d4=None, checkpoint_every=16, tol=1e-12, max_iter=30, method="theta",
                       opts=None, steps=None):
    """sum_i b_i (x_T,i - x_0,i) after T steps of dt -- or after the steps whose sizes `steps` lists (then T
    and dt are ignored; reference_schedule gives the reference's own) -- from sorted positions (x0, y) at
    stress tau (a 0-d tensor to differentiate by it): the reference's implicit step (method "theta", with
    ThetaOptions opts) or backward Euler ("backward_euler", Newton to tol).  Only every
    `checkpoint_every`-th state is kept for the backward pass; the steps in between are recomputed."""
    if method not in ("theta", "backward_euler"):
        raise ValueError(f"unknown method {method!r}")
    d4 = d4 if d4 is not None else _pkg().default_context()
    if not torch.is_tensor(tau):
        tau = torch.tensor(float(tau), dtype=torch.float64, device=x0.device)
    sizes = list(steps) if steps is not None else [dt] * T
    x = x0
    done = 0
    seg = checkpoint_every if checkpoint_every > 0 else len(sizes)
    while done < len(sizes):
        chunk = sizes[done:done + seg]
        if checkpoint_every > 0 and torch.is_grad_enabled():
            x = checkpoint(_segment, x, y, tau, chunk, dt, d4, tol, max_iter, method, opts, use_reentrant=False)
        else:
            x = _segment(x, y, tau, chunk, dt, d4, tol, max_iter, method, opts)
        done += len(chunk)
    return (_signs(x0.numel(), x0.device) * (x - x0)).sum()


def reference_schedule(x, y, tau, t_end, d4=None, opts=None, max_steps=1000000, progress=None, max_accepted=None,
                       **sim_params):
    """The step sizes the reference's loop accepts from sorted positions (x, y) at constant stress tau until
    its time passes t_end, or until it has accepted max_accepted steps: Simulation::run with its step
    doubling, error estimate and step-size control (simulation.cpp:88-276, precision_handler.cpp:34-63), run
    by the library's host Simulation (d4_sim, its passes on the GPU).  sim_params as 2d4_amd.sim (init_step_size, max_step_size, prec, dipole_prec, solver).
    Returns (sizes, final sorted positions of that run); accumulated_strain(..., steps=sizes) replays the
    accepted steps as a differentiable map.  progress(iterations, accepted, sim_time), if given, is called
    after every iteration."""
    import numpy as np
    simmod = import_module("2d4_amd.sim")
    d4 = d4 if d4 is not None else _pkg().default_context()
    opts = opts if opts is not None else ThetaOptions()
    xn = np.asarray(x.detach().cpu() if torch.is_tensor(x) else x, dtype=np.float64)
    yn = np.asarray(y.detach().cpu() if torch.is_tensor(y) else y, dtype=np.float64)
    n = xn.size
    b = np.where(np.arange(n) < n // 2, 1.0, -1.0)
    extra = {}
    if opts.cutoff_mult is not None:
        extra["cutoff_mult"] = float(opts.cutoff_mult)
    sim = simmod.Simulation.from_arrays(d4, xn, yn, b, heaviside=opts.heaviside, wfunc=opts.wfunc,
                                        stress=("fixed", float(tau), 0.0), time_limit=float(t_end), **extra,
                                        **sim_params)
    try:
        start = sim.positions()
        if not (np.array_equal(start[:, 0], xn) and np.array_equal(start[:, 1], yn)):
            raise ValueError("positions must be in the reference's sorted order (b = +1 first, y descending)")
        sizes = []
        iterations = 0
        while True:
            st = sim.stats()
            h = st["step_size"]  # the size this iteration tries
            accepted, stopped = sim.step()
            if stopped:
                break
            iterations += 1
            if accepted:
                sizes.append(h)
                if len(sizes) > max_steps:
                    raise RuntimeError(f"more than {max_steps} accepted steps before t = {t_end}")
            if progress is not None:
                progress(iterations, len(sizes), st["sim_time"] + (h if accepted else 0.0))
            if max_accepted is not None and len(sizes) >= max_accepted:
                break
        return sizes, sim.positions()
    finally:
        sim.close()


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
