"""C5: the strain gradient through T = 256 implicit steps at N = 2048 uniform positions, tau = 0.2,
checkpoints every 16 steps, against central differences with h = 1e-6 on 32 random coordinates and on tau.

The steps are the reference's own (adjoint.theta_step: two half steps of simulation.cpp:98-154) with the
sizes its loop accepts from this start (adjoint.reference_schedule: Simulation::run's step doubling, error
estimate and step-size control, default precision, steps at most 1e-3).  C5's fixed dt = 1e-3 is not
differentiable from uniform positions (tests/test_oracle.py::test_c5_fixed_steps_from_uniform_positions_are_
not_differentiable): the reference's control takes steps of 1e-9 ... 1e-6 here, so 256 of them reach
t ~ 5e-5.  The gradient is that of the map with the step sequence held fixed, and so are the differences."""
from importlib import import_module

import numpy as np
import pytest

from helpers import c5_config, wrap

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
adj = import_module("2d4_amd.adjoint")

N, T, TAU, H, EVERY = 2048, 256, 0.2, 1e-6, 16
NCOORD, GROUPS = 32, 4


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


@pytest.fixture(scope="module")
def c5(ctx):
    xy = c5_config(N)
    x, y = xy[:, 0].copy(), xy[:, 1].copy()
    sizes, end = adj.reference_schedule(x, y, TAU, 1.0, ctx, max_step_size=1e-3, solver=1, max_accepted=T)
    x0 = _dev(x).requires_grad_()
    yy = _dev(y).requires_grad_()
    tau = torch.tensor(TAU, dtype=torch.float64, device="cuda:0", requires_grad=True)
    L = adj.accumulated_strain(x0, yy, tau, 0, 0.0, ctx, checkpoint_every=EVERY, steps=sizes)
    L.backward()
    d = xy[:, None, :] - xy[None, :, :]
    d -= np.rint(d)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    coords = np.random.default_rng(2049).choice(2 * N, NCOORD, replace=False)
    return dict(x=x, y=y, sizes=sizes, end=end, L=L.item(), gx=x0.grad.cpu().numpy(), gy=yy.grad.cpu().numpy(),
                gt=tau.grad.item(), nearest=r.min(axis=1), coords=coords)


def _loss(ctx, c5, x, y, tau=TAU):
    with torch.no_grad():
        return adj.accumulated_strain(_dev(x), _dev(y), tau, 0, 0.0, ctx, checkpoint_every=0,
                                      steps=c5["sizes"]).item()


def test_c5_schedule_and_replay(ctx, c5):
    """256 accepted steps of the reference's loop; the differentiable map replays them to the loop's end."""
    sizes = c5["sizes"]
    print(f"{len(sizes)} accepted steps from {min(sizes):.3e} to {max(sizes):.3e}, t = {sum(sizes):.4e}, "
          f"L = {c5['L']:.10e}")
    assert len(sizes) == T and max(sizes) <= 1e-3
    with torch.no_grad():
        xt = _dev(c5["x"])
        for h in sizes:
            xt = adj.theta_step(xt, _dev(c5["y"]), torch.tensor(TAU, dtype=torch.float64, device="cuda:0"), h, ctx)
    dev = np.abs(wrap(xt.cpu().numpy() - c5["end"][:, 0])).max()
    print(f"replay vs the loop's end: max |dx| {dev:.3e}")
    assert dev <= 1e-9
    assert np.isfinite(c5["gx"]).all() and np.isfinite(c5["gy"]).all() and np.isfinite(c5["gt"])


@pytest.mark.parametrize("group", range(GROUPS))
def test_c5_gradient_matches_central_differences(ctx, c5, group):
    """Eight of the 32 coordinates per group: d(h) = (L(z + h e_k) - L(z - h e_k)) / 2h against the adjoint.
    Where a pair of the trajectory comes close, the map's third derivative makes d(h) itself miss by more
    than the tolerance; there d(h/2) must move toward the adjoint at second order and the extrapolation
    (4 d(h/2) - d(h)) / 3 must meet it."""
    gx, gy = c5["gx"], c5["gy"]
    gmax = max(np.abs(gx).max(), np.abs(gy).max())
    per = NCOORD // GROUPS
    x, y = c5["x"], c5["y"]

    def central(col, k, h):
        p_, m_ = (x if col == 0 else y).copy(), (x if col == 0 else y).copy()
        p_[k] += h
        m_[k] -= h
        if col == 0:
            return (_loss(ctx, c5, p_, y) - _loss(ctx, c5, m_, y)) / (2 * h)
        return (_loss(ctx, c5, x, p_) - _loss(ctx, c5, x, m_)) / (2 * h)

    for kk in c5["coords"][group * per:(group + 1) * per]:
        col, k = divmod(int(kk), N)
        g = (gx if col == 0 else gy)[k]
        tol = 1e-5 * max(abs(g), 1e-3 * gmax)
        fd = central(col, k, H)
        msg = (f"{'xy'[col]}[{k}] nearest {c5['nearest'][k]:.2e}: adjoint {g:.10e} fd {fd:.10e} "
               f"|diff| {abs(fd - g):.2e} (tol {tol:.2e})")
        if abs(fd - g) <= tol:
            print(msg)
            continue
        fd2 = central(col, k, H / 2)
        rich = (4 * fd2 - fd) / 3
        print(msg + f"; h/2: |diff| {abs(fd2 - g):.2e}, extrapolated |diff| {abs(rich - g):.2e}")
        assert abs(fd2 - g) <= abs(fd - g) / 2.5, (col, k)
        assert abs(rich - g) <= tol, (col, k)


def test_c5_stress_gradient_matches_central_difference(ctx, c5):
    fd = (_loss(ctx, c5, c5["x"], c5["y"], TAU + H) - _loss(ctx, c5, c5["x"], c5["y"], TAU - H)) / (2 * H)
    print(f"tau: adjoint {c5['gt']:.10e} fd {fd:.10e}")
    assert abs(fd - c5["gt"]) <= 1e-5 * abs(c5["gt"])
