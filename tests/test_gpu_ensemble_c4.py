"""The C4 workload itself: 256 realizations, N_r = 2 U{128..2048} (even N in [280, 4088] for this seed),
uniform positions, heaviside cutoff 0.05 (-u 0.05 sqrt(N) -h), tau = 0.05 (-r 0 -s 0.05), default
precision control, the GPU LU -- the ensemble bench.py measures -- for two lockstep iterations.

Checked at full size: every realization's invariants (y untouched, x in [-1/2, 1/2), the accumulated
strain equal to sum_i b_i (x_end - x_0) modulo whole cells, one step taken per iteration); eight
sampled realizations, the smallest and the largest among them, bit for bit against the plain
one-realization path; and the smallest realization's first step against the oracle's loop (orc_run)."""
from importlib import import_module

import numpy as np
import pytest

from helpers import wrap

pytestmark = pytest.mark.gpu

ens = import_module("2d4_amd.ensemble")
simmod = import_module("2d4_amd.sim")
dconf = import_module("2d4_amd.dconf")


def _c4(R=256, seed=4242):
    rng = np.random.default_rng(seed)
    sizes = 2 * rng.integers(128, 2049, R)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    T = int(off[-1])
    x = rng.uniform(-0.5, 0.5, T)
    y = rng.uniform(-0.5, 0.5, T)
    b = np.concatenate([np.where(np.arange(n) % 2 == 0, -1.0, 1.0) for n in sizes])
    params = [dict(cutoff_mult=0.05 * float(np.sqrt(n)), heaviside=True, stress=("fixed", 0.05, 0.0), solver=1)
              for n in sizes]
    return sizes, off, x, y, b, params


def test_c4_ensemble_two_iterations(ctx, orc):
    sizes, off, x, y, b, params = _c4()
    assert sizes.min() == 280 and sizes.max() == 4088 and len(sizes) == 256
    print(f"[c4] R=256 sum N={off[-1]}", flush=True)
    e = ens.Ensemble(ctx, off, x, y, b, params)
    assert e.step(2) == 256
    pos = e.positions()
    st = e.stats()
    e.close()
    for r, n in enumerate(sizes):
        s = slice(off[r], off[r + 1])
        x0 = dconf.sort_order(y[s], b[s])
        start = np.stack([x[s][x0], y[s][x0]], axis=1)
        p = pos[s]
        np.testing.assert_array_equal(p[:, 1], start[:, 1])
        assert np.all((p[:, 0] >= -0.5) & (p[:, 0] < 0.5))
        bs = np.where(np.arange(n) < n // 2, 1.0, -1.0)
        gap = st[r]["strain"] - (bs * wrap(p[:, 0] - start[:, 0])).sum()
        assert abs(gap - round(gap)) < 1e-9, (r, gap)
        assert st[r]["successful_steps"] + st[r]["failed_steps"] == 2
    # bit for bit against the plain path on eight realizations, the smallest and largest among them
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 0, 50, 101, 150, 200, 255})
    for r in pick:
        s = slice(off[r], off[r + 1])
        print(f"[c4] plain realization {r} n={sizes[r]}", flush=True)
        sim = simmod.Simulation.from_arrays(ctx, x[s], y[s], b[s], **params[r])
        sim.step()
        sim.step()
        np.testing.assert_array_equal(sim.positions(), pos[s])
        pst = sim.stats()
        sim.close()
        for k in ("sim_time", "step_size", "strain", "successful_steps", "failed_steps"):
            assert pst[k] == st[r][k], (r, k)
    # the smallest realization's first iteration against the oracle's loop (dense LU, the same step)
    r = int(np.argmin(sizes))
    s = slice(off[r], off[r + 1])
    o = dconf.sort_order(y[s], b[s])
    xy0 = np.stack([x[s][o], y[s][o]], axis=1)
    kw = dict(cutoff_mult=params[r]["cutoff_mult"], heaviside=1, stress_kind=1, stress_init=0.05, stress_rate=0.0,
              step_limit=0, time_limit=0.0)
    sim = simmod.Simulation.from_arrays(ctx, x[s], y[s], b[s], **params[r])
    for _ in range(6):  # iterations up to the first accepted step, as orc_run with step_limit 0 runs them
        acc, _ = sim.step()
        if acc:
            break
    one = sim.positions()
    gst = sim.stats()
    sim.close()
    ref, rst = orc.run(xy0, **kw)
    dev = dconf.max_periodic_dx(one, ref)
    print(f"[c4] smallest realization n={sizes[r]}: first accepted step ({gst['failed_steps']} rejected before) "
          f"vs the oracle's loop {dev:.3e}", flush=True)
    assert (rst.ok_steps, rst.failed_steps) == (gst["successful_steps"], gst["failed_steps"]) == (1, rst.failed_steps)
    assert dev <= 1e-9


def test_c4_device_step_matches_host_driven_ensemble(ctx):
    # the same workload as one captured graph per lockstep iteration (include/d4_dstep.h: dense step matrices,
    # the LU in a fixed order by |y| inside the band that order fixes for the heaviside radius) against the
    # host-driven ensemble (CSC, per-pattern reverse Cuthill-McKee order): another LU order, so the same
    # accept/reject decisions and positions and strains to rounding, not bit for bit
    dstep = import_module("2d4_amd.dstep")
    sizes, off, x, y, b, params = _c4()
    e = ens.Ensemble(ctx, off, x, y, b, params)
    e.step(2)
    pos_h, st_h = e.positions(), e.stats()
    e.close()
    print("[c4] device step", flush=True)
    s = dstep.DeviceSimulation(ctx, off, x, y, b, [{k: v for k, v in p.items() if k != "solver"} for p in params])
    s.run(2)
    pos_d, st_d = s.positions(), s.stats()
    s.close()
    worst = 0.0
    for r in range(len(sizes)):
        sl = slice(off[r], off[r + 1])
        assert (st_d[r]["successful_steps"], st_d[r]["failed_steps"]) == \
            (st_h[r]["successful_steps"], st_h[r]["failed_steps"]), r
        assert st_d[r]["sim_time"] == pytest.approx(st_h[r]["sim_time"], rel=1e-9, abs=1e-300)
        assert st_d[r]["strain"] == pytest.approx(st_h[r]["strain"], rel=1e-8, abs=1e-12)
        np.testing.assert_array_equal(pos_d[sl, 1], pos_h[sl, 1])
        worst = max(worst, float(np.abs(wrap(pos_d[sl, 0] - pos_h[sl, 0])).max()))
    print(f"[c4] device step vs host-driven ensemble: max |dx| {worst:.3e}", flush=True)
    assert worst <= 1e-9


def test_c4_device_step_200_iterations(ctx):
    """C4 as stated: 200 lockstep iterations of all 256 realizations through the device-resident step (one
    captured graph per iteration). Every realization takes 200 loop iterations (accepted + failed), keeps y,
    keeps x normalised, and its strain is sum_i b_i (x_end - x_0) modulo whole cells; the factorizations are
    three per iteration, the rare ones with row interchanges included."""
    dstep = import_module("2d4_amd.dstep")
    sizes, off, x, y, b, params = _c4()
    s = dstep.DeviceSimulation(ctx, off, x, y, b, [{k: v for k, v in q.items() if k != "solver"} for q in params])
    for k in range(10):
        s.run(20)
        print(f"[c4 device] {20 * (k + 1)} iterations", flush=True)
    pos = s.positions()
    st = s.stats()
    s.close()
    for r, n in enumerate(sizes):
        sl = slice(off[r], off[r + 1])
        o = dconf.sort_order(y[sl], b[sl])
        start = np.stack([x[sl][o], y[sl][o]], axis=1)
        p = pos[sl]
        np.testing.assert_array_equal(p[:, 1], start[:, 1])
        assert np.all((p[:, 0] >= -0.5) & (p[:, 0] < 0.5))
        bs = np.where(np.arange(n) < n // 2, 1.0, -1.0)
        gap = st[r]["strain"] - (bs * wrap(p[:, 0] - start[:, 0])).sum()
        assert abs(gap - round(gap)) < 1e-9, (r, gap)
        assert st[r]["successful_steps"] + st[r]["failed_steps"] == 200
        assert st[r]["factorizations"] == 600
    pivoted = sum(q["factorizations"] - q["unpivoted_factorizations"] for q in st)
    accepted = sum(q["successful_steps"] for q in st)
    print(f"[c4 device] {accepted} accepted steps, {pivoted} factorizations with row interchanges", flush=True)
