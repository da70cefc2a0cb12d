"""The implicit step resident on the device (include/d4_dstep.h) against the oracle's loop (orc_run):
C1 and C2 step by step from the oracle's own states, C1's whole run to 1e-9 for as long as the
reference's two builds agree within it, a tapered cyclic-load run with the dipole criterion, graph
replays against the same launches enqueued eagerly, and a batch of realizations against each one alone."""
import os
from importlib import import_module

import numpy as np
import pytest

from test_gpu_sim import C1, C2, FIX, TRAJ_TOL, _oracle_kw, _random, _sorted_xy

pytestmark = pytest.mark.gpu

dconf = import_module("2d4_amd.dconf")
dstep = import_module("2d4_amd.dstep")


def _one(ctx, xy, b_sorted, **kw):
    """a one-realization device step from positions already in the sorted order"""
    return dstep.DeviceSimulation(ctx, [0, len(xy)], xy[:, 0], xy[:, 1], b_sorted, kw)


def _bs(n):
    return np.where(np.arange(n) < n // 2, 1.0, -1.0)


def test_c1_each_step_from_oracle_trajectory(ctx, orc):
    xyb = dconf.read_dconf(FIX)
    state = _sorted_xy(xyb[:, 0], xyb[:, 1], xyb[:, 2])
    b = _bs(len(state))
    one = dict(C1, step_limit=0)
    worst = 0.0
    for k in range(100):
        nxt, _ = orc.run(state, **_oracle_kw(one))
        s = _one(ctx, state, b, **one)
        assert s.run(3) == 0  # the step limit stops it after its one accepted step
        worst = max(worst, dconf.max_periodic_dx(s.positions(), nxt))
        assert s.stats()[0]["successful_steps"] == 1
        s.close()
        state = nxt
    print(f"[dstep C1] worst single-step deviation {worst:.3e}", flush=True)
    assert worst <= TRAJ_TOL


def test_c1_whole_run_within_the_reference_flag_spread(ctx, orc):
    import oracle
    fast = oracle.Oracle(fast=True)
    xyb = dconf.read_dconf(FIX)
    xy0 = _sorted_xy(xyb[:, 0], xyb[:, 1], xyb[:, 2])
    s = _one(ctx, xy0, _bs(len(xy0)), **C1)
    okw = _oracle_kw(C1)
    dev, flags = [], []
    for k in range(1, 101):
        s.run(1)
        a, _ = orc.run(xy0, **dict(okw, step_limit=k - 1))
        f, _ = fast.run(xy0, **dict(okw, step_limit=k - 1))
        dev.append(dconf.max_periodic_dx(s.positions(), a))
        flags.append(dconf.max_periodic_dx(f, a))
    st = s.stats()[0]
    s.close()
    K = next((k for k, d in enumerate(flags) if d > TRAJ_TOL), len(flags))
    print(f"[dstep C1] whole run: reference builds within 1e-9 for {K} steps; device step vs -O2 after steps "
          f"1..{K}: max {max(dev[:K]):.3e}; after 10/100 steps {dev[9]:.2e} {dev[99]:.2e}", flush=True)
    assert st["successful_steps"] == 100 and st["failed_steps"] == 0
    assert max(dev[:K]) <= TRAJ_TOL


def test_c2_heaviside_n1024_each_step(ctx, orc):
    x, y, b = _random(1024, 2024)
    state = _sorted_xy(x, y, b)
    bs = _bs(len(state))
    h = 1e-6
    worst = 0.0
    for _ in range(10):
        one = dict(C2, init_step_size=h, step_limit=0)
        nxt, rst = orc.run(state, **_oracle_kw(one))
        s = _one(ctx, state, bs, **one)
        s.run(rst.ok_steps + rst.failed_steps)
        st = s.stats()[0]
        assert (st["successful_steps"], st["failed_steps"]) == (rst.ok_steps, rst.failed_steps)
        worst = max(worst, dconf.max_periodic_dx(s.positions(), nxt))
        assert abs(st["step_size"] - rst.step_size) <= 1e-12 * rst.step_size
        s.close()
        state, h = nxt, rst.step_size
    print(f"[dstep C2] worst single-step deviation {worst:.3e}", flush=True)
    assert worst <= TRAJ_TOL


def test_taper_dipole_cyclic_n256(ctx, orc):
    x, y, b = _random(256, 5)
    kw = dict(cutoff_mult=0.9, wfunc="p", stress=("cyclic", 0.05, 2.0, 0.01), prec=1e-6, dipole_prec=0.1,
              step_limit=40, max_step_size=1e-3)
    s = dstep.DeviceSimulation(ctx, [0, 256], x, y, b, kw)
    s.run(200)
    st = s.stats()[0]
    ref, rst = orc.run(_sorted_xy(x, y, b), **_oracle_kw(kw))
    dev = dconf.max_periodic_dx(s.positions(), ref)
    print(f"[dstep taper] steps {st['successful_steps']}+{st['failed_steps']} (oracle {rst.ok_steps}+"
          f"{rst.failed_steps}) max periodic dx {dev:.3e} strain {st['strain']:.6e} / {rst.strain:.6e}", flush=True)
    assert (st["successful_steps"], st["failed_steps"]) == (rst.ok_steps, rst.failed_steps)
    assert dev <= TRAJ_TOL
    assert abs(st["strain"] - rst.strain) <= 1e-9 * (abs(rst.strain) + 1e-12)
    s.close()


def _batch_inputs():
    sizes = [64, 300, 130, 1024, 96]
    rng = np.random.default_rng(41)
    xs, ys, bs, params = [], [], [], []
    for k, n in enumerate(sizes):
        xs.append(rng.uniform(-0.5, 0.5, n))
        ys.append(rng.uniform(-0.5, 0.5, n))
        bs.append(np.where(np.arange(n) % 2 == 0, -1.0, 1.0))
        params.append(dict(cutoff_mult=0.05 * np.sqrt(n), heaviside=True, stress=("fixed", 0.05, 0.0))
                      if k % 2 == 0 else
                      dict(cutoff_mult=0.9, wfunc="m", stress=("cyclic", 0.05, 2.0, 0.01), prec=1e-6,
                           dipole_prec=0.1, max_step_size=1e-3, init_step_size=1e-5))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    return sizes, off, np.concatenate(xs), np.concatenate(ys), np.concatenate(bs), params


def test_graph_replays_equal_eager_launches(ctx):
    """Each replay of the captured iteration against the same launches enqueued one by one, with host
    reads and synchronizations between the replays: the same bits every time."""
    sizes, off, x, y, b, params = _batch_inputs()
    g = dstep.DeviceSimulation(ctx, off, x, y, b, params, use_graph=True)
    e = dstep.DeviceSimulation(ctx, off, x, y, b, params, use_graph=False)
    for k in range(5):
        g.run(1)
        e.run(1)
        np.testing.assert_array_equal(g.positions(), e.positions())
        assert g.stats() == e.stats(), k
    g.close()
    e.close()


def test_batch_equals_each_realization_alone(ctx):
    sizes, off, x, y, b, params = _batch_inputs()
    s = dstep.DeviceSimulation(ctx, off, x, y, b, params)
    assert s.run(6) == len(sizes)
    pos = s.positions()
    st = s.stats()
    s.close()
    for r in range(len(sizes)):
        sl = slice(off[r], off[r + 1])
        one = dstep.DeviceSimulation(ctx, [0, sizes[r]], x[sl], y[sl], b[sl], params[r])
        one.run(6)
        np.testing.assert_array_equal(one.positions(), pos[sl])
        assert one.stats()[0] == st[r], r
        one.close()


def test_step_limits_stop_realizations(ctx):
    n = 64
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-0.5, 0.5, 2 * n), rng.uniform(-0.5, 0.5, 2 * n)
    b = np.tile(np.where(np.arange(n) % 2 == 0, -1.0, 1.0), 2)
    params = [dict(init_step_size=1e-4, max_step_size=1e-4, prec=1e10, dipole_prec=0.0, step_limit=1),
              dict(init_step_size=1e-4, max_step_size=1e-4, prec=1e10, dipole_prec=0.0, step_limit=3)]
    s = dstep.DeviceSimulation(ctx, [0, n, 2 * n], x, y, b, params)
    assert s.run(2) == 1  # the first realization has taken its 2 accepted steps (interrupt() uses >)
    assert s.run(10) == 0
    assert [t["successful_steps"] for t in s.stats()] == [2, 4]
    s.close()
