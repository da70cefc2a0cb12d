"""Batched ragged passes and the lockstep ensemble (include/d4_batch.h) against
the single-realization path: every realization's results must be bit-identical."""
from importlib import import_module

import numpy as np
import pytest

from helpers import random_config

pytestmark = pytest.mark.gpu

ens = import_module("2d4_amd.ensemble")
simmod = import_module("2d4_amd.sim")


def _ragged(sizes, seed):
    parts = [random_config(n, seed + k) for k, n in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    return off, np.concatenate(parts)


def test_batch_speed_pass_equals_single(ctx):
    sizes = [1, 2, 63, 64, 65, 300, 130]
    off, xy = _ragged(sizes, 11)
    tau = np.linspace(-0.1, 0.2, len(sizes))
    v = ens.batch_speed_pass(ctx, off, xy, tau)
    for r in range(len(sizes)):
        s = slice(off[r], off[r + 1])
        np.testing.assert_array_equal(v[s], ctx.speed_pass(xy[s], tau[r]))


def test_batch_speed_pass_tolerance_equals_single(ctx):
    sizes = [64, 129, 200]
    off, xy = _ragged(sizes, 3)
    caps = np.array([1e-3, 0.0, 4e-4])
    tol0 = np.concatenate([np.full(n, c if c > 0 else 1.0) for n, c in zip(sizes, caps)])
    v, tol = ens.batch_speed_pass(ctx, off, xy, 0.05, tol=tol0.copy(), tol_cap2=caps)
    for r in range(len(sizes)):
        s = slice(off[r], off[r + 1])
        t1 = tol0[s].copy()
        v1 = ctx.speed_pass(xy[s], 0.05, tol=t1 if caps[r] > 0 else None, tol_cap2=caps[r])
        np.testing.assert_array_equal(v[s], v1)
        np.testing.assert_array_equal(tol[s], t1)


def test_batch_jacobian_pass_equals_single(ctx, pkg):
    sizes = [2, 64, 65, 257, 300]
    off, xy = _ragged(sizes, 21)
    kinds = [dict(cutoff_mult=0.05 * np.sqrt(2), heaviside=True, wfunc="c"),
             dict(cutoff_mult=0.9, heaviside=False, wfunc="p"),
             dict(cutoff_mult=1e20, heaviside=False, wfunc="m"),
             dict(cutoff_mult=0.05 * np.sqrt(257), heaviside=True, wfunc="c"),
             dict(cutoff_mult=0.7, heaviside=False, wfunc="c")]
    params = [pkg.jac_params(h=1e-3 * (r + 1), tau_a=0.01 * r, tau_b=0.02 * r, n=n, **kinds[r])
              for r, n in enumerate(sizes)]
    vA, vB, (Ap, Ai, Ax), idx, dvec = ens.batch_jacobian_pass(ctx, off, xy, params)
    for r, n in enumerate(sizes):
        s = slice(off[r], off[r + 1])
        one = ctx.jacobian_pass(xy[s], params[r])
        base = Ap[off[r]]
        np.testing.assert_array_equal(vA[s], one["vA"])
        np.testing.assert_array_equal(vB[s], one["vB"])
        np.testing.assert_array_equal(dvec[s], one["dvec"])
        np.testing.assert_array_equal(Ap[off[r]:off[r + 1] + 1] - base, one["Ap"])
        np.testing.assert_array_equal(Ai[base:Ap[off[r + 1]]] - off[r], one["Ai"])
        np.testing.assert_array_equal(Ax[base:Ap[off[r + 1]]], one["Ax"])
        np.testing.assert_array_equal(idx[s] - base, one["indexes"])


def _ens_inputs(sizes, seed):
    rng = np.random.default_rng(seed)
    xs, ys, bs = [], [], []
    for n in sizes:
        xs.append(rng.uniform(-0.5, 0.5, n))
        ys.append(rng.uniform(-0.5, 0.5, n))
        bs.append(np.where(np.arange(n) % 2 == 0, -1.0, 1.0))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    return off, np.concatenate(xs), np.concatenate(ys), np.concatenate(bs)


def _params(sizes):
    out = []
    for k, n in enumerate(sizes):
        if k % 3 == 2:  # Gaussian taper, dipole criterion, cyclic load, 'p' weights
            out.append(dict(cutoff_mult=0.9, wfunc="p", stress=("cyclic", 0.05, 2.0, 0.01), prec=1e-6,
                            dipole_prec=0.1, max_step_size=1e-3))
        else:  # -r 0 -s 0.05 -u 0.05 sqrt(N) -h
            out.append(dict(cutoff_mult=0.05 * np.sqrt(n), heaviside=True, stress=("fixed", 0.05, 0.0),
                            init_step_size=1e-5))
    return out


def _plain(ctx, off, x, y, b, params, steps):
    res = []
    for r in range(len(params)):
        s = slice(off[r], off[r + 1])
        sim = simmod.Simulation.from_arrays(ctx, x[s], y[s], b[s], **params[r])
        for _ in range(steps):
            sim.step()
        res.append((sim.positions(), sim.stats()))
        sim.close()
    return res


def _check_equal(e, off, plain):
    pos = e.positions()
    st = e.stats()
    for r, (p, s) in enumerate(plain):
        np.testing.assert_array_equal(pos[off[r]:off[r + 1]], p)
        for k in ("sim_time", "step_size", "strain", "successful_steps", "failed_steps", "pair_passes",
                  "factorizations", "unpivoted_factorizations"):
            assert st[r][k] == s[k], (r, k, st[r][k], s[k])


def test_ensemble_equals_plain_path(ctx):
    sizes = [64, 128, 130, 200, 256, 96]
    off, x, y, b = _ens_inputs(sizes, 5)
    params = _params(sizes)
    e = ens.Ensemble(ctx, off, x, y, b, params)
    assert e.step(6) == len(sizes)
    _check_equal(e, off, _plain(ctx, off, x, y, b, params, 6))
    e.close()


def test_ensemble_gpu_dense_solver_equals_plain_path(ctx):
    sizes = [64, 300, 130, 96]
    off, x, y, b = _ens_inputs(sizes, 7)
    params = [dict(p, solver=1) for p in _params(sizes)]
    e = ens.Ensemble(ctx, off, x, y, b, params)
    e.step(4)
    _check_equal(e, off, _plain(ctx, off, x, y, b, params, 4))
    e.close()


def test_ensemble_gpu_dense_solver_padded_orders_equal_plain_path(ctx):
    # realizations over several padded orders (320 .. 1344) so that batched LU
    # calls hold different matrices than the plain path's one-matrix calls
    sizes = [700, 1300, 260, 520, 1030, 264, 1290]
    off, x, y, b = _ens_inputs(sizes, 21)
    params = [dict(cutoff_mult=0.05 * np.sqrt(n), heaviside=True, stress=("fixed", 0.05, 0.0), solver=1)
              for n in sizes]
    e = ens.Ensemble(ctx, off, x, y, b, params)
    e.step(4)
    _check_equal(e, off, _plain(ctx, off, x, y, b, params, 4))
    st = e.stats()
    fact = sum(s["factorizations"] for s in st)
    unpiv = sum(s["unpivoted_factorizations"] for s in st)
    print(f"factorizations {fact}, unpivoted {unpiv}", flush=True)
    assert 0 < unpiv <= fact
    e.close()


def test_ensemble_launch_shapes_do_not_change_results(ctx):
    sizes = [70, 190, 64, 128]
    off, x, y, b = _ens_inputs(sizes, 9)
    params = _params(sizes)
    plain = _plain(ctx, off, x, y, b, params, 4)
    for cfg in (dict(pass_grid=0, tile_order=0, use_graph=0), dict(pass_grid=512, tile_order=2, use_graph=1),
                dict(pass_grid=0, tile_order=1, use_graph=1, host_threads=1)):
        e = ens.Ensemble(ctx, off, x, y, b, params)
        e.set_config(**cfg)
        e.step(4)
        _check_equal(e, off, plain)
        e.close()


def test_ensemble_autotune_then_step(ctx):
    sizes = [300, 256, 64, 400]
    off, x, y, b = _ens_inputs(sizes, 13)
    params = _params(sizes)
    e = ens.Ensemble(ctx, off, x, y, b, params)
    e.step(1)
    cfg, ms = e.autotune()
    assert ms > 0 and cfg["tile_order"] in (0, 1, 2)
    e.step(2)
    _check_equal(e, off, _plain(ctx, off, x, y, b, params, 3))
    t = e.timing()
    assert t["total_ms"] >= t["pass_ms"] > 0
    e.close()


def test_ensemble_step_limit_stops_realizations(ctx):
    sizes = [64, 80]
    off, x, y, b = _ens_inputs(sizes, 17)
    params = [dict(init_step_size=1e-4, max_step_size=1e-4, prec=1e10, dipole_prec=0.0, step_limit=1),
              dict(init_step_size=1e-4, max_step_size=1e-4, prec=1e10, dipole_prec=0.0, step_limit=3)]
    e = ens.Ensemble(ctx, off, x, y, b, params)
    assert e.step(2) == 2
    assert e.step(1) == 1   # the first has taken its 2 accepted steps (interrupt() uses >)
    assert e.step(5) == 0
    st = e.stats()
    assert [s["successful_steps"] for s in st] == [2, 4]


def test_ensemble_over_32768_columns_equals_plain_path(ctx):
    # more than 16 x 2048 columns in one batched pass: the CSC scan of the
    # batched Jacobian spreads over many workgroups (the plain path's is one)
    sizes = [4096, 4000, 3500, 4096, 3800, 4096, 3900, 4096, 2600]
    assert sum(sizes) > 32768
    off, x, y, b = _ens_inputs(sizes, 31)
    params = [dict(cutoff_mult=0.05 * np.sqrt(n), heaviside=True, stress=("fixed", 0.05, 0.0), solver=1)
              for n in sizes]
    e = ens.Ensemble(ctx, off, x, y, b, params)
    e.step(2)
    _check_equal(e, off, _plain(ctx, off, x, y, b, params, 2))
    e.close()


def test_batch_jacobian_pass_growing_layout_on_shared_context(ctx, pkg):
    """A batched Jacobian pass over far more columns than the first one on the same context: the device
    CSC is sized for the new layout's diagonals (total) beside the pair list, and each realization still
    equals the single-realization pass."""
    for sizes, seed, rc in (([64, 100], 41, 0.05), ([4000] * 36, 43, 0.002)):
        off, xy = _ragged(sizes, seed)
        params = [pkg.jac_params(h=1e-3, tau_a=0.0, tau_b=0.05, n=n, cutoff_mult=rc * np.sqrt(n), heaviside=True)
                  for n in sizes]
        print(f"[batch grow] R={len(sizes)} total={off[-1]}", flush=True)
        vA, vB, (Ap, Ai, Ax), idx, dvec = ens.batch_jacobian_pass(ctx, off, xy, params)
        for r in sorted({0, len(sizes) // 2, len(sizes) - 1}):
            s = slice(off[r], off[r + 1])
            one = ctx.jacobian_pass(xy[s], params[r])
            base = Ap[off[r]]
            np.testing.assert_array_equal(vA[s], one["vA"])
            np.testing.assert_array_equal(Ap[off[r]:off[r + 1] + 1] - base, one["Ap"])
            np.testing.assert_array_equal(Ai[base:Ap[off[r + 1]]] - off[r], one["Ai"])
            np.testing.assert_array_equal(Ax[base:Ap[off[r + 1]]], one["Ax"])
            np.testing.assert_array_equal(dvec[s], one["dvec"])


def test_ensemble_host_thread_change_keeps_gpu_solver(ctx):
    """Changing the host thread count after the batched GPU LU exists: the LU's host loops run on the new
    pool, and the results stay those of the plain path."""
    sizes = [300, 130, 96, 200]
    off, x, y, b = _ens_inputs(sizes, 19)
    params = [dict(p, solver=1) for p in _params(sizes)]
    e = ens.Ensemble(ctx, off, x, y, b, params)
    e.step(1)
    e.set_config(host_threads=2)
    e.step(2)
    e.set_config(host_threads=3)
    e.step(1)
    _check_equal(e, off, _plain(ctx, off, x, y, b, params, 4))
    e.close()


def test_batch_passes_against_the_oracle(ctx, orc, pkg):
    """Each realization of a batched speed pass and batched Jacobian pass against the oracle's
    single-realization passes (orc_speed_pass, orc_jacobian_pass), to the bounds of the plain passes:
    per row the sum of its pairs' bounds (B), per CSC entry the bound from the pairs' d sigma bounds;
    the CSC structure bit for bit."""
    from helpers import jacobian_allowance, speed_allowance
    sizes = [2, 63, 200, 301, 64, 512]
    off, xy = _ragged(sizes, 71)
    tau = np.linspace(0.0, 0.1, len(sizes))
    v = ens.batch_speed_pass(ctx, off, xy, tau)
    kinds = [dict(cutoff_mult=1e20, heaviside=False, wfunc="c"), dict(cutoff_mult=0.05 * np.sqrt(63), heaviside=True),
             dict(cutoff_mult=0.9, heaviside=False, wfunc="m"), dict(cutoff_mult=0.1 * np.sqrt(301), heaviside=True),
             dict(cutoff_mult=1e20, heaviside=False, wfunc="p"), dict(cutoff_mult=0.05 * np.sqrt(512), heaviside=True)]
    params = [pkg.jac_params(h=1e-3, tau_a=0.02, tau_b=0.05, n=n, **kinds[r]) for r, n in enumerate(sizes)]
    vA, vB, (Ap, Ai, Ax), idx, dvec = ens.batch_jacobian_pass(ctx, off, xy, params)
    for r, n in enumerate(sizes):
        s = slice(off[r], off[r + 1])
        x = xy[s]
        allow = speed_allowance(orc, x)
        assert np.all(np.abs(v[s] - orc.speed_pass(x, tau[r])) <= allow + 1e-15), r
        p = params[r]
        R = orc.jacobian_pass(x, p.h, p.tau_a, p.tau_b, p.cutoff, p.cut_sq, p.one_per_cut_sq, bool(p.heaviside),
                              p.wfunc.decode())
        base = Ap[off[r]]
        np.testing.assert_array_equal(Ap[off[r]:off[r + 1] + 1] - base, R["Ap"])
        np.testing.assert_array_equal(Ai[base:Ap[off[r + 1]]] - off[r], R["Ai"])
        np.testing.assert_array_equal(idx[s] - base, R["indexes"])
        bound, dbound = jacobian_allowance(orc, x, p.h, R)
        assert np.all(np.abs(Ax[base:Ap[off[r + 1]]] - R["Ax"]) <= bound), r
        assert np.all(np.abs(dvec[s] - R["dvec"]) <= dbound), r
        assert np.all(np.abs(vA[s] - R["vA"]) <= allow + 1e-15) and np.all(np.abs(vB[s] - R["vB"]) <= allow + 1e-15)
