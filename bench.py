"""Benchmark of the periodic pair-interaction path on one MI355X.

Headline (one JSON line on stdout): FP64 pair evaluations per second of the
Jacobian + speed pass at N = 16384 -- speeds for all N plus the full dense
dv_i/dx_j (as M = V + diag(s) = -h dv/dx, column-major), every ordered pair
(i, j), i != j, counted once: N (N - 1) = 268 419 072 per pass.  Inputs are
resident in HBM before the timed region; a step is one such pass.

Also reported: the pair kernel's roofline (algorithmic FP64 operations per
launch over its HIP-event-timed duration against the FP64 vector peak), the
CPU baseline (the oracle's C restatement, timed on this host), the speed-only
pass and the reference's final row-scaled matrix.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector (= FP64 matrix) peak
PEAK_HBM_GBS = 8000.0
# algorithmic FP64 operations per unordered pair, counted from the reference
# source (+ - * / = 1, exp2 and cos = 1 each, shared subexpressions once per function):
OPS_JAC_PAIR = 390          # setup 7 + Field::xy 125 + Field::xy_diff_x 248 + assembly 8 + 2 transcendental
OPS_SPEED_PAIR = 138        # setup 7 + Field::xy 125 + assembly 4 + 2 transcendental


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sorted_config(n, seed):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(-0.5, 0.5, (n, 2))
    b = np.where(np.arange(n) % 2 == 1, 1.0, -1.0)
    return np.ascontiguousarray(xy[np.argsort(-(xy[:, 1] + b), kind="stable")])


def cpu_baseline(n_single, n_omp, min_seconds):
    """The oracle (oracle/liboracle_fast.so: reference flags, OpenMP) on this host's cores."""
    import oracle
    import ctypes as C
    orc = oracle.Oracle(fast=True)
    out = {}
    for label, n, fn in (("1core", n_single, "orc_jacobian_dense"), ("omp", n_omp, "orc_jacobian_dense_omp")):
        xy = sorted_config(n, 5)
        vA, vB, d = np.empty(n), np.empty(n), np.empty(n)
        M = np.empty((n, n))
        p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
        f = getattr(orc.L, fn)

        def run():
            if fn == "orc_jacobian_dense":
                f(n, p(xy), 1e-6, 0.0, 0.1, b"c", p(vA), p(vB), p(M), p(d), None, 0.0)
            else:
                f(n, p(xy), 1e-6, 0.0, 0.1, b"c", p(vA), p(vB), p(M), p(d))
        run()  # warm-up
        times = []
        t_all = time.perf_counter()
        while len(times) < 5 or time.perf_counter() - t_all < min_seconds:
            t0 = time.perf_counter()
            run()
            times.append(time.perf_counter() - t0)
            if len(times) >= 50:
                break
        med = float(np.median(times))
        out[label] = dict(n=n, seconds=med, reps=len(times), pair_evals_per_s=n * (n - 1) / med,
                          threads=1 if label == "1core" else orc.L.orc_num_threads())
    return out


def ensemble_bench(pkg, ctx, steps, R, seed=4242, plain=True, solver=1, graphs=True):
    """Batched ensemble vs the plain one-realization-at-a-time GPU path on the
    same realizations, same number of lockstep iterations, same solver.

    Workload: R realizations, N_r = 2 U{128..2048} (even N in [256, 4096]),
    uniform positions, b alternating, heaviside cutoff 0.05 (-u 0.05 sqrt(N) -h),
    tau = 0.05 (-r 0 -s 0.05), default precision control, linear solves as a
    dense LU on the GPU.  A step is one loop iteration (stages I-IV) of one
    realization; the first iteration (with run()'s initial speed pass) is
    untimed in both paths.  plain=False measures the batched path alone; solver=0
    puts every realization's linear solves on the host sparse LU instead."""
    ens = importlib.import_module("2d4_amd.ensemble")
    simmod = importlib.import_module("2d4_amd.sim")
    rng = np.random.default_rng(seed)
    sizes = 2 * rng.integers(128, 2049, R)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    T = int(off[-1])
    x = rng.uniform(-0.5, 0.5, T)
    y = rng.uniform(-0.5, 0.5, T)
    b = np.concatenate([np.where(np.arange(n) % 2 == 0, -1.0, 1.0) for n in sizes])
    params = [dict(cutoff_mult=0.05 * float(np.sqrt(n)), heaviside=True, stress=("fixed", 0.05, 0.0), solver=solver)
              for n in sizes]
    log(f"[bench] ensemble R={R} sum N={T}: batched")
    e = ens.Ensemble(ctx, off, x, y, b, params)
    e.step(1)
    cfg, tune_ms = e.autotune()
    log(f"[bench] ensemble autotuned {cfg} ({tune_ms:.2f} ms per speed+Jacobian pass)")
    if not graphs:  # every launch its own dispatch (hardware counters are not collected inside a graph)
        e.set_config(use_graph=0)
        cfg = e.config()
    t_before = e.timing()
    t_batched = 0.0
    done = 0
    while done < steps:  # chunks of at most 20 iterations, one progress line each
        chunk = min(20, steps - done)
        t0 = time.perf_counter()
        e.step(chunk)
        t_batched += time.perf_counter() - t0
        done += chunk
        if steps > 20:
            log(f"[bench] ensemble: {done}/{steps} lockstep iterations, {t_batched:.1f} s")
    t_after = e.timing()
    pos_b = e.positions()
    st_b = e.stats()
    e.close()
    rs = R * steps
    if not plain:
        return {"metric": "ensemble steps/s (one loop iteration of one realization = one step), batched only",
                "batched_steps_per_s": rs / t_batched, "R": R, "sum_n": T, "steps_timed": steps,
                "batched_s": t_batched, "batched_split_ms": {k: t_after[k] - t_before[k] for k in t_after},
                "factorizations": int(sum(s["factorizations"] for s in st_b)),
                "unpivoted_factorizations": int(sum(s["unpivoted_factorizations"] for s in st_b)),
                "launch_config": cfg}
    log("[bench] ensemble: plain path, one realization at a time")
    t_plain = 0.0
    equal = True
    for r in range(R):
        sl = slice(off[r], off[r + 1])
        sim = simmod.Simulation.from_arrays(ctx, x[sl], y[sl], b[sl], **params[r])
        sim.step()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.step()
        t_plain += time.perf_counter() - t0
        equal &= bool(np.array_equal(sim.positions(), pos_b[sl]))
        sim.close()
    return {"metric": "ensemble steps/s (one loop iteration of one realization = one step), batched vs plain",
            "batched_steps_per_s": rs / t_batched, "plain_steps_per_s": rs / t_plain,
            "ratio": t_plain / t_batched, "R": R, "sum_n": T, "n_min": int(sizes.min()), "n_max": int(sizes.max()),
            "steps_timed": steps, "batched_s": t_batched, "plain_s": t_plain,
            "batched_split_ms": {k: t_after[k] - t_before[k] for k in t_after},
            "accepted": int(sum(s["successful_steps"] for s in st_b)),
            "failed": int(sum(s["failed_steps"] for s in st_b)),
            "factorizations": int(sum(s["factorizations"] for s in st_b)),
            "unpivoted_factorizations": int(sum(s["unpivoted_factorizations"] for s in st_b)),
            "launch_config": cfg,
            "solver": "dense LU on the GPU, both paths: threshold pivoting u = 0.001 (blocked LU without interchanges, "
                      "one rocSOLVER getrf per matrix where a multiplier exceeds 1000), triangular-solve kernel",
            "bitwise_equal_to_plain": equal}


def implicit_step_bench(pkg, ctx, gpu_steps=50, cpu_steps=10, with_cpu=True, seed=77, solvers=(0, 1)):
    """ms per implicit step (one loop iteration of Simulation::run, stages I-IV) of the plain
    drop-in on the N = 1024 workload: uniform positions, heaviside cutoff 0.05 (-u 1.6 -h),
    tau = 0.1 constant (-r 0 -s 0.1), default precision control.

    GPU: pair passes on the GPU with the linear solves on the host (sparse LU, UMFPACK's place)
    and with the GPU dense LU.  CPU: the same host loop with the oracle's single-threaded passes
    and the host sparse LU (tests/csrc/libsim_cpu_fast.so, built with the reference's flags).
    The roofline fraction counts the algorithmic pair work of the step: 138 FP64 ops per pair
    for each of the two Jacobian and three speed passes, plus 252 per pair inside the cutoff
    for each Jacobian pass."""
    import ctypes as C
    simmod = importlib.import_module("2d4_amd.sim")
    n = 1024
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, n)
    y = rng.uniform(-0.5, 0.5, n)
    b = np.where(np.arange(n) % 2 == 0, -1.0, 1.0)
    params = dict(cutoff_mult=0.05 * np.sqrt(n), heaviside=True, stress=("fixed", 0.1, 0.0))
    pairs = n * (n - 1) // 2
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    dx -= np.rint(dx)
    dy -= np.rint(dy)
    inside = int(np.count_nonzero(np.triu(dx * dx + dy * dy < 0.05 ** 2, 1)))
    ops = pairs * 5 * 138 + inside * 2 * 252

    def timed(make, steps):
        sim = make()
        sim.step()  # the first iteration includes run()'s initial speed pass
        st0 = sim.stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.step()
        dt = (time.perf_counter() - t0) / steps
        st = sim.stats()
        sim.close()
        st["pass_ms_per_step"] = (st["pass_ms"] - st0["pass_ms"]) / steps
        st["solve_ms_per_step"] = (st["solve_ms"] - st0["solve_ms"]) / steps
        return dt, st

    out = {"workload": "N=1024 uniform, heaviside cutoff 0.05, tau=0.1, adaptive (the C2 configuration)",
           "pairs": pairs, "pairs_inside_cutoff": inside, "algorithmic_ops_per_step": ops}
    for label, solver in (("gpu_passes_host_lu", 0), ("gpu_passes_gpu_lu", 1)):
        if solver not in solvers:
            continue
        dt, st = timed(lambda: simmod.Simulation.from_arrays(ctx, x, y, b, solver=solver, **params), gpu_steps)
        out[label] = {"ms_per_step": dt * 1e3, "steps": gpu_steps, "accepted": st["successful_steps"],
                      "failed": st["failed_steps"], "fraction_of_fp64_peak": ops / dt / 78.6e12,
                      "pair_passes_ms_per_step": st["pass_ms_per_step"],
                      "linear_solves_ms_per_step": st["solve_ms_per_step"]}
    shim = os.path.join(ROOT, "tests", "csrc", "libsim_cpu_fast.so")
    if with_cpu and os.path.exists(shim):
        extra = importlib.import_module("2d4_amd._bind_extra")
        L = C.CDLL(shim)
        extra.bind(L)
        dt, st = timed(lambda: simmod.Simulation._create(L, C.c_void_p(1), x, y, b, **params), cpu_steps)
        out["cpu"] = {"ms_per_step": dt * 1e3, "steps": cpu_steps, "cores": 1, "kind": "port",
                      "accepted": st["successful_steps"], "failed": st["failed_steps"],
                      "pair_passes_ms_per_step": st["pass_ms_per_step"],
                      "linear_solves_ms_per_step": st["solve_ms_per_step"],
                      "sample": "the host loop (d4_sim.cpp) with the oracle's single-threaded passes and the host "
                                "sparse LU, -Ofast -ffast-math -march=x86-64-v3, as the reference runs (one core)"}
    return out


def device_step_bench(pkg, ctx, iters=32):
    """The implicit step resident on the device (include/d4_dstep.h: every stage a launch, one iteration
    a HIP graph replay) on the N = 1024 workload of implicit_step_bench (heaviside cutoff 0.05,
    tau = 0.1, adaptive), for one realization and for a batch of 64, beside the host-driven batched
    ensemble (include/d4_batch.h, GPU LU) on the same 64 realizations."""
    dstep = importlib.import_module("2d4_amd.dstep")
    ens = importlib.import_module("2d4_amd.ensemble")
    n = 1024
    out = {"workload": "N=1024 per realization, uniform, heaviside cutoff 0.05, tau=0.1, adaptive (C2)",
           "matrix": "dense n x n per realization; GPU dense LU (u = 0.001 threshold pivoting) in a fixed order by |y| "
                     "inside the band envelope that order fixes for the heaviside storage radius"}
    for R in (1, 64):
        rng = np.random.default_rng(77)
        x = rng.uniform(-0.5, 0.5, R * n)
        y = rng.uniform(-0.5, 0.5, R * n)
        b = np.tile(np.where(np.arange(n) % 2 == 0, -1.0, 1.0), R)
        off = (np.arange(R + 1) * n).astype(np.int32)
        params = dict(cutoff_mult=0.05 * np.sqrt(n), heaviside=True, stress=("fixed", 0.1, 0.0))
        s = dstep.DeviceSimulation(ctx, off, x, y, b, params)
        s.run(2)
        t0 = time.perf_counter()
        s.run(iters)
        t = time.perf_counter() - t0
        st = s.stats()
        s.set_graph(False)
        t1 = time.perf_counter()
        s.run(8)
        te = (time.perf_counter() - t1) / 8
        s.close()
        row = {"ms_per_iteration": t * 1e3 / iters, "realization_steps_per_s": R * iters / t,
               "eager_ms_per_iteration": te * 1e3, "iterations": iters,
               "accepted": int(sum(q["successful_steps"] for q in st)), "failed": int(sum(q["failed_steps"] for q in st))}
        if R > 1:
            e = ens.Ensemble(ctx, off, x, y, b, [dict(params, solver=1) for _ in range(R)])
            e.step(2)
            t0 = time.perf_counter()
            e.step(8)
            row["host_driven_ensemble_realization_steps_per_s"] = R * 8 / (time.perf_counter() - t0)
            e.close()
        out[f"R{R}"] = row
    # the bench ensemble's own workload (ensemble_bench: R = 256, N_r even in [256, 4096], heaviside cutoff 0.05,
    # tau = 0.05, seed 4242) as one captured graph per lockstep iteration
    R, seed = 256, 4242
    rng = np.random.default_rng(seed)
    sizes = 2 * rng.integers(128, 2049, R)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    T = int(off[-1])
    x = rng.uniform(-0.5, 0.5, T)
    y = rng.uniform(-0.5, 0.5, T)
    b = np.concatenate([np.where(np.arange(m) % 2 == 0, -1.0, 1.0) for m in sizes])
    params = [dict(cutoff_mult=0.05 * float(np.sqrt(m)), heaviside=True, stress=("fixed", 0.05, 0.0)) for m in sizes]
    t0 = time.perf_counter()
    s = dstep.DeviceSimulation(ctx, off, x, y, b, params)
    tc = time.perf_counter() - t0
    s.run(1)  # the first iteration (graph capture) untimed, as the ensemble's
    t0 = time.perf_counter()
    s.run(2)
    t = time.perf_counter() - t0
    st = s.stats()
    s.close()
    out["ensemble_R256"] = {"workload": "the bench ensemble (R=256, N_r even in [256, 4096], heaviside cutoff 0.05, "
                                        "tau=0.05, seed 4242)", "create_s": tc, "iterations": 2,
                            "ms_per_iteration": t * 1e3 / 2, "realization_steps_per_s": R * 2 / t,
                            "accepted": int(sum(q["successful_steps"] for q in st)),
                            "failed": int(sum(q["failed_steps"] for q in st))}
    log(f"[bench] device step, bench ensemble: {out['ensemble_R256']}")
    return out


def c5_reference_step_check(T=4, n=2048, eps=1e-9, coord=7, seed=2048):
    """The C5 configuration (N = 2048 uniform, tau = 0.2, fixed dt = 1e-3: -P 1e10 -p 0 -M 1e-3
    --initial-stepsize 1e-3) run through the reference's own step (the device-resident loop, pinned to
    the oracle's loop step by step): the accumulated strain, the largest wrapped position change per
    step and the central difference of the strain in x_coord, with the start and its two perturbations
    by +-eps run as one batch of three realizations."""
    dstep = importlib.import_module("2d4_amd.dstep")
    pkg = importlib.import_module("2d4_amd")
    ctx = pkg.Context(0)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, n)
    y = rng.uniform(-0.5, 0.5, n)
    b = np.where(np.arange(n) % 2 == 0, -1.0, 1.0)
    xs = np.concatenate([x, x, x])
    xs[n + coord] += eps
    xs[2 * n + coord] -= eps
    p = dict(init_step_size=1e-3, max_step_size=1e-3, prec=1e10, dipole_prec=0.0, stress=("fixed", 0.2, 0.0))
    s = dstep.DeviceSimulation(ctx, [0, n, 2 * n, 3 * n], xs, np.tile(y, 3), np.tile(b, 3), p)
    prev = s.positions()[:n, 0]
    rows = []
    t0 = time.perf_counter()
    for k in range(1, T + 1):
        s.run(1)
        pos = s.positions()
        st = s.stats()
        a = pos[:n, 0]
        step = float(np.abs((a - prev + 0.5) % 1 - 0.5).max())
        dev = float(np.abs((pos[n:2 * n, 0] - a + 0.5) % 1 - 0.5).max())
        rows.append({"step": k, "strain": st[0]["strain"], "max_wrapped_step_change": step,
                     "max_deviation_of_the_perturbed_run": dev,
                     "central_difference_dstrain_dx": (st[1]["strain"] - st[2]["strain"]) / (2 * eps),
                     "unpivoted_factorizations": st[0]["unpivoted_factorizations"],
                     "factorizations": st[0]["factorizations"], "seconds": time.perf_counter() - t0})
        log(f"[bench] C5 reference step {k}: {rows[-1]}")
        prev = a
    s.close()
    return {"n": n, "tau": 0.2, "dt": 1e-3, "eps": eps, "coordinate": coord, "steps": rows}


def adjoint_bench(n, T, dt, every, method="theta"):
    """Strain gradient through implicit steps (2d4_amd/adjoint.py) at N = n: T steps of dt of the reference's
    step ("theta") or of backward Euler ("backward_euler"), or ("adaptive") the steps the reference's own
    loop accepts until time T dt with its step at most dt (adjoint.reference_schedule: default precision,
    GPU dense LU).  Times the forward pass, the checkpointed forward and the backward."""
    import torch
    adj = importlib.import_module("2d4_amd.adjoint")
    pkg = importlib.import_module("2d4_amd")
    ctx = pkg.Context(0)
    x, y = adj.separated_config(n, n, 0.5 / np.sqrt(n), device="cuda:0")
    tau = torch.tensor(0.2, dtype=torch.float64, device="cuda:0")
    step_method = "backward_euler" if method == "backward_euler" else "theta"
    sizes, schedule = [dt] * T, {}
    if method == "adaptive":
        t0 = time.perf_counter()
        last = [t0]

        def progress(it, acc, t):
            if time.perf_counter() - last[0] > 10:
                last[0] = time.perf_counter()
                log(f"[bench] adjoint schedule: {it} iterations, {acc} accepted, t = {t:.4e} of {T * dt:.4e}")

        sizes, _ = adj.reference_schedule(x, y, 0.2, T * dt, ctx, max_step_size=dt, solver=1, progress=progress)
        schedule = {"accepted_steps": len(sizes), "min_step": min(sizes), "max_step": max(sizes),
                    "schedule_s": time.perf_counter() - t0}
        log(f"[bench] adjoint schedule: {schedule}")
    adj.reset_stats()
    t0 = time.perf_counter()
    with torch.no_grad():
        xs = x
        for k in range(0, len(sizes), every):
            xs = adj._segment(xs, y, tau, sizes[k:k + every], dt, ctx, 1e-12, 30, step_method, None)
            log(f"[bench] adjoint forward {min(k + every, len(sizes))} steps: {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    fwd_stats = dict(adj.stats)
    x0 = x.clone().requires_grad_()
    yy = y.clone().requires_grad_()
    tt = tau.clone().requires_grad_()
    L = adj.accumulated_strain(x0, yy, tt, T, dt, ctx, checkpoint_every=every, method=step_method, steps=sizes)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    return {"n": n, "T": T, "dt": dt, "checkpoint_every": every, "tau": 0.2, "method": method, **schedule,
            "steps": len(sizes), "forward_s": t1 - t0, "forward_checkpointed_s": t2 - t1, "backward_s": t3 - t2,
            "ms_per_step_forward": (t1 - t0) * 1e3 / len(sizes), "ms_per_step_backward": (t3 - t2) * 1e3 / len(sizes),
            "forward_stats": fwd_stats, "L": L.item(), "max_abs_dL_dx0": x0.grad.abs().max().item(),
            "max_abs_dL_dy": yy.grad.abs().max().item(), "dL_dtau": tt.grad.item(),
            "config": "uniform positions with a minimum separation of 0.5/sqrt(N), sorted"}


def c5_adjoint_bench(n=2048, T=256, ncoord=32, seed=2048, h=1e-6):
    """C5 at N = n uniform positions (x then y from numpy's generator with this seed, b alternating, sorted),
    tau = 0.2: the strain gradient through T steps of the reference's implicit step with the sizes its own loop
    accepts from this start (adjoint.reference_schedule: default precision, steps at most 1e-3, GPU dense LU),
    checkpoints every 16 steps, against central differences with step h on ncoord random coordinates and on
    tau.  Fixed steps of 1e-3 are not differentiable from this start (--c5-check)."""
    import torch
    adj = importlib.import_module("2d4_amd.adjoint")
    pkg = importlib.import_module("2d4_amd")
    ctx = pkg.Context(0)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, n)
    y = rng.uniform(-0.5, 0.5, n)
    b = np.where(np.arange(n) % 2 == 1, 1.0, -1.0)
    o = np.argsort(-(y + b), kind="stable")
    x, y = x[o].copy(), y[o].copy()
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")  # noqa: E731
    t0 = time.perf_counter()
    sizes, _ = adj.reference_schedule(x, y, 0.2, 1.0, ctx, max_step_size=1e-3, solver=1, max_accepted=T)
    t_sched = time.perf_counter() - t0
    log(f"[bench] C5 schedule: {len(sizes)} steps, {min(sizes):.3e} ... {max(sizes):.3e}, {t_sched:.1f} s")

    def loss(xa, ya, tau=0.2):
        with torch.no_grad():
            return adj.accumulated_strain(dev(xa), dev(ya), tau, 0, 0.0, ctx, checkpoint_every=0, steps=sizes).item()

    t0 = time.perf_counter()
    L0 = loss(x, y)
    t_fwd = time.perf_counter() - t0
    x0, yy = dev(x).requires_grad_(), dev(y).requires_grad_()
    tt = torch.tensor(0.2, dtype=torch.float64, device="cuda:0", requires_grad=True)
    t_ck0 = time.perf_counter()
    L = adj.accumulated_strain(x0, yy, tt, 0, 0.0, ctx, checkpoint_every=16, steps=sizes)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    gx, gy, gt = x0.grad.cpu().numpy(), yy.grad.cpu().numpy(), tt.grad.item()
    gmax = max(np.abs(gx).max(), np.abs(gy).max())
    coords = np.random.default_rng(seed + 1).choice(2 * n, ncoord, replace=False)
    rows, worst = [], 0.0
    t0 = time.perf_counter()
    for kk in coords:
        col, k = divmod(int(kk), n)
        arr = x if col == 0 else y
        p_, m_ = arr.copy(), arr.copy()
        p_[k] += h
        m_[k] -= h
        fd = ((loss(p_, y) - loss(m_, y)) if col == 0 else (loss(x, p_) - loss(x, m_))) / (2 * h)
        g = float((gx if col == 0 else gy)[k])
        rel = abs(fd - g) / max(abs(g), 1e-3 * gmax)
        worst = max(worst, rel)
        rows.append({"coord": f"{'xy'[col]}[{k}]", "adjoint": g, "fd": fd, "rel": rel})
        if rel > 1e-5:  # the difference at h/2 and the extrapolation (4 d(h/2) - d(h)) / 3
            p_[k] -= h / 2
            m_[k] += h / 2
            fd2 = ((loss(p_, y) - loss(m_, y)) if col == 0 else (loss(x, p_) - loss(x, m_))) / h
            rows[-1].update(fd_half_h=fd2, rel_half_h=abs(fd2 - g) / max(abs(g), 1e-3 * gmax),
                            rel_extrapolated=abs((4 * fd2 - fd) / 3 - g) / max(abs(g), 1e-3 * gmax))
        log(f"[bench] C5 {rows[-1]}")
    fdt = (loss(x, y, 0.2 + h) - loss(x, y, 0.2 - h)) / (2 * h)
    t_fd = time.perf_counter() - t0
    return {"n": n, "T": len(sizes), "tau": 0.2, "checkpoint_every": 16, "h": h, "coordinates": ncoord,
            "config": f"uniform positions (numpy default_rng({seed}), x then y), sorted; step sizes of the "
                      "reference's loop (default precision, at most 1e-3)",
            "min_step": min(sizes), "max_step": max(sizes), "sim_time": float(sum(sizes)),
            "schedule_s": t_sched, "forward_s": t_fwd, "ms_per_step_forward": t_fwd * 1e3 / len(sizes),
            "forward_checkpointed_s": t1 - t_ck0, "backward_s": t2 - t1,
            "ms_per_step_backward": (t2 - t1) * 1e3 / len(sizes), "fd_s": t_fd, "L": L0,
            "max_abs_grad": gmax, "dL_dtau": gt, "dL_dtau_fd": fdt, "tau_rel": abs(fdt - gt) / abs(gt),
            "worst_rel_vs_fd": worst, "checks": rows}


def c4_device_run(iterations=200, R=256, seed=4242):
    """The C4 workload (the bench ensemble: R realizations, N_r even in [256, 4096], heaviside cutoff 0.05,
    tau = 0.05, seed 4242) through the device-resident step for `iterations` lockstep iterations, graph capture
    and the first iteration included in the timed region; progress every 20 iterations."""
    dstep = importlib.import_module("2d4_amd.dstep")
    pkg = importlib.import_module("2d4_amd")
    ctx = pkg.Context(0)
    rng = np.random.default_rng(seed)
    sizes = 2 * rng.integers(128, 2049, R)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    T = int(off[-1])
    x = rng.uniform(-0.5, 0.5, T)
    y = rng.uniform(-0.5, 0.5, T)
    b = np.concatenate([np.where(np.arange(m) % 2 == 0, -1.0, 1.0) for m in sizes])
    params = [dict(cutoff_mult=0.05 * float(np.sqrt(m)), heaviside=True, stress=("fixed", 0.05, 0.0)) for m in sizes]
    t0 = time.perf_counter()
    s = dstep.DeviceSimulation(ctx, off, x, y, b, params)
    done = 0
    while done < iterations:
        k = min(20, iterations - done)
        s.run(k)
        done += k
        log(f"[bench] C4 device step: {done}/{iterations} iterations, {time.perf_counter() - t0:.1f} s")
    t = time.perf_counter() - t0
    st = s.stats()
    s.close()
    fac = sum(q["factorizations"] for q in st)
    unp = sum(q["unpivoted_factorizations"] for q in st)
    return {"workload": f"C4: R={R}, N_r even in [256, 4096] (sum {T}), heaviside cutoff 0.05, tau=0.05, seed {seed}",
            "iterations": iterations, "seconds": t, "realization_steps_per_s": R * iterations / t,
            "accepted": int(sum(q["successful_steps"] for q in st)), "failed": int(sum(q["failed_steps"] for q in st)),
            "factorizations": int(fac), "pivoted_factorizations": int(fac - unp)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--cpu-seconds", type=float, default=6.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the speed-only and scaled-matrix side measurements")
    ap.add_argument("--ens-steps", type=int, default=2, help="timed lockstep iterations of the ensemble")
    ap.add_argument("--ens-r", type=int, default=256, help="realizations in the ensemble")
    ap.add_argument("--ens-solver", type=int, default=1, choices=(0, 1),
                    help="with --ens-only: 1 = GPU dense LU (default), 0 = host sparse LU for every realization")
    ap.add_argument("--no-ens", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ens-no-graph", action="store_true",
                    help="with --ens-only: launch every pass without hipGraph replays (for hardware counters)")
    ap.add_argument("--no-step", action="store_true", help="skip the N=1024 implicit-step measurement")
    ap.add_argument("--dstep-only", action="store_true",
                    help="only the device-resident step measurement (C2 at 1 and 64 realizations, the bench ensemble)")
    ap.add_argument("--step-only", action="store_true",
                    help="only the N=1024 implicit-step measurement, without its CPU leg (for profiling)")
    ap.add_argument("--step-solver", type=int, default=-1, choices=(-1, 0, 1),
                    help="with --step-only: 0 = host sparse LU, 1 = GPU dense LU, -1 = both")
    ap.add_argument("--ens-only", action="store_true",
                    help="only the batched ensemble (no plain comparison, no pair-pass headline), for profiling")
    ap.add_argument("--adjoint", type=str, default="",
                    help="only the adjoint measurement: N,T,dt[,method] (e.g. 2048,256,1e-3; method theta: T steps of "
                         "dt of the reference's step, adaptive: the steps the reference's loop accepts until T dt, "
                         "backward_euler), checkpoints every 16 steps")
    ap.add_argument("--c4-dstep", type=int, default=0,
                    help="only the C4 workload through the device-resident step for this many lockstep iterations")
    ap.add_argument("--c5-adjoint", type=str, default="",
                    help="only C5's strain gradient: N,T,ncoord (e.g. 2048,256,32) through T steps of the reference's "
                         "own accepted sizes from uniform positions, against central differences on ncoord coordinates")
    ap.add_argument("--c5-check", type=int, default=0,
                    help="only the C5 configuration through the reference's own step for this many steps")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): this library runs each pass on one "
              "GPU and does not shard it across several", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calib_copy_mib:
        nel = args.calib_copy_mib * (1 << 20) // 8
        src = torch.rand(nel, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": nel * 8}), flush=True)
        return
    if args.c4_dstep:
        print(json.dumps({"c4_device_step": c4_device_run(args.c4_dstep)}), flush=True)
        return
    if args.c5_adjoint:
        n_, T_, k_ = (int(v) for v in args.c5_adjoint.split(","))
        print(json.dumps({"c5_adjoint": c5_adjoint_bench(n_, T_, k_)}), flush=True)
        return
    if args.c5_check:
        print(json.dumps({"c5_reference_step": c5_reference_step_check(args.c5_check)}), flush=True)
        return
    if args.adjoint:
        parts = args.adjoint.split(",")
        method = parts[3] if len(parts) > 3 else "theta"
        print(json.dumps({"adjoint": adjoint_bench(int(parts[0]), int(parts[1]), float(parts[2]), 16, method)}),
              flush=True)
        return
    pkg = importlib.import_module("2d4_amd")
    dev = torch.device("cuda:0")
    if args.step_only:
        sctx = pkg.Context(0)
        solvers = (0, 1) if args.step_solver < 0 else (args.step_solver,)
        print(json.dumps({"implicit_step": implicit_step_bench(pkg, sctx, with_cpu=False, solvers=solvers)}),
              flush=True)
        return
    if args.dstep_only:
        print(json.dumps({"device_step": device_step_bench(pkg, pkg.Context(0))}), flush=True)
        return
    if args.ens_only:
        ectx = pkg.Context(0)
        print(json.dumps({"ensemble": ensemble_bench(pkg, ectx, args.ens_steps, args.ens_r, plain=False, solver=args.ens_solver,
                                                     graphs=not args.ens_no_graph)}), flush=True)
        return
    n = args.n
    ctx = pkg.Context(0)
    xy = torch.from_numpy(sorted_config(n, 1234)).to(dev)
    M = torch.empty((n, n), dtype=torch.float64, device=dev)
    p = pkg.jac_params(h=1e-6, tau_a=0.0, tau_b=0.1, n=n, wfunc="c")
    pairs = n * (n - 1)
    unordered = pairs // 2

    log(f"[bench] n={n} warmup={args.warmup}")
    for _ in range(args.warmup):
        ctx.jacobian_dense_t(xy, p, M=M)
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    ctx.pair_kernel_time()
    log(f"[bench] timing {args.steps} passes")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.jacobian_dense_t(xy, p, M=M)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    k_ms, k_launches = ctx.pair_kernel_time()
    ctx.set_profiling(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    k_avg_ms = k_ms / max(k_launches, 1)
    value = pairs / (ms_step * 1e-3)
    achieved = OPS_JAC_PAIR * unordered / (k_avg_ms * 1e-3) / 1e12

    extra = {}
    if not args.no_extra:
        v = torch.empty(n, dtype=torch.float64, device=dev)
        for _ in range(2):
            ctx.speed_pass_t(xy, 0.1, out=v)
        torch.cuda.synchronize()
        ctx.set_profiling(True)
        ctx.pair_kernel_time()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.speed_pass_t(xy, 0.1, out=v)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s_ms, s_l = ctx.pair_kernel_time()
        ctx.set_profiling(False)
        s_kavg = s_ms / max(s_l, 1)
        extra["speed_pass"] = dict(ms_per_pass=(t1 - t0) * 1e3 / args.steps, kernel_ms=s_kavg,
                                   pair_evals_per_s=pairs / ((t1 - t0) / args.steps),
                                   kernel_tflops=OPS_SPEED_PAIR * unordered / (s_kavg * 1e-3) / 1e12)
        # the reference's final matrix A = I + diag((1+d)/2) M (simulation.cpp:516-522)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.jacobian_dense_t(xy, p, M=M, scale=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        extra["scaled_matrix_pass"] = dict(ms_per_pass=(t1 - t0) * 1e3 / args.steps,
                                           pair_evals_per_s=pairs / ((t1 - t0) / args.steps))
        # the tabulated-field variant: 1024 x 1024 FP64 tables of sigma and d sigma, bilinear
        tab_s, tab_d = ctx.table_build_t(1024)
        for name, fn in (("table_speed_pass", lambda: ctx.speed_pass_table_t(xy, tab_s, 0.1, out=v)),
                         ("table_jacobian_pass", lambda: ctx.jacobian_dense_table_t(xy, p, tab_s, tab_d, M=M))):
            fn()
            torch.cuda.synchronize()
            ctx.set_profiling(True)
            ctx.pair_kernel_time()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            k_ms2, k_l2 = ctx.pair_kernel_time()
            ctx.set_profiling(False)
            extra[name] = dict(ms_per_pass=(t1 - t0) * 1e3 / args.steps, kernel_ms=k_ms2 / max(k_l2, 1),
                               pair_evals_per_s=pairs / ((t1 - t0) / args.steps), table="1024 x 1024 FP64")
        del tab_s, tab_d
        # the same pass through the C ABI with host buffers (d4_jacobian_dense): positions up, then speeds,
        # weights and the n x n matrix down into the caller's pageable memory; reported beside `value`, not as it
        xy_h = xy.cpu().numpy()
        Mh = np.empty((n, n))
        ctx.jacobian_dense(xy_h, p, out=Mh)  # touches the host pages once
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.jacobian_dense(xy_h, p, out=Mh)
        t1 = time.perf_counter()
        extra["host_buffer_pass"] = dict(ms_per_pass=(t1 - t0) * 1e3 / reps, passes=reps,
                                         pair_evals_per_s=pairs / ((t1 - t0) / reps),
                                         matrix_bytes_to_host=8 * n * n)
        del Mh, xy_h
    del M
    torch.cuda.empty_cache()
    ensemble = None
    if not args.no_ens:
        ensemble = ensemble_bench(pkg, ctx, args.ens_steps, args.ens_r)
    step = None
    dstep_res = None
    if not args.no_step:
        log("[bench] implicit step, N=1024 (GPU, then the CPU loop)")
        step = implicit_step_bench(pkg, ctx, with_cpu=not args.no_cpu)
        log("[bench] device-resident implicit step, N=1024, 1 and 64 realizations; the bench ensemble")
        dstep_res = device_step_bench(pkg, ctx)
        if ensemble is not None and "plain_steps_per_s" in ensemble and args.ens_r == 256:
            # the same workload as one captured graph per lockstep iteration (the device-resident step)
            g = dstep_res["ensemble_R256"]["realization_steps_per_s"]
            ensemble["graph_step_steps_per_s"] = g
            ensemble["graph_step_ratio_vs_plain"] = g / ensemble["plain_steps_per_s"]

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_pair_dense_n16384.json")
    if os.path.exists(tpath) and n == 16384:
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, this host)")
        c = cpu_baseline(2048, 4096, args.cpu_seconds)
        cpu = {"value": c["omp"]["pair_evals_per_s"], "unit": "pair evals/s", "cores": c["omp"]["threads"],
               "kind": "port",
               "sample": (f"oracle dense Jacobian+speed pass (liboracle_fast.so, reference flags -Ofast "
                          f"-ffast-math), N={c['omp']['n']} with OpenMP over rows (each pair evaluated per row), "
                          f"median of {c['omp']['reps']} on {cpu_model()}; single-thread i<j loop as the "
                          f"reference runs it: N={c['1core']['n']}, "
                          f"{c['1core']['pair_evals_per_s']:.4g} pair evals/s"),
               "single_thread_value": c["1core"]["pair_evals_per_s"]}

    line = {
        "metric": "FP64 pair evals/s (\u03c3+\u2202\u03c3) at N=16384; ensemble steps/s batched vs plain",
        "value": value,
        "unit": "pair evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: uniform random positions in the unit periodic cell, sorted as SimulationData sorts",
        "config": {"workload": f"Jacobian+speed pass, N={n} (half b=+1, half b=-1), unlimited cutoff, dense "
                               f"M = V + diag(s) column-major, speeds at two stresses, weights d",
                   "n": n, "pair_evals_per_step": pairs, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "pipe": "FP64 VALU (MI355X FP64 vector peak)", "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": traffic, "kernel": "pair_pass_kernel<kModeJacDense>",
                     "kernel_avg_ms": k_avg_ms, "ops_per_launch": OPS_JAC_PAIR * unordered},
        "cpu_baseline": cpu,
        "ensemble": ensemble,
        "implicit_step": step,
        "device_step": dstep_res,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
