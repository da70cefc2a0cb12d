"""The GPU dense LU of the simulation's linear solves (d4_lu_solve_batch,
include/d4_sim.h; UMFPACK's place, simulation.cpp:635-659) against numpy's
LAPACK solve, its threshold pivoting rule (u = 0.001, UMFPACK's symmetric
strategy: the diagonal is kept while every multiplier is at most 1000 in
magnitude), singular matrices, invalid input, the independence of a
system's bits from the rest of its batch, and the reordering of cutoff
patterns into a narrow envelope whose zero blocks the factorization skips."""
from importlib import import_module

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
simmod = import_module("2d4_amd.sim")


def _csc(A):
    Ap, Ai, Ax = [0], [], []
    for c in range(A.shape[1]):
        rows = np.nonzero(A[:, c])[0]
        Ai.extend(rows.tolist())
        Ax.extend(A[rows, c].tolist())
        Ap.append(len(Ai))
    return np.array(Ap, np.int32), np.array(Ai, np.int32), np.array(Ax, np.float64)


def _sys(A, b):
    return (*_csc(A), b)


def _dominant(n, rng, density=0.2):
    A = rng.uniform(-1, 1, (n, n)) * (rng.uniform(0, 1, (n, n)) < density) * (0.5 / n)
    return A + np.eye(n)


def _needs_pivoting(n, rng):
    # first-stage multipliers |a_i0 / a_00| far above 1000
    A = rng.standard_normal((n, n))
    A[np.diag_indices(n)] *= 1e-7
    return A


def _rel_err(x, ref):
    return float(np.max(np.abs(x - ref)) / max(1.0, float(np.max(np.abs(ref)))))


def _backward(A, x, b):
    r = np.linalg.norm(A @ x - b, np.inf)
    return float(r / (np.linalg.norm(A, np.inf) * np.linalg.norm(x, np.inf) + np.linalg.norm(b, np.inf)))


def test_dominant_systems_match_lapack(ctx):
    rng = np.random.default_rng(1)
    sizes = [1, 2, 63, 64, 65, 255, 256, 257, 300, 700, 1100]
    mats = [_dominant(n, rng) for n in sizes]
    bs = [rng.uniform(-1, 1, n) for n in sizes]
    xs, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(A, b) for A, b in zip(mats, bs)])
    assert not sing.any() and unpiv.all()
    for A, b, x in zip(mats, bs, xs):
        err = _rel_err(x, np.linalg.solve(A, b))
        assert err <= 1e-12, (A.shape, err)


def test_threshold_rule_at_the_boundary(ctx):
    # the multiplier of the first stage is exactly 1000 (kept) or 1000.5
    # (pivoted); in the 3x3 case the second stage's pivot 1e-4 gives 1e4
    keep = np.array([[1.0, 0.5], [1000.0, 1.0]])
    swap = np.array([[1.0, 0.5], [1000.5, 1.0]])
    late = np.array([[1.0, 0.0, 0.0], [0.0, 1e-4, 1.0], [0.0, 1.0, 1.0]])
    mats = [keep, swap, late]
    bs = [np.array([1.0, 2.0]), np.array([1.0, 2.0]), np.array([1.0, 2.0, 3.0])]
    xs, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(A, b) for A, b in zip(mats, bs)])
    assert unpiv.tolist() == [True, False, False]
    assert not sing.any()
    for A, b, x in zip(mats, bs, xs):
        assert _rel_err(x, np.linalg.solve(A, b)) <= 1e-15


def test_pivoting_systems_match_lapack(ctx):
    rng = np.random.default_rng(2)
    mats = [_needs_pivoting(n, rng) for n in (50, 300, 600, 1300)]
    z = rng.standard_normal((257, 257))
    np.fill_diagonal(z, 0.0)  # zero pivots without row interchanges
    mats += [z, np.array([[0.0, 1.0], [1.0, 0.0]])]
    bs = [rng.standard_normal(A.shape[0]) for A in mats]
    xs, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(A, b) for A, b in zip(mats, bs)])
    assert not sing.any() and not unpiv.any()
    for A, b, x in zip(mats, bs, xs):
        bw = _backward(A, x, b)
        assert bw <= 1e-14, (A.shape, bw)
        fw = _rel_err(x, np.linalg.solve(A, b))
        assert fw <= 1e-15 * np.linalg.cond(A, np.inf) * A.shape[0], (A.shape, fw)


def test_general_matrices_within_backward_bound(ctx):
    # random matrices that the rule may keep without interchanges (multipliers
    # below 1000, growth up to a few thousand): backward stable to 1e-12
    rng = np.random.default_rng(6)
    mats = [rng.standard_normal((n, n)) for n in (50, 50, 100, 100, 300, 300, 600)]
    bs = [rng.standard_normal(A.shape[0]) for A in mats]
    xs, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(A, b) for A, b in zip(mats, bs)])
    assert not sing.any()
    print(f"kept without interchanges: {int(unpiv.sum())} of {len(mats)}", flush=True)
    for A, b, x in zip(mats, bs, xs):
        bw = _backward(A, x, b)
        assert bw <= 1e-12, (A.shape, bw)


def test_singular_matrix_is_reported(ctx):
    A = _dominant(40, np.random.default_rng(3))
    A[:, 17] = 0.0
    _, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(A, np.ones(40)), _sys(np.eye(3), np.ones(3))])
    assert sing.tolist() == [True, False]
    assert unpiv.tolist() == [False, True]


def test_batch_composition_does_not_change_bits(ctx):
    rng = np.random.default_rng(4)
    mats = [_dominant(n, rng) for n in (300, 1030, 64)] + [_needs_pivoting(n, rng) for n in (520, 70)]
    systems = [_sys(A, rng.uniform(-1, 1, A.shape[0])) for A in mats]
    together, _, unpiv = simmod.lu_solve_batch(ctx, systems)
    assert unpiv.tolist() == [True, True, True, False, False]
    alone = [simmod.lu_solve_batch(ctx, [s])[0][0] for s in systems]
    extra = [_sys(_dominant(n, rng), np.ones(n)) for n in (290, 1100)]
    mixed, _, _ = simmod.lu_solve_batch(ctx, extra[:1] + systems[::-1] + extra[1:])
    mixed = mixed[1:-1][::-1]
    for a, t, m in zip(alone, together, mixed):
        np.testing.assert_array_equal(a, t)
        np.testing.assert_array_equal(a, m)


def test_invalid_matrices_are_rejected(ctx):
    Ap, Ai, Ax = _csc(np.eye(4))
    bad_row = Ai.copy()
    bad_row[2] = 4
    with pytest.raises(RuntimeError):
        simmod.lu_solve_batch(ctx, [(Ap, bad_row, Ax, np.ones(4))])
    bad_start = Ap.copy() + 1
    with pytest.raises(RuntimeError):
        simmod.lu_solve_batch(ctx, [(bad_start, Ai, Ax, np.ones(4))])
    # the context stays usable
    xs, _, _ = simmod.lu_solve_batch(ctx, [(Ap, Ai, Ax, np.arange(4.0))])
    np.testing.assert_array_equal(xs[0], np.arange(4.0))


def test_dense_device_matrix_matches_csc_path(ctx):
    # the same matrices as a dense CUDA tensor (row-major, and read transposed)
    # give the bits of the CSC entry point, and solve against numpy's LAPACK
    import torch
    rng = np.random.default_rng(9)
    for A in (_dominant(300, rng), _needs_pivoting(130, rng)):
        n = A.shape[0]
        b = rng.standard_normal(n)
        (x_csc,), _, (unpiv,) = simmod.lu_solve_batch(ctx, [_sys(A, b)])
        At = torch.from_numpy(A).to("cuda:0")
        bt = torch.from_numpy(b).to("cuda:0")
        sing, u = ctx.lu_factor_dense_t(At)
        assert not sing and u == bool(unpiv)
        x = ctx.lu_solve_dense_t(bt).cpu().numpy()
        np.testing.assert_array_equal(x, x_csc)
        assert _backward(A, x, b) <= 1e-14
        # the transpose of the same storage: A^T x = b
        ctx.lu_factor_dense_t(At, transpose=True)
        xt = ctx.lu_solve_dense_t(bt).cpu().numpy()
        assert _backward(A.T, xt, b) <= 1e-14


def test_large_and_empty_systems_in_one_batch(ctx):
    # order 4096 (64 block steps of the factorization) beside an empty system
    # and a 1 x 1 one; the empty system's solution is empty
    rng = np.random.default_rng(10)
    big = _dominant(4096, rng, density=0.05)
    b_big = rng.uniform(-1, 1, 4096)
    empty = (np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), np.zeros(0))
    one = (np.array([0, 1], np.int32), np.array([0], np.int32), np.array([4.0]), np.array([2.0]))
    xs, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(big, b_big), empty, one])
    assert not sing.any() and unpiv.all()
    assert xs[1].size == 0
    assert xs[2][0] == 0.5
    assert _rel_err(xs[0], np.linalg.solve(big, b_big)) <= 1e-12
    assert _backward(big, xs[0], b_big) <= 1e-13  # the scale of n u = 4096 * 1.1e-16


def _cutoff_matrix(n, rng, rc=0.05, scale=0.5):
    # a step matrix's shape: the identity plus entries on the heaviside-cutoff
    # pattern of n random points in the periodic unit square, rows scaled so
    # that the off-diagonal sum stays below `scale`
    x, y = rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n)
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    dx -= np.rint(dx)
    dy -= np.rint(dy)
    P = (dx * dx + dy * dy < rc * rc) & ~np.eye(n, dtype=bool)
    V = np.where(P, rng.uniform(-1, 1, (n, n)), 0.0)
    return np.eye(n) + V * (scale / max(1.0, np.abs(V).sum(1).max()))


def test_cutoff_systems_reordered_match_lapack(ctx):
    # cutoff patterns are factored in reverse Cuthill-McKee order with the
    # blocks outside the envelope skipped; the orders cover 1 to 64 blocks
    rng = np.random.default_rng(12)
    sizes = [90, 300, 1000, 2100, 4000]
    mats = [_cutoff_matrix(n, rng) for n in sizes]
    bs = [rng.uniform(-1, 1, n) for n in sizes]
    xs, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(A, b) for A, b in zip(mats, bs)])
    assert not sing.any() and unpiv.all()
    for A, b, x in zip(mats, bs, xs):
        assert _rel_err(x, np.linalg.solve(A, b)) <= 1e-12, A.shape
        assert _backward(A, x, b) <= 1e-15 * A.shape[0], A.shape
    # the bits of each system do not depend on its batch
    alone = [simmod.lu_solve_batch(ctx, [_sys(A, b)])[0][0] for A, b in zip(mats, bs)]
    for a, t in zip(alone, xs):
        np.testing.assert_array_equal(a, t)


def test_envelope_skipping_keeps_the_bits_of_the_full_factorization(ctx):
    # a matrix already narrow in its own order keeps that order; the blocks the
    # envelope skips are zero, so the CSC path (skipping) and the dense device
    # path (every block) give the same bits
    import torch
    rng = np.random.default_rng(13)
    n = 700
    i, j = np.indices((n, n))
    A = np.where(np.abs(i - j) <= 40, rng.uniform(-1, 1, (n, n)), 0.0) * 0.012 + np.eye(n)
    b = rng.standard_normal(n)
    (x_csc,), _, (unpiv,) = simmod.lu_solve_batch(ctx, [_sys(A, b)])
    assert unpiv
    sing, u = ctx.lu_factor_dense_t(torch.from_numpy(A).to("cuda:0"))
    assert not sing and u
    x = ctx.lu_solve_dense_t(torch.from_numpy(b).to("cuda:0")).cpu().numpy()
    np.testing.assert_array_equal(x, x_csc)
    assert _rel_err(x, np.linalg.solve(A, b)) <= 1e-13


def test_reordered_matrix_agrees_with_its_natural_order(ctx):
    # the same cutoff matrix through the CSC path (reordered) and the dense
    # device path (natural order): different rounding, the same solution
    import torch
    rng = np.random.default_rng(14)
    A = _cutoff_matrix(1500, rng)
    b = rng.standard_normal(1500)
    (x_re,), _, _ = simmod.lu_solve_batch(ctx, [_sys(A, b)])
    ctx.lu_factor_dense_t(torch.from_numpy(A).to("cuda:0"))
    x_nat = ctx.lu_solve_dense_t(torch.from_numpy(b).to("cuda:0")).cpu().numpy()
    assert _rel_err(x_re, x_nat) <= 1e-13


def test_reordered_matrix_needing_interchanges(ctx):
    # a cutoff pattern with a diagonal far below its off-diagonal entries: the
    # multiplier check fails and the reordered matrix is refactored whole with
    # partial pivoting; the solution comes back in the caller's order
    rng = np.random.default_rng(15)
    A = _cutoff_matrix(800, rng, scale=1.0)
    A[np.diag_indices(800)] = rng.uniform(0.5, 1.0, 800) * 1e-6
    b = rng.standard_normal(800)
    good = _cutoff_matrix(600, rng)
    b2 = rng.standard_normal(600)
    xs, sing, unpiv = simmod.lu_solve_batch(ctx, [_sys(A, b), _sys(good, b2)])
    assert not sing.any()
    assert unpiv.tolist() == [False, True]
    assert _backward(A, xs[0], b) <= 1e-13
    assert _rel_err(xs[1], np.linalg.solve(good, b2)) <= 1e-12


def test_order_kept_for_a_repeated_pattern(ctx):
    # a slot whose pattern repeats keeps its order; the values are new each time,
    # and the results equal those of a context that computes the order afresh
    import importlib
    pkg = importlib.import_module("2d4_amd")
    rng = np.random.default_rng(16)
    A = _cutoff_matrix(1200, rng)
    P = A != 0
    A2 = np.where(P, A * rng.uniform(0.5, 1.5, A.shape), 0.0)
    b = rng.uniform(-1, 1, 1200)
    simmod.lu_solve_batch(ctx, [_sys(A, b)])
    (x_again,), _, _ = simmod.lu_solve_batch(ctx, [_sys(A2, b)])
    fresh = pkg.Context(0)
    try:
        (x_fresh,), _, _ = simmod.lu_solve_batch(fresh, [_sys(A2, b)])
    finally:
        fresh.close()
    np.testing.assert_array_equal(x_again, x_fresh)
    assert _rel_err(x_again, np.linalg.solve(A2, b)) <= 1e-12
    # a different pattern in the same slot is ordered afresh
    B = _cutoff_matrix(1200, rng)
    (xb,), _, _ = simmod.lu_solve_batch(ctx, [_sys(B, b)])
    assert _rel_err(xb, np.linalg.solve(B, b)) <= 1e-12
