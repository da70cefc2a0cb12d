"""The host sparse LU that stands where the reference calls UMFPACK
(simulation.cpp:635-659), checked against numpy's dense LAPACK solve."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import random_config, wrap

SHIM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "libsolver_shim.so")


@pytest.fixture(scope="module")
def ts():
    if not os.path.exists(SHIM):
        pytest.skip("tests/csrc/libsolver_shim.so not built (__graft_entry__.build())")
    L = C.CDLL(SHIM)
    L.ts_solve.argtypes = [C.c_int] + [C.c_void_p] * 5 + [C.POINTER(C.c_longlong)]
    return L


def _csc(A):
    n = A.shape[0]
    Ap, Ai, Ax = [0], [], []
    for c in range(n):
        rows = np.flatnonzero(A[:, c])
        Ai += rows.tolist()
        Ax += A[rows, c].tolist()
        Ap.append(len(Ai))
    return np.array(Ap, np.int32), np.array(Ai, np.int32), np.array(Ax)


def _solve(ts, A, b):
    Ap, Ai, Ax = _csc(A)
    x = np.empty(A.shape[0])
    fill = C.c_longlong(0)
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    rc = ts.ts_solve(A.shape[0], p(Ap), p(Ai), p(Ax), p(b), p(x), C.byref(fill))
    return rc, x, fill.value


def _geometric_system(n, rc, seed, h=1e-3):
    """The structure of an implicit-step matrix: I + diag + couplings within rc, random signs."""
    xy = random_config(n, seed)
    rng = np.random.default_rng(seed)
    A = np.eye(n)
    for i in range(n):
        d = wrap(xy[i] - xy)
        near = np.flatnonzero((d ** 2).sum(1) < rc * rc)
        for j in near:
            if j > i:
                v = h * rng.normal() / max((d[j] ** 2).sum(), 1e-4)
                A[i, j] += v
                A[j, i] += v
    A += np.diag(np.abs(A).sum(1) * rng.uniform(0.0, 0.5, n))
    return A


@pytest.mark.parametrize("n,rc", [(50, 1.0), (300, 0.08), (1024, 0.05), (2000, 0.05)])
def test_sparse_lu_matches_dense_solve(ts, n, rc):
    A = _geometric_system(n, rc, n)
    b = np.random.default_rng(1).normal(size=n)
    rc_, x, fill = _solve(ts, A, b)
    assert rc_ == 0
    x0 = np.linalg.solve(A, b)
    assert np.abs(x - x0).max() <= 1e-12 * np.abs(x0).max()
    assert np.abs(A @ x - b).max() <= 1e-13 * (np.abs(A).sum(1).max() * np.abs(x).max())


def test_sparse_lu_pivots_off_the_diagonal(ts):
    """Zero and tiny diagonals force row pivoting (threshold 0.1)."""
    n = 400
    A = _geometric_system(n, 0.1, 5)
    np.fill_diagonal(A, 0.0)
    A[np.arange(0, n, 7), np.arange(0, n, 7)] = 1e-14
    A += np.diag(np.where(np.arange(n) % 3 == 0, 0.0, 1.0))
    if abs(np.linalg.det(A[:50, :50])) == 0:
        pytest.skip("degenerate sample")
    b = np.random.default_rng(2).normal(size=n)
    rc_, x, _ = _solve(ts, A, b)
    x0 = np.linalg.solve(A, b)
    cond = np.linalg.cond(A)
    assert np.abs(x - x0).max() <= 1e-15 * cond * np.abs(x0).max() * 10


def test_sparse_lu_reports_singular(ts):
    A = _geometric_system(200, 0.1, 6)
    A[:, 17] = 0.0
    A[17, :] = 0.0
    rc_, x, _ = _solve(ts, A, np.ones(200))
    assert rc_ == 1


# --- the GPU dense LU's symmetric order and block envelope (host code) -------

def _band(ts, A, bs=64):
    ts.ts_band_order.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.POINTER(C.c_longlong)]
    n = A.shape[0]
    npad = (n + bs - 1) // bs * bs
    nb = npad // bs
    Ap, Ai, _ = _csc(A)
    ordv = np.empty(n, np.int32)
    first, last, top = (np.empty(nb, np.int32) for _ in range(3))
    tiles = C.c_longlong(0)
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    re = ts.ts_band_order(n, npad, bs, p(Ap), p(Ai), p(ordv), p(first), p(last), p(top), C.byref(tiles))
    return bool(re), ordv, first, last, top, tiles.value


def _cutoff_pattern(n, rng, rc=0.05):
    # the heaviside-cutoff Jacobian pattern of n random points in the periodic unit square
    x, y = rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n)
    dx, dy = wrap(x[:, None] - x[None, :]), wrap(y[:, None] - y[None, :])
    return dx * dx + dy * dy < rc * rc


def _lu_nopivot(A):
    A = A.copy()
    for k in range(A.shape[0] - 1):
        A[k + 1:, k] /= A[k, k]
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    return A


def _outside_envelope(F, last, bs=64):
    # nonzeros of the factors in blocks the envelope declares zero
    nb = len(last)
    bad = 0
    for p in range(nb):
        for b in range(last[p] + 1, nb):
            bad += np.count_nonzero(F[b * bs:(b + 1) * bs, p * bs:(p + 1) * bs])
            bad += np.count_nonzero(F[p * bs:(p + 1) * bs, b * bs:(b + 1) * bs])
    return bad


@pytest.mark.parametrize("n", [300, 1000])
def test_cutoff_pattern_is_reordered_into_a_narrow_envelope(ts, n):
    rng = np.random.default_rng(n)
    P = _cutoff_pattern(n, rng)
    re, ordv, first, last, top, tiles = _band(ts, P.astype(float))
    nb = len(last)
    assert re
    assert np.array_equal(np.sort(ordv), np.arange(n))
    dense_tiles = sum((nb - p - 1) ** 2 for p in range(nb))
    assert tiles * 8 <= dense_tiles, (tiles, dense_tiles)
    assert tiles == sum((last[p] - p) ** 2 for p in range(nb))
    # envelope invariants: last nondecreasing and >= its block, top <= first <= block
    assert np.all(np.diff(last) >= 0) and np.all(last >= np.arange(nb))
    assert np.all(top <= first) and np.all(first <= np.arange(nb))
    # elimination without interchanges of the reordered matrix (padded with
    # the identity) leaves every block past the envelope exactly zero
    vals = np.where(P, rng.uniform(-1, 1, P.shape), 0.0) / P.sum(1).max()
    A = vals + 2.0 * np.eye(n)
    npad = nb * 64
    Ar = np.eye(npad)
    Ar[:n, :n] = A[np.ix_(ordv, ordv)]
    assert _outside_envelope(_lu_nopivot(Ar), last) == 0


@pytest.mark.parametrize("n,rc", [(400, 0.05), (2048, 0.05), (1000, 0.12)])
def test_key_band_envelope_holds_for_any_x(ts, n, rc):
    # rows sorted by |y|: the envelope from the keys alone (the device-resident step fixes it once per run)
    # contains the envelope of the heaviside-cutoff pattern for every x, and the no-interchange LU of such a
    # matrix stays inside it; and it costs about what reverse Cuthill-McKee's per-pattern envelope does
    ts.ts_key_band_envelope.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_double] + [C.c_void_p] * 3 + \
        [C.POINTER(C.c_longlong)]
    rng = np.random.default_rng(n)
    y = rng.uniform(-0.5, 0.5, n)
    ordv = np.argsort(np.abs(y), kind="stable")
    key = np.ascontiguousarray(np.abs(y)[ordv])
    bs = 64
    npad = (n + bs - 1) // bs * bs
    nb = npad // bs
    first, last, top = (np.empty(nb, np.int32) for _ in range(3))
    tiles = C.c_longlong(0)
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    ts.ts_key_band_envelope(n, npad, bs, p(key), rc * (1 + 1e-12), p(first), p(last), p(top), C.byref(tiles))
    assert np.all(np.diff(last) >= 0) and np.all(last >= np.arange(nb)) and np.all(top <= first)
    for trial in range(3):
        x = rng.uniform(-0.5, 0.5, n)
        dx, dy = wrap(x[:, None] - x[None, :]), wrap(y[:, None] - y[None, :])
        P = (dx * dx + dy * dy < rc * rc)[np.ix_(ordv, ordv)]
        i, j = np.nonzero(P)
        lo, hi = np.minimum(i, j), np.maximum(i, j)
        assert np.all(last[lo // bs] >= hi // bs), "an entry outside the fixed envelope"
        if trial == 0 and n <= 1000:
            vals = np.where(P, rng.uniform(-1, 1, P.shape), 0.0) / P.sum(1).max()
            Ar = np.eye(npad)
            Ar[:n, :n] = vals + 2.0 * np.eye(n)
            assert _outside_envelope(_lu_nopivot(Ar), last) == 0
    if n >= 2048:
        _, _, _, _, _, rcm_tiles = _band(ts, P.astype(float))
        assert tiles.value <= 1.25 * rcm_tiles, (tiles.value, rcm_tiles)


def test_narrow_natural_order_is_kept(ts):
    # a band of half-width 40 in its own order: reverse Cuthill-McKee cannot halve the tiles
    n = 700
    i, j = np.indices((n, n))
    A = np.where(np.abs(i - j) <= 40, 0.01, 0.0) + np.eye(n)
    re, ordv, first, last, top, tiles = _band(ts, A)
    assert not re and np.array_equal(ordv, np.arange(n))
    assert list(last[:-1]) == [b + 1 for b in range(len(last) - 1)]
    npad = len(last) * 64
    Ap = np.eye(npad)
    Ap[:n, :n] = A
    assert _outside_envelope(_lu_nopivot(Ap), last) == 0


def test_rcm_order_is_a_permutation_per_component(ts):
    # two components and an isolated node: every node appears once
    ts.ts_rcm_order.argtypes = [C.c_int] + [C.c_void_p] * 3
    A = np.eye(9)
    for a, b in [(0, 3), (3, 5), (1, 2), (2, 7), (7, 8)]:
        A[a, b] = A[b, a] = 1.0
    Ap, Ai, _ = _csc(A)
    o = np.empty(9, np.int32)
    ts.ts_rcm_order(9, C.c_void_p(Ap.ctypes.data), C.c_void_p(Ai.ctypes.data), C.c_void_p(o.ctypes.data))
    assert sorted(o.tolist()) == list(range(9))
    # a path graph is numbered end to end: neighbours stay adjacent
    pos = np.empty(9, int)
    pos[o] = np.arange(9)
    for a, b in [(0, 3), (3, 5), (1, 2), (2, 7), (7, 8)]:
        assert abs(pos[a] - pos[b]) == 1


def _refactor(ts, A1, A2, b):
    ts.ts_refactor.argtypes = [C.c_int] + [C.c_void_p] * 7
    Ap, Ai, Ax1 = _csc(A1)
    Ap2, Ai2, Ax2 = _csc(A2)
    assert np.array_equal(Ap, Ap2) and np.array_equal(Ai, Ai2)
    x = np.empty(A1.shape[0])
    rep = C.c_longlong(-1)
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    rc = ts.ts_refactor(A1.shape[0], p(Ap), p(Ai), p(Ax1), p(Ax2), p(b), p(x), C.byref(rep))
    return rc, x, rep.value


def test_refactorization_replays_the_reach_bit_for_bit(ts):
    # a cutoff pattern factored twice: new values with the same pivot choices
    # replay the first factorization's reach and give a fresh factorization's bits
    rng = np.random.default_rng(21)
    n = 400
    P = _cutoff_pattern(n, rng, rc=0.08)
    A1 = np.where(P, rng.uniform(-1, 1, P.shape), 0.0) * 0.05 + np.eye(n)
    A2 = np.where(P, A1 * rng.uniform(0.9, 1.1, P.shape), 0.0)
    b = rng.uniform(-1, 1, n)
    rc, x, rep = _refactor(ts, A1, A2, b)
    assert rc == 0 and rep == 1
    rc_f, x_fresh, _ = _solve(ts, A2, b)
    np.testing.assert_array_equal(x, x_fresh)
    assert np.max(np.abs(A2 @ x - b)) <= 1e-13


def test_refactorization_falls_back_when_a_pivot_changes(ts):
    # the second matrix's diagonal is too small to keep: the pivot sequence
    # changes, the replay stops and the full factorization runs
    rng = np.random.default_rng(22)
    n = 400
    P = _cutoff_pattern(n, rng, rc=0.08)
    A1 = np.where(P, rng.uniform(-1, 1, P.shape), 0.0) * 0.05 + np.eye(n)
    A2 = A1.copy()
    A2[np.diag_indices(n)] = 1e-9
    b = rng.uniform(-1, 1, n)
    rc, x, rep = _refactor(ts, A1, A2, b)
    assert rep == 0
    rc_f, x_fresh, _ = _solve(ts, A2, b)
    assert rc == rc_f
    np.testing.assert_array_equal(x, x_fresh)


@pytest.mark.parametrize("moved,orders", [(2, 1), (60, 2)])
def test_column_order_kept_while_the_pattern_drifts_by_at_most_one_percent(ts, moved, orders):
    # between two Jacobians a few pairs cross the cutoff: the order of the first
    # pattern stays while the patterns differ in <= 1 % of the entries
    ts.ts_two_patterns.argtypes = [C.c_int] + [C.c_void_p] * 8 + [C.POINTER(C.c_longlong)] * 2
    rng = np.random.default_rng(23)
    n = 500
    P = _cutoff_pattern(n, rng, rc=0.08)
    A1 = np.where(P, rng.uniform(-1, 1, P.shape), 0.0) * 0.05 + np.eye(n)
    A2 = A1.copy()
    off = np.argwhere(np.triu(P, 1))
    for i, j in off[rng.choice(len(off), moved // 2, replace=False)]:  # symmetric pairs leave the pattern
        A2[i, j] = A2[j, i] = 0.0
    b = rng.uniform(-1, 1, n)
    c1, c2 = _csc(A1), _csc(A2)
    removed = len(c1[1]) - len(c2[1])
    assert removed == moved and (100 * removed <= len(c1[1])) == (orders == 1)
    x = np.empty(n)
    nord, nrep = C.c_longlong(0), C.c_longlong(0)
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    rc = ts.ts_two_patterns(n, p(c1[0]), p(c1[1]), p(c1[2]), p(c2[0]), p(c2[1]), p(c2[2]), p(b), p(x),
                            C.byref(nord), C.byref(nrep))
    assert rc == 0 and nord.value == orders and nrep.value == 0
    assert np.max(np.abs(x - np.linalg.solve(A2, b))) <= 1e-12


@pytest.mark.parametrize("n", [70, 300, 1000])
def test_symmetric_envelope_pass_equals_the_general_one(ts, n):
    """block_envelope's pass for symmetric patterns (one gather per column, what the GPU LU uses for the step
    matrices) gives the general pass's envelope, in the natural order and in random orders; on an
    unsymmetric pattern the general pass still sees entries the symmetric one would miss."""
    ts.ts_block_envelope.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 3 + [C.c_int] + \
        [C.c_void_p] * 3 + [C.POINTER(C.c_longlong)]
    rng = np.random.default_rng(n)
    bs = 64
    npad = (n + bs - 1) // bs * bs
    nb = npad // bs
    p = lambda a: C.c_void_p(a.ctypes.data) if a is not None else None  # noqa: E731

    def env(A, inv, sym):
        Ap, Ai, _ = _csc(A)
        out = [np.empty(nb, np.int32) for _ in range(3)]
        tiles = C.c_longlong(0)
        ts.ts_block_envelope(n, npad, bs, p(Ap), p(Ai), p(inv), sym, p(out[0]), p(out[1]), p(out[2]),
                             C.byref(tiles))
        return [o.tolist() for o in out] + [tiles.value]

    A = (_cutoff_pattern(n, rng) | np.eye(n, dtype=bool)).astype(float)
    for inv in (None, rng.permutation(n).astype(np.int32), rng.permutation(n).astype(np.int32)):
        assert env(A, inv, 1) == env(A, inv, 0)
    U = np.eye(n)
    U[n - 1, 0] = 1.0  # one entry far below the diagonal, without its mirror
    assert env(U, None, 0) != env(U, None, 1)
