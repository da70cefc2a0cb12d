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
