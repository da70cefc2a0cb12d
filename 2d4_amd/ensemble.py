"""Batched ragged ensembles (include/d4_batch.h).

Realization r is the slice off[r]:off[r+1] of concatenated arrays.  The
batched passes give, per realization, exactly the results of the
single-realization entry points; an Ensemble steps R host simulations
(2d4_amd.sim semantics each) in lockstep with one launch per pair pass for
all of them and the per-realization host work on a thread pool.
"""
import ctypes as C

import numpy as np

from ._bind_extra import EnsConfig, SimStats
from .sim import make_params


def _off(off):
    off = np.ascontiguousarray(off, dtype=np.int32)
    if off.ndim != 1 or off.size < 2 or off[0] != 0 or np.any(np.diff(off) < 0):
        raise ValueError("offsets must be 0, ..., total and non-decreasing")
    return off


def batch_speed_pass(ctx, off, xy, tau, tol=None, tol_cap2=None):
    """Speeds of every realization (d4_speed_pass semantics per slice)."""
    off = _off(off)
    R, T = off.size - 1, int(off[-1])
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(T, 2)
    tau = np.ascontiguousarray(np.broadcast_to(np.asarray(tau, dtype=np.float64), (R,)))
    v = np.empty(T)
    caps = None
    if tol is not None:
        tol = np.ascontiguousarray(tol, dtype=np.float64)
        caps = np.ascontiguousarray(np.broadcast_to(np.asarray(tol_cap2, dtype=np.float64), (R,)))
    ctx.check(ctx.L.d4_batch_speed_pass(ctx.h, R, off.ctypes.data, xy.ctypes.data, tau.ctypes.data, v.ctypes.data,
                                        None if tol is None else tol.ctypes.data,
                                        None if caps is None else caps.ctypes.data))
    return v if tol is None else (v, tol)


def batch_jacobian_pass(ctx, off, xy, params, tol=None, cap=None):
    """Jacobian pass of every realization; params: one JacParams per realization.
    Returns vA, vB, (Ap, Ai, Ax) block-diagonal with global indices, indexes, dvec[, tol]."""
    from importlib import import_module
    JacParams = import_module("2d4_amd").JacParams
    off = _off(off)
    R, T = off.size - 1, int(off[-1])
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(T, 2)
    parr = (JacParams * R)(*params)
    vA, vB, dvec = np.empty(T), np.empty(T), np.empty(T)
    Ap = np.empty(T + 1, dtype=np.int32)
    idx = np.empty(T, dtype=np.int32)
    if tol is not None:
        tol = np.ascontiguousarray(tol, dtype=np.float64)
    cap = cap or T * 16 + 16
    for _ in range(2):
        Ai = np.empty(cap, dtype=np.int32)
        Ax = np.empty(cap)
        nnz = C.c_int(0)
        rc = ctx.L.d4_batch_jacobian_pass(ctx.h, R, off.ctypes.data, xy.ctypes.data, C.cast(parr, C.c_void_p),
                                          vA.ctypes.data, vB.ctypes.data, Ap.ctypes.data, Ai.ctypes.data,
                                          Ax.ctypes.data, cap, C.byref(nnz), idx.ctypes.data, dvec.ctypes.data,
                                          None if tol is None else tol.ctypes.data)
        if rc == -3:  # D4_ERR_CAPACITY
            cap = nnz.value
            continue
        ctx.check(rc)
        break
    n = nnz.value
    out = (vA, vB, (Ap, Ai[:n].copy(), Ax[:n].copy()), idx, dvec)
    return out if tol is None else out + (tol,)


class Ensemble:
    def __init__(self, ctx, off, x, y, b, params=None):
        """params: one dict of 2d4_amd.sim keyword parameters for all
        realizations, or a list of R such dicts."""
        self.ctx, self.L = ctx, ctx.L
        self.off = _off(off)
        R = self.off.size - 1
        x, y, b = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, b))
        if not (x.shape == y.shape == b.shape == (int(self.off[-1]),)):
            raise ValueError("x, y, b must have off[-1] entries")
        per = isinstance(params, (list, tuple))
        if per and len(params) != R:
            raise ValueError("params needs one entry per realization")
        plist = [make_params(self.L, **(p or {})) for p in (params if per else [params or {}])]
        parr = (type(plist[0]) * len(plist))(*plist)
        h = C.c_void_p()
        rc = self.L.d4_ens_create(ctx.h, R, self.off.ctypes.data, x.ctypes.data, y.ctypes.data, b.ctypes.data,
                                  parr, 1 if per else 0, C.byref(h))
        if rc != 0:
            raise ValueError(f"d4_ens_create failed ({rc}): {ctx.L.d4_last_error(ctx.h).decode(errors='replace')}")
        self.h, self.R = h, R

    def close(self):
        if getattr(self, "h", None):
            self.L.d4_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"ensemble error {rc}: {self.L.d4_ens_last_error(self.h).decode(errors='replace')}")

    def step(self, nsteps=1):
        """nsteps lockstep iterations; returns how many realizations took part in the last."""
        na = C.c_int(0)
        self._check(self.L.d4_ens_step(self.h, nsteps, C.byref(na)))
        return na.value

    def positions(self, original_order=False):
        xy = np.empty((int(self.off[-1]), 2))
        self._check(self.L.d4_ens_positions(self.h, xy.ctypes.data, 1 if original_order else 0))
        return xy

    def stats(self):
        st = (SimStats * self.R)()
        self._check(self.L.d4_ens_stats(self.h, C.cast(st, C.c_void_p)))
        return [{k: getattr(s, k) for k, _ in SimStats._fields_} for s in st]

    def config(self):
        c = EnsConfig()
        self._check(self.L.d4_ens_get_config(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in EnsConfig._fields_}

    def set_config(self, **kw):
        c = EnsConfig(**{**self.config(), **kw})
        self._check(self.L.d4_ens_set_config(self.h, C.byref(c)))

    def autotune(self):
        c, ms = EnsConfig(), C.c_double(0)
        self._check(self.L.d4_ens_autotune(self.h, C.byref(c), C.byref(ms)))
        return {k: getattr(c, k) for k, _ in EnsConfig._fields_}, ms.value

    def timing(self):
        t = np.zeros(4)
        self._check(self.L.d4_ens_timing(self.h, t.ctypes.data))
        return {"pass_ms": t[0], "solve_ms": t[1], "host_ms": t[2], "total_ms": t[3]}
