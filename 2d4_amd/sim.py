"""Host simulation over lib2d4 (include/d4_sim.h).

The reference's SimulationData + Simulation::run (simulation_data.cpp,
simulation.cpp:30-281): the adaptive step-doubled implicit step, with every
O(N^2) pair pass on the GPU and the sparse linear solves on the host.
Keyword parameters use the reference's command-line meaning
(project_parser.cpp:28-68):

  init_step_size   --initial-stepsize      max_step_size  -M (<= 0: none)
  prec             -P                      dipole_prec    -p (0: off)
  cutoff_mult      -u                      heaviside      -h
  wfunc            -w ('c', 'p', 'm')      time_limit     -t (<= 0: none)
  step_limit       -C (< 0: none)          strain_limit   (<= 0: none)
  stress           None (zero stress), ("fixed", init, rate), ("cyclic", init,
                   rate, period) or ("increasing", init, rate, period, amp)
                   (stress_protocol.cpp:13-73)
  speed_threshold  cutoff-change threshold (simulation.cpp:215-219; <= 0: none)
"""
import ctypes as C

import numpy as np

from ._bind_extra import SimParams, SimStats

_STRESS = {"none": 0, "fixed": 1, "cyclic": 2, "increasing": 3}


def make_params(L, **kw):
    p = SimParams()
    L.d4_sim_default_params(C.byref(p))
    stress = kw.pop("stress", None)
    if stress is not None:
        kind, *vals = stress
        p.stress_kind = _STRESS[kind]
        for name, v in zip(("stress_init", "stress_rate", "stress_period", "stress_amp"), vals):
            setattr(p, name, float(v))
    for k, v in kw.items():
        if k == "wfunc":
            v = v.encode() if isinstance(v, str) else v
        elif k == "heaviside":
            v = 1 if v else 0
        if not hasattr(p, k):
            raise TypeError(f"unknown simulation parameter {k!r}")
        setattr(p, k, v)
    return p


class Simulation:
    """A d4_sim: positions sorted as the reference sorts them (b = +1 first,
    y descending within each sign, simulation_data.cpp:116)."""

    def __init__(self, L, ctx_handle, handle, n, owner=None):
        self.L, self._ctx_h, self.h, self.n = L, ctx_handle, handle, n
        self._owner = owner  # keeps the context alive

    @classmethod
    def from_arrays(cls, ctx, x, y, b, **params):
        return cls._create(ctx.L, ctx.h, x, y, b, owner=ctx, **params)

    @classmethod
    def from_file(cls, ctx, path, **params):
        return cls._create_file(ctx.L, ctx.h, path, owner=ctx, **params)

    @classmethod
    def _create(cls, L, ctx_h, x, y, b, owner=None, **params):
        x, y, b = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, b))
        if not (x.shape == y.shape == b.shape and x.ndim == 1):
            raise ValueError("x, y and b must be 1-D arrays of one length")
        p = make_params(L, **params)
        h = C.c_void_p()
        rc = L.d4_sim_create(ctx_h, len(x), x.ctypes.data, y.ctypes.data, b.ctypes.data, C.byref(p), C.byref(h))
        if rc != 0:
            raise ValueError(f"d4_sim_create failed ({rc}): the input is not a valid configuration "
                             "(integer Burgers vectors summing to 0)")
        return cls(L, ctx_h, h, len(x), owner)

    @classmethod
    def _create_file(cls, L, ctx_h, path, owner=None, **params):
        p = make_params(L, **params)
        h = C.c_void_p()
        rc = L.d4_sim_create_from_file(ctx_h, str(path).encode(), C.byref(p), C.byref(h))
        if rc != 0:
            raise ValueError(f"d4_sim_create_from_file failed ({rc}) for {path}")
        n = sum(1 for line in open(path) if line.strip())
        return cls(L, ctx_h, h, n, owner)

    def close(self):
        if getattr(self, "h", None):
            self.L.d4_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"simulation error {rc}: {self.L.d4_sim_last_error(self.h).decode(errors='replace')}")

    def run(self):
        """Simulation::run until interrupt() holds; returns stats()."""
        self._check(self.L.d4_sim_run(self.h))
        return self.stats()

    def step(self):
        """One loop iteration; returns (accepted, stopped)."""
        acc, stop = C.c_int(0), C.c_int(0)
        self._check(self.L.d4_sim_step(self.h, C.byref(acc), C.byref(stop)))
        return bool(acc.value), bool(stop.value)

    def positions(self, original_order=False):
        xy = np.empty((self.n, 2))
        self._check(self.L.d4_sim_positions(self.h, xy.ctypes.data, 1 if original_order else 0))
        return xy

    def stats(self):
        st = SimStats()
        self._check(self.L.d4_sim_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in SimStats._fields_}

    def write_dconf(self, path):
        self._check(self.L.d4_sim_write_dconf(self.h, str(path).encode()))


def lu_solve_batch(ctx, systems):
    """d4_lu_solve_batch (include/d4_sim.h): the GPU dense LU of solver 1 on K
    independent systems.  systems: [(Ap, Ai, Ax, b), ...] with Ap the n + 1
    column pointers from 0, Ai the rows and Ax the values of a CSC matrix.
    Returns (xs, singular, unpivoted): the K solutions and two boolean arrays."""
    K = len(systems)
    keep = []

    def arr(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data

    n = np.array([len(s[0]) - 1 for s in systems], dtype=np.int32)
    xs = [np.zeros(k) for k in n]
    ptrs = {name: (C.c_void_p * max(K, 1))() for name in ("Ap", "Ai", "Ax", "b", "x")}
    for k, (Ap, Ai, Ax, b) in enumerate(systems):
        ptrs["Ap"][k] = arr(Ap, np.int32)
        ptrs["Ai"][k] = arr(Ai, np.int32)
        ptrs["Ax"][k] = arr(Ax, np.float64)
        ptrs["b"][k] = arr(b, np.float64)
        ptrs["x"][k] = xs[k].ctypes.data
    singular = np.zeros(max(K, 1), dtype=np.uint8)
    unpivoted = np.zeros(max(K, 1), dtype=np.uint8)
    rc = ctx.L.d4_lu_solve_batch(ctx.h, K, n.ctypes.data, C.cast(ptrs["Ap"], C.c_void_p),
                                 C.cast(ptrs["Ai"], C.c_void_p), C.cast(ptrs["Ax"], C.c_void_p),
                                 C.cast(ptrs["b"], C.c_void_p), C.cast(ptrs["x"], C.c_void_p),
                                 singular.ctypes.data, unpivoted.ctypes.data)
    if rc:
        raise RuntimeError(f"d4_lu_solve_batch: {ctx.L.d4_last_error(ctx.h).decode()} (code {rc})")
    return xs, singular[:K].astype(bool), unpivoted[:K].astype(bool)
