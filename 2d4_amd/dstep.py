"""The implicit step resident on the device (include/d4_dstep.h): Simulation::run's loop
(simulation.cpp:88-276) for one or many realizations, every stage a kernel launch, one iteration
captured as a HIP graph and replayed.  Parameters as 2d4_amd.sim (the solver parameter is ignored:
the step matrix is dense and factored by the GPU dense LU)."""
import ctypes as C

import numpy as np

from ._bind_extra import SimStats
from .ensemble import _off
from .sim import make_params


class DeviceSimulation:
    def __init__(self, ctx, off, x, y, b, params=None, use_graph=True):
        """Realization r = the slice off[r]:off[r+1] of x, y, b; params: one dict of 2d4_amd.sim keyword
        parameters for all realizations, or a list of R of them."""
        self.ctx, self.L = ctx, ctx.L
        self.off = _off(off)
        R = self.off.size - 1
        x, y, b = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, b))
        if not (x.shape == y.shape == b.shape == (int(self.off[-1]),)):
            raise ValueError("x, y, b must have off[-1] entries")
        per = isinstance(params, (list, tuple))
        if per and len(params) != R:
            raise ValueError("params needs one entry per realization")
        plist = [make_params(self.L, **dict(p or {})) for p in (params if per else [params or {}])]
        parr = (type(plist[0]) * len(plist))(*plist)
        h = C.c_void_p()
        rc = self.L.d4_dstep_create(ctx.h, R, self.off.ctypes.data, x.ctypes.data, y.ctypes.data, b.ctypes.data,
                                    C.cast(parr, C.c_void_p), 1 if per else 0, C.byref(h))
        if rc != 0:
            raise ValueError(f"d4_dstep_create failed ({rc}): {ctx.L.d4_last_error(ctx.h).decode(errors='replace')}")
        self.h, self.R = h, R
        if not use_graph:
            self._check(self.L.d4_dstep_set_graph(self.h, 0))

    def close(self):
        if getattr(self, "h", None):
            self.L.d4_dstep_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"device step error {rc}: {self.L.d4_dstep_last_error(self.h).decode(errors='replace')}")

    def set_graph(self, use_graph):
        self._check(self.L.d4_dstep_set_graph(self.h, 1 if use_graph else 0))

    def run(self, iterations=1):
        """`iterations` loop iterations of every running realization; returns how many still run."""
        na = C.c_int(0)
        self._check(self.L.d4_dstep_run(self.h, int(iterations), C.byref(na)))
        return na.value

    def positions(self, original_order=False):
        xy = np.empty((int(self.off[-1]), 2))
        self._check(self.L.d4_dstep_positions(self.h, xy.ctypes.data, 1 if original_order else 0))
        return xy

    def stats(self):
        st = (SimStats * self.R)()
        self._check(self.L.d4_dstep_stats(self.h, C.cast(st, C.c_void_p)))
        return [{k: getattr(s, k) for k, _ in SimStats._fields_} for s in st]
