"""The oracle: ctypes access to oracle/liboracle.so (the CPU restatement of the
reference path) and, when built, oracle/_ref (the reference's own sources).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product (2d4_amd).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = C.POINTER
pd = P(C.c_double)
vp = C.c_void_p


class RunParams(C.Structure):
    _fields_ = [("dt0", C.c_double), ("max_dt", C.c_double), ("prec", C.c_double), ("dipole_prec", C.c_double),
                ("cutoff_mult", C.c_double), ("heaviside", C.c_int), ("wfunc", C.c_char),
                ("stress_kind", C.c_int), ("stress_init", C.c_double), ("stress_rate", C.c_double),
                ("stress_period", C.c_double), ("stress_amp", C.c_double), ("time_limit", C.c_double),
                ("step_limit", C.c_long), ("strain_limit", C.c_double), ("speed_threshold", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("sim_time", C.c_double), ("step_size", C.c_double), ("strain", C.c_double),
                ("ok_steps", C.c_long), ("failed_steps", C.c_long), ("last_max_ratio", C.c_double)]


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def build(quiet=True):
    import subprocess
    r = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


class Oracle:
    def __init__(self, fast=False, path=None):
        """fast: the build with the reference's flags; path: another build of d4_oracle.c (tests only)."""
        if path is None:
            path = os.path.join(HERE, "liboracle_fast.so" if fast else "liboracle.so")
            if not os.path.exists(path):
                build()
        L = C.CDLL(path)
        d, i, ch = C.c_double, C.c_int, C.c_char
        sig = {
            "orc_normalize": (d, [d]), "orc_weight": (d, [d, ch]), "orc_weight_inv": (d, [d, ch]),
            "orc_xy4": (d, [d, d, d, d]), "orc_dxy4": (d, [d, d, d, d, i]), "orc_xy": (d, [d, d]),
            "orc_dxy": (d, [d, d]), "orc_dyxy4": (d, [d, d, d, d]),
            "orc_table_bilinear": (d, [vp, i, d, d]), "orc_table_build": (None, [i, vp, vp]),
            "orc_speed_pass_table": (None, [i, vp, vp, i, d, vp]),
            "orc_field_inputs": (None, [d, d, pd, pd]),
            "orc_field_batch": (None, [C.c_int64, vp, vp, vp, vp, vp, vp, i]),
            "orc_inputs_batch": (None, [C.c_int64, vp, vp, vp, vp]),
            "orc_speed_pass": (None, [i, vp, d, vp, vp, d]),
            "orc_jacobian_pass": (i, [i, vp, d, d, d, d, d, d, i, ch, vp, vp, vp, vp, vp, i, vp, vp, vp, d]),
            "orc_jacobian_from_prev": (None, [i, vp, vp, vp, vp, ch, d, d, d, d, d, vp]),
            "orc_jacobian_dense": (None, [i, vp, d, d, d, ch, vp, vp, vp, vp, vp, d]),
            "orc_cutoff": (None, [i, d, i, pd, pd, pd]),
            "orc_ext_stress": (d, [i, d, d, d, d, d]),
            "orc_run": (i, [i, vp, P(RunParams), P(RunStats)]),
            "orc_read_dconf": (i, [C.c_char_p, vp, vp, i]),
            "orc_speed_pass_omp": (None, [i, vp, d, vp]),
            "orc_jacobian_dense_omp": (None, [i, vp, d, d, d, ch, vp, vp, vp, vp]),
            "orc_num_threads": (i, []),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        self.L = L

    # field
    def field(self, dx, dy, e=None, c=None, exact=False):
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        n = dx.size
        s, ds = np.empty(n), np.empty(n)
        e = None if e is None else np.ascontiguousarray(e, dtype=np.float64)
        c = None if c is None else np.ascontiguousarray(c, dtype=np.float64)
        self.L.orc_field_batch(n, _ptr(dx), _ptr(dy), _ptr(e), _ptr(c), _ptr(s), _ptr(ds), 1 if exact else 0)
        return s, ds

    def inputs(self, dx, dy):
        """exp2(PI2LN2INV dx), cos(2 pi dy) with this library's libm (simulation.cpp:456-457)"""
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        e, c = np.empty(dx.size), np.empty(dx.size)
        self.L.orc_inputs_batch(dx.size, _ptr(dx), _ptr(dy), _ptr(e), _ptr(c))
        return e, c

    # passes
    def speed_pass(self, xy, tau=0.0, tol=None, cap2=0.0):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        n = xy.shape[0]
        v = np.empty(n)
        self.L.orc_speed_pass(n, _ptr(xy), tau, _ptr(v), _ptr(tol), cap2)
        return v

    def jacobian_pass(self, xy, h, tau_a, tau_b, cutoff, cut_sq, one_per, heaviside, wfunc, tol=None, cap2=0.0):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        n = xy.shape[0]
        cap = n * n
        vA, vB, dvec = np.empty(n), np.empty(n), np.empty(n)
        Ap = np.zeros(n + 1, dtype=np.int32)
        Ai = np.zeros(cap, dtype=np.int32)
        Ax = np.zeros(cap)
        idx = np.zeros(n, dtype=np.int32)
        nz = self.L.orc_jacobian_pass(n, _ptr(xy), h, tau_a, tau_b, cutoff, cut_sq, one_per, 1 if heaviside else 0,
                                      wfunc.encode(), _ptr(vA), _ptr(vB), _ptr(Ap), _ptr(Ai), _ptr(Ax), cap,
                                      _ptr(idx), _ptr(dvec), _ptr(tol), cap2)
        assert nz >= 0
        return dict(vA=vA, vB=vB, Ap=Ap, Ai=Ai[:nz].copy(), Ax=Ax[:nz].copy(), indexes=idx, dvec=dvec, nnz=nz)

    def jacobian_from_prev(self, Ap, Ai, Ax, dvec, wfunc, base_t, new_t, old_t, tau_old, tau_new, speed):
        n = dvec.size
        self.L.orc_jacobian_from_prev(n, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(dvec), wfunc.encode(), base_t, new_t,
                                      old_t, tau_old, tau_new, _ptr(speed))

    def jacobian_dense(self, xy, h, tau_a, tau_b, wfunc="c", tol=None, cap2=0.0):
        """returns vA, vB, M (row, col) = V + diag(s), dvec"""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        n = xy.shape[0]
        vA, vB, dvec = np.empty(n), np.empty(n), np.empty(n)
        M = np.empty((n, n))
        self.L.orc_jacobian_dense(n, _ptr(xy), h, tau_a, tau_b, wfunc.encode(), _ptr(vA), _ptr(vB), _ptr(M),
                                  _ptr(dvec), _ptr(tol), cap2)
        return vA, vB, M.T.copy(), dvec

    def cutoff(self, n, mult, heaviside):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.L.orc_cutoff(n, mult, 1 if heaviside else 0, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def run(self, xy, **kw):
        p = RunParams(dt0=1e-6, max_dt=0.0, prec=1e-5, dipole_prec=0.05, cutoff_mult=1e20, heaviside=0,
                      wfunc=b"c", stress_kind=0, stress_init=0.0, stress_rate=0.0, stress_period=0.0,
                      stress_amp=0.0, time_limit=0.0, step_limit=-1, strain_limit=0.0, speed_threshold=0.0)
        for k, v in kw.items():
            setattr(p, k, v.encode() if isinstance(v, str) else v)
        xy = np.array(xy, dtype=np.float64).reshape(-1, 2)
        st = RunStats()
        self.L.orc_run(xy.shape[0], _ptr(xy), C.byref(p), C.byref(st))
        return xy, st

    def read_dconf(self, path, cap=1 << 20):
        xy = np.empty((cap, 2))
        order = np.empty(cap, dtype=np.uint32)
        n = self.L.orc_read_dconf(path.encode(), _ptr(xy), _ptr(order), cap)
        if n < 0:
            raise ValueError(f"orc_read_dconf failed ({n})")
        return xy[:n].copy(), order[:n].copy()


class Ref:
    """The reference's own compiled sources (oracle/_ref); only where /root/reference was present."""

    def __init__(self, flavour="o2"):
        path = os.path.join(HERE, "_ref", f"libref_{flavour}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        d, i, ch = C.c_double, C.c_int, C.c_char
        sig = {
            "ref_xy": (d, [d, d]), "ref_dxy": (d, [d, d]), "ref_xy4": (d, [d, d, d, d]),
            "ref_dxy4": (d, [d, d, d, d]), "ref_inputs": (None, [d, d, pd, pd]),
            "ref_field_batch": (None, [C.c_long, vp, vp, vp, vp, vp, vp, vp, vp]),
            "ref_normalize": (d, [d]), "ref_weight": (d, [d, ch]), "ref_weight_inv": (d, [d, ch]),
            "ref_tolerance_replay": (i, [d, d, C.c_uint, C.c_long, vp, vp, vp]),
            "ref_new_step_size": (d, [d, d, C.c_uint, d, d]),
            "ref_read_dconf": (i, [C.c_char_p, d, i, vp, vp, vp, i, vp]),
            "ref_write_dconf": (i, [C.c_char_p, vp, C.c_char_p]),
            "ref_ext_stress": (d, [i, d, d, d, d, d]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        self.L = L

    def field_batch(self, dx, dy):
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        n = dx.size
        e, c = np.empty(n), np.empty(n)
        for k in range(n):
            ee, cc = C.c_double(), C.c_double()
            self.L.ref_inputs(dx[k], dy[k], C.byref(ee), C.byref(cc))
            e[k], c[k] = ee.value, cc.value
        xy2, dxy2, xy4, dxy4 = (np.empty(n) for _ in range(4))
        self.L.ref_field_batch(n, _ptr(dx), _ptr(dy), _ptr(e), _ptr(c), _ptr(xy2), _ptr(dxy2), _ptr(xy4), _ptr(dxy4))
        return dict(e=e, c=c, xy2=xy2, dxy2=dxy2, xy4=xy4, dxy4=dxy4)

    def read_dconf(self, path, mult=1e20, heaviside=False, cap=1 << 20):
        xy = np.empty((cap, 2))
        order = np.empty(cap, dtype=np.uint32)
        b = np.empty(cap, dtype=np.int32)
        cut = np.empty(3)
        n = self.L.ref_read_dconf(path.encode(), mult, 1 if heaviside else 0, _ptr(xy), _ptr(order), _ptr(b), cap,
                                  _ptr(cut))
        return xy[:n].copy(), order[:n].copy(), b[:n].copy(), cut


def ref_available(flavour="o2"):
    return os.path.exists(os.path.join(HERE, "_ref", f"libref_{flavour}.so"))
