"""2d4_amd -- MI355X (gfx950) implementation of 2D4's periodic pair-interaction path.

Python mirror of the C ABI in include/d4.h (ctypes over 2d4_amd/lib/lib2d4.so).
The host-side semantics follow the reference's Field / Simulation interface:

  field_eval          sdddstCore::Field::xy / xy_diff_x       (Field.cpp:295-346)
  speed_pass          Simulation::calculateSpeedsAtTime       (simulation.cpp:287-333)
  jacobian_pass       Simulation::calcJacobianAndSpeedsAtTimes (simulation.cpp:345-527)
  jacobian_rescale    Simulation::calcJacobianAndSpeedsFromPrev (simulation.cpp:546-608)
  jacobian_dense      the Jacobian pass with an unlimited cutoff, dense output
  vjp                 adjoint of the speed pass (no reference counterpart)

Every call runs on the GPU through lib2d4; there is no CPU fallback: if the
library is missing or no GPU is present, calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "lib2d4.so")

D4_OK = 0
D4_ERR_INVALID = -1
D4_ERR_HIP = -2
D4_ERR_CAPACITY = -3
D4_ERR_NO_DEVICE = -4

_P = C.POINTER
_d = C.c_double
_pd = _P(C.c_double)
_pi = _P(C.c_int)
_vp = C.c_void_p


class D4Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lib2d4 error {code}: {msg}")
        self.code = code


class JacParams(C.Structure):
    """d4_jac_params (include/d4.h); fields as SimulationData names them (simulation_data.h:181-197)."""
    _fields_ = [("h", _d), ("tau_a", _d), ("tau_b", _d), ("cutoff", _d), ("cut_sq", _d),
                ("one_per_cut_sq", _d), ("heaviside", C.c_int), ("wfunc", C.c_char), ("tol_cap2", _d),
                ("exact", C.c_int)]


# symbol -> (restype, argtypes); the set include/d4.h declares
SIGNATURES = {
    "d4_ctx_create": (C.c_int, [C.c_int, _P(_vp)]),
    "d4_ctx_destroy": (None, [_vp]),
    "d4_last_error": (C.c_char_p, [_vp]),
    "d4_ctx_device": (C.c_int, [_vp]),
    "d4_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "d4_ctx_use_default_stream": (C.c_int, [_vp]),
    "d4_ctx_synchronize": (C.c_int, [_vp]),
    "d4_ctx_set_profiling": (C.c_int, [_vp, C.c_int]),
    "d4_ctx_pair_kernel_time": (C.c_int, [_vp, _pd, _pi]),
    "d4_dev_alloc": (C.c_int, [_vp, C.c_int64, _P(_vp)]),
    "d4_dev_free": (C.c_int, [_vp, _vp]),
    "d4_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_int64]),
    "d4_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_int64]),
    "d4_field_eval": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
    "d4_field_eval_dev": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
    "d4_speed_pass": (C.c_int, [_vp, C.c_int, _vp, _d, _vp, _vp, _d]),
    "d4_speed_pass_dev": (C.c_int, [_vp, C.c_int, _vp, _d, _vp, _vp, _d]),
    "d4_speed_pass_dev_noalloc": (C.c_int, [_vp, C.c_int, _vp, _d, _vp, _vp, _d]),
    "d4_jacobian_pass": (C.c_int, [_vp, C.c_int, _vp, _P(JacParams), _vp, _vp, _vp, _vp, _vp, C.c_int,
                                   _pi, _vp, _vp, _vp]),
    "d4_jacobian_rescale": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, C.c_char, _d, _d, _vp]),
    "d4_jacobian_dense_dev": (C.c_int, [_vp, C.c_int, _vp, _P(JacParams), _vp, _vp, _vp, C.c_int64, _vp, _vp]),
    "d4_jacobian_dense_scale_dev": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, _vp]),
    "d4_jacobian_dense_cut_dev": (C.c_int, [_vp, C.c_int, _vp, _P(JacParams), _vp, _vp, _vp, C.c_int64, _vp, _vp]),
    "d4_theta_pair_vjp_dev": (C.c_int, [_vp, C.c_int, _vp, _P(JacParams), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d4_jacobian_dense": (C.c_int, [_vp, C.c_int, _vp, _P(JacParams), _vp, _vp, _vp, _vp, _vp, C.c_int]),
    "d4_vjp_dev": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _pd, C.c_int]),
    "d4_table_build_dev": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "d4_table_eval_dev": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp]),
    "d4_speed_pass_table_dev": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _d, _vp]),
    "d4_jacobian_dense_table_dev": (C.c_int, [_vp, C.c_int, _vp, _P(JacParams), _vp, _vp, C.c_int, _vp, _vp, _vp,
                                              C.c_int64, _vp]),
}

_lib = None


def lib():
    """Loads lib2d4.so (built by __graft_entry__.build()); raises if it is missing.

    torch is imported first when available: it bundles its own HIP runtime
    (soname libamdhip64.so.7), and loading it before lib2d4 makes the dynamic
    linker bind lib2d4 to that same runtime, so one process holds one HIP
    runtime and device pointers pass freely between torch and lib2d4."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        from . import _bind_extra
        _bind_extra.bind(L)
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """Owns a d4_ctx (device, stream, device buffers)."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.d4_ctx_create(device, C.byref(h))
        if rc != D4_OK:
            raise D4Error(rc, "d4_ctx_create failed (no usable GPU?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.d4_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != D4_OK:
            raise D4Error(rc, self.L.d4_last_error(self.h).decode(errors="replace"))
        return rc

    def set_stream(self, stream_handle):
        self.check(self.L.d4_ctx_set_stream(self.h, C.c_void_p(stream_handle) if stream_handle else None))

    def use_torch_stream(self):
        """Orders the context's work on torch's current stream (its default stream
        is the device's legacy NULL stream, which a non-blocking stream would not
        wait for)."""
        import torch
        h = torch.cuda.current_stream().cuda_stream
        if h:
            self.set_stream(h)
        else:
            self.check(self.L.d4_ctx_use_default_stream(self.h))

    def synchronize(self):
        self.check(self.L.d4_ctx_synchronize(self.h))

    def set_profiling(self, on=True):
        self.check(self.L.d4_ctx_set_profiling(self.h, 1 if on else 0))

    def pair_kernel_time(self):
        """(summed ms, launches) of the pair kernels since the last call (profiling on)."""
        ms, k = C.c_double(0.0), C.c_int(0)
        self.check(self.L.d4_ctx_pair_kernel_time(self.h, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    # ---------------------------------------------------------- numpy API
    def field_eval(self, dx, dy, e2pix=None, c2piy=None, exact=False, want_sigma=True, want_dsigma=True):
        """Field::xy / xy_diff_x per pair (Field.cpp:295-346); returns (sxy, dsxy)."""
        dx, dy = _f64(dx), _f64(dy)
        n = dx.size
        e = _f64(e2pix) if e2pix is not None else None
        c = _f64(c2piy) if c2piy is not None else None
        s = np.empty(n) if want_sigma else None
        ds = np.empty(n) if want_dsigma else None
        self.check(self.L.d4_field_eval(self.h, n, _ptr(dx), _ptr(dy), _ptr(e), _ptr(c), _ptr(s), _ptr(ds),
                                        1 if exact else 0))
        return s, ds

    def speed_pass(self, xy, tau=0.0, tol=None, tol_cap2=0.0):
        """calculateSpeedsAtTime (simulation.cpp:287-333); xy is (n, 2) sorted; tol updated in place."""
        xy = _f64(xy).reshape(-1, 2)
        n = xy.shape[0]
        v = np.empty(n)
        if tol is not None:
            assert tol.dtype == np.float64 and tol.flags.c_contiguous and tol.size == n
        self.check(self.L.d4_speed_pass(self.h, n, _ptr(xy), float(tau), _ptr(v), _ptr(tol), float(tol_cap2)))
        return v

    def jacobian_pass(self, xy, params, tol=None, cap=None):
        """calcJacobianAndSpeedsAtTimes (simulation.cpp:345-527) -> dict with vA, vB, Ap, Ai, Ax, indexes, dvec."""
        xy = _f64(xy).reshape(-1, 2)
        n = xy.shape[0]
        vA, vB, dvec = np.empty(n), np.empty(n), np.empty(n)
        Ap = np.empty(n + 1, dtype=np.int32)
        idx = np.empty(n, dtype=np.int32)
        nnz = C.c_int(0)
        cap = int(cap) if cap is not None else 0
        for _ in range(2):
            Ai = np.empty(max(cap, 1), dtype=np.int32)
            Ax = np.empty(max(cap, 1))
            rc = self.L.d4_jacobian_pass(self.h, n, _ptr(xy), C.byref(params), _ptr(vA), _ptr(vB), _ptr(Ap),
                                         _ptr(Ai), _ptr(Ax), cap, C.byref(nnz), _ptr(idx), _ptr(dvec), _ptr(tol))
            if rc == D4_ERR_CAPACITY:
                cap = nnz.value
                continue
            self.check(rc)
            break
        k = nnz.value
        return dict(vA=vA, vB=vB, Ap=Ap, Ai=Ai[:k].copy(), Ax=Ax[:k].copy(), indexes=idx, dvec=dvec, nnz=k)

    def jacobian_rescale(self, Ap, Ai, Ax, dvec, wfunc, factor, dtau, speed=None):
        """calcJacobianAndSpeedsFromPrev (simulation.cpp:546-608), arrays updated in place."""
        n = dvec.size
        for a in (Ax, dvec) + ((speed,) if speed is not None else ()):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        Ap = np.ascontiguousarray(Ap, dtype=np.int32)
        Ai = np.ascontiguousarray(Ai, dtype=np.int32)
        self.check(self.L.d4_jacobian_rescale(self.h, n, _ptr(Ap), _ptr(Ai), _ptr(Ax), _ptr(dvec),
                                              wfunc.encode() if isinstance(wfunc, str) else wfunc,
                                              float(factor), float(dtau), _ptr(speed)))

    def jacobian_dense(self, xy, params, tol=None, scale=False, out=None):
        """Dense Jacobian pass: (vA, vB, M, dvec) with M in (row, col) orientation.  With out, an (n, n)
        C-contiguous float64 array, the library writes the matrix there column-major (out[c, r] = M[r, c])
        and out itself is returned, without the transposed copy."""
        xy = _f64(xy).reshape(-1, 2)
        n = xy.shape[0]
        vA, vB, dvec = np.empty(n), np.empty(n), np.empty(n)
        if out is not None:
            if out.shape != (n, n) or out.dtype != np.float64 or not out.flags.c_contiguous:
                raise ValueError("out must be an (n, n) C-contiguous float64 array")
            M = out
        else:
            M = np.empty((n, n))
        self.check(self.L.d4_jacobian_dense(self.h, n, _ptr(xy), C.byref(params), _ptr(vA), _ptr(vB), _ptr(M),
                                            _ptr(dvec), _ptr(tol), 1 if scale else 0))
        return vA, vB, (out if out is not None else M.T.copy()), dvec

    # ---------------------------------------------------------- torch API (device tensors)
    def speed_pass_t(self, xy, tau=0.0, tol=None, tol_cap2=0.0, out=None):
        import torch
        n = xy.shape[0]
        v = out if out is not None else torch.empty(n, dtype=torch.float64, device=xy.device)
        self.use_torch_stream()
        self.check(self.L.d4_speed_pass_dev(self.h, n, xy.data_ptr(), float(tau), v.data_ptr(),
                                            tol.data_ptr() if tol is not None else None, float(tol_cap2)))
        return v

    def jacobian_dense_t(self, xy, params, M=None, tol=None, scale=False, cutoff=False):
        """(vA, vB, M, d): M[col, row] = V + diag(s) (the row scaling and +1 too with scale=True); with
        cutoff=True V keeps only the pairs inside the storage radius, tapered, as the reference stores them
        (d4_jacobian_dense_cut_dev), else every pair (d4_jacobian_dense_dev)."""
        import torch
        n = xy.shape[0]
        dev = xy.device
        vA = torch.empty(n, dtype=torch.float64, device=dev)
        vB = torch.empty(n, dtype=torch.float64, device=dev)
        d = torch.empty(n, dtype=torch.float64, device=dev)
        if M is None:
            M = torch.empty((n, n), dtype=torch.float64, device=dev)  # M[col, row]
        self.use_torch_stream()
        fn = self.L.d4_jacobian_dense_cut_dev if cutoff else self.L.d4_jacobian_dense_dev
        self.check(fn(self.h, n, xy.data_ptr(), C.byref(params), vA.data_ptr(), vB.data_ptr(), M.data_ptr(), n,
                      d.data_ptr(), tol.data_ptr() if tol is not None else None))
        if scale:
            self.dense_scale_t(M, d)
        return vA, vB, M, d

    def dense_scale_t(self, M, d):
        """M[col, row] -> the reference's final matrix I + diag((1+d)/2) M, in place."""
        n = M.shape[0]
        self.use_torch_stream()
        self.check(self.L.d4_jacobian_dense_scale_dev(self.h, n, M.data_ptr(), M.stride(0), d.data_ptr()))
        return M

    def theta_pair_vjp_t(self, xy, params, a, c, u, w, t, want_dy=True):
        """(gx, gy): the pair term of the implicit step's vector-Jacobian product (d4_theta_pair_vjp_dev)."""
        import torch
        n = xy.shape[0]
        gx = torch.empty(n, dtype=torch.float64, device=xy.device)
        gy = torch.empty(n, dtype=torch.float64, device=xy.device) if want_dy else None
        vs = [v.contiguous() for v in (a, c, u, w, t)]
        self.use_torch_stream()
        self.check(self.L.d4_theta_pair_vjp_dev(self.h, n, xy.data_ptr(), C.byref(params),
                                                *[v.data_ptr() for v in vs], gx.data_ptr(),
                                                gy.data_ptr() if gy is not None else None))
        return gx, gy

    def vjp_t(self, xy, w, exact=True, want_dy=True):
        """(gx, gy, gtau) of L = sum_i w_i v_i(x, y, tau)."""
        import torch
        n = xy.shape[0]
        gx = torch.empty(n, dtype=torch.float64, device=xy.device)
        gy = torch.empty(n, dtype=torch.float64, device=xy.device) if want_dy else None
        gt = C.c_double(0.0)
        self.use_torch_stream()
        self.check(self.L.d4_vjp_dev(self.h, n, xy.data_ptr(), w.data_ptr(), gx.data_ptr(),
                                     gy.data_ptr() if gy is not None else None, C.byref(gt), 1 if exact else 0))
        return gx, gy, gt.value


def _tp(t):
    return t.data_ptr() if t is not None else None


def _table_methods():
    """Tabulated-field calls on device tensors (added to Context below)."""

    def table_build_t(self, n=1024, device="cuda:0", with_derivative=True):
        """(tab_s, tab_d): n x n node tables of Field::xy and xy_diff_x over [-1/2, 1/2)^2."""
        import torch
        tab_s = torch.empty((n, n), dtype=torch.float64, device=device)
        tab_d = torch.empty((n, n), dtype=torch.float64, device=device) if with_derivative else None
        self.use_torch_stream()
        self.check(self.L.d4_table_build_dev(self.h, n, tab_s.data_ptr(), _tp(tab_d)))
        return tab_s, tab_d

    def table_eval_t(self, dx, dy, tab_s, tab_d=None):
        import torch
        m = dx.numel()
        s = torch.empty(m, dtype=torch.float64, device=dx.device)
        d = torch.empty(m, dtype=torch.float64, device=dx.device) if tab_d is not None else None
        self.use_torch_stream()
        self.check(self.L.d4_table_eval_dev(self.h, m, dx.data_ptr(), dy.data_ptr(), tab_s.data_ptr(), _tp(tab_d),
                                            tab_s.shape[0], s.data_ptr(), _tp(d)))
        return s, d

    def speed_pass_table_t(self, xy, tab_s, tau=0.0, out=None):
        import torch
        n = xy.shape[0]
        v = out if out is not None else torch.empty(n, dtype=torch.float64, device=xy.device)
        self.use_torch_stream()
        self.check(self.L.d4_speed_pass_table_dev(self.h, n, xy.data_ptr(), tab_s.data_ptr(), tab_s.shape[0],
                                                  float(tau), v.data_ptr()))
        return v

    def jacobian_dense_table_t(self, xy, params, tab_s, tab_d, M=None):
        import torch
        n = xy.shape[0]
        dev = xy.device
        vA, vB, d = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3))
        if M is None:
            M = torch.empty((n, n), dtype=torch.float64, device=dev)  # M[col, row]
        self.use_torch_stream()
        self.check(self.L.d4_jacobian_dense_table_dev(self.h, n, xy.data_ptr(), C.byref(params), tab_s.data_ptr(),
                                                      tab_d.data_ptr(), tab_s.shape[0], vA.data_ptr(),
                                                      vB.data_ptr(), M.data_ptr(), n, d.data_ptr()))
        return vA, vB, M, d

    return [table_build_t, table_eval_t, speed_pass_table_t, jacobian_dense_table_t]


for _f in _table_methods():
    setattr(Context, _f.__name__, _f)


def _lu_methods():
    """The simulation's GPU dense LU on a dense device matrix (added to Context below)."""

    def lu_factor_dense_t(self, A, transpose=False):
        """Factors the (n, n) float64 CUDA tensor A (row-major), or its transpose; returns
        (singular, unpivoted).  Replaces the context's previous factorization."""
        n = A.shape[0]
        if A.dim() != 2 or A.shape[1] != n or A.stride(1) != 1:
            raise ValueError("A must be a square row-major (n, n) tensor")
        ld = A.stride(0)
        rs, cs = (1, ld) if transpose else (ld, 1)
        s, u = C.c_int(0), C.c_int(0)
        self.check(self.L.d4_lu_factor_dense_dev(self.h, n, A.data_ptr(), rs, cs, C.byref(s), C.byref(u)))
        return bool(s.value), bool(u.value)

    def lu_solve_dense_t(self, b, out=None):
        """x = A^-1 b for the matrix of the last lu_factor_dense_t (b: float64 CUDA vector)."""
        import torch
        b = b.contiguous()
        x = out if out is not None else torch.empty_like(b)
        self.check(self.L.d4_lu_solve_dense_dev(self.h, b.numel(), b.data_ptr(), x.data_ptr()))
        return x

    return [lu_factor_dense_t, lu_solve_dense_t]


for _f in _lu_methods():
    setattr(Context, _f.__name__, _f)


def jac_params(h=0.0, tau_a=0.0, tau_b=0.0, n=None, cutoff_mult=None, heaviside=False, wfunc="c", tol_cap2=0.0,
               cutoff=None, exact=False):
    """Builds d4_jac_params; cutoff from the multiplier as SimulationData::updateCutOff (simulation_data.cpp:253-262)."""
    if cutoff is None:
        if cutoff_mult is None:
            cutoff = 1e20 / np.sqrt(n if n else 1)
        else:
            m = cutoff_mult
            if not heaviside and cutoff_mult >= np.sqrt(2 * n) / 12:
                m = 1e20
            cutoff = m / np.sqrt(n)
    cut_sq = cutoff * cutoff
    return JacParams(h=h, tau_a=tau_a, tau_b=tau_b, cutoff=cutoff, cut_sq=cut_sq, one_per_cut_sq=1.0 / cut_sq,
                     heaviside=1 if heaviside else 0, wfunc=wfunc.encode(), tol_cap2=tol_cap2,
                     exact=1 if exact else 0)


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
