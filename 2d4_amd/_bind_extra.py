"""ctypes signatures of the entry points declared outside include/d4.h
(include/d4_sim.h and later headers); bound when those symbols exist."""
import ctypes as C

_P = C.POINTER
_d = C.c_double
_vp = C.c_void_p
_pi = _P(C.c_int)


class SimParams(C.Structure):
    """d4_sim_params (include/d4_sim.h)."""
    _fields_ = [("init_step_size", _d), ("max_step_size", _d), ("prec", _d), ("dipole_prec", _d),
                ("cutoff_mult", _d), ("heaviside", C.c_int), ("wfunc", C.c_char), ("stress_kind", C.c_int),
                ("stress_init", _d), ("stress_rate", _d), ("stress_period", _d), ("stress_amp", _d),
                ("time_limit", _d), ("step_limit", C.c_int64), ("strain_limit", _d), ("speed_threshold", _d),
                ("solver", C.c_int)]


class SimStats(C.Structure):
    """d4_sim_stats (include/d4_sim.h)."""
    _fields_ = [("sim_time", _d), ("step_size", _d), ("strain", _d), ("last_max_error_ratio_sq", _d),
                ("successful_steps", C.c_int64), ("failed_steps", C.c_int64), ("pair_passes", C.c_int64),
                ("factorizations", C.c_int64), ("singular_solves", C.c_int64),
                ("unpivoted_factorizations", C.c_int64), ("pass_ms", _d), ("solve_ms", _d)]


class EnsConfig(C.Structure):
    """d4_ens_config (include/d4_batch.h)."""
    _fields_ = [("pass_grid", C.c_int), ("tile_order", C.c_int), ("use_graph", C.c_int), ("host_threads", C.c_int)]


SIGNATURES = {
    "d4_sim_default_params": (None, [_P(SimParams)]),
    "d4_sim_create": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _P(SimParams), _P(_vp)]),
    "d4_sim_create_from_file": (C.c_int, [_vp, C.c_char_p, _P(SimParams), _P(_vp)]),
    "d4_sim_destroy": (None, [_vp]),
    "d4_sim_last_error": (C.c_char_p, [_vp]),
    "d4_sim_run": (C.c_int, [_vp]),
    "d4_sim_step": (C.c_int, [_vp, _pi, _pi]),
    "d4_sim_positions": (C.c_int, [_vp, _vp, C.c_int]),
    "d4_sim_get_stats": (C.c_int, [_vp, _P(SimStats)]),
    "d4_sim_write_dconf": (C.c_int, [_vp, C.c_char_p]),
    "d4_lu_solve_batch": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d4_lu_factor_dense_dev": (C.c_int, [_vp, C.c_int, _vp, C.c_longlong, C.c_longlong, _P(C.c_int), _P(C.c_int)]),
    "d4_lu_solve_dense_dev": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    # include/d4_batch.h
    "d4_batch_speed_pass": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d4_batch_jacobian_pass": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _pi,
                                         _vp, _vp, _vp]),
    "d4_ens_create": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _P(SimParams), C.c_int, _P(_vp)]),
    "d4_ens_destroy": (None, [_vp]),
    "d4_ens_last_error": (C.c_char_p, [_vp]),
    "d4_ens_step": (C.c_int, [_vp, C.c_int, _pi]),
    "d4_ens_positions": (C.c_int, [_vp, _vp, C.c_int]),
    "d4_ens_stats": (C.c_int, [_vp, _vp]),
    "d4_ens_set_config": (C.c_int, [_vp, _P(EnsConfig)]),
    "d4_ens_get_config": (C.c_int, [_vp, _P(EnsConfig)]),
    "d4_ens_autotune": (C.c_int, [_vp, _P(EnsConfig), _P(_d)]),
    "d4_ens_timing": (C.c_int, [_vp, _vp]),
    # include/d4_dstep.h
    "d4_dstep_create": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, _P(_vp)]),
    "d4_dstep_destroy": (None, [_vp]),
    "d4_dstep_last_error": (C.c_char_p, [_vp]),
    "d4_dstep_set_graph": (C.c_int, [_vp, C.c_int]),
    "d4_dstep_run": (C.c_int, [_vp, C.c_int, _pi]),
    "d4_dstep_positions": (C.c_int, [_vp, _vp, C.c_int]),
    "d4_dstep_stats": (C.c_int, [_vp, _vp]),
}


def bind(L):
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name, None)
        if f is not None:
            f.restype = res
            f.argtypes = args
