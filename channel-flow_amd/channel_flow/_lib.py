"""Loader for the in-tree HIP library ``lib/libcfhip.so`` (C-ABI in
include/channel_flow.h).  There is no CPU fallback: importing a solver class
when the library is missing raises immediately."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libcfhip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)


class CfParams(ctypes.Structure):
    """Mirror of ``cf_params`` (include/channel_flow.h)."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("cfl", ctypes.c_double), ("nu", ctypes.c_double),
        ("dPdx", ctypes.c_double), ("dPdz", ctypes.c_double),
        ("omega_z", ctypes.c_double),
        ("x_mass_cte", ctypes.c_int), ("y_mass_cte", ctypes.c_int),
        ("les_model", ctypes.c_int), ("iwall_model", ctypes.c_int),
        ("dirichlet_nu_t", ctypes.c_int), ("jref_eqwm", ctypes.c_int),
        ("cs", ctypes.c_double),
        ("eqwm_ncv", ctypes.c_int), ("eqwm_max_iter", ctypes.c_int),
        ("eqwm_stretch", ctypes.c_double), ("eqwm_tol", ctypes.c_double),
        ("tau_mean_eqwm", ctypes.c_double),
        ("wall_surrogate", ctypes.c_int), ("device", ctypes.c_int),
        ("u_centerline", ctypes.c_double),
    ]


# name -> (restype, argtypes); every symbol declared in include/channel_flow.h
SIGNATURES = {
    "cf_last_error": (ctypes.c_char_p, []),
    "cf_create": (ctypes.c_int, [ctypes.POINTER(CfParams), _dp, _dp, _dp,
                                 ctypes.POINTER(ctypes.c_void_p)]),
    "cf_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_set_state": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "cf_get_state": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "cf_get_pressure": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_compute_pressure": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_set_nu_t": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_get_nu_t": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_get_avg_nu_t": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_get_wall": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "cf_set_tau_init": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_set_scalars": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_get_scalars": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_init_mass_flow": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_apply_boundary_conditions": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_compute_dt": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_rk3_step": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_run_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "cf_projection_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "cf_compute_rhs": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]),
    "cf_compute_eddy_viscosity": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_compute_wall_model": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip]),
    "cf_mean_mass_flow_u": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_mean_mass_flow_v": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_mean_centerline_u": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_check_divergence": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "cf_statistics": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "cf_eqwm_batch": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     _dp, _ip]),
    "cf_eqwm_batch_noalloc": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             _dp, _ip]),
    "cf_mlp_create": (ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(ctypes.c_void_p)]),
    "cf_mlp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_mlp_nparams": (ctypes.c_int, []),
    "cf_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "cf_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "cf_mlp_get_norm": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp]),
    "cf_mlp_set_norm": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp]),
    "cf_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_float, ctypes.c_float, _fp]),
    "cf_mlp_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _fp, _fp]),
    "cf_mlp_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, _ip]),
    "cf_set_wall_surrogate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "cf_wall_fallbacks": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_stats_accumulate": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_stats_window": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "cf_stats_reset": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_footprint": (ctypes.c_longlong, [ctypes.POINTER(CfParams), _dp, _dp, _dp]),
    "cf_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(CfParams), _dp, _dp, _dp,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "cf_batch_member": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "cf_batch_run_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "cf_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "cf_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CfParams), _dp, _dp, _dp, _dp, _dp,
                                        _dp, ctypes.c_int, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_longlong)]),
    "cf_server_set_batching": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_longlong]),
    "cf_server_start": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_server_cancel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    "cf_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "cf_server_status": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _ip, _ip]),
    "cf_server_result": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, _dp, _dp]),
    "cf_server_case_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp, ctypes.c_int]),
    "cf_server_step_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _fp, ctypes.c_int]),
    "cf_server_summary": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
    "cf_timing_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "cf_timing_report": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    "cf_timing_reset": (ctypes.c_int, [ctypes.c_void_p]),
    "cf_device_bytes": (ctypes.c_longlong, [ctypes.c_void_p]),
    "cf_sync": (ctypes.c_int, [ctypes.c_void_p]),
}

_lib = None


def load(path=None):
    """Load libcfhip.so and bind every C-ABI symbol; raise if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            "libcfhip.so not found at %s: build it with channel-flow_amd/build.sh "
            "(there is no CPU fallback)" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = load().cf_last_error()
        raise RuntimeError("%s failed: %s" % (what, msg.decode() if msg else "unknown error"))


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def fptr(a):
    return a.ctypes.data_as(_fp)


def precision_code(name):
    """cf_params.precision of a case's precision name: 0 "fp64", 1 "mixed"."""
    codes = {"fp64": 0, "mixed": 1}
    if name not in codes:
        raise ValueError("precision must be 'fp64' or 'mixed', not %r" % (name,))
    return codes[name]
