"""ctypes binding of liblekf_hip.so (the C ABI declared in include/lekf.h).

The product path has no CPU fallback: if the HIP library is missing this
module raises on import of the binding, and every compute entry point
returns LEKF_ERR_NODEV when no GPU is present.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_float, c_int, c_long, c_uint32, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblekf_hip.so")

LEKF_OK = 0
LEKF_ERR_EIGEN = 2
LEKF_WARN_RANKDEF = 4
LEKF_ERR_SIZE = 99
LEKF_ERR_HIP = 100
LEKF_ERR_ARG = 101
LEKF_ERR_NODEV = 102

LEKF_PREC_F64 = 0
LEKF_PREC_F32 = 1
LEKF_PREC_BF16X3 = 2
PREC = {"f64": LEKF_PREC_F64, "f32": LEKF_PREC_F32, "bf16x3": LEKF_PREC_BF16X3}

LEKF_INFL_ADAPTIVE = 0
LEKF_INFL_FIXED = 1

LEKF_F_WLOC = 1
LEKF_F_MINFL = 2
LEKF_F_FIXED_INFL = 4
LEKF_F_TRANSM = 8
LEKF_F_APPLY = 16
LEKF_F_MIXED = 32
LEKF_F_BF16X3 = 64

LEKF_PT_RANKDEF = 1

LEKF_RELAX_NONE = 0
LEKF_RELAX_RTPP = 1
LEKF_RELAX_RTPS = 2
RELAX = {"none": LEKF_RELAX_NONE, "rtpp": LEKF_RELAX_RTPP, "rtps": LEKF_RELAX_RTPS}

LEKF_TUNE_GRAM_WAVES = 0
LEKF_TUNE_GRAM_WS_MB = 1
LEKF_TUNE_SEL_PER_CU = 2
LEKF_TUNE_EIG_SPLIT = 3
LEKF_TUNE_SPLIT_RCAP = 4
LEKF_TUNE_MIXED_KAPPA = 5
LEKF_NOALLOC_BYTES = 256 << 20
LEKF_NOALLOC_MAX_JOBS = 1024
LEKF_PT_NOCONV = 2
LEKF_PT_FP64 = 4
LEKF_NVMAX = 16
LEKF_KMAX = 256
LEKF_KSPLIT = 95

_dp = POINTER(c_double)
_ip = POINTER(c_int)
_lp = POINTER(c_long)


class LekfJob(Structure):
    _fields_ = [
        ("pt_begin", c_long), ("pt0", c_long), ("npt", c_int), ("k", c_int), ("flags", c_int),
        ("nij", c_int),
        ("nv3d", c_int), ("nv2d", c_int), ("minfl", c_double), ("infl_fixed", c_double),
        ("hdxf", c_void_p), ("y_so", c_long), ("y_sm", c_long), ("hdxf32", c_void_p),
        ("rdiag", c_void_p), ("dep", c_void_p), ("loc_off", c_void_p), ("loc_cnt", c_void_p),
        ("loc_idx", c_void_p), ("loc_rho", c_void_p), ("loc_rdiag", c_void_p),
        ("infl_in", c_void_p), ("infl_out", c_void_p), ("trans", c_void_p), ("transm", c_void_p),
        ("pao", c_void_p), ("status", c_void_p), ("x3b", c_void_p), ("x3a", c_void_p),
        ("x2b", c_void_p), ("x2a", c_void_p), ("upd_mask", c_void_p),
    ]


class GlobalCfg(Structure):
    _fields_ = [
        ("nlon", c_int), ("nlat", c_int), ("nlev", c_int), ("lon", _dp), ("lat", _dp),
        ("sig", _dp), ("ne", c_int), ("nv3d", c_int), ("nv2d", c_int), ("ips", c_int),
        ("iq", c_int), ("q_top_lev", c_int), ("sigma_obs", c_double), ("sigma_obsv", c_double),
        ("gross_error", c_double), ("var_local", c_double * 7), ("prec", c_int),
        ("infl_mode", c_int), ("infl_value", c_double), ("nobs_max", c_int),
    ]


class L96Cfg(Structure):
    _fields_ = [("nx", c_int), ("ne", c_int), ("xlocal", c_double), ("prec", c_int),
                ("infl_mode", c_int), ("infl_value", c_double)]


class RegionalCfg(Structure):
    _fields_ = [
        ("nx", c_int), ("ny", c_int), ("nlev", c_int), ("dx", c_double), ("dy", c_double),
        ("zlev", _dp), ("ne", c_int), ("nv3d", c_int), ("sigma_h", c_double),
        ("sigma_v", c_double), ("nobs_max", c_int), ("prec", c_int), ("infl_mode", c_int),
        ("infl_value", c_double), ("minfl", c_double), ("list_budget", c_long),
    ]


class RadarCfg(Structure):
    _fields_ = [("var", c_int * 11), ("method", c_int), ("use_terminal_velocity", c_int),
                ("min_ref_dbz", c_double), ("low_ref_shift", c_double), ("zmax", c_double)]


class ScaleQcCfg(Structure):
    _fields_ = ([(k, c_int) for k in ("use_radar_ref", "use_radar_vr", "use_obserr_radar_ref",
                                      "use_obserr_radar_vr", "min_radar_ref_member",
                                      "min_radar_ref_member_obsref")] +
                [(k, c_double) for k in ("radar_ref_thres_dbz", "min_radar_ref_dbz",
                                         "low_ref_shift", "obserr_radar_ref", "obserr_radar_vr",
                                         "gross_error", "gross_error_rain",
                                         "gross_error_radar_ref", "gross_error_radar_vr",
                                         "gross_error_radar_prh", "gross_error_tcx",
                                         "gross_error_tcy", "gross_error_tcp")])


# namelist defaults of the departure and QC step (scale/common/common_nml.f90:83-91,
# 140-146,165-174)
SCALE_QC_DEFAULTS = dict(use_radar_ref=1, use_radar_vr=1, use_obserr_radar_ref=0,
                         use_obserr_radar_vr=0, min_radar_ref_member=1,
                         min_radar_ref_member_obsref=1, radar_ref_thres_dbz=15.0,
                         min_radar_ref_dbz=0.0, low_ref_shift=0.0, obserr_radar_ref=5.0,
                         obserr_radar_vr=3.0, gross_error=5.0, gross_error_rain=-1.0,
                         gross_error_radar_ref=-1.0, gross_error_radar_vr=-1.0,
                         gross_error_radar_prh=-1.0, gross_error_tcx=-1.0, gross_error_tcy=-1.0,
                         gross_error_tcp=-1.0)

# scale/common/common_obs_scale.f90:71-72
ID_RADAR_REF, ID_RADAR_VR = 4001, 4002

_lib = None


def lib():
    """Load liblekf_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m letkf_amd.build` "
            "(there is no CPU fallback for the LETKF kernels)")
    L = ctypes.CDLL(LIB_PATH)
    L.lekf_version.restype = c_char_p
    L.lekf_last_error.restype = c_char_p
    L.lekf_device_count.restype = c_int
    L.lekf_core_f64.restype = c_int
    L.lekf_core_f64.argtypes = [c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                c_int, _dp]
    L.lekf_core_batch.restype = c_int
    L.lekf_core_batch.argtypes = [POINTER(LekfJob), c_int, c_int, c_void_p]
    L.lekf_core_batch_noalloc.restype = c_int
    L.lekf_core_batch_noalloc.argtypes = [POINTER(LekfJob), c_int, c_int, c_void_p]
    L.lekf_obs_local_l96.restype = c_int
    L.lekf_obs_local_l96.argtypes = [c_int, c_int, c_void_p, c_double, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]
    L.lekf_global_create.restype = c_void_p
    L.lekf_global_create.argtypes = [POINTER(GlobalCfg), _ip]
    L.lekf_l96_create.restype = c_void_p
    L.lekf_l96_create.argtypes = [POINTER(L96Cfg), _ip]
    if hasattr(L, "lekf_regional_create"):
        L.lekf_regional_create.restype = c_void_p
        L.lekf_regional_create.argtypes = [POINTER(RegionalCfg), _ip]
        L.lekf_set_obs_regional.restype = c_int
        L.lekf_set_obs_regional.argtypes = [c_void_p, c_int] + [_dp] * 5 + [_ip, _ip, _dp]
    L.lekf_destroy.restype = None
    L.lekf_destroy.argtypes = [c_void_p]
    L.lekf_set_ensemble.restype = c_int
    L.lekf_set_ensemble.argtypes = [c_void_p, _dp, _dp]
    L.lekf_set_infl.restype = c_int
    L.lekf_set_infl.argtypes = [c_void_p, _dp, _dp]
    L.lekf_set_obs_global.restype = c_int
    L.lekf_set_obs_global.argtypes = [c_void_p, c_int] + [_dp] * 6 + [_ip, _dp]
    L.lekf_set_obs_global_hx.restype = c_int
    L.lekf_set_obs_global_hx.argtypes = [c_void_p, c_int] + [_dp] * 6
    L.lekf_obsope_global.restype = c_int
    L.lekf_obsope_global.argtypes = [c_void_p, c_int] + [_dp] * 6 + [_ip, _dp, _dp, _dp]
    L.lekf_set_obs_l96.restype = c_int
    L.lekf_set_obs_l96.argtypes = [c_void_p, c_int, _ip, _dp, _dp, _dp]
    L.lekf_das_run.restype = c_int
    L.lekf_das_run.argtypes = [c_void_p]
    L.lekf_set_graph.restype = c_int
    L.lekf_set_graph.argtypes = [c_void_p, c_int]
    L.lekf_das_select.restype = c_int
    L.lekf_das_select.argtypes = [c_void_p]
    L.lekf_das_solve.restype = c_int
    L.lekf_das_solve.argtypes = [c_void_p]
    L.lekf_set_stream.restype = c_int
    L.lekf_set_stream.argtypes = [c_void_p, c_void_p]
    L.lekf_das_run_many.restype = c_int
    L.lekf_das_run_many.argtypes = [POINTER(c_void_p), c_int]
    L.lekf_sync.restype = c_int
    L.lekf_sync.argtypes = [c_void_p]
    L.lekf_get_analysis.restype = c_int
    L.lekf_get_analysis.argtypes = [c_void_p, _dp, _dp]
    L.lekf_get_infl.restype = c_int
    L.lekf_get_infl.argtypes = [c_void_p, _dp, _dp]
    L.lekf_get_nobsl.restype = c_int
    L.lekf_get_nobsl.argtypes = [c_void_p, _ip]
    L.lekf_get_local.restype = c_int
    L.lekf_get_local.argtypes = [c_void_p, c_long, c_int, _ip, _dp]
    L.lekf_get_obs_index.restype = c_int
    L.lekf_get_obs_index.argtypes = [c_void_p, _ip, _ip, _ip]
    L.lekf_get_flags.restype = c_int
    L.lekf_get_flags.argtypes = [c_void_p, _lp, _lp]
    L.lekf_get_fp64_count.restype = c_int
    L.lekf_get_fp64_count.argtypes = [c_void_p, _lp]
    L.lekf_cycle_swap.restype = c_int
    L.lekf_cycle_swap.argtypes = [c_void_p]
    L.lekf_device_state.restype = c_int
    L.lekf_device_state.argtypes = [c_void_p] + [POINTER(c_void_p)] * 4
    L.lekf_device_bytes.restype = c_long
    L.lekf_device_bytes.argtypes = [c_void_p]
    L.lekf_stream.restype = c_void_p
    L.lekf_stream.argtypes = [c_void_p]
    L.lekf_dev_malloc.restype = c_void_p
    L.lekf_dev_malloc.argtypes = [ctypes.c_size_t]
    L.lekf_dev_free.restype = c_int
    L.lekf_dev_free.argtypes = [c_void_p]
    L.lekf_memcpy.restype = c_int
    L.lekf_memcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_int, c_void_p]
    L.lekf_memset.restype = c_int
    L.lekf_memset.argtypes = [c_void_p, c_int, ctypes.c_size_t, c_void_p]
    L.lekf_stream_create.restype = c_void_p
    L.lekf_stream_create.argtypes = []
    L.lekf_stream_destroy.restype = c_int
    L.lekf_stream_destroy.argtypes = [c_void_p]
    L.lekf_stream_sync.restype = c_int
    L.lekf_stream_sync.argtypes = [c_void_p]
    L.lekf_event_create.restype = c_void_p
    L.lekf_event_create.argtypes = []
    L.lekf_event_destroy.restype = c_int
    L.lekf_event_destroy.argtypes = [c_void_p]
    L.lekf_event_record.restype = c_int
    L.lekf_event_record.argtypes = [c_void_p, c_void_p]
    L.lekf_event_elapsed_ms.restype = c_int
    L.lekf_event_elapsed_ms.argtypes = [c_void_p, c_void_p, POINTER(c_float)]
    L.lekf_event_query.restype = c_int
    L.lekf_event_query.argtypes = [c_void_p]
    L.lekf_event_sync.restype = c_int
    L.lekf_event_sync.argtypes = [c_void_p]
    L.lekf_mem_info.restype = c_int
    L.lekf_mem_info.argtypes = [POINTER(ctypes.c_size_t), POINTER(ctypes.c_size_t)]
    L.lekf_solve_workspace_bytes.restype = c_long
    L.lekf_solve_workspace_bytes.argtypes = [c_long, c_int]
    L.lekf_das_run_batch.restype = c_int
    L.lekf_das_run_batch.argtypes = [POINTER(c_void_p), c_int, c_void_p]
    L.lekf_l96_forecast.restype = c_int
    L.lekf_l96_forecast.argtypes = [c_void_p, c_int, c_double, c_double]
    L.lekf_set_relax.restype = c_int
    L.lekf_set_relax.argtypes = [c_void_p, c_int, c_double, c_double, c_int, c_double]
    L.lekf_set_heights.restype = c_int
    L.lekf_set_heights.argtypes = [c_void_p, _dp]
    L.lekf_obsope_radar.restype = c_int
    L.lekf_obsope_radar.argtypes = [c_void_p, POINTER(RadarCfg), c_int] + [_dp] * 10 + [_ip, _dp]
    L.lekf_obs_departure_scale.restype = c_int
    L.lekf_obs_departure_scale.argtypes = ([c_void_p, POINTER(ScaleQcCfg), c_int] + [_dp] * 3 +
                                           [_ip] + [_dp] * 4 + [_ip])
    L.lekf_set_obs_radar.restype = c_int
    L.lekf_set_obs_radar.argtypes = ([c_void_p, POINTER(RadarCfg), POINTER(ScaleQcCfg), c_int] +
                                     [_dp] * 11 + [_ip, _ip])
    L.lekf_tune_set.restype = c_int
    L.lekf_tune_set.argtypes = [c_int, c_long]
    L.lekf_tune_get.restype = c_long
    L.lekf_tune_get.argtypes = [c_int]
    L.lekf_ktime_enable.argtypes = [c_int]
    L.lekf_ktime_read.argtypes = [c_char_p, POINTER(c_double), POINTER(c_long)]
    _lib = L
    return L


# every symbol include/lekf.h declares (checked by the CPU test suite)
EXPORTED = [
    "lekf_core_f64", "lekf_core_batch", "lekf_core_batch_noalloc", "lekf_obs_local_l96", "lekf_global_create",
    "lekf_l96_create", "lekf_regional_create", "lekf_destroy", "lekf_set_ensemble",
    "lekf_set_infl", "lekf_set_obs_global", "lekf_set_obs_l96", "lekf_set_obs_regional",
    "lekf_das_run", "lekf_das_select", "lekf_das_solve", "lekf_set_stream", "lekf_das_run_many", "lekf_sync", "lekf_get_analysis", "lekf_get_infl",
    "lekf_get_nobsl", "lekf_get_local", "lekf_get_obs_index", "lekf_get_flags", "lekf_get_fp64_count",
    "lekf_cycle_swap", "lekf_device_state", "lekf_device_bytes", "lekf_stream",
    "lekf_device_count", "lekf_version", "lekf_last_error", "lekf_dev_malloc", "lekf_dev_free",
    "lekf_memcpy", "lekf_memset", "lekf_stream_create", "lekf_stream_destroy", "lekf_stream_sync",
    "lekf_event_create", "lekf_event_destroy", "lekf_event_record", "lekf_event_elapsed_ms",
    "lekf_event_query", "lekf_event_sync", "lekf_das_run_batch", "lekf_mem_info",
    "lekf_solve_workspace_bytes",
    "lekf_set_obs_global_hx", "lekf_obsope_global", "lekf_set_graph", "lekf_tune_set",
    "lekf_tune_get", "lekf_ktime_enable", "lekf_ktime_read", "lekf_set_relax", "lekf_l96_forecast", "lekf_set_heights",
    "lekf_obsope_radar", "lekf_obs_departure_scale", "lekf_set_obs_radar",
]


class LetkfError(RuntimeError):
    """Nonzero status of the C ABI (the reference would STOP with this code)."""

    def __init__(self, code, where=""):
        msg = ""
        try:
            msg = lib().lekf_last_error().decode()
        except Exception:
            pass
        super().__init__(f"{where}: status {code} {msg}".strip())
        self.code = code


def check(code, where=""):
    if code != LEKF_OK:
        raise LetkfError(code, where)
    return code


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)
