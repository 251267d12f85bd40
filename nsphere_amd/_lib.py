"""ctypes binding of libnsphere_gpu.so (the C-ABI declared in include/nsphere_gpu.h).

The product path has no fallback: if the native library is missing or was
built for another target, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsphere_gpu.so")

NS_OK, NS_ERR_ARG, NS_ERR_HIP, NS_ERR_NOMEM, NS_ERR_STATE, NS_ERR_NODEV = range(6)
NS_OPT_TIEBREAK, NS_OPT_PROFILE, NS_OPT_INTEGRATOR, NS_OPT_BUDGET, NS_OPT_PRECISION, NS_OPT_SORT_TILE = 1, 2, 3, 4, 5, 6
NS_OPT_SORT_ALG = 7
NS_OPT_GRAPH = 8
NS_OPT_LANE_LEVELS, NS_OPT_DEEP_BLOCKS = 9, 10
NS_OPT_OVERLAP = 11
NS_OPT_PROFILE_KERNELS = 12
NS_OPT_SORT_KEYS = 13
NS_OPT_GRAPH_STEPS = 14
SORT_REDUCE_SCAN, SORT_ONESWEEP = 0, 1
FORM_PLAIN, FORM_BINNED, FORM_BINNED_FAST = 0, 1, 2
NS_LVL_MASK, NS_LVL_NOCONV, NS_LVL_LC = 0x3F, 0x40, 0x80
NS_MAX_KERNEL_STATS = 32


class NativeLibraryMissing(ImportError):
    pass


class NsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class StepParams(ctypes.Structure):
    """Mirror of ns_step_params (method-5 constants, nsphere.c:5777-5782)."""

    _fields_ = [
        ("dt", ctypes.c_double),
        ("G", ctypes.c_double),
        ("halo_mass", ctypes.c_double),
        ("vel_conv_sq", ctypes.c_double),
        ("radius_tol", ctypes.c_double),
        ("velocity_tol", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("max_subdiv", ctypes.c_int32),
        ("out_type", ctypes.c_int32),
        ("N_taumin", ctypes.c_int32),
        ("identity_gravity", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

    def replace(self, **kw) -> "StepParams":
        p = StepParams()
        ctypes.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class DensityInfo(ctypes.Structure):
    _fields_ = [("kept", ctypes.c_int64), ("decimated", ctypes.c_int32), ("grid", ctypes.c_int32),
                ("taps", ctypes.c_int32), ("sigma", ctypes.c_double), ("r_min", ctypes.c_double),
                ("r_max", ctypes.c_double)]


class TuneConfig(ctypes.Structure):
    """Mirror of ns_tune_config."""

    _fields_ = [("sort_alg", ctypes.c_int32), ("sort_threads", ctypes.c_int32), ("sort_items", ctypes.c_int32),
                ("lane_levels", ctypes.c_int32), ("deep_blocks", ctypes.c_int32), ("sort_keys", ctypes.c_int32),
                ("ms_per_step", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class TuneResult(ctypes.Structure):
    """Mirror of ns_tune_result."""

    _fields_ = [("best", TuneConfig), ("ntried", ctypes.c_int32), ("tried", TuneConfig * 32)]


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is not built; run `python nsphere_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    c = ctypes
    vp, i64, i32, dbl = c.c_void_p, c.c_int64, c.c_int, c.c_double
    pd = c.POINTER(c.c_double)
    sig = {
        "ns_gpu_last_error": (c.c_char_p, []),
        "ns_gpu_version": (c.c_char_p, []),
        "ns_step_params_default": (None, [c.POINTER(StepParams)]),
        "ns_step_params_unit": (None, [c.POINTER(StepParams)]),
        "ns_gpu_ctx_create": (i32, [i64, i32, c.POINTER(vp)]),
        "ns_gpu_ctx_destroy": (i32, [vp]),
        "ns_gpu_set_option": (i32, [vp, i32, i64]),
        "ns_gpu_npts": (i64, [vp]),
        "ns_gpu_upload": (i32, [vp, c.POINTER(pd)]),
        "ns_gpu_download": (i32, [vp, c.POINTER(pd)]),
        "ns_gpu_upload_soa": (i32, [vp, vp, vp, vp, vp]),
        "ns_gpu_download_soa": (i32, [vp, vp, vp, vp, vp]),
        "ns_gpu_device_state": (i32, [vp, c.POINTER(vp), c.POINTER(vp), c.POINTER(vp), c.POINTER(vp)]),
        "ns_gpu_sort_particles": (i32, [vp]),
        "ns_gpu_inverse_map": (i32, [vp, vp]),
        "ns_gpu_permutation": (i32, [vp, vp]),
        "ns_gpu_permute_column": (i32, [vp, vp, vp]),
        "ns_gpu_sort_particles_noalloc": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, i64, vp]),
        "ns_gpu_integrate": (i32, [vp, c.POINTER(StepParams)]),
        "ns_gpu_step": (i32, [vp, c.POINTER(StepParams)]),
        "ns_gpu_run": (i32, [vp, c.POINTER(StepParams), i64]),
        "ns_gpu_graph_prepare": (i32, [vp, c.POINTER(StepParams), i64]),
        "ns_gpu_sync": (i32, [vp]),
        "ns_gpu_cap_hits": (i32, [vp, c.POINTER(c.c_uint64), i32]),
        "ns_gpu_levels": (i32, [vp, vp]),
        "ns_gpu_energy": (i32, [vp, c.POINTER(StepParams), pd, pd]),
        "ns_gpu_kernel_stats": (i32, [vp, c.POINTER(c.c_int), c.POINTER(c.c_char_p),
                                      c.POINTER(c.c_int64), c.POINTER(c.c_double)]),
        "ns_gpu_kernel_stats_reset": (i32, [vp]),
        "ns_gpu_force_evals": (i32, [vp, c.POINTER(c.c_uint64)]),
        "ns_gpu_chain_steps": (i32, [vp, c.POINTER(c.c_uint64)]),
        "ns_gpu_chain_floor": (i32, [i32, i64, c.POINTER(c.c_double)]),
        "ns_gpu_set_masses": (i32, [vp, vp]),
        "ns_gpu_set_halos": (i32, [vp, i32, vp, vp]),
        "ns_gpu_autotune": (i32, [vp, c.POINTER(StepParams), i32, c.POINTER(TuneResult)]),
        "ns_mlp_create": (i32, [i32, c.c_uint64, c.POINTER(vp)]),
        "ns_mlp_destroy": (i32, [vp]),
        "ns_mlp_train": (i32, [vp, vp, vp, i64, i32, i64, c.c_float, c.POINTER(c.c_float)]),
        "ns_mlp_gradient": (i32, [vp, vp, vp, i64, vp, c.POINTER(c.c_float)]),
        "ns_mlp_predict": (i32, [vp, vp, i64, vp]),
        "ns_mlp_get_params": (i32, [vp, vp, vp, vp]),
        "ns_mlp_set_params": (i32, [vp, vp, vp, vp]),
        "ns_gpu_set_mlp": (i32, [vp, vp]),
        "ns_gpu_snapshot_records": (i32, [vp, vp]),
        "ns_gpu_potential": (i32, [vp, c.POINTER(StepParams), vp, c.POINTER(c.c_double)]),
        "ns_gpu_density": (i32, [vp, c.POINTER(StepParams), vp, c.POINTER(DensityInfo)]),
        "ns_gpu_rank_files": (i32, [vp, c.POINTER(StepParams), c.c_char_p, c.c_char_p]),
        "ns_gpu_generate_ic": (i32, [vp, i32, c.c_uint64, c.c_double, c.c_double, c.c_double]),
        "ns_gpu_upload_ic": (i32, [vp, vp, c.c_double]),
        "ns_ic_file_write": (i32, [c.c_char_p, i32, vp]),
        "ns_ic_file_read": (i32, [c.c_char_p, c.POINTER(i32), vp]),
        "ns_gpu_gather_ids": (i32, [vp, vp, i32, vp, vp, vp, vp]),
        "ns_apd_writer_open": (i32, [vp, c.c_char_p, i32, c.POINTER(vp)]),
        "ns_apd_writer_write": (i32, [vp]),
        "ns_apd_writer_close": (i32, [vp, c.POINTER(i64)]),
        "ns_gpu_features": (i32, [vp, c.POINTER(StepParams), vp]),
        "ns_gpu_enclosed_mass": (i32, [vp, vp]),
        "ns_ic_generate": (i32, [i32, i64, c.c_uint64, dbl, dbl, dbl, vp, vp, vp, vp]),
        "ns_philox4x32_10": (None, [c.POINTER(c.c_uint32), c.POINTER(c.c_uint32), c.POINTER(c.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc: int) -> None:
    if rc != NS_OK:
        msg = lib.ns_gpu_last_error()
        raise NsError(rc, msg.decode() if msg else "unknown error")


def exported_symbols() -> list[str]:
    return [name for name in dir(lib) if name.startswith("ns_")]
