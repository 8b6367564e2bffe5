"""ctypes declarations of the C ABI in include/lumagpu.h.

The library is loaded from luma_amd/lib/libluma_amd.so (built in-tree by luma_amd.build).
There is no fallback: if the library is missing or fails to load, importing this module
raises, so nothing can silently run on the CPU in its place.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libluma_amd.so"

OK = 0
EINVAL, EHIP, ENOMEM, EUNSUPPORTED, ESTATE, EPHYS = -1, -2, -3, -4, -5, -6
ERROR_NAMES = {EINVAL: "EINVAL", EHIP: "EHIP", ENOMEM: "ENOMEM", EUNSUPPORTED: "EUNSUPPORTED",
               ESTATE: "ESTATE", EPHYS: "EPHYS"}

FP64, FP32S, FP16S, FP64T, FP32T = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": FP64, "fp32s": FP32S, "fp16s": FP16S, "fp64t": FP64T, "fp32t": FP32T}
NUM_MOMENTS = 5


class GridDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int), ("Q", C.c_int),
        ("N", C.c_int), ("M", C.c_int), ("K", C.c_int),
        ("precision", C.c_int), ("regularised", C.c_int), ("forced", C.c_int),
        ("fz_compat", C.c_int), ("level", C.c_int),
        ("omega", C.c_double), ("rho_in", C.c_double), ("rho0", C.c_double),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("wall_thk", C.c_double * 6),
        ("coarse_lims", C.c_int * 6),
        ("dh", C.c_double), ("origin", C.c_double * 3), ("dt", C.c_double), ("dm", C.c_double),
    ]


class AdjResult(C.Structure):
    _fields_ = [("J", C.c_double), ("dJ_domega", C.c_double), ("dJ_dinlet", C.c_double)]


class LumaGpuError(RuntimeError):
    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} failed with {ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

# name -> argtypes (all return int unless listed in _RESTYPES)
_PROTOS = {
    "lumagpu_grid_create": [C.POINTER(GridDesc), C.POINTER(_vp)],
    "lumagpu_grid_destroy": [_vp],
    "lumagpu_grid_add_child": [_vp, _vp],
    "lumagpu_set_labels": [_vp, _ip, _dp, _dp, _dp],
    "lumagpu_set_inlet": [_vp, _dp, _dp, _dp],
    "lumagpu_init_equilibrium": [_vp, _dp, _dp],
    "lumagpu_upload_f": [_vp, _dp],
    "lumagpu_upload_macro": [_vp, _dp, _dp],
    "lumagpu_download_f": [_vp, _dp],
    "lumagpu_download_macro": [_vp, _dp, _dp],
    "lumagpu_set_force": [_vp, _dp],
    "lumagpu_step": [_vp, C.c_int],
    "lumagpu_stream_bc_macro": [_vp, C.c_int],
    "lumagpu_force_collide_swap": [_vp],
    "lumagpu_reset_forces": [_vp],
    "lumagpu_ibm_apply": [_vp],
    "lumagpu_ibm_interpolate": [_vp],
    "lumagpu_ibm_compute_force": [_vp],
    "lumagpu_ibm_spread": [_vp],
    "lumagpu_ibm_update_macro": [_vp],
    "lumagpu_run": [_vp, C.c_int],
    "lumagpu_run_timed": [_vp, C.c_int, _dp, _dp],
    "lumagpu_set_store_macro": [_vp, C.c_int],
    "lumagpu_set_zslab": [_vp, C.c_int],
    "lumagpu_get_time": [_vp, _ip],
    "lumagpu_sync": [_vp],
    "lumagpu_moments": [_vp, _dp],
    "lumagpu_moments_noalloc": [_vp, _vp],
    "lumagpu_set_positions": [_vp, _dp, _dp, _dp],
    "lumagpu_ibm_add_body": [_vp, C.c_int, _dp, C.c_int, _ip],
    "lumagpu_ibm_initialise": [_vp],
    "lumagpu_ibm_add_filament": [_vp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_double, C.c_double, _ip],
    "lumagpu_ibm_subiterations": [_vp, _ip, _dp],
    "lumagpu_ibm_set_velocity": [_vp, C.c_int, _dp],
    "lumagpu_ibm_set_epsilon": [_vp, C.c_int, _dp],
    "lumagpu_ibm_get_markers": [_vp, C.c_int, _dp],
    "lumagpu_ibm_body_force": [_vp, C.c_int, _dp],
    "lumagpu_write_restart": [_vp, C.c_char_p, C.c_int, C.c_int],
    "lumagpu_read_restart": [_vp, C.c_char_p, C.c_int],
    "lumagpu_io_lite": [_vp, C.c_char_p, C.c_int, C.c_int],
    "lumagpu_ibm_write_ld": [_vp, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int],
    "lumagpu_ibm_write_tip": [_vp, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int],
    "lumagpu_ibm_count": [_vp, _ip, C.c_int, _ip],
    "lumagpu_ibm_read_geometry": [C.POINTER(_vp), C.c_int, C.c_char_p, _dp, _ip],
    "lumagpu_ibm_write_restart": [C.POINTER(_vp), C.c_int, C.c_char_p],
    "lumagpu_download_tangent": [_vp, _dp, _dp, _dp],
    "lumagpu_read_inlet_profile": [_vp, C.c_char_p],
    "lumagpu_read_initial_velocity": [_vp, C.c_char_p, _dp],
    "lumagpu_set_gravity": [_vp, C.c_double, C.c_int],
    "lumagpu_set_velocity_ramp": [_vp, C.c_double],
    "lumagpu_set_pressure_delta": [_vp, C.c_double],
    "lumagpu_ibm_set_universal_epsilon": [_vp, C.c_int],
    "lumagpu_set_reynolds_ramp": [_vp, C.c_double, C.c_double],
    "lumagpu_set_temporal_blocking": [_vp, C.c_int, C.c_int],
    "lumagpu_get_omega": [_vp, _dp],
    "lumagpu_tangent_seed_omega": [_vp, C.c_double],
    "lumagpu_ibm_body_force_tangent": [_vp, C.c_int, _dp, _dp],
    "lumagpu_ibm_get_marker_tangents": [_vp, C.c_int, _dp],
    "lumagpu_ibm_seed_youngs": [_vp, C.c_int, C.c_double],
    "lumagpu_ibm_subiteration_log": [_vp, _ip, C.c_int, _ip],
    "lumagpu_ibm_pin_subiterations": [_vp, _ip, C.c_int],
    "lumagpu_ibm_read_restart": [C.POINTER(_vp), C.c_int, C.c_char_p],
    "lumagpu_checkpoint_save": [_vp, C.POINTER(_vp)],
    "lumagpu_checkpoint_restore": [_vp, _vp],
    "lumagpu_checkpoint_free": [_vp],
    "lumagpu_adj_run": [_vp, C.c_int, C.c_int, _dp, C.c_int, C.POINTER(AdjResult), _dp, _dp],
    "lumagpu_last_error": [],
    "lumagpu_abi_version": [],
    "lumagpu_footprint": [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
}
_RESTYPES = {"lumagpu_last_error": C.c_char_p}

_lib = None


def load() -> C.CDLL:
    """Load libluma_amd.so once; raise if it is absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m luma_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, args in _PROTOS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, C.c_int)
        _lib = lib
    return _lib


def exported_symbols():
    return list(_PROTOS)


def check(rc: int, fn: str) -> None:
    if rc != OK:
        msg = load().lumagpu_last_error()
        raise LumaGpuError(rc, fn, msg.decode() if msg else "")


def call(name: str, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc
