"""ctypes binding of libfases_core.so (include/fases_core.h).

The library is loaded from the package directory. There is no fallback: if it is
missing the import error says how to build it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfases_core.so"

FASES_OK = 0
FASES_E_ARG = -1
FASES_E_IO = -2
FASES_E_PARSE = -3
FASES_E_HIP = -4
FASES_E_NOMEM = -5
FASES_E_STATE = -6
FASES_E_RANGE = -7
FASES_E_NONFINITE = -8

FASES_FP64, FASES_FP32, FASES_FP16 = 0, 1, 2
FASES_SRC_PCS, FASES_SRC_RICKER, FASES_SRC_SINE = 0, 1, 2

PRECISIONS = {"fp64": FASES_FP64, "fp32": FASES_FP32, "fp16": FASES_FP16}


class fases_params(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("precision", C.c_int),
        ("h", C.c_double),
        ("dt", C.c_double),
        ("c", C.c_double),
        ("rho", C.c_double),
        ("tol", C.c_double),
        ("pml_keys", C.c_float * 6),
        ("pml_cells", C.c_int),
        ("pml_sigma_max", C.c_double),
        ("pml_kappa_max", C.c_double),
        ("pml_alpha_max", C.c_double),
        ("fp16_rescale_every", C.c_int),
        ("use_graph", C.c_int),
        ("kernel_variant", C.c_int),
        ("tile_z", C.c_int),
        ("adjoint_variant", C.c_int),
        ("tile_rows", C.c_int),
        ("source_patch", C.c_int),
        ("tile_cols", C.c_int),
    ]


class fases_ga_params(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("ism_order", C.c_int),
        ("n_rays", C.c_int),
        ("rt_max_order", C.c_int),
        ("recv_radius", C.c_float),
        ("c", C.c_float),
        ("fs", C.c_float),
        ("Nt", C.c_int),
    ]


class FasesError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fases error {code}: {msg}")
        self.code = code


_P = C.POINTER
_dbl = _P(C.c_double)
_flt = _P(C.c_float)
_i64 = _P(C.c_int64)
_i32 = _P(C.c_int32)
_u8 = _P(C.c_uint8)

# name -> (restype, argtypes)
SIGNATURES = {
    "fases_params_default": (C.c_int, [_P(fases_params)]),
    "fases_read_config": (C.c_int, [C.c_char_p, _flt, C.c_int, _P(C.c_int)]),
    "fases_read_poles": (C.c_int, [C.c_char_p, _P(C.c_int), _P(C.c_int)]),
    "fases_read_fit": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int] + [_dbl] * 7),
    "fases_read_sources": (C.c_int, [C.c_char_p, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                     _i64, _flt, _flt, _flt, _P(C.c_int)]),
    "fases_read_receivers": (C.c_int, [C.c_char_p, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                       _i64, _flt, _P(C.c_int)]),
    "fases_source_waveform": (C.c_int, [C.c_int, C.c_float, C.c_double, C.c_float, C.c_int64, _dbl]),
    "fases_source_pref": (C.c_int, [_dbl, C.c_int64, C.c_float, C.c_float, C.c_float, _dbl, _dbl]),
    "fases_write_f32": (C.c_int, [C.c_char_p, _dbl, C.c_int64]),
    "fases_scene_load_obj": (C.c_int, [C.c_char_p, C.c_float, _P(C.c_void_p)]),
    "fases_scene_shoebox": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_float, _P(C.c_void_p)]),
    "fases_scene_load_obj_dev": (C.c_int, [C.c_char_p, C.c_float, C.c_int, _P(C.c_void_p)]),
    "fases_scene_load_obj_classified": (C.c_int, [C.c_char_p, C.c_float, _flt, _P(C.c_uint8), C.c_int, C.c_int,
                                                  C.c_int, _P(C.c_void_p)]),
    "fases_scene_shoebox_dev": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, _P(C.c_void_p)]),
    "fases_scene_info": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int), _P(C.c_int), _P(C.c_int), _flt]),
    "fases_scene_group_name": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int]),
    "fases_scene_get": (C.c_int, [C.c_void_p, _flt, _u8]),
    "fases_scene_free": (None, [C.c_void_p]),
    "fases_boundary_map": (C.c_int, [_u8, C.c_int64, _i32, _i64]),
    "fases_create": (C.c_int, [_P(fases_params), _P(C.c_void_p)]),
    "fases_destroy": (C.c_int, [C.c_void_p]),
    "fases_last_error": (C.c_char_p, [C.c_void_p]),
    "fases_set_geometry": (C.c_int, [C.c_void_p, _flt, _u8, C.c_int, C.c_int, C.c_int]),
    "fases_set_materials": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int] + [_dbl] * 7),
    "fases_set_sources": (C.c_int, [C.c_void_p, C.c_int, _i64, _dbl, C.c_int, _dbl, C.c_int64, _dbl]),
    "fases_set_receivers": (C.c_int, [C.c_void_p, C.c_int, _i64]),
    "fases_set_fields": (C.c_int, [C.c_void_p, _dbl, _dbl, _dbl, _dbl]),
    "fases_get_fields": (C.c_int, [C.c_void_p, _dbl, _dbl, _dbl, _dbl]),
    "fases_reset": (C.c_int, [C.c_void_p]),
    "fases_run": (C.c_int, [C.c_void_p, C.c_int64, _dbl, _dbl, _dbl, _dbl]),
    "fases_info": (C.c_int, [C.c_void_p, _i64, _i64, _i64, _i64]),
    "fases_device_count": (C.c_int, [_P(C.c_int)]),
    "fases_get_slices": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _flt]),
    "fases_get_slices_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _flt]),
    "fases_scene_group_areas": (C.c_int, [C.c_void_p, _flt, _flt]),
    "fases_air_attenuation": (C.c_int, [C.c_float, C.c_float, C.c_float, _flt, C.c_int, _flt]),
    "fases_read_mat": (C.c_int, [C.c_char_p, _flt, _P(C.c_int)]),
    "fases_estimate_t60": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_int, _flt, _flt, _flt, _flt, _flt]),
    "fases_tile_config": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int), _P(C.c_int)]),
    "fases_hbm_copy": (C.c_int, [C.c_int, C.c_int64, C.c_int, _dbl]),
    "fases_ga_params_default": (C.c_int, [_P(fases_ga_params)]),
    "fases_ga_run": (C.c_int, [C.c_void_p, _P(fases_ga_params), C.c_int, _flt, _flt, C.c_int, _flt, _flt, _flt, _flt,
                               _P(C.c_void_p)]),
    "fases_ga_free": (None, [C.c_void_p]),
    "fases_ga_info": (C.c_int, [C.c_void_p, _P(C.c_int), _i64, _i64, _dbl, _dbl, _dbl]),
    "fases_ga_direct": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _flt]),
    "fases_ga_images": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _P(C.c_int), _P(C.c_int), _flt, _flt, _flt,
                                  _P(C.c_int), _P(C.c_int), _flt, _flt]),
    "fases_ga_energy": (C.c_int, [C.c_void_p, _flt, _flt]),
    "fases_ga_band_irs": (C.c_int, [C.c_void_p, C.c_int, _flt, _flt]),
    "fases_ga_arrivals": (C.c_int, [C.c_void_p, C.c_int, _P(C.c_int), _flt]),
    "fases_ga_dump_rays": (C.c_int, [C.c_void_p, C.c_int, _flt, C.c_int, C.c_int, C.c_char_p]),
    "fases_pml_tables": (C.c_int, [C.c_void_p, C.c_int, _P(C.c_int32), _dbl, _dbl, _dbl]),
    "fases_batch_run": (C.c_int, [_P(C.c_void_p), C.c_int, _i64, _P(_P(C.c_double)), _P(C.c_double)]),
    "fases_adjoint_forward": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, _P(C.c_double)]),
    "fases_adjoint": (C.c_int, [C.c_void_p, _P(C.c_double), _P(C.c_double), _P(C.c_double), _P(C.c_double)]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libfases_core.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m fases_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, ctx=None) -> None:
    if rc != FASES_OK:
        msg = load().fases_last_error(ctx)
        raise FasesError(rc, msg.decode() if msg else "")
