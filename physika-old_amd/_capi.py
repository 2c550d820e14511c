"""ctypes binding of the C ABI in include/pbf.h (libpbf.so).

The library is required: loading fails loudly when libpbf.so has not been
built (``python __graft_entry__.py build``); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
_LIB = _PKG / "_build" / "libpbf.so"

# PBF_ABI_VERSION of the include/pbf.h this table mirrors; lib() refuses a
# library built from another version (its entry points may take other arguments)
ABI_VERSION = 9

PBF_OK = 0
PBF_ERR_INVALID = -1
PBF_ERR_HIP = -2
PBF_ERR_CAPACITY = -3
PBF_ERR_ALLOC = -4
PBF_ERR_UNSUPPORTED = -5

PBF_FLAG_VISCOSITY = 1
PBF_FLAG_BOX = 2
PBF_FLAG_EAGER = 4


class pbf_params(C.Structure):
    _fields_ = [
        ("smoothing_length", C.c_float),
        ("sampling_distance", C.c_float),
        ("rest_density", C.c_float),
        ("dt", C.c_float),
        ("gravity", C.c_float * 3),
        ("iterations", C.c_int),
        ("lambda_rest", C.c_float),
        ("lambda_eps", C.c_float),
        ("corr_scale", C.c_float),
        ("pair_eps", C.c_float),
        ("scene_lo", C.c_float * 3),
        ("scene_hi", C.c_float * 3),
        ("mask_words", C.c_int),
        ("list_slots", C.c_int),
        ("viscosity", C.c_float),
        ("viscosity_iterations", C.c_int),
        ("box_lo", C.c_float * 3),
        ("box_hi", C.c_float * 3),
        ("box_spacing", C.c_float),
        ("normal_friction", C.c_float),
        ("tangent_friction", C.c_float),
        ("flags", C.c_uint),
        ("stage_min_particles", C.c_int),
    ]


class pbf_stats(C.Structure):
    _fields_ = [
        ("neighbor_pairs", C.c_longlong),
        ("out_of_grid", C.c_longlong),
        ("nonfinite", C.c_longlong),
        ("mask_words_needed", C.c_longlong),
        ("mask_words_reserved", C.c_longlong),
        ("mask_fallback_waves", C.c_longlong),
        ("list_slots_needed", C.c_longlong),
        ("list_slots_reserved", C.c_longlong),
        ("cells_used", C.c_longlong),
        ("candidates", C.c_longlong),
        ("staged_waves", C.c_longlong),
        ("grid_dims", C.c_int * 3),
        ("grid_lo", C.c_float * 3),
        ("cell_size", C.c_float),
        ("density_factor", C.c_float),
        ("particle_mass", C.c_float),
        ("neighbor_threshold", C.c_float),
        ("graph_captures", C.c_longlong),
        ("graph_replays", C.c_longlong),
        ("pass_lanes", C.c_int),
    ]


PASS_NAMES = ("sort", "neighbors", "density_lambda", "displacement")


class pbf_profile(C.Structure):
    _fields_ = [("ms", C.c_double * 4), ("launches", C.c_longlong * 4)]


class pbf_tuning(C.Structure):
    _fields_ = [("pass_block", C.c_int), ("stage_min_particles", C.c_int), ("persistent", C.c_int),
                ("pass_lanes", C.c_int)]


class pbf_tune_report(C.Structure):
    _fields_ = [
        ("best", pbf_tuning),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
        ("candidates", C.c_int),
        ("from_cache", C.c_int),
        ("key", C.c_char * 128),
    ]


# every symbol include/pbf.h declares, with its ctypes signature
_VP = C.c_void_p
_CTX = C.c_void_p
SIGNATURES = {
    "pbf_abi_version": (C.c_int, []),
    "pbf_status_string": (C.c_char_p, [C.c_int]),
    "pbf_params_default": (None, [C.POINTER(pbf_params)]),
    "pbf_create": (C.c_int, [C.POINTER(pbf_params), C.POINTER(_CTX)]),
    "pbf_destroy": (C.c_int, [_CTX]),
    "pbf_set_params": (C.c_int, [_CTX, C.POINTER(pbf_params)]),
    "pbf_get_params": (C.c_int, [_CTX, C.POINTER(pbf_params)]),
    "pbf_reserve": (C.c_int, [_CTX, C.c_int]),
    "pbf_reserve_batch": (C.c_int, [_CTX, C.c_int, C.c_int]),
    "pbf_neighbor_query": (C.c_int, [_CTX, _VP, C.c_int, _VP, _VP, C.c_longlong, C.POINTER(C.c_longlong), _VP]),
    "pbf_neighbor_query_fixed": (C.c_int, [_CTX, _VP, C.c_int, C.c_int, _VP, _VP, _VP]),
    "pbf_integrate": (C.c_int, [_CTX, _VP, _VP, _VP, C.c_int, _VP]),
    "pbf_density": (C.c_int, [_CTX, _VP, _VP, _VP, C.c_longlong, _VP, C.c_int, _VP]),
    "pbf_pbd_constrain": (C.c_int, [_CTX, _VP, _VP, _VP, _VP, C.c_longlong, _VP, C.c_int, _VP]),
    "pbf_pbd_constrain_mass": (C.c_int, [_CTX, _VP, _VP, _VP, _VP, C.c_longlong, _VP, _VP, C.c_int, _VP]),
    "pbf_pbd_constrain_noalloc": (C.c_int, [_CTX, _VP, _VP, _VP, _VP, C.c_longlong, _VP, C.c_int, _VP]),
    "pbf_step": (C.c_int, [_CTX, _VP, _VP, _VP, C.c_int, C.c_int, _VP]),
    "pbf_batch_step": (C.c_int, [_CTX, _VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_int, _VP]),
    "pbf_rollout_forward": (C.c_int, [_CTX, _VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP]),
    "pbf_rollout_backward": (C.c_int, [_CTX, _VP, _VP, _VP]),
    "pbf_rollout_forward_f64": (C.c_int, [_CTX, _VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP]),
    "pbf_rollout_backward_f64": (C.c_int, [_CTX, _VP, _VP, _VP]),
    "pbf_step_neighbors": (C.c_int, [_CTX, _VP, _VP, C.c_longlong, C.POINTER(C.c_longlong), _VP]),
    "pbf_viscosity": (C.c_int, [_CTX, _VP, _VP, _VP, _VP, C.c_longlong, C.c_int, _VP]),
    "pbf_box_constrain": (C.c_int, [_CTX, _VP, _VP, C.c_int, _VP]),
    "pbf_get_stats": (C.c_int, [_CTX, C.POINTER(pbf_stats)]),
    "pbf_set_profiling": (C.c_int, [_CTX, C.c_int]),
    "pbf_get_profile": (C.c_int, [_CTX, C.POINTER(pbf_profile)]),
    "pbf_set_tuning": (C.c_int, [_CTX, C.POINTER(pbf_tuning)]),
    "pbf_get_tuning": (C.c_int, [_CTX, C.POINTER(pbf_tuning)]),
    "pbf_autotune": (C.c_int, [_CTX, _VP, _VP, _VP, C.c_int, C.c_int, C.c_int, C.c_char_p,
                               C.POINTER(pbf_tune_report), _VP]),
    "pbf_last_error": (C.c_char_p, [_CTX]),
}


class PBFError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"pbf status {status}: {message}")
        self.status = status


def library_path(must_exist: bool = True) -> Path:
    if must_exist and not _LIB.exists():
        raise RuntimeError(f"{_LIB} is not built; run `python __graft_entry__.py build` (no CPU fallback exists)")
    return _LIB


_lib = None


def lib():
    """The loaded libpbf.so with typed entry points."""
    global _lib
    if _lib is None:
        L = C.CDLL(str(library_path(True)))
        L.pbf_abi_version.restype = C.c_int
        L.pbf_abi_version.argtypes = []
        if L.pbf_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{_LIB} has ABI version {L.pbf_abi_version()}, this binding expects {ABI_VERSION}; "
                               "rebuild it (`python __graft_entry__.py build`)")
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_native = None
_NATIVE = _LIB.parent / ("_pbf_torch" + __import__("sysconfig").get_config_var("EXT_SUFFIX"))


def native():
    """The pybind11 layer (_pbf_torch, csrc/pbf_torch.cpp): the reverse mode as
    a torch C++ autograd Function over the same libpbf.so. Loaded after lib(),
    so the library it links is the checked one."""
    global _native
    if _native is None:
        lib()
        if not _NATIVE.exists():
            raise RuntimeError(f"{_NATIVE} is not built; run `python __graft_entry__.py build`")
        import importlib.util
        import torch  # noqa: F401  (libtorch before the extension)
        spec = importlib.util.spec_from_file_location("_pbf_torch", _NATIVE)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        if mod.abi_version() != ABI_VERSION:
            raise RuntimeError(f"{_NATIVE} links a library of ABI version {mod.abi_version()}, expected {ABI_VERSION}")
        _native = mod
    return _native


def check(status: int, ctx=None) -> None:
    if status != PBF_OK:
        msg = lib().pbf_last_error(ctx).decode() if ctx else lib().pbf_status_string(status).decode()
        raise PBFError(status, msg or lib().pbf_status_string(status).decode())


def default_params() -> pbf_params:
    p = pbf_params()
    lib().pbf_params_default(C.byref(p))
    return p


def params_from(**kw) -> pbf_params:
    """Default parameters with keyword overrides (vector fields take 3-sequences)."""
    p = default_params()
    for k, v in kw.items():
        if k in ("gravity", "scene_lo", "scene_hi", "box_lo", "box_hi"):
            arr = getattr(p, k)
            for i in range(3):
                arr[i] = float(v[i])
        elif not hasattr(p, k):
            raise KeyError(k)
        else:
            setattr(p, k, v)
    return p


def params_dict(p: pbf_params) -> dict:
    out = {}
    for name, _ in pbf_params._fields_:
        v = getattr(p, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out
