"""ctypes binding of ``include/watercube.h`` (the library's C ABI).

Loading fails loudly when ``libwatercube.so`` has not been built: there is no CPU fallback for
any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libwatercube.so"

WC_OK = 0
WC_ERR_INVALID_ARG = -1
WC_ERR_HIP = -2
WC_ERR_NO_MEMORY = -3
WC_ERR_NO_DEVICE = -4
WC_ERR_CAPACITY = -5
WC_ERR_DEVICE_FAULT = -6

WC_EOS_CUBIC = 0
WC_EOS_LINEAR = 1

WC_PRESET_REFERENCE_DEFAULT = 0
WC_PRESET_DAM_BREAK_16K = 1
WC_PRESET_BLOCK_1M = 2
WC_PRESET_BLOCK_SPHERE_17M = 3

WC_NUM_PASSES = 6
WC_NOALLOC_MAX_PARTICLES = 262144
WC_NOALLOC_MAX_CELLS = 262144

# struct Particle (src/core/util.h:29-35) as a numpy record: 32 bytes, offsets 0/12/16/28.
PARTICLE_DTYPE = np.dtype(
    {
        "names": ["position", "density", "velocity", "pressure"],
        "formats": [("<f4", (3,)), "<f4", ("<f4", (3,)), "<f4"],
        "offsets": [0, 12, 16, 28],
        "itemsize": 32,
    }
)


class wc_params(C.Structure):
    _fields_ = [
        ("num_particles", C.c_uint32),
        ("grid_res", C.c_int32),
        ("size", C.c_float),
        ("bin_size", C.c_float),
        ("kernel_radius", C.c_float),
        ("particle_mass", C.c_float),
        ("rest_density", C.c_float),
        ("rest_pressure", C.c_float),
        ("stiffness", C.c_float),
        ("viscosity", C.c_float),
        ("gravity", C.c_float * 3),
        ("poly6_const", C.c_float),
        ("spiky_const", C.c_float),
        ("visc_const", C.c_float),
        ("wall_damping", C.c_float),
        ("border", C.c_float),
        ("max_speed", C.c_float),
        ("eos", C.c_int32),
        ("wall_density", C.c_int32),
        ("wall_force", C.c_int32),
        ("mouse_force", C.c_int32),
        ("camera_position", C.c_float * 3),
        ("mouse_ray_direction", C.c_float * 3),
        ("time_scale", C.c_float),
        ("particle_radius", C.c_float),
        ("reserved", C.c_uint32 * 4),
    ]

    def copy(self) -> "wc_params":
        out = wc_params()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(wc_params))
        return out

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if isinstance(v, C.Array) else v
        return d


class wc_device_view(C.Structure):
    _fields_ = [
        ("posrho", C.c_void_p * 2),
        ("velp", C.c_void_p * 2),
        ("cell_range", C.c_void_p),
        ("sorted", C.c_void_p),
        ("stream", C.c_void_p),
        ("num_particles", C.c_uint32),
        ("num_cells", C.c_uint32),
    ]


class wc_closure_stats(C.Structure):
    _fields_ = [("macro_steps", C.c_uint32), ("fallbacks", C.c_uint32), ("wall_ms", C.c_double),
                ("max_density_jump", C.c_double), ("max_accepted_jump", C.c_double), ("peak_density", C.c_double)]


class wc_tuning(C.Structure):
    _fields_ = [("density_variant", C.c_int32), ("update_variant", C.c_int32), ("neighbour_list", C.c_int32)]


_P = C.POINTER
_u32p = _P(C.c_uint32)
_f32p = _P(C.c_float)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol include/watercube.h declares.
SIGNATURES = {
    "wc_params_preset": (C.c_int, [_P(wc_params), C.c_int]),
    "wc_params_derive": (C.c_int, [_P(wc_params)]),
    "wc_scene_block": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, _f32p, C.c_float, C.c_float, _f32p, C.c_uint64]),
    "wc_scene_sphere": (C.c_int, [_vp, C.c_uint32, _f32p, C.c_float, C.c_float, _f32p, C.c_uint64]),
    "wc_scene_reference": (C.c_int, [_vp, C.c_uint32, C.c_float, C.c_uint64]),
    "wc_create": (C.c_int, [_P(wc_params), _P(_vp)]),
    "wc_destroy": (C.c_int, [_vp]),
    "wc_set_params": (C.c_int, [_vp, _P(wc_params)]),
    "wc_get_params": (C.c_int, [_vp, _P(wc_params)]),
    "wc_upload": (C.c_int, [_vp, _vp, C.c_uint32]),
    "wc_download": (C.c_int, [_vp, C.c_int, _vp, C.c_uint32]),
    "wc_sort_run": (C.c_int, [_vp]),
    "wc_density": (C.c_int, [_vp]),
    "wc_update": (C.c_int, [_vp, C.c_float]),
    "wc_advect": (C.c_int, [_vp, C.c_float]),
    "wc_gravity_direction": (C.c_int, [_P(C.c_float), _P(C.c_float)]),
    "wc_step": (C.c_int, [_vp, C.c_float, C.c_uint32]),
    "wc_step_noalloc": (C.c_int, [_P(wc_params), C.c_float, _vp, _vp, _u32p, _vp]),
    "wc_synchronize": (C.c_int, [_vp]),
    "wc_create_batch": (C.c_int, [_P(wc_params), C.c_uint32, _P(_vp)]),
    "wc_num_scenes": (C.c_int, [_vp]),
    "wc_scene_upload": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint32]),
    "wc_scene_download": (C.c_int, [_vp, C.c_uint32, C.c_int, _vp, C.c_uint32]),
    "wc_scene_get_cells": (C.c_int, [_vp, C.c_uint32, _u32p, _u32p, C.c_uint32]),
    "wc_set_scene_dt": (C.c_int, [_vp, _f32p, C.c_uint32]),
    "wc_step_scenes": (C.c_int, [_vp, C.c_uint32]),
    "wc_scene_latency": (C.c_int, [_vp, _P(C.c_double), C.c_uint32]),
    "wc_get_counts": (C.c_int, [_vp, _u32p, C.c_uint32]),
    "wc_get_offsets": (C.c_int, [_vp, _u32p, C.c_uint32]),
    "wc_get_cell_keys": (C.c_int, [_vp, _u32p, C.c_uint32]),
    "wc_get_sorted": (C.c_int, [_vp, _u32p, C.c_uint32]),
    "wc_device_view_get": (C.c_int, [_vp, _P(wc_device_view)]),
    "wc_profile_steps": (C.c_int, [_vp, C.c_float, C.c_uint32, _P(C.c_double)]),
    "wc_pass_name": (C.c_char_p, [C.c_int]),
    "wc_num_variants": (C.c_int, []),
    "wc_variant_info": (C.c_int, [C.c_int, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_int32)]),
    "wc_get_neighbour_counts": (C.c_int, [_vp, _u32p, C.c_uint32]),
    "wc_get_tuning": (C.c_int, [_vp, _P(wc_tuning)]),
    "wc_set_tuning": (C.c_int, [_vp, _P(wc_tuning)]),
    "wc_tune": (C.c_int, [_vp, C.c_float, C.c_uint32, _P(wc_tuning), _P(C.c_double), _P(C.c_double),
                          _P(C.c_double), _P(C.c_double)]),
    "wc_mlp_create": (C.c_int, [C.c_uint32, C.c_uint64, _P(_vp)]),
    "wc_mlp_destroy": (C.c_int, [_vp]),
    "wc_mlp_num_params": (C.c_int, []),
    "wc_mlp_get_params": (C.c_int, [_vp, _f32p, C.c_uint32]),
    "wc_mlp_set_params": (C.c_int, [_vp, _f32p, C.c_uint32]),
    "wc_mlp_predict": (C.c_int, [_vp, _vp, _vp, C.c_uint32]),
    "wc_mlp_forward": (C.c_int, [_vp, _vp, _vp, C.c_uint32]),
    "wc_mlp_gradient": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _f32p, _P(C.c_double)]),
    "wc_mlp_train_step": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_float, _P(C.c_double)]),
    "wc_mlp_profile_predict": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, _P(C.c_double)]),
    "wc_mlp_profile_train": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_float, C.c_uint32, _P(C.c_double)]),
    "wc_device_alloc": (C.c_int, [C.c_size_t, _P(_vp)]),
    "wc_device_free": (C.c_int, [_vp]),
    "wc_device_upload": (C.c_int, [_vp, _vp, C.c_size_t]),
    "wc_device_download": (C.c_int, [_vp, _vp, C.c_size_t]),
    "wc_track_ids": (C.c_int, [_vp, C.c_int]),
    "wc_get_ids": (C.c_int, [_vp, C.c_int, _u32p, C.c_uint32]),
    "wc_max_density": (C.c_int, [_vp, _f32p]),
    "wc_closure_features": (C.c_int, [_vp, C.c_float, _u32p, C.c_uint32, _vp]),
    "wc_closure_infer": (C.c_int, [_vp, _vp, C.c_float, _vp]),
    "wc_closure_update": (C.c_int, [_vp, C.c_float, _vp, C.c_float, _vp]),
    "wc_closure_set_smoothing": (C.c_int, [_vp, C.c_int]),
    "wc_closure_set_integrator": (C.c_int, [_vp, C.c_int]),
    "wc_closure_run": (C.c_int, [_vp, _vp, C.c_float, C.c_uint32, C.c_uint32, C.c_float, _P(wc_closure_stats)]),
    "wc_last_error": (C.c_char_p, []),
    "wc_abi_version": (C.c_int, []),
    "wc_target_arch": (C.c_char_p, []),
}

_lib = None


class WaterCubeError(RuntimeError):
    def __init__(self, code: int, where: str, text: str):
        super().__init__(f"{where} failed with {code}: {text}")
        self.code = code


def lib() -> C.CDLL:
    """Load libwatercube.so (built by ``python -m watercube_amd.build``)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m watercube_amd.build` "
                "(there is no CPU fallback)"
            )
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, where: str) -> None:
    if rc != WC_OK:
        raise WaterCubeError(rc, where, lib().wc_last_error().decode(errors="replace"))


def f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


def u32ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_u32p)


def variants() -> list[tuple[int, int, int, int]]:
    """(threads per workgroup, LDS list entries per lane, slots in flight, per-plane drain) of
    each variant of the neighbour passes."""
    out = []
    for v in range(lib().wc_num_variants()):
        b, c, k, d = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().wc_variant_info(v, C.byref(b), C.byref(c), C.byref(k), C.byref(d)), "wc_variant_info")
        out.append((b.value, c.value, k.value, d.value))
    return out


def preset(which: int) -> wc_params:
    p = wc_params()
    check(lib().wc_params_preset(C.byref(p), which), "wc_params_preset")
    return p


def derive(p: wc_params) -> wc_params:
    check(lib().wc_params_derive(C.byref(p)), "wc_params_derive")
    return p
