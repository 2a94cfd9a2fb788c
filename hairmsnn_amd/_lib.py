"""ctypes bindings to libhmsnn.so (the C ABI declared in include/hmsnn.h).

The product path has no CPU fallback: if the HIP library is missing or a GPU
is absent, calls raise HmsnnError.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libhmsnn.so"

HMSNN_OK = 0
HMSNN_EINVAL = 1
HMSNN_EIO = 2
STATUS_NAMES = {1: "EINVAL", 2: "EIO", 3: "EDEVICE", 4: "ENOMEM", 5: "EUNSUPPORTED"}


class HmsnnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{STATUS_NAMES.get(code, code)}] {msg}")
        self.code = code


class HairParams(C.Structure):
    _fields_ = [("beta_m", C.c_float), ("beta_n", C.c_float), ("alpha", C.c_float), ("gain", C.c_float * 4)]


class BsdfQuery(C.Structure):
    _fields_ = [("wo", C.c_float * 3), ("wi", C.c_float * 3), ("h", C.c_float), ("sigma_a", C.c_float * 3)]


class Info(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("path_v1", C.c_int), ("path_v2", C.c_int),
        ("num_segments", C.c_int), ("num_strands", C.c_int), ("bvh_nodes", C.c_int), ("bvh_depth", C.c_int),
        ("env_width", C.c_int), ("env_height", C.c_int),
        ("num_dir_lights", C.c_int), ("has_env", C.c_int), ("has_sphere", C.c_int), ("mis", C.c_int),
        ("env_pdf", C.c_int),
        ("scene_scale", C.c_float), ("sigma_a", C.c_float * 3), ("beta_m", C.c_float), ("beta_n", C.c_float),
        ("alpha", C.c_float), ("gain", C.c_float * 4), ("env_scale", C.c_float), ("env_rot_phi", C.c_float),
        ("sphere", C.c_float * 4), ("sphere_albedo", C.c_float * 3), ("sphere_alpha", C.c_float),
        ("camera", C.c_float * 12), ("dir_from", (C.c_float * 3) * 8), ("dir_emit", (C.c_float * 3) * 8),
        ("num_triangles", C.c_int), ("num_curves", C.c_int),
    ]


class MsnnStats(C.Structure):
    _fields_ = [
        ("gbuffer_ms", C.c_double), ("train_ms", C.c_double), ("infer_ms", C.c_double), ("render_ms", C.c_double),
        ("total_ms", C.c_double), ("loss", C.c_double), ("frames", C.c_int), ("train_steps", C.c_int),
        ("setup_ms", C.c_double), ("warmup_ms", C.c_double),
    ]


class NrcStats(C.Structure):
    _fields_ = [
        ("gbuffer_ms", C.c_double), ("infer_ms", C.c_double), ("render_ms", C.c_double), ("train_ms", C.c_double),
        ("total_ms", C.c_double), ("frames", C.c_int), ("train_steps", C.c_int), ("cache_queries", C.c_int),
        ("train_records_used", C.c_int),
    ]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double

# name -> (restype, argtypes)
_SIGS = {
    "hmsnn_last_error": (C.c_char_p, []),
    "hmsnn_version": (_I, []),
    "hmsnn_device_count": (_I, [C.POINTER(_I), C.c_char_p, _I]),
    "hmsnn_dev_alloc": (_I, [C.c_size_t, C.POINTER(_P)]),
    "hmsnn_dev_free": (_I, [_P]),
    "hmsnn_memcpy_h2d": (_I, [_P, _P, C.c_size_t]),
    "hmsnn_memcpy_d2h": (_I, [_P, _P, C.c_size_t]),
    "hmsnn_dev_synchronize": (_I, []),
    "hmsnn_hair_bsdf_eval_dev": (_I, [C.POINTER(HairParams), _P, _I, _P, _P, _P]),
    "hmsnn_hair_bsdf_sample_dev": (_I, [C.POINTER(HairParams), _P, _P, _I, _P, _P, _P, _P, _P]),
    "hmsnn_parse_scene": (_I, [C.c_char_p, C.POINTER(Info)]),
    "hmsnn_get_string": (_I, [_P, _I, C.c_char_p, _I]),
    "hmsnn_generate_host": (_I, [C.c_char_p, _P, C.c_size_t, C.POINTER(_I), _P, C.c_size_t, C.POINTER(_I),
                                 C.POINTER(_I)]),
    "hmsnn_create": (_I, [C.c_char_p, C.POINTER(_P)]),
    "hmsnn_create_from_json": (_I, [C.c_char_p, C.c_char_p, C.POINTER(_P)]),
    "hmsnn_destroy": (_I, [_P]),
    "hmsnn_get_info": (_I, [_P, C.POINTER(Info)]),
    "hmsnn_export_segments": (_I, [_P, _P]),
    "hmsnn_export_curves": (_I, [_P, _P]),
    "hmsnn_generate_curves_host": (_I, [C.c_char_p, _P, C.c_size_t, C.POINTER(_I)]),
    "hmsnn_export_mesh": (_I, [_P, _P, _P, _P]),
    "hmsnn_obj_load": (_I, [C.c_char_p, C.POINTER(_I), _P, _P, _P, _I]),
    "hmsnn_export_env": (_I, [_P, _P, _P, _P, _P, _P]),
    "hmsnn_set_strand_sigma_a": (_I, [_P, _P, _I]),
    "hmsnn_set_seed_offset": (_I, [_P, _I]),
    "hmsnn_set_pass_concurrency": (_I, [_P, _I]),
    "hmsnn_set_pt_schedule": (_I, [_P, _I]),
    "hmsnn_pt_reset": (_I, [_P]),
    "hmsnn_pt_passes": (_I, [_P, _I]),
    "hmsnn_ctx_synchronize": (_I, [_P]),
    "hmsnn_pt_accum_count": (_I, [_P, C.POINTER(_I)]),
    "hmsnn_get_image": (_I, [_P, _I, _P]),
    "hmsnn_render_path_tracing": (_I, [_P, _I, _P]),
    "hmsnn_get_last_pass": (_I, [_P, _P]),
    "hmsnn_pt_trace_stats": (_I, [_P, _I, C.POINTER(_D)]),
    "hmsnn_pt_time_passes": (_I, [_P, _I, C.POINTER(_D)]),
    "hmsnn_trace_rays_dev": (_I, [_P, _P, _I, _I, _P, _P, _P]),
    "hmsnn_msnn_begin": (_I, [_P, _I]),
    "hmsnn_msnn_frames": (_I, [_P, _I, C.POINTER(MsnnStats)]),
    "hmsnn_msnn_get_image": (_I, [_P, _I, _P]),
    "hmsnn_msnn_save_weights": (_I, [_P, C.c_char_p]),
    "hmsnn_msnn_get_sizes": (_I, [_P, _P]),
    "hmsnn_msnn_get_buffer": (_I, [_P, _I, _P, C.c_size_t]),
    "hmsnn_render_hair_msnn": (_I, [_P, _I, _I, _P, C.POINTER(MsnnStats)]),
    "hmsnn_nrc_begin": (_I, [_P]),
    "hmsnn_nrc_set_params": (_I, [_P, _F, _I]),
    "hmsnn_nrc_frames": (_I, [_P, _I, C.POINTER(NrcStats)]),
    "hmsnn_nrc_get_image": (_I, [_P, _P]),
    "hmsnn_nrc_get_sizes": (_I, [_P, _P]),
    "hmsnn_nrc_get_buffer": (_I, [_P, _I, _P, C.c_size_t]),
    "hmsnn_nrc_get_mlp_weights": (_I, [_P, _P, _I]),
    "hmsnn_nrc_set_mlp_weights": (_I, [_P, _P, _I]),
    "hmsnn_nrc_save_weights": (_I, [_P, C.c_char_p]),
    "hmsnn_render_nrc": (_I, [_P, _I, _P, C.POINTER(NrcStats)]),
    "hmsnn_mlp_create": (_I, [_I, C.c_uint64, C.POINTER(_P)]),
    "hmsnn_mlp_destroy": (_I, [_P]),
    "hmsnn_mlp_get_weights": (_I, [_P, _P, _I]),
    "hmsnn_mlp_set_weights": (_I, [_P, _P, _I]),
    "hmsnn_mlp_save": (_I, [_P, C.c_char_p]),
    "hmsnn_mlp_load": (_I, [_P, C.c_char_p]),
    "hmsnn_mlp_encode_dev": (_I, [_P, _I, _P, _P]),
    "hmsnn_mlp_forward_dev": (_I, [_P, _P, _I, _P, _P]),
    "hmsnn_mlp_get_fp8_scales": (_I, [_P, _P]),
    "hmsnn_mlp_backward_dev": (_I, [_P, _P, _P, _I, _I, _P, C.POINTER(_F), _P]),
    "hmsnn_mlp_train_step_dev": (_I, [_P, _P, _P, _I, _I, _F, C.POINTER(_F), _P]),
    "hmsnn_mlp_train_step_dev_noalloc": (_I, [_P, _P, _P, _I, _I, _F, _P, _P]),
    "hmsnn_save_exr": (_I, [C.c_char_p, _P, _I, _I]),
    "hmsnn_exr_size": (_I, [C.c_char_p, C.POINTER(_I), C.POINTER(_I)]),
    "hmsnn_load_exr": (_I, [C.c_char_p, _P, _I, _I]),
    "hmsnn_set_hair_params": (_I, [_P, _F, _F, _F]),
    "hmsnn_adjoint_render": (_I, [_P, _I, _P, _P, C.POINTER(_D), _P]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libhmsnn.so from the package directory; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("HMSNN_LIB", LIB_PATH))
    if not path.exists():
        raise HmsnnError(3, f"HIP library not built: {path} is missing (run __graft_entry__.build())")
    lib = C.CDLL(str(path))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != HMSNN_OK:
        msg = load().hmsnn_last_error().decode(errors="replace")
        raise HmsnnError(rc, msg)


def declared_symbols() -> list[str]:
    return sorted(_SIGS)
