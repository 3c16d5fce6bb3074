"""ctypes binding of libajisai_gpu.so (the C-ABI declared in include/ajisai_gpu.h).

The library is built in-tree by ``__graft_entry__.build()``. Loading fails
loudly when it is missing: there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libajisai_gpu.so")

AJG_OK, AJG_ERR_HIP, AJG_ERR_OOM, AJG_ERR_INVALID, AJG_ERR_CANCELLED, AJG_ERR_INTERNAL = range(6)
GEOM_QUAD, GEOM_CUBE, GEOM_MESH = 0, 1, 2
MAT_DIFFUSE, MAT_DISNEY = 0, 1
LIGHT_DIRECTIONAL, LIGHT_ENV = 0, 1

STATUS_NAMES = {0: "AJG_OK", 1: "AJG_ERR_HIP", 2: "AJG_ERR_OOM", 3: "AJG_ERR_INVALID",
                4: "AJG_ERR_CANCELLED", 5: "AJG_ERR_INTERNAL"}


class AjgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class MeshView(C.Structure):
    _fields_ = [("position_buffer", C.c_void_p), ("position_offset", C.c_size_t),
                ("position_stride", C.c_size_t), ("position_size", C.c_size_t),
                ("index_buffer", C.c_void_p), ("index_offset", C.c_size_t),
                ("index_stride", C.c_size_t), ("index_size", C.c_size_t)]


class AccelInfo(C.Structure):
    _fields_ = [("tri_count", C.c_uint32), ("node_count", C.c_uint32),
                ("iterations", C.c_uint32), ("depth", C.c_uint32),
                ("device_bytes", C.c_uint64), ("scene_box", C.c_float * 6)]


class Primitive(C.Structure):
    _fields_ = [("view", MeshView), ("kind", C.c_uint32), ("two_sided", C.c_uint32),
                ("material", C.c_uint32), ("reserved", C.c_uint32),
                ("normal_and_vs", C.c_void_p), ("emission", C.c_float * 3),
                ("reserved2", C.c_uint32)]


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("base_color", C.c_float * 3),
                ("subsurface", C.c_float), ("metallic", C.c_float), ("specular", C.c_float),
                ("specular_tint", C.c_float), ("roughness", C.c_float),
                ("anisotropic", C.c_float), ("sheen", C.c_float), ("sheen_tint", C.c_float),
                ("clearcoat", C.c_float), ("clearcoat_gloss", C.c_float)]


class Light(C.Structure):
    _fields_ = [("type", C.c_uint32), ("dir", C.c_float * 3), ("angle", C.c_float),
                ("radiance", C.c_float * 3), ("env_w", C.c_uint32), ("env_h", C.c_uint32),
                ("env_texels", C.c_void_p), ("rot", C.c_float), ("scaler", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("raster2camera", C.c_float * 16), ("camera2world", C.c_float * 16),
                ("resolution", C.c_float * 2), ("lens_radius", C.c_float),
                ("focal_distance", C.c_float)]


class PTArgs(C.Structure):
    _fields_ = [("spp", C.c_int32), ("tile_size", C.c_int32), ("min_bounces", C.c_int32),
                ("max_bounces", C.c_int32), ("cont_prob", C.c_float), ("use_mis", C.c_int32),
                ("seed_mode", C.c_int32), ("seed_base", C.c_uint32),
                ("env_cdf_fp64", C.c_int32), ("max_slots", C.c_int32), ("flags", C.c_uint32)]


RENDER_TIMING = 1
RENDER_COUNT = 2
RENDER_FINITE_PDF = 4
RENDER_INLINE = 8


class Film(C.Structure):
    _fields_ = [("sum", C.c_void_p), ("sum_sq", C.c_void_p), ("count", C.c_void_p),
                ("per_sample", C.c_void_p)]


class RenderStats(C.Structure):
    _fields_ = [("closest_queries", C.c_uint64), ("any_queries", C.c_uint64),
                ("paths", C.c_uint64), ("iterations", C.c_uint32), ("slots", C.c_uint32),
                ("ms_total", C.c_float), ("ms_trace_closest", C.c_float),
                ("ms_trace_any", C.c_float), ("ms_shade", C.c_float),
                ("launches_closest", C.c_uint32), ("launches_any", C.c_uint32),
                ("launches_shade", C.c_uint32), ("reserved", C.c_uint32),
                ("nodes_fetched", C.c_uint64), ("tris_tested", C.c_uint64),
                ("nodes_fetched_any", C.c_uint64), ("tris_tested_any", C.c_uint64)]


class SppmArgs(C.Structure):
    _fields_ = [("forward_max_depth", C.c_int32), ("init_radius", C.c_float),
                ("iteration_count", C.c_int32), ("photons_per_iteration", C.c_int32),
                ("photon_min_depth", C.c_int32), ("photon_max_depth", C.c_int32),
                ("photon_cont_prob", C.c_float), ("update_alpha", C.c_float),
                ("photon_seed_base", C.c_uint32), ("flags", C.c_uint32)]


class SppmStats(C.Structure):
    _fields_ = [("photons", C.c_uint64), ("deposits", C.c_uint64), ("grid_entries", C.c_uint64),
                ("visible_points", C.c_uint32), ("pad", C.c_uint32), ("ms_total", C.c_float),
                ("ms_visible_points", C.c_float), ("ms_photons", C.c_float),
                ("ms_grid", C.c_float)]


SPPM_PIXEL_DTYPE = np.dtype([("pos", "<f4", 3), ("beta", "<f4", 3), ("direct", "<f4", 3),
                             ("tau", "<f4", 3), ("radius", "<f4"), ("n", "<f4"),
                             ("valid", "<u4"), ("pad", "<u4")])


class VcmArgs(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("base_radius", C.c_float),
                ("radius_alpha", C.c_float), ("use_vm", C.c_int32), ("use_vc", C.c_int32),
                ("min_path_length", C.c_int32), ("max_path_length", C.c_int32),
                ("cont_prob", C.c_float), ("seed_base", C.c_uint32), ("flags", C.c_uint32)]


class VcmStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("light_vertices", C.c_uint64), ("ms_total", C.c_float),
                ("ms_light", C.c_float), ("ms_grid", C.c_float), ("ms_camera", C.c_float)]


class GuideArgs(C.Structure):
    _fields_ = [("alpha", C.c_float), ("train", C.c_int32), ("batch", C.c_uint32),
                ("lr", C.c_float), ("train_slots", C.c_uint32), ("max_steps", C.c_uint32),
                ("steps_per_batch", C.c_uint32), ("warmup_passes", C.c_uint32),
                ("warmup_spp", C.c_uint32), ("warmup_steps_per_batch", C.c_uint32)]


class GuideStats(C.Structure):
    _fields_ = [("train_steps", C.c_uint32), ("warmup_steps", C.c_uint32), ("records", C.c_uint64),
                ("ms_infer", C.c_float), ("ms_train", C.c_float), ("last_loss", C.c_float),
                ("ms_warmup", C.c_float)]


class Instance(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("material", C.c_uint32), ("object2world", C.c_float * 12)]


INSTANCE_KEEP_MATERIAL = 0xFFFFFFFF


class Filter(C.Structure):
    _fields_ = [("type", C.c_uint32), ("radius", C.c_float * 2), ("alpha", C.c_float)]


FILTER_BOX, FILTER_GAUSSIAN = 0, 1


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("args", PTArgs), ("film", Film)]


class BatchArgs(C.Structure):
    _fields_ = [("max_slots", C.c_int32), ("samples_per_item", C.c_int32),
                ("trav_blocks_per_cu", C.c_int32), ("lds_stack", C.c_int32),
                ("iters_per_check", C.c_int32), ("use_graph", C.c_int32), ("flags", C.c_uint32),
                ("pixel_tile", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("ms_total", C.c_float), ("iterations", C.c_uint32), ("slots", C.c_uint32),
                ("graph_launches", C.c_uint32), ("paths", C.c_uint64),
                ("closest_queries", C.c_uint64), ("any_queries", C.c_uint64)]


RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("tri_id", "<u4"), ("u", "<f4"), ("v", "<f4")])
NODE_DTYPE = np.dtype([("mn", "<f4", 3), ("mx", "<f4", 3), ("count", "<u4"), ("first", "<u4")])

_P = C.c_void_p
_SIGS = {
    "ajg_abi_version": ([], C.c_int),
    "ajg_ctx_create": ([C.c_int, C.POINTER(_P)], C.c_int),
    "ajg_ctx_destroy": ([_P], C.c_int),
    "ajg_last_error": ([_P], C.c_char_p),
    "ajg_ctx_inject_failure": ([_P, C.c_int, C.c_int], C.c_int),
    "ajg_accel_build": ([_P, C.POINTER(MeshView), C.c_uint32, C.c_uint32, C.POINTER(_P)], C.c_int),
    "ajg_accel_build_soup": ([_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.POINTER(_P)], C.c_int),
    "ajg_accel_destroy": ([_P], C.c_int),
    "ajg_accel_get_info": ([_P, C.POINTER(AccelInfo)], C.c_int),
    "ajg_accel_download": ([_P, _P, _P], C.c_int),
    "ajg_accel_trace_closest": ([_P, _P, C.c_uint32, _P, _P], C.c_int),
    "ajg_accel_trace_any": ([_P, _P, C.c_uint32, _P, _P], C.c_int),
    "ajg_accel_trace_counts": ([_P, _P, C.c_uint32, _P, _P, _P], C.c_int),
    "ajg_scene_create": ([_P, C.POINTER(Primitive), C.c_uint32, C.POINTER(Material), C.c_uint32,
                          C.POINTER(Light), C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(_P)],
                         C.c_int),
    "ajg_scene_destroy": ([_P], C.c_int),
    "ajg_scene_create_instanced": ([_P, C.POINTER(Primitive), C.c_uint32, C.POINTER(Instance),
                                    C.c_uint32, C.POINTER(Material), C.c_uint32, C.POINTER(Light),
                                    C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(_P)], C.c_int),
    "ajg_scene_instances": ([_P, _P, _P], C.c_int),
    "ajg_scene_trace_closest": ([_P, _P, C.c_uint32, _P, _P, _P], C.c_int),
    "ajg_scene_trace_any": ([_P, _P, C.c_uint32, _P, _P], C.c_int),
    "ajg_scene_get_accel": ([_P, C.POINTER(_P)], C.c_int),
    "ajg_scene_env_tables": ([_P, C.c_uint32, _P, _P, _P, _P], C.c_int),
    "ajg_render_pt": ([_P, C.POINTER(Camera), C.POINTER(PTArgs), C.POINTER(Film),
                       C.POINTER(RenderStats), _P], C.c_int),
    "ajg_render_sppm": ([_P, C.POINTER(Camera), C.POINTER(SppmArgs), _P, _P,
                         C.POINTER(SppmStats), _P], C.c_int),
    "ajg_render_vcm": ([_P, C.POINTER(Camera), C.POINTER(VcmArgs), _P, _P,
                        C.POINTER(VcmStats), _P], C.c_int),
    "ajg_film_resolve": ([_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P, _P], C.c_int),
    "ajg_film_resolve_filter": ([_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Filter), _P,
                                 _P, _P, _P], C.c_int),
    "ajg_image_srgb8": ([_P, C.c_uint32, C.c_uint32, _P, C.c_int32, _P, _P], C.c_int),
    "ajg_camera_setup": ([C.POINTER(Camera), _P, _P, _P, C.c_float, C.c_float, C.c_float,
                          C.c_float, C.c_float], C.c_int),
    "ajg_render_pt_guided": ([_P, C.POINTER(Camera), C.POINTER(PTArgs), _P, C.POINTER(GuideArgs),
                              C.POINTER(Film), C.POINTER(RenderStats), C.POINTER(GuideStats), _P],
                             C.c_int),
    "ajg_render_batch": ([_P, C.POINTER(Job), C.c_uint32, C.POINTER(BatchArgs),
                          C.POINTER(BatchStats), _P, _P], C.c_int),
    "ajg_batch_autotune": ([_P, C.POINTER(Job), C.c_uint32, C.c_uint32, C.POINTER(BatchArgs),
                            C.POINTER(C.c_float)], C.c_int),
    "ajg_guide_param_count": ([], C.c_int),
    "ajg_guide_create": ([_P, C.c_uint32, _P, C.POINTER(_P)], C.c_int),
    "ajg_guide_destroy": ([_P], C.c_int),
    "ajg_guide_get_params": ([_P, _P], C.c_int),
    "ajg_guide_set_params": ([_P, _P], C.c_int),
    "ajg_guide_forward": ([_P, _P, _P, C.c_uint32, _P, _P, _P], C.c_int),
    "ajg_guide_train_step": ([_P, _P, C.c_uint32, C.c_float, _P, _P, _P], C.c_int),
    "ajg_guide_train_step_noalloc": ([_P, _P, C.c_uint32, C.c_float, _P], C.c_int),
    "ajg_ctx_set_cancel": ([_P, C.c_int], C.c_int),
}

_lib = None


def exported_symbols():
    """Names of the C-ABI entry points this binding declares."""
    return sorted(_SIGS)


def load(path: str = LIB_PATH):
    """Load libajisai_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first "
                           "(the HIP extension is required; there is no CPU fallback)")
    # torch ships its own HIP runtime under the same soname as the system one
    # this library links; whichever is loaded first serves both, and torch
    # finds no GPU on a runtime other than its own. Load torch's first where
    # torch is installed; without it the library takes the system runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(status, ctx=None, what=""):
    if status != AJG_OK:
        msg = ""
        if ctx is not None and _lib is not None:
            raw = _lib.ajg_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise AjgError(status, f"{what}: {msg}")
