"""ctypes binding of the C ABI in include/gbr.h (libgbr.so).

The library is built in-tree by habitat-sim_amd/build.sh into
habitat_sim_amd/_lib/libgbr.so.  Loading fails loudly if it is missing; there
is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libgbr.so")

OK = 0
ERR_ARGUMENT, ERR_RANGE, ERR_NOT_FOUND, ERR_IO, ERR_SIZE = 1, 2, 3, 4, 5
ERR_DEVICE, ERR_OOM, ERR_UNSUPPORTED, ERR_STATE = 6, 7, 8, 9

FLAG_NO_TEXTURES = 0x1
STANDALONE_QUIET_LOG = 0x1
FILE_WHOLE = 0x1
FILE_GENERATE_MIPMAP = 0x2
PLANE_COLOR, PLANE_DEPTH, PLANE_LINEAR_DEPTH, PLANE_SEMANTIC, PLANE_ALL = 0x1, 0x2, 0x4, 0x8, 0xF
FORMAT_RGBA8, FORMAT_RGB8 = 0, 1
MEMORY_HOST, MEMORY_DEVICE = 0, 1
CULL_NONE, CULL_FRUSTUM, CULL_HIZ, CULL_LEARNED = 0, 1, 2, 3
LIGHT_DIRECTIONAL, LIGHT_POINT, MAX_LIGHTS = 0, 1, 64
REDWOOD_MODEL_FLOATS = 80 * 80 * 5
TIMING_STAGES = ("nodes", "cull", "setup1", "clip1", "raster1", "hiz", "setup2", "clip2",
                 "raster2", "resolve")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)


class Config(ctypes.Structure):
    _fields_ = [("tile_size", ctypes.c_int32 * 2), ("tile_count", ctypes.c_int32 * 2),
                ("flags", ctypes.c_uint32), ("max_light_count", ctypes.c_uint32),
                ("ambient_factor", ctypes.c_float), ("device", ctypes.c_int32),
                ("standalone_flags", ctypes.c_uint32), ("planes", ctypes.c_uint32),
                ("top_down", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 7)]


class SceneStatsInfo(ctypes.Structure):
    _fields_ = [("node_count", ctypes.c_uint64), ("draw_count", ctypes.c_uint64),
                ("draw_batch_count", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [("bin_size", ctypes.c_int32), ("chunk_triangles", ctypes.c_int32),
                ("bin_capacity_factor", ctypes.c_float), ("min_bin_capacity", ctypes.c_int32),
                ("cull_mode", ctypes.c_int32), ("queue_grain", ctypes.c_int32),
                ("timing", ctypes.c_int32), ("raster_mode", ctypes.c_int32),
                ("proposal_threshold", ctypes.c_float), ("lane_pixels", ctypes.c_int32),
                ("graph", ctypes.c_int32), ("proposal_history", ctypes.c_int32),
                ("heavy_entries", ctypes.c_int32)]


class AutotuneParams(ctypes.Structure):
    _fields_ = [("warmup", ctypes.c_int32), ("frames", ctypes.c_int32),
                ("passes", ctypes.c_int32)]


class AutotuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int32), ("initial_ms", ctypes.c_float),
                ("best_ms", ctypes.c_float), ("best", Tuning)]


VIS_FEATURES = 32
VIS_PARAMS = 74497
VIS_HIDDEN = 256
VIS_FEATURES_POSE, VIS_FEATURES_HISTORY = 0, 1


class VisFitParams(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("learning_rate", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("pos_weight", ctypes.c_float), ("reserved", ctypes.c_int32),
                ("seed", ctypes.c_uint64)]


class VisFitReport(ctypes.Structure):
    _fields_ = [("loss_first", ctypes.c_float), ("loss_last", ctypes.c_float),
                ("rows", ctypes.c_uint64), ("positives", ctypes.c_uint64)]


class FrameStatsInfo(ctypes.Structure):
    _fields_ = [("work_items", ctypes.c_uint64), ("triangles", ctypes.c_uint64),
                ("bin_entries", ctypes.c_uint64), ("overflowed_bins", ctypes.c_uint64),
                ("raster_items", ctypes.c_uint64), ("items_phase1", ctypes.c_uint64),
                ("items_candidates", ctypes.c_uint64), ("items_phase2", ctypes.c_uint64),
                ("triangles_setup", ctypes.c_uint64), ("vertices_setup", ctypes.c_uint64),
                ("triangles_phase2", ctypes.c_uint64), ("vertices_phase2", ctypes.c_uint64),
                ("raster_pixels", ctypes.c_uint64), ("fragments", ctypes.c_uint64),
                ("triangles_clipped", ctypes.c_uint64), ("clip_fan_triangles", ctypes.c_uint64),
                ("items_scored", ctypes.c_uint64)]


class MeshDesc(ctypes.Structure):
    _fields_ = [("vertex_count", ctypes.c_uint32), ("positions", ctypes.c_void_p),
                ("texcoords", ctypes.c_void_p), ("colors", ctypes.c_void_p),
                ("colors_u8", ctypes.c_uint32), ("texcoords_half", ctypes.c_uint32),
                ("index_count", ctypes.c_uint32), ("indices", ctypes.c_void_p),
                ("source_index_size", ctypes.c_uint32), ("normals", ctypes.c_void_p)]


class TextureDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("layers", ctypes.c_uint32), ("levels", ctypes.c_uint32),
                ("texels", ctypes.c_void_p), ("min_filter", ctypes.c_uint32),
                ("mag_filter", ctypes.c_uint32), ("wrap_s", ctypes.c_uint32),
                ("wrap_t", ctypes.c_uint32), ("compressed", ctypes.c_uint32)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("flat", ctypes.c_uint32), ("base_color", ctypes.c_float * 4),
                ("texture", ctypes.c_int32), ("layer", ctypes.c_uint32),
                ("texture_matrix", ctypes.c_float * 9)]


class NodeDesc(ctypes.Structure):
    _fields_ = [("parent", ctypes.c_int32), ("transform", ctypes.c_float * 16),
                ("name", ctypes.c_char_p)]


class MeshEntryDesc(ctypes.Structure):
    _fields_ = [("node", ctypes.c_uint32), ("mesh", ctypes.c_uint32),
                ("material", ctypes.c_int32), ("has_view", ctypes.c_uint32),
                ("view_index_offset_bytes", ctypes.c_uint32),
                ("view_index_count", ctypes.c_uint32), ("view_material", ctypes.c_int32)]


class FileDesc(ctypes.Structure):
    _fields_ = [("mesh_count", ctypes.c_uint32), ("meshes", ctypes.POINTER(MeshDesc)),
                ("texture_count", ctypes.c_uint32), ("textures", ctypes.POINTER(TextureDesc)),
                ("material_count", ctypes.c_uint32), ("materials", ctypes.POINTER(MaterialDesc)),
                ("node_count", ctypes.c_uint32), ("nodes", ctypes.POINTER(NodeDesc)),
                ("mesh_entry_count", ctypes.c_uint32),
                ("mesh_entries", ctypes.POINTER(MeshEntryDesc)),
                ("has_scene", ctypes.c_uint32)]


R = ctypes.c_void_p  # gbr_renderer*
S = ctypes.c_int32   # gbr_status

# name -> (restype, argtypes); the complete exported surface of include/gbr.h
SIGNATURES = {
    "gbr_abi_version": (ctypes.c_int32, []),
    "gbr_last_error": (ctypes.c_char_p, []),
    "gbr_config_init": (S, [ctypes.POINTER(Config)]),
    "gbr_tuning_init": (S, [ctypes.POINTER(Tuning)]),
    "gbr_create": (S, [ctypes.POINTER(Config), ctypes.POINTER(R)]),
    "gbr_destroy": (S, [R]),
    "gbr_get_config": (S, [R, ctypes.POINTER(Config)]),
    "gbr_scene_count": (S, [R, ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_add_file": (S, [R, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p]),
    "gbr_add_file_data": (S, [R, ctypes.POINTER(FileDesc), ctypes.c_uint32, ctypes.c_char_p,
                              ctypes.c_char_p]),
    "gbr_has_node_hierarchy": (S, [R, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32)]),
    "gbr_add_node_hierarchy": (S, [R, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_add_node_hierarchy_semantic": (S, [R, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_void_p,
                                            ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_add_node_hierarchies": (S, [R, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gbr_add_empty_node": (S, [R, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_add_light": (S, [R, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                          ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_clear": (S, [R, ctypes.c_uint32]),
    "gbr_set_scene_viewport": (S, [R, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32]),
    "gbr_scene_viewport": (S, [R, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]),
    "gbr_set_scene_planes": (S, [R, ctypes.c_uint32, ctypes.c_uint32]),
    "gbr_clear_lights": (S, [R, ctypes.c_uint32]),
    "gbr_light_colors": (S, [R, ctypes.c_uint32, ctypes.POINTER(c_float_p),
                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_light_ranges": (S, [R, ctypes.c_uint32, ctypes.POINTER(c_float_p),
                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_camera": (S, [R, ctypes.c_uint32, ctypes.c_void_p]),
    "gbr_camera_depth_unprojection": (S, [R, ctypes.c_uint32, ctypes.c_void_p]),
    "gbr_update_camera": (S, [R, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    "gbr_update_cameras": (S, [R, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                               ctypes.c_void_p]),
    "gbr_transformations": (S, [R, ctypes.c_uint32, ctypes.POINTER(c_float_p),
                                ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_mark_transformations_dirty": (S, [R, ctypes.c_uint32]),
    "gbr_scene_stats": (S, [R, ctypes.c_uint32, ctypes.POINTER(SceneStatsInfo)]),
    "gbr_draw": (S, [R]),
    "gbr_synchronize": (S, [R]),
    "gbr_stream": (S, [R, ctypes.POINTER(ctypes.c_void_p)]),
    "gbr_color_image_into": (S, [R, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_int32, ctypes.c_int32]),
    "gbr_depth_image_into": (S, [R, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_int32]),
    "gbr_linear_depth_image_into": (S, [R, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_int32]),
    "gbr_semantic_image_into": (S, [R, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_int32]),
    "gbr_autotune_params_init": (S, [ctypes.POINTER(AutotuneParams)]),
    "gbr_autotune": (S, [R, ctypes.POINTER(AutotuneParams), ctypes.POINTER(AutotuneReport)]),
    "gbr_color_device_pointer": (S, [R, ctypes.POINTER(ctypes.c_void_p)]),
    "gbr_depth_device_pointer": (S, [R, ctypes.POINTER(ctypes.c_void_p)]),
    "gbr_plane_device_pointer": (S, [R, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
    "gbr_calculate_depth_unprojection": (S, [ctypes.c_void_p, ctypes.c_void_p]),
    "gbr_unproject_depth": (S, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                ctypes.c_int64]),
    "gbr_get_tuning": (S, [R, ctypes.POINTER(Tuning)]),
    "gbr_set_tuning": (S, [R, ctypes.POINTER(Tuning)]),
    "gbr_frame_stats": (S, [R, ctypes.POINTER(FrameStatsInfo)]),
    "gbr_frame_timings": (S, [R, ctypes.POINTER(ctypes.c_float)]),
    "gbr_set_node_hierarchy_semantic": (S, [R, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32]),
    "gbr_vis_fit_params_init": (S, [ctypes.POINTER(VisFitParams)]),
    "gbr_vis_init": (S, [R, ctypes.c_uint64]),
    "gbr_vis_collect": (S, [R, ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_vis_clear_rows": (S, [R]),
    "gbr_vis_set_feature_set": (S, [R, ctypes.c_int32]),
    "gbr_vis_feature_set": (S, [R, ctypes.POINTER(ctypes.c_int32)]),
    "gbr_vis_rows": (S, [R, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_vis_fit": (S, [R, ctypes.POINTER(VisFitParams), ctypes.POINTER(VisFitReport)]),
    "gbr_vis_get_weights": (S, [R, ctypes.c_void_p]),
    "gbr_vis_set_weights": (S, [R, ctypes.c_void_p]),
    "gbr_vis_loss_grad": (S, [R, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_float,
                              ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]),
    "gbr_vis_predict": (S, [R, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "gbr_frame_visibility": (S, [R, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    # Redwood depth noise
    "gbr_redwood_create": (S, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64,
                               ctypes.POINTER(ctypes.c_void_p)]),
    "gbr_redwood_destroy": (S, [ctypes.c_void_p]),
    "gbr_redwood_simulate": (S, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "gbr_redwood_simulate_host": (S, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_void_p]),
    "gbr_redwood_simulate_host_noalloc": (S, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_void_p]),
    "gbr_redwood_simulate_normals": (S, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "gbr_redwood_calls": (S, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "gbr_redwood_random_bits": (S, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
}

_lib = None


class GbrError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[gbr status {status}] {message}")
        self.status = status


def lib():
    """Load libgbr.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"native library missing: {LIB_PATH}; run habitat-sim_amd/build.sh "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = lib().gbr_last_error()
        raise GbrError(status, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
