"""ctypes mirror of include/lrt_hip.h and include/lrt_synth.h.

Loads ``liblrt_hip.so`` from this package directory. There is no fallback: a
missing or stale library raises ImportError with the build command.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblrt_hip.so")

LRT_OK = 0
STATUS_NAMES = {
    0: "Ok", 1: "Unknown", 2: "NotFound", 3: "InvalidArgument", 4: "Unsupported", 5: "IoFailure",
    6: "DeviceFailure", 7: "ShaderFailure", 8: "PluginFailure", 9: "OutOfMemory", 10: "Cancelled",
    11: "InternalError",
}

DEBUG_COUNTS, DEBUG_TOUCHED, DEBUG_OFFSETS, DEBUG_PROJ, DEBUG_KEYS, DEBUG_VALUES, DEBUG_RANGES, \
    DEBUG_LAST_COUNT, DEBUG_FINAL_T, DEBUG_RECTS, DEBUG_KERNEL_MS, DEBUG_TILE_ORDER, DEBUG_WALK_STEPS = range(13)


class Lens(C.Structure):
    _fields_ = [
        ("orthographic", C.c_int32),
        ("focal", C.c_double),
        ("haperture", C.c_double),
        ("vaperture", C.c_double),
        ("near_z", C.c_double),
        ("far_z", C.c_double),
        ("window_translate", C.c_double * 2),
        ("window_scale", C.c_double * 2),
        ("window_roll", C.c_double),
    ]


class CameraC(C.Structure):
    _fields_ = [("camera_to_world", C.c_double * 16), ("lens", Lens)]


class ProjectionC(C.Structure):
    _fields_ = [
        ("world_to_view", C.c_double * 16),
        ("focal_x", C.c_double), ("focal_y", C.c_double),
        ("centre_x", C.c_double), ("centre_y", C.c_double),
        ("near_z", C.c_double), ("far_z", C.c_double),
        ("orthographic", C.c_int32),
        ("eye_world", C.c_double * 3),
    ]


class Splats(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("shape", C.c_void_p),
        ("sh", C.c_void_p),
        ("filter3d", C.c_void_p),
        ("count", C.c_uint32),
        ("rest_per_colour", C.c_uint32),
        ("sh_words", C.c_uint32),
    ]


class SplatEdit(C.Structure):
    _fields_ = [
        ("active", C.c_int32), ("shape", C.c_uint32), ("mode", C.c_uint32), ("invert", C.c_int32),
        ("centre", C.c_float * 3), ("size", C.c_float * 3), ("tint", C.c_float * 3),
        ("saturation", C.c_float), ("brightness", C.c_float), ("opacity", C.c_float),
        ("min_opacity", C.c_float), ("max_scale", C.c_float),
    ]


class SplatInstanceC(C.Structure):
    _fields_ = [("splats", C.POINTER(Splats)), ("object_to_world", C.c_double * 16), ("edit", SplatEdit)]


class PointsC(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("colours", C.c_void_p), ("count", C.c_uint32)]


class PointStyleC(C.Structure):
    _fields_ = [("size_mode", C.c_uint32), ("size", C.c_float), ("constant_colour", C.c_int32),
                ("colour", C.c_float * 3)]


class PointInstanceC(C.Structure):
    _fields_ = [("points", C.POINTER(PointsC)), ("object_to_world", C.c_double * 16), ("style", PointStyleC)]


class RenderSettingsC(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("antialias", C.c_int32), ("max_sh_degree", C.c_uint32),
        ("background", C.c_float * 4),
        ("depth_mode", C.c_int32), ("depth_threshold", C.c_float),
        ("linearise", C.c_int32), ("time_stages", C.c_int32),
        ("dilation", C.c_float), ("filter3d", C.c_int32), ("blend_batch", C.c_uint32), ("tile_order", C.c_int32),
        ("persistent_blend", C.c_int32), ("use_graph", C.c_int32), ("blend_segment", C.c_uint32),
        ("keep_walk_state", C.c_int32), ("tile_size", C.c_uint32), ("radix_bits", C.c_uint32),
    ]


class RenderTargetsC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("colour", C.c_void_p), ("depth", C.c_void_p)]


class ViewC(C.Structure):
    _fields_ = [("projection", ProjectionC), ("targets", RenderTargetsC)]


class FrameStatsC(C.Structure):
    _fields_ = [
        ("splats", C.c_uint32), ("visible", C.c_uint32), ("pairs", C.c_uint32),
        ("project_ms", C.c_double), ("depth_sort_ms", C.c_double), ("counts_ms", C.c_double),
        ("emit_ms", C.c_double), ("tile_sort_ms", C.c_double), ("blend_ms", C.c_double),
        ("total_ms", C.c_double), ("key_bits", C.c_uint32), ("sort_passes", C.c_uint32),
    ]


class DiffSplatsC(C.Structure):
    _fields_ = [
        ("means", C.c_void_p), ("log_scales", C.c_void_p), ("quats", C.c_void_p),
        ("opacity_logits", C.c_void_p), ("sh", C.c_void_p), ("filter3d", C.c_void_p),
        ("count", C.c_uint32), ("rest_per_colour", C.c_uint32),
    ]


class DiffGradsC(C.Structure):
    _fields_ = [
        ("means", C.c_void_p), ("log_scales", C.c_void_p), ("quats", C.c_void_p),
        ("opacity_logits", C.c_void_p), ("sh", C.c_void_p),
    ]


class LodBuildSettingsC(C.Structure):
    _fields_ = [("max_group_fraction", C.c_float), ("coarsest_level", C.c_uint32)]


class LodInfoC(C.Structure):
    _fields_ = [("count", C.c_uint32), ("rest_per_colour", C.c_uint32), ("sh_words", C.c_uint32),
                ("levels", C.c_uint32), ("merged", C.c_uint32), ("level", C.c_uint32 * 10),
                ("first", C.c_uint32 * 10), ("groups", C.c_uint32 * 10), ("bounds_lo", C.c_float * 3),
                ("extent", C.c_float)]


class LodInstanceC(C.Structure):
    _fields_ = [("cloud", C.c_void_p), ("object_to_world", C.c_double * 16), ("edit", SplatEdit),
                ("has_threshold", C.c_int32), ("threshold", C.c_float)]


class CutStatsC(C.Structure):
    _fields_ = [("splats", C.c_uint32), ("merged", C.c_uint32), ("available", C.c_uint32)]


class SceneSpecC(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32), ("degree", C.c_uint32), ("seed", C.c_uint64),
        ("offset", C.c_float * 3),
        ("log_scale_mean", C.c_float), ("log_scale_sd", C.c_float),
        ("opacity_logit_mean", C.c_float), ("opacity_logit_sd", C.c_float),
        ("sh_sd", C.c_float), ("clusters", C.c_uint32),
        ("box_half", C.c_float), ("cluster_sd_lo", C.c_float), ("cluster_sd_hi", C.c_float),
        ("near_fraction", C.c_float), ("near_log_scale_mean", C.c_float), ("near_log_scale_sd", C.c_float),
    ]


# Every symbol include/lrt_hip.h and include/lrt_synth.h declare, with its signature.
ENCODING_FIELDS = ("floats_per_record", "x", "y", "z", "opacity", "scale0", "scale1", "scale2", "rot_w", "rot_x",
                   "rot_y", "rot_z", "dc0", "dc1", "dc2", "rest_base", "rest_per_colour", "file_per_colour",
                   "rest_colour_outer", "opacity_mode", "scale_mode", "colour_mode", "rotation_mode", "rest_mode",
                   "fractional_bits", "flip_yz")


class SplatEncodingC(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ENCODING_FIELDS]


class CloudInfoC(C.Structure):
    _fields_ = [("declared", C.c_uint32), ("count", C.c_uint32), ("rest_per_colour", C.c_uint32),
                ("sh_words", C.c_uint32), ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3),
                ("upload_ms", C.c_float), ("decode_ms", C.c_float)]


SIGNATURES = {
    "lrt_lens_default": (None, [C.POINTER(Lens)]),
    "lrt_camera_looking_at": (None, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(CameraC)]),
    "lrt_projection_for": (C.c_int, [C.POINTER(CameraC), C.c_uint32, C.c_uint32, C.POINTER(ProjectionC)]),
    "lrt_render_settings_default": (None, [C.POINTER(RenderSettingsC)]),
    "lrt_splat_edit_default": (None, [C.POINTER(SplatEdit)]),
    "lrt_context_create": (C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "lrt_context_destroy": (None, [C.c_void_p]),
    "lrt_context_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lrt_last_error": (C.c_char_p, [C.c_void_p]),
    "lrt_raster_render": (C.c_int, [C.c_void_p, C.POINTER(ProjectionC), C.POINTER(SplatInstanceC), C.c_uint32,
                                    C.POINTER(RenderSettingsC), C.POINTER(RenderTargetsC), C.POINTER(FrameStatsC)]),
    "lrt_point_style_default": (None, [C.POINTER(PointStyleC)]),
    "lrt_raster_render_composite": (C.c_int, [C.c_void_p, C.POINTER(ProjectionC), C.POINTER(SplatInstanceC),
                                              C.c_uint32, C.POINTER(PointInstanceC), C.c_uint32,
                                              C.POINTER(RenderTargetsC), C.POINTER(RenderSettingsC),
                                              C.POINTER(RenderTargetsC), C.POINTER(FrameStatsC)]),
    "lrt_reference_render": (C.c_int, [C.c_void_p, C.POINTER(ProjectionC), C.POINTER(SplatInstanceC), C.c_uint32,
                                       C.POINTER(RenderSettingsC), C.c_void_p, C.c_void_p, C.c_int32]),
    "lrt_reference_render_pixels": (C.c_int, [C.c_void_p, C.POINTER(ProjectionC), C.POINTER(SplatInstanceC),
                                              C.c_uint32, C.POINTER(RenderSettingsC), C.c_void_p, C.c_uint32,
                                              C.c_void_p, C.c_void_p, C.c_int32]),
    "lrt_raster_render_views": (C.c_int, [C.c_void_p, C.POINTER(ViewC), C.c_uint32, C.POINTER(SplatInstanceC),
                                          C.c_uint32, C.POINTER(RenderSettingsC), C.POINTER(FrameStatsC)]),
    "lrt_autotune_views": (C.c_int, [C.c_void_p, C.POINTER(ViewC), C.c_uint32, C.POINTER(SplatInstanceC), C.c_uint32,
                                     C.POINTER(RenderSettingsC), C.POINTER(C.c_float)]),
    "lrt_raster_render_diff": (C.c_int, [C.c_void_p, C.POINTER(ProjectionC), C.POINTER(DiffSplatsC),
                                         C.POINTER(C.c_double), C.POINTER(RenderSettingsC),
                                         C.POINTER(RenderTargetsC), C.POINTER(FrameStatsC)]),
    "lrt_raster_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DiffGradsC), C.POINTER(C.c_float)]),
    "lrt_raster_render_diff_views": (C.c_int, [C.c_void_p, C.POINTER(ViewC), C.c_uint32, C.POINTER(DiffSplatsC),
                                               C.POINTER(C.c_double), C.POINTER(RenderSettingsC),
                                               C.POINTER(FrameStatsC)]),
    "lrt_raster_backward_views": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(DiffGradsC),
                                            C.POINTER(C.c_float)]),
    "lrt_raster_debug": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "lrt_frame_params": (C.c_int, [C.POINTER(ProjectionC), C.POINTER(SplatInstanceC), C.POINTER(RenderSettingsC),
                                   C.c_uint32, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "lrt_prefix_sum": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "lrt_prefix_sum_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "lrt_radix_sort": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                 C.c_uint32, C.POINTER(C.c_int32)]),
    "lrt_radix_sort_digits": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.POINTER(C.c_int32)]),
    "lrt_context_synchronize": (C.c_int, [C.c_void_p]),
    "lrt_copy_to_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "lrt_lod_build_settings_default": (None, [C.POINTER(LodBuildSettingsC)]),
    "lrt_lod_build": (C.c_int, [C.c_void_p, C.POINTER(Splats), C.POINTER(LodBuildSettingsC), C.POINTER(C.c_void_p)]),
    "lrt_lod_destroy": (None, [C.c_void_p]),
    "lrt_lod_get_info": (C.c_int, [C.c_void_p, C.POINTER(LodInfoC)]),
    "lrt_lod_splats": (C.c_int, [C.c_void_p, C.POINTER(Splats), C.POINTER(C.c_void_p)]),
    "lrt_lod_gaussians": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Splats), C.POINTER(C.c_void_p)]),
    "lrt_lod_select": (C.c_int, [C.c_void_p, C.POINTER(ProjectionC), C.POINTER(LodInstanceC), C.c_uint32, C.c_float,
                                 C.POINTER(SplatInstanceC), C.POINTER(C.c_uint32), C.POINTER(CutStatsC)]),
    "lrt_write_exr": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int32]),
    "lrt_splats_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "lrt_raw_splats_create": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(SplatEncodingC), C.POINTER(C.c_void_p)]),
    "lrt_raw_splats_destroy": (None, [C.c_void_p]),
    "lrt_raw_splats_count": (C.c_uint32, [C.c_void_p]),
    "lrt_raw_splats_encoding": (C.c_int, [C.c_void_p, C.POINTER(SplatEncodingC)]),
    "lrt_raw_splats_records": (C.POINTER(C.c_float), [C.c_void_p]),
    "lrt_cloud_sh_words": (C.c_uint32, [C.c_void_p, C.c_uint32]),
    "lrt_cloud_decode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(CloudInfoC)]),
    "lrt_synth_scene": (C.c_int, [C.POINTER(SceneSpecC), C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lrt_pack_splats": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
}


def _load() -> C.CDLL:
    # torch ships its own libamdhip64.so.7. Loading it first makes the library's
    # DT_NEEDED resolve to that same runtime, so torch's streams and allocations
    # and the library's kernels live in one HIP runtime, not two.
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or `make -C lucabrtrender_amd/csrc`) first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class LrtError(RuntimeError):
    """A failed lrt_* call: `code` is the lrt_status (the reference's ErrorCode)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code
        self.name = STATUS_NAMES.get(code, str(code))
