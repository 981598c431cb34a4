"""ctypes bindings of the oracle (oracle/liblrt_oracle.so): the CPU restatement
of the reference rasteriser, used only as the checker."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liblrt_oracle.so")


class OrCamera(C.Structure):
    _fields_ = [("camera_to_world", C.c_double * 16), ("orthographic", C.c_int32), ("focal", C.c_double),
                ("haperture", C.c_double), ("vaperture", C.c_double), ("near_z", C.c_double),
                ("far_z", C.c_double), ("window_translate", C.c_double * 2), ("window_scale", C.c_double * 2),
                ("window_roll", C.c_double)]


class OrProjection(C.Structure):
    _fields_ = [("world_to_view", C.c_double * 16), ("focal_x", C.c_double), ("focal_y", C.c_double),
                ("centre_x", C.c_double), ("centre_y", C.c_double), ("near_z", C.c_double),
                ("far_z", C.c_double), ("orthographic", C.c_int32), ("eye_world", C.c_double * 3)]


class OrEdit(C.Structure):
    _fields_ = [("active", C.c_uint32), ("shape", C.c_uint32), ("mode", C.c_uint32), ("invert", C.c_uint32),
                ("centre", C.c_float * 3), ("size", C.c_float * 3), ("tint", C.c_float * 3),
                ("saturation", C.c_float), ("brightness", C.c_float), ("opacity", C.c_float),
                ("min_opacity", C.c_float), ("max_scale", C.c_float)]


class OrFrame(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("tiles_x", C.c_uint32), ("tiles_y", C.c_uint32),
                ("tile_shift", C.c_uint32), ("focal_x", C.c_float), ("focal_y", C.c_float), ("centre_x", C.c_float), ("centre_y", C.c_float),
                ("near_z", C.c_float), ("far_z", C.c_float), ("lim_lo_x", C.c_float), ("lim_hi_x", C.c_float),
                ("lim_lo_y", C.c_float), ("lim_hi_y", C.c_float), ("orthographic", C.c_uint32),
                ("antialias", C.c_uint32), ("m", C.c_float * 12), ("eye", C.c_float * 3), ("count", C.c_uint32),
                ("base", C.c_uint32), ("rest_per_colour", C.c_uint32), ("sh_words", C.c_uint32),
                ("sh_limit", C.c_uint32), ("tile_bits", C.c_uint32), ("depth_mode", C.c_uint32),
                ("depth_threshold", C.c_float), ("bg", C.c_float * 4), ("linearise", C.c_uint32),
                ("dilation", C.c_float), ("use_filter3d", C.c_uint32), ("edit", OrEdit)]


class OrSettings(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("antialias", C.c_int32),
                ("max_sh_degree", C.c_uint32), ("background", C.c_float * 4), ("depth_mode", C.c_int32),
                ("depth_threshold", C.c_float), ("linearise", C.c_int32), ("dilation", C.c_float),
                ("filter3d", C.c_int32), ("tile_size", C.c_uint32)]


class OrEncoding(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("floats_per_record", "x", "y", "z", "opacity", "scale0", "scale1",
                                          "scale2", "rot_w", "rot_x", "rot_y", "rot_z", "dc0", "dc1", "dc2",
                                          "rest_base", "file_per_colour", "rest_colour_outer", "opacity_mode",
                                          "scale_mode", "colour_mode", "rotation_mode", "rest_mode",
                                          "fractional_bits", "flip_yz")]


class OrCloud(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("shape", C.c_void_p), ("sh", C.c_void_p), ("filter3d", C.c_void_p)]


class OrDiffCloud(C.Structure):
    _fields_ = [("means", C.c_void_p), ("log_scales", C.c_void_p), ("quats", C.c_void_p), ("logits", C.c_void_p),
                ("sh", C.c_void_p), ("filter3d", C.c_void_p), ("count", C.c_uint32), ("rest_per_colour", C.c_uint32)]


class OrPointStyle(C.Structure):
    _fields_ = [("size_mode", C.c_uint32), ("size", C.c_float), ("constant_colour", C.c_int32),
                ("colour", C.c_float * 3)]


class OrPoints(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("colours", C.c_void_p), ("style", OrPointStyle)]


class OrStats(C.Structure):
    _fields_ = [("splats", C.c_uint32), ("visible", C.c_uint32), ("pairs", C.c_uint32)]


P = C.c_void_p
_SIGS = {
    "or_expf": (C.c_float, [C.c_float]),
    "or_logf": (C.c_float, [C.c_float]),
    "or_half_to_float": (C.c_float, [C.c_uint16]),
    "or_float_to_half": (C.c_uint16, [C.c_float]),
    "or_encode_quaternion": (C.c_uint32, [C.POINTER(C.c_float)]),
    "or_decode_quaternion": (None, [C.c_uint32, C.POINTER(C.c_float)]),
    "or_pack_halves": (C.c_uint32, [C.c_float, C.c_float]),
    "or_decode_records": (C.c_uint32, [P, C.c_uint32, C.POINTER(OrEncoding), C.c_uint32, P, P, P, C.c_uint32]),
    "or_look_at": (None, [C.POINTER(C.c_double)] * 4),
    "or_mat_mul": (None, [C.POINTER(C.c_double)] * 3),
    "or_inverse_affine": (None, [C.POINTER(C.c_double)] * 2),
    "or_translation": (None, [C.POINTER(C.c_double)] * 2),
    "or_scaling": (None, [C.POINTER(C.c_double)] * 2),
    "or_rotation_y": (None, [C.c_double, C.POINTER(C.c_double)]),
    "or_rotation_z": (None, [C.c_double, C.POINTER(C.c_double)]),
    "or_projection_for": (None, [C.POINTER(OrCamera), C.c_uint32, C.c_uint32, C.POINTER(OrProjection)]),
    "or_bits_for": (C.c_uint32, [C.c_uint32]),
    "or_frame_for": (None, [C.POINTER(OrProjection), C.POINTER(OrSettings), C.POINTER(C.c_double),
                            C.POINTER(OrEdit), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                            C.POINTER(OrFrame)]),
    "or_project": (None, [C.POINTER(OrFrame), P, P, P, P, P, P, P, P, P, C.c_int]),
    "or_scan_exclusive": (C.c_uint64, [P, P, C.c_uint32]),
    "or_radix_sort": (None, [P, P, C.c_uint32, C.c_uint32]),
    "or_raster": (C.c_uint32, [C.POINTER(OrFrame), C.POINTER(OrCloud), C.c_uint32, P, P, P, P, P, C.c_uint32,
                               P, P, P, P, P, C.POINTER(OrStats), C.c_int]),
    "or_blend": (None, [C.POINTER(OrFrame), P, P, P, P, P, P, P, C.c_int]),
    "or_points_project": (None, [C.POINTER(OrFrame), C.POINTER(OrPointStyle), P, P, P, P, P, P, P, C.c_int]),
    "or_raster_composite": (C.c_uint32, [C.POINTER(OrFrame), C.POINTER(OrCloud), C.c_uint32, C.POINTER(OrFrame),
                                         C.POINTER(OrPoints), C.c_uint32, P, P, P, P, P, P, P, C.c_uint32, P, P, P,
                                         C.POINTER(OrStats), C.c_int]),
    "or_blend_composite": (None, [C.POINTER(OrFrame), P, P, P, P, P, P, P, C.c_int]),
    "or_reference_render_points_f32": (None, [C.POINTER(OrFrame), C.POINTER(OrCloud), C.c_uint32,
                                              C.POINTER(OrFrame), C.POINTER(OrPoints), C.c_uint32, P, P, C.c_int]),
    "or_reference_render_f32": (None, [C.POINTER(OrFrame), C.POINTER(OrCloud), C.c_uint32, P, P, C.c_uint32, P, P,
                                       C.c_int, C.c_int]),
    "or_reference_render_f64": (None, [C.POINTER(OrFrame), C.POINTER(OrCloud), C.c_uint32, P, P, C.c_uint32, P, P,
                                       C.c_int, C.c_int]),
    "or_morton30": (C.c_uint32, [C.c_float, C.c_float, C.c_float]),
    "or_lod_morton": (None, [P, C.c_uint32, C.POINTER(C.c_float), C.c_float, P, P]),
    "or_lod_groups": (C.c_uint32, [P, C.c_uint32, C.c_uint32, P, P, P]),
    "or_lod_leaf_moments": (None, [P, P, P, C.c_uint32, P, C.c_uint32, C.c_uint32, C.c_uint32, P]),
    "or_lod_merge_moments": (None, [P, C.c_uint32, P, C.c_uint32, P, C.c_uint32, P]),
    "or_lod_finalize": (None, [P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P]),
    "or_lod_projected_edge": (C.c_float, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "or_lod_cut_groups": (None, [C.c_void_p, P, C.c_uint32, C.c_uint32, C.c_int, P]),
    "or_lod_cut_splats": (None, [C.c_void_p, P, C.c_uint32, P, C.c_uint32, P]),
    "or_compare_images": (None, [P, P, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint64)]),
    "or_srgb_to_linear": (C.c_float, [C.c_float]),
    "or_diff_render_f64": (None, [C.POINTER(OrFrame), C.POINTER(OrDiffCloud), C.c_void_p, C.c_void_p, C.c_uint32,
                                  C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
}


def load(path: str = ORACLE_PATH):
    """The oracle at `path` with its signatures set (a planted-mistake build of
    the same source loads the same way)."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `make -C oracle`")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def dbl(values) -> "C.Array":
    a = [float(v) for v in np.asarray(values, dtype=np.float64).reshape(-1)]
    return (C.c_double * len(a))(*a)


def ptr(a: np.ndarray):
    return a.ctypes.data if a is not None else None


def camera(camera_to_world, focal=50.0, orthographic=False, haperture=24.576, near_z=0.1, far_z=10000.0,
           window_translate=(0.0, 0.0), window_scale=(1.0, 1.0), window_roll=0.0) -> OrCamera:
    c = OrCamera()
    c.camera_to_world[:] = [float(v) for v in np.asarray(camera_to_world, dtype=np.float64).reshape(16)]
    c.orthographic = 1 if orthographic else 0
    c.focal, c.haperture, c.vaperture, c.near_z, c.far_z = focal, haperture, 18.672, near_z, far_z
    c.window_translate[:] = list(window_translate)
    c.window_scale[:] = list(window_scale)
    c.window_roll = window_roll
    return c


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    out = (C.c_double * 16)()
    lib.or_look_at(dbl(eye), dbl(target), dbl(up), out)
    return np.array(out[:], dtype=np.float64)


def projection_for(cam: OrCamera, width: int, height: int) -> OrProjection:
    out = OrProjection()
    lib.or_projection_for(C.byref(cam), width, height, C.byref(out))
    return out


def settings(width, height, antialias=True, max_sh_degree=3, background=(0, 0, 0, 0), depth_mode=0,
             depth_threshold=0.5, linearise=True, dilation=0.3, filter3d=False, tile_size=16) -> OrSettings:
    s = OrSettings()
    s.width, s.height, s.antialias, s.max_sh_degree = width, height, 1 if antialias else 0, max_sh_degree
    s.background[:] = [float(v) for v in background]
    s.depth_mode, s.depth_threshold, s.linearise = depth_mode, depth_threshold, 1 if linearise else 0
    s.dilation, s.filter3d = dilation, 1 if filter3d else 0
    s.tile_size = tile_size
    return s


def edit(active=False, shape=0, mode=2, invert=False, centre=(0, 0, 0), size=(1, 1, 1), tint=(1, 1, 1),
         saturation=1.0, brightness=1.0, opacity=1.0, min_opacity=0.0, max_scale=0.0) -> OrEdit:
    e = OrEdit()
    e.active, e.shape, e.mode, e.invert = 1 if active else 0, shape, mode, 1 if invert else 0
    e.centre[:], e.size[:], e.tint[:] = [float(v) for v in centre], [float(v) for v in size], \
        [float(v) for v in tint]
    e.saturation, e.brightness, e.opacity = saturation, brightness, opacity
    e.min_opacity, e.max_scale = min_opacity, max_scale
    return e


def frames_for(projection: OrProjection, s: OrSettings, instances):
    """instances: list of (scene, object_to_world, OrEdit or None). Returns OrFrame array and clouds."""
    n = len(instances)
    frames = (OrFrame * max(1, n))()
    clouds = (OrCloud * max(1, n))()
    base = 0
    for k, (scene, o2w, ed) in enumerate(instances):
        lib.or_frame_for(C.byref(projection), C.byref(s), dbl(o2w), C.byref(ed) if ed is not None else None,
                         scene.count, base, scene.rest_per_colour, scene.sh_words, C.byref(frames[k]))
        clouds[k].positions = ptr(scene.positions)
        clouds[k].shape = ptr(scene.shape)
        clouds[k].sh = ptr(scene.sh) if scene.sh is not None else None
        clouds[k].filter3d = ptr(scene.filter3d) if scene.filter3d is not None else None
        base += scene.count
    return frames, clouds, n


def point_style(size_mode=0, size=0.01, constant_colour=False, colour=(1.0, 1.0, 1.0)) -> OrPointStyle:
    st = OrPointStyle()
    st.size_mode, st.size, st.constant_colour = size_mode, size, 1 if constant_colour else 0
    st.colour[:] = list(colour)
    return st


def point_frames_for(projection: OrProjection, s: OrSettings, instances, base: int):
    """instances: list of (positions float32 [N, 4], colours uint32 [N, 2] or None,
    object_to_world, OrPointStyle); their slots from `base`. Returns OrFrame and
    OrPoints arrays and the count."""
    n = len(instances)
    frames = (OrFrame * max(1, n))()
    points = (OrPoints * max(1, n))()
    for k, (positions, colours, o2w, style) in enumerate(instances):
        count = int(positions.shape[0])
        lib.or_frame_for(C.byref(projection), C.byref(s), dbl(o2w), None, count, base, 0, 1, C.byref(frames[k]))
        points[k].positions = ptr(positions)
        points[k].colours = ptr(colours) if colours is not None else None
        points[k].style = style
        base += count
    return frames, points, n


def raster_composite(frames, clouds, n, pframes, points, npts, under_colour=None, under_depth=None, threads=0):
    """The whole frame with point instances and an under-layer, restated."""
    total = sum(frames[k].count for k in range(n)) + sum(pframes[k].count for k in range(npts))
    f0 = frames[0] if n > 0 or npts == 0 else pframes[0]
    tiles = f0.tiles_x * f0.tiles_y
    proj = np.zeros((max(total, 1), 12), np.float32)
    touched = np.zeros(max(total, 1), np.uint32)
    keys = np.zeros(max(total, 1), np.uint32)
    ranges = np.zeros(tiles * 2, np.uint32)
    colour = np.zeros((f0.height, f0.width, 4), np.float32)
    depth = np.zeros((f0.height, f0.width), np.float32)
    stats = OrStats()
    uc = np.ascontiguousarray(under_colour, dtype=np.float32) if under_colour is not None else None
    ud = np.ascontiguousarray(under_depth, dtype=np.float32) if under_depth is not None else None
    pairs = lib.or_raster_composite(frames, clouds, n, pframes, points, npts, ptr(uc), ptr(ud), ptr(proj),
                                    ptr(touched), ptr(keys), None, None, 0, None, None, None, C.byref(stats), threads)
    tiles_out = np.zeros(max(pairs, 1), np.uint32)
    splats_out = np.zeros(max(pairs, 1), np.uint32)
    lib.or_raster_composite(frames, clouds, n, pframes, points, npts, ptr(uc), ptr(ud), ptr(proj), ptr(touched),
                            ptr(keys), ptr(tiles_out), ptr(splats_out), pairs, ptr(ranges), ptr(colour), ptr(depth),
                            C.byref(stats), threads)
    return dict(proj=proj[:total], touched=touched[:total], depth_keys=keys[:total], pair_tiles=tiles_out[:pairs],
                pair_splats=splats_out[:pairs], ranges=ranges.reshape(-1, 2), colour=colour, depth=depth,
                visible=stats.visible, pairs=pairs, splats=stats.splats)


def reference_points(frames, clouds, n, pframes, points, npts, threads=0):
    """The FP32 brute force with point instances as discs: whole images."""
    f0 = frames[0] if n > 0 or npts == 0 else pframes[0]
    colour = np.zeros((f0.height, f0.width, 4), np.float32)
    depth = np.zeros((f0.height, f0.width), np.float32)
    lib.or_reference_render_points_f32(frames, clouds, n, pframes, points, npts, ptr(colour), ptr(depth), threads)
    return colour, depth


def raster(frames, clouds, n, threads=0, want_image=True, using=None):
    """The reference's whole frame restated: dict of host arrays. `using`: another
    build of the oracle (a planted mistake) instead of the shipped one."""
    lib_ = using or lib
    total = sum(frames[k].count for k in range(n))
    f0 = frames[0]
    tiles = f0.tiles_x * f0.tiles_y
    px = f0.width * f0.height
    proj = np.zeros((max(total, 1), 12), np.float32)
    touched = np.zeros(max(total, 1), np.uint32)
    keys = np.zeros(max(total, 1), np.uint32)
    ranges = np.zeros(tiles * 2, np.uint32)
    colour = np.zeros((f0.height, f0.width, 4), np.float32) if want_image else None
    depth = np.zeros((f0.height, f0.width), np.float32) if want_image else None
    last = np.zeros((f0.height, f0.width), np.uint32) if want_image else None
    final_t = np.zeros((f0.height, f0.width), np.float32) if want_image else None
    stats = OrStats()
    # First pass sizes the pairs.
    pairs = lib_.or_raster(frames, clouds, n, ptr(proj), ptr(touched), ptr(keys), None, None, 0, None, None, None,
                          None, None, C.byref(stats), threads)
    tiles_out = np.zeros(max(pairs, 1), np.uint32)
    splats_out = np.zeros(max(pairs, 1), np.uint32)
    lib_.or_raster(frames, clouds, n, ptr(proj), ptr(touched), ptr(keys), ptr(tiles_out), ptr(splats_out), pairs,
                  ptr(ranges), ptr(colour), ptr(depth), ptr(last), ptr(final_t), C.byref(stats), threads)
    del px
    return dict(proj=proj[:total], touched=touched[:total], depth_keys=keys[:total], pair_tiles=tiles_out[:pairs],
                pair_splats=splats_out[:pairs], ranges=ranges.reshape(-1, 2), colour=colour, depth=depth,
                last_count=last, final_t=final_t, visible=stats.visible, pairs=pairs, splats=stats.splats)


def reference(frames, clouds, n, f64=False, pixels=None, threads=0, keyed=False):
    """The brute force; pixels: None (whole image) or (xs, ys) uint32 arrays; keyed:
    order by the rasteriser's 24-bit depth key instead of exact view z."""
    f0 = frames[0]
    dt = np.float64 if f64 else np.float32
    fn = lib.or_reference_render_f64 if f64 else lib.or_reference_render_f32
    if pixels is None:
        colour = np.zeros((f0.height, f0.width, 4), dt)
        depth = np.zeros((f0.height, f0.width), dt)
        fn(frames, clouds, n, None, None, 0, ptr(colour), ptr(depth), 1 if keyed else 0, threads)
    else:
        xs = np.ascontiguousarray(pixels[0], dtype=np.uint32)
        ys = np.ascontiguousarray(pixels[1], dtype=np.uint32)
        colour = np.zeros((xs.size, 4), dt)
        depth = np.zeros(xs.size, dt)
        fn(frames, clouds, n, ptr(xs), ptr(ys), xs.size, ptr(colour), ptr(depth), 1 if keyed else 0, threads)
    return colour, depth


def compare_images(a: np.ndarray, b: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    h, w = a.shape[0], a.shape[1]
    p99, mx, over2 = C.c_uint32(), C.c_uint32(), C.c_uint64()
    lib.or_compare_images(ptr(a), ptr(b), w, h, C.byref(p99), C.byref(mx), C.byref(over2))
    return p99.value, mx.value, over2.value


def linear_error(a: np.ndarray, b: np.ndarray):
    """p99 and max of the absolute per-channel difference in linear units."""
    d = np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)).reshape(-1)
    return float(np.quantile(d, 0.99)), float(d.max())


def diff_frame(projection: OrProjection, s: OrSettings, count: int, rest_per_colour: int,
               object_to_world=np.eye(4)) -> OrFrame:
    """The frame one FP32-parameter cloud is drawn with."""
    f = OrFrame()
    words = 1 if rest_per_colour == 0 else (rest_per_colour * 3 + 1) // 2
    lib.or_frame_for(C.byref(projection), C.byref(s), dbl(object_to_world), None, count, 0, rest_per_colour, words,
                     C.byref(f))
    return f


def diff_render_f64(frame: OrFrame, params: dict, filter3d=None, threads=0) -> np.ndarray:
    """or_diff_render_f64 of float32 parameter arrays: the FP64 image [H, W, 4]."""
    arrs = {k: np.ascontiguousarray(params[k], dtype=np.float32) for k in
            ("means", "log_scales", "quats", "opacity_logits", "sh")}
    f3 = np.ascontiguousarray(filter3d, dtype=np.float32) if filter3d is not None else None
    cloud = OrDiffCloud(ptr(arrs["means"]), ptr(arrs["log_scales"]), ptr(arrs["quats"]), ptr(arrs["opacity_logits"]),
                        ptr(arrs["sh"]), ptr(f3), arrs["means"].shape[0], arrs["sh"].shape[1] - 1)
    colour = np.zeros((frame.height, frame.width, 4), np.float64)
    lib.or_diff_render_f64(C.byref(frame), C.byref(cloud), None, None, 0, ptr(colour), None, 1, threads)
    return colour


def diff_render_f64_pixels(frame: OrFrame, params: dict, px, py, keyed=True, filter3d=None, threads=0) -> np.ndarray:
    """or_diff_render_f64 at the pixels (px[k], py[k]) only: FP64 colours [len(px), 4]."""
    arrs = {k: np.ascontiguousarray(params[k], dtype=np.float32) for k in
            ("means", "log_scales", "quats", "opacity_logits", "sh")}
    f3 = np.ascontiguousarray(filter3d, dtype=np.float32) if filter3d is not None else None
    cloud = OrDiffCloud(ptr(arrs["means"]), ptr(arrs["log_scales"]), ptr(arrs["quats"]), ptr(arrs["opacity_logits"]),
                        ptr(arrs["sh"]), ptr(f3), arrs["means"].shape[0], arrs["sh"].shape[1] - 1)
    xs = np.ascontiguousarray(px, dtype=np.uint32)
    ys = np.ascontiguousarray(py, dtype=np.uint32)
    colour = np.zeros((xs.size, 4), np.float64)
    lib.or_diff_render_f64(C.byref(frame), C.byref(cloud), ptr(xs), ptr(ys), xs.size, ptr(colour), None,
                           1 if keyed else 0, threads)
    return colour


# ---- levels of detail ---------------------------------------------------------------

class OrCutParams(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("extent", C.c_float), ("eye", C.c_float * 3), ("pixels_per_unit", C.c_float),
                ("threshold", C.c_float), ("orthographic", C.c_uint32)]


def lod_build(positions, shape, sh, rest_per_colour, max_group_fraction=0.5, coarsest_level=1):
    """LodBuilder::build (modules/lod/src/Lod.cpp:81-283) over the oracle's stages:
    bounds and extent, Morton keys and the stable 30-bit sort, the splats in
    that order, the finest level worth storing, then from it up each stored
    level's groups, moments and Gaussians. Returns a dict of host arrays."""
    n = positions.shape[0]
    assert n > 0
    keep = rest_per_colour
    words = 1 if keep == 0 else (keep * 3 + 1) // 2
    lo = positions[:, :3].min(axis=0).astype(np.float32)
    hi = positions[:, :3].max(axis=0).astype(np.float32)
    extent = np.float32(0.0)
    for k in range(3):
        extent = max(extent, np.float32(hi[k] - lo[k]))
    extent = np.float32(max(extent, np.float32(1e-6))) * np.float32(1.0001)
    keys = np.zeros(n, np.uint64)
    order = np.zeros(n, np.uint32)
    lo_c = (C.c_float * 3)(*[float(v) for v in lo])
    lib.or_lod_morton(ptr(np.ascontiguousarray(positions, np.float32)), n, lo_c, float(extent), ptr(keys), ptr(order))
    lib.or_radix_sort(ptr(keys), ptr(order), n, 30)
    pos = np.ascontiguousarray(positions[order], np.float32)
    shp = np.ascontiguousarray(shape[order], np.uint32)
    shs = np.ascontiguousarray(sh[order], np.uint32) if keep else np.zeros((n, 1), np.uint32)

    def groups_at(r):
        group = np.zeros(n, np.uint32)
        starts = np.zeros(n, np.uint32)
        cells = np.zeros(n, np.uint32)
        g = lib.or_lod_groups(ptr(keys), n, r, ptr(group), ptr(starts), ptr(cells))
        return group, starts[:g].copy(), cells[:g].copy()

    coarsest = min(max(coarsest_level, 1), 10)
    finest = coarsest
    levels = {}
    for r in range(coarsest, 11):
        levels[r] = groups_at(r)
        if np.float32(len(levels[r][1])) > np.float32(max_group_fraction) * np.float32(n) and r > coarsest:
            del levels[r]
            break
        finest = r
    stride = 13 + keep * 3
    stored = []
    fine_moments = None
    for r in range(finest, coarsest - 1, -1):
        group, starts, cells = levels[r]
        gcount = len(starts)
        moments = np.zeros((gcount, stride), np.float32)
        if r == finest:
            lib.or_lod_leaf_moments(ptr(pos), ptr(shp), ptr(shs), n, ptr(starts), gcount, keep, words, ptr(moments))
        else:
            fine_group = levels[r + 1][0]
            lib.or_lod_merge_moments(ptr(starts), gcount, ptr(fine_group), len(levels[r + 1][1]), ptr(fine_moments),
                                     stride, ptr(moments))
        gp = np.zeros((gcount, 4), np.float32)
        gs = np.zeros((gcount, 4), np.uint32)
        gh = np.zeros((gcount, words), np.uint32)
        lib.or_lod_finalize(ptr(moments), gcount, keep, words, ptr(gp), ptr(gs), ptr(gh))
        stored.append(dict(level=r, positions=gp, shape=gs, sh=gh if keep else None, cells=cells, starts=starts,
                           group=group, moments=moments))
        fine_moments = moments
    stored.reverse()
    return dict(count=n, keep=keep, sh_words=words, lo=lo, extent=np.float32(extent), keys=keys, order=order,
                positions=pos, shape=shp, sh=shs if keep else None, levels=stored,
                groups=stored[-1]["group"])


def lod_cut(lod, eye, pixels_per_unit, threshold, orthographic=False):
    """CutSelector::select for one instance held in memory (Lod.cpp:350-547):
    every level's selected Gaussians (coarsest first), then the selected splats
    in Morton order. Returns (positions, shape, sh, merged drawn, splats drawn)."""
    p = OrCutParams()
    p.lo[:] = [float(v) for v in lod["lo"]]
    p.extent = float(lod["extent"])
    p.eye[:] = [float(v) for v in eye]
    p.pixels_per_unit = float(pixels_per_unit)
    p.threshold = float(threshold)
    p.orthographic = 1 if orthographic else 0
    parts = []
    for k, level in enumerate(lod["levels"]):
        sel = np.zeros(len(level["cells"]), np.uint32)
        lib.or_lod_cut_groups(C.byref(p), ptr(level["cells"]), len(level["cells"]), level["level"], 1 if k == 0 else 0,
                              ptr(sel))
        parts.append((sel.astype(bool), level["positions"], level["shape"], level["sh"]))
    finest = lod["levels"][-1]
    sel = np.zeros(lod["count"], np.uint32)
    lib.or_lod_cut_splats(C.byref(p), ptr(lod["groups"]), lod["count"], ptr(finest["cells"]), finest["level"], ptr(sel))
    parts.append((sel.astype(bool), lod["positions"], lod["shape"], lod["sh"]))
    positions = np.concatenate([pp[s] for s, pp, _, _ in parts])
    shape = np.concatenate([ss[s] for s, _, ss, _ in parts])
    sh = np.concatenate([hh[s] for s, _, _, hh in parts]) if lod["keep"] else None
    merged = int(sum(s.sum() for s, _, _, _ in parts[:-1]))
    return positions, shape, sh, merged, int(parts[-1][0].sum())
