"""Python mirror of the reference's rasteriser interface over the C ABI.

Names and argument meanings follow lucabRTrender's lrt::render
(modules/render/include/lrt/render/{Camera,TileRasterizer}.h): ``Camera.looking_at``,
``Lens``, ``projection_for``, ``RenderSettings``, ``SplatInstance``,
``TileRasterizer.render`` returning ``FrameStats``, and errors carrying the
reference's ErrorCode names. Device memory and streams come from torch; every
computation runs in liblrt_hip.so.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._abi import LrtError, lib

IDENTITY = tuple(float(v) for v in np.eye(4).reshape(-1))


def _mat(m) -> "C.Array":
    a = np.asarray(m, dtype=np.float64).reshape(16)
    return (C.c_double * 16)(*a.tolist())


@dataclass
class Lens:
    """render::Lens (Camera.h:23-44); defaults are the reference's."""
    orthographic: bool = False
    focal: float = 50.0
    haperture: float = 24.576
    vaperture: float = 18.672
    near_z: float = 0.1
    far_z: float = 10000.0
    window_translate: tuple = (0.0, 0.0)
    window_scale: tuple = (1.0, 1.0)
    window_roll: float = 0.0

    def to_c(self) -> _abi.Lens:
        out = _abi.Lens()
        out.orthographic = 1 if self.orthographic else 0
        out.focal, out.haperture, out.vaperture = self.focal, self.haperture, self.vaperture
        out.near_z, out.far_z = self.near_z, self.far_z
        out.window_translate[:] = list(self.window_translate)
        out.window_scale[:] = list(self.window_scale)
        out.window_roll = self.window_roll
        return out


@dataclass
class Camera:
    """render::Camera (Camera.h:46-53)."""
    camera_to_world: tuple = IDENTITY
    lens: Lens = field(default_factory=Lens)

    @staticmethod
    def looking_at(eye, target, up=(0.0, 1.0, 0.0)) -> "Camera":
        c = _abi.CameraC()
        lib.lrt_camera_looking_at((C.c_double * 3)(*eye), (C.c_double * 3)(*target), (C.c_double * 3)(*up),
                                  C.byref(c))
        return Camera(tuple(c.camera_to_world), Lens())

    def to_c(self) -> _abi.CameraC:
        c = _abi.CameraC()
        c.camera_to_world[:] = list(np.asarray(self.camera_to_world, dtype=np.float64).reshape(16))
        c.lens = self.lens.to_c()
        return c


def projection_for(camera: Camera, width: int, height: int) -> _abi.ProjectionC:
    """projectionFor (modules/render/src/Camera.cpp:20-48)."""
    out = _abi.ProjectionC()
    code = lib.lrt_projection_for(C.byref(camera.to_c()), width, height, C.byref(out))
    if code != _abi.LRT_OK:
        raise LrtError(code, "projection_for")
    return out


@dataclass
class SplatEdit:
    """render::SplatEdit (TileRasterizer.h:29-50)."""
    active: bool = False
    shape: int = 0
    mode: int = 2
    centre: tuple = (0.0, 0.0, 0.0)
    size: tuple = (1.0, 1.0, 1.0)
    tint: tuple = (1.0, 1.0, 1.0)
    saturation: float = 1.0
    brightness: float = 1.0
    opacity: float = 1.0
    min_opacity: float = 0.0
    max_scale: float = 0.0
    invert: bool = False

    def to_c(self) -> _abi.SplatEdit:
        e = _abi.SplatEdit()
        e.active = 1 if self.active else 0
        e.shape, e.mode, e.invert = self.shape, self.mode, 1 if self.invert else 0
        e.centre[:], e.size[:], e.tint[:] = list(self.centre), list(self.size), list(self.tint)
        e.saturation, e.brightness, e.opacity = self.saturation, self.brightness, self.opacity
        e.min_opacity, e.max_scale = self.min_opacity, self.max_scale
        return e


@dataclass
class RenderSettings:
    """render::RenderSettings (TileRasterizer.h:105-118) + dilation / filter3d."""
    width: int = 1920
    height: int = 1080
    antialias: bool = True
    max_sh_degree: int = 3
    background: tuple = (0.0, 0.0, 0.0, 0.0)
    depth_mode: int = 0
    depth_threshold: float = 0.5
    linearise: bool = True
    time_stages: bool = False
    dilation: float = 0.3
    filter3d: bool = False
    blend_batch: int = 256          # records each 8x8 quadrant's LDS list holds: 64, 128 or 256
    tile_order: int = 1             # 0 tile index order, 1 heaviest tiles first (the default: shorter tail)
    persistent_blend: bool = False  # batches: a persistent blend taking quadrants from a queue
    use_graph: bool = False         # batches: capture the launch sequence once and replay it
    blend_segment: int = 0          # a tile with more pairs is split across waves and composited (0: never)
    keep_walk_state: bool = False   # keep each pixel's last contributor and final transmittance (debug)
    tile_size: int = 16             # pixels a square tile's side: 8, 16 or 32 (differentiable frames: 16)
    radix_bits: int = 0             # the sort's widest digit: 8, 9 or 10 (0: 9 where that saves a pass, else 8)

    def to_c(self) -> _abi.RenderSettingsC:
        s = _abi.RenderSettingsC()
        s.width, s.height = self.width, self.height
        s.antialias = 1 if self.antialias else 0
        s.max_sh_degree = self.max_sh_degree
        s.background[:] = list(self.background)
        s.depth_mode, s.depth_threshold = self.depth_mode, self.depth_threshold
        s.linearise = 1 if self.linearise else 0
        s.time_stages = 1 if self.time_stages else 0
        s.dilation = self.dilation
        s.filter3d = 1 if self.filter3d else 0
        s.blend_batch = self.blend_batch
        s.tile_order = self.tile_order
        s.persistent_blend = 1 if self.persistent_blend else 0
        s.use_graph = 1 if self.use_graph else 0
        s.blend_segment = self.blend_segment
        s.keep_walk_state = 1 if self.keep_walk_state else 0
        s.tile_size = self.tile_size
        s.radix_bits = self.radix_bits
        return s


class GpuSplats:
    """scene::GpuSplats over torch device tensors (GpuClouds.h:41-76).

    positions: float32 [N, 4]; shape: int32/uint32 [N, 4]; sh: [N, sh_words] or
    None at degree 0; filter3d: float32 [N] or None.
    """

    def __init__(self, positions, shape, sh=None, rest_per_colour: int = 0, filter3d=None):
        import torch
        n = int(positions.shape[0])
        if positions.dtype != torch.float32 or tuple(positions.shape) != (n, 4) or not positions.is_contiguous():
            raise LrtError(3, "positions must be a contiguous float32 [N, 4] tensor")
        if tuple(shape.shape) != (n, 4) or shape.element_size() != 4 or not shape.is_contiguous():
            raise LrtError(3, "shape must be a contiguous 32-bit [N, 4] tensor")
        words = 1 if rest_per_colour == 0 else (rest_per_colour * 3 + 1) // 2
        if rest_per_colour:
            if sh is None or tuple(sh.shape) != (n, words) or sh.element_size() != 4 or not sh.is_contiguous():
                raise LrtError(3, f"sh must be a contiguous 32-bit [N, {words}] tensor")
        if filter3d is not None and (tuple(filter3d.shape) != (n,) or filter3d.dtype != torch.float32):
            raise LrtError(3, "filter3d must be float32 [N]")
        for t in (positions, shape, sh, filter3d):
            if t is not None and not t.is_cuda:
                raise LrtError(3, "splat tensors must live on the GPU")
        self.positions, self.shape, self.sh, self.filter3d = positions, shape, sh, filter3d
        self.count, self.rest_per_colour, self.sh_words = n, rest_per_colour, words
        self._c = _abi.Splats(positions.data_ptr(), shape.data_ptr(), sh.data_ptr() if sh is not None else None,
                              filter3d.data_ptr() if filter3d is not None else None, n, rest_per_colour, words)

    @property
    def c(self) -> _abi.Splats:
        return self._c


class DiffSplats:
    """A cloud of FP32 parameters for the differentiable frame, over torch device tensors.

    means [N, 3] (the cloud's own space), log_scales [N, 3], quats [N, 4] (x, y, z, w;
    normalised on use), opacity_logits [N], sh [N, 1 + rest, 3] (dc first, rgb per
    basis; rest 0, 3, 8 or 15), filter3d [N] or None. Drawn as the packed cloud
    would be: scale exp(log s), opacity sigmoid(logit), colour max(0.5 + C0 dc +
    rest(dir), 0).
    """

    def __init__(self, means, log_scales, quats, opacity_logits, sh, filter3d=None):
        import torch
        n = int(means.shape[0])
        rest = int(sh.shape[1]) - 1 if sh.dim() == 3 else -1
        want = {"means": (means, (n, 3)), "log_scales": (log_scales, (n, 3)), "quats": (quats, (n, 4)),
                "opacity_logits": (opacity_logits, (n,)), "sh": (sh, (n, rest + 1, 3))}
        if filter3d is not None:
            want["filter3d"] = (filter3d, (n,))
        for name, (t, shape) in want.items():
            if t.dtype != torch.float32 or tuple(t.shape) != shape or not t.is_contiguous() or not t.is_cuda:
                raise LrtError(3, f"{name} must be a contiguous float32 CUDA tensor of shape {shape}")
        if rest not in (0, 3, 8, 15):
            raise LrtError(3, "sh must hold 1, 4, 9 or 16 coefficients per colour")
        self.means, self.log_scales, self.quats = means, log_scales, quats
        self.opacity_logits, self.sh, self.filter3d = opacity_logits, sh, filter3d
        self.count, self.rest_per_colour = n, rest
        self._c = _abi.DiffSplatsC(means.data_ptr(), log_scales.data_ptr(), quats.data_ptr(),
                                   opacity_logits.data_ptr(), sh.data_ptr(),
                                   filter3d.data_ptr() if filter3d is not None else None, n, rest)

    @property
    def c(self) -> _abi.DiffSplatsC:
        return self._c


class GpuPoints:
    """scene::GpuPoints over torch device tensors (GpuClouds.h:78-85).

    positions: float32 [N, 4] (x, y, z, 1); colours: 32-bit [N, 2] (f16 r | g << 16,
    b | a << 16, linear) or None when every instance's style gives a constant colour.
    """

    def __init__(self, positions, colours=None):
        import torch
        n = int(positions.shape[0])
        if positions.dtype != torch.float32 or tuple(positions.shape) != (n, 4) or not positions.is_contiguous():
            raise LrtError(3, "positions must be a contiguous float32 [N, 4] tensor")
        if colours is not None and (tuple(colours.shape) != (n, 2) or colours.element_size() != 4
                                    or not colours.is_contiguous()):
            raise LrtError(3, "colours must be a contiguous 32-bit [N, 2] tensor")
        for t in (positions, colours):
            if t is not None and not t.is_cuda:
                raise LrtError(3, "point tensors must live on the GPU")
        self.positions, self.colours, self.count = positions, colours, n
        self._c = _abi.PointsC(positions.data_ptr(), colours.data_ptr() if colours is not None else None, n)

    @property
    def c(self) -> _abi.PointsC:
        return self._c


def pack_points(xyz, rgb=None):
    """Host arrays of a point cloud as points_decode.slang:32-41 packs them:
    positions float32 [N, 4] (x, y, z, 1) and colours uint32 [N, 2] (f16 r | g,
    b | 1.0), from xyz [N, 3] and linear rgb [N, 3] (None: white)."""
    xyz = np.asarray(xyz, dtype=np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    positions = np.concatenate([xyz, np.ones((n, 1), np.float32)], axis=1)
    rgb = np.ones((n, 3), np.float32) if rgb is None else np.asarray(rgb, dtype=np.float32).reshape(n, 3)
    with np.errstate(over="ignore"):
        halves = np.concatenate([rgb, np.ones((n, 1), np.float32)], axis=1).astype(np.float16).view(np.uint16)
    h = halves.astype(np.uint32)
    colours = np.stack([h[:, 0] | (h[:, 1] << 16), h[:, 2] | (h[:, 3] << 16)], axis=1)
    return positions, colours


@dataclass
class PointStyle:
    """render::PointStyle (modules/render/include/lrt/render/Points.h:21-33), the
    fields the disc route reads: size_mode WORLD (0) or PIXELS (1), size (a
    diameter), constant_colour and colour."""
    WORLD = 0
    PIXELS = 1
    size_mode: int = 0
    size: float = 0.01
    constant_colour: bool = False
    colour: tuple = (1.0, 1.0, 1.0)

    def to_c(self) -> _abi.PointStyleC:
        c = _abi.PointStyleC()
        c.size_mode = self.size_mode
        c.size = self.size
        c.constant_colour = 1 if self.constant_colour else 0
        c.colour[:] = [float(v) for v in self.colour]
        return c


@dataclass
class PointInstance:
    """render::PointInstance (Points.h:35-39): points drawn as opaque discs."""
    points: Optional[GpuPoints]
    object_to_world: tuple = IDENTITY
    style: PointStyle = field(default_factory=PointStyle)


@dataclass
class SplatInstance:
    """render::SplatInstance (TileRasterizer.h:52-67), without relighting."""
    splats: Optional[GpuSplats]
    object_to_world: tuple = IDENTITY
    edit: SplatEdit = field(default_factory=SplatEdit)


@dataclass
class FrameStats:
    """render::FrameStats (TileRasterizer.h:120-126)."""
    splats: int = 0
    visible: int = 0
    pairs: int = 0
    project_ms: float = 0.0
    depth_sort_ms: float = 0.0
    counts_ms: float = 0.0
    emit_ms: float = 0.0
    tile_sort_ms: float = 0.0
    blend_ms: float = 0.0
    total_ms: float = 0.0
    key_bits: int = 0       # the sort's virtual key: tile bits + the frame's depth-key range bits
    sort_passes: int = 0    # eight key bits a pass


class RenderTargets:
    """render::RenderTargets: colour float4 [H, W, 4] and depth [H, W], bottom row first."""

    def __init__(self, width: int, height: int, device="cuda"):
        import torch
        self.width, self.height = width, height
        self.colour = torch.empty((height, width, 4), dtype=torch.float32, device=device)
        self.depth = torch.empty((height, width), dtype=torch.float32, device=device)


@dataclass
class View:
    """One camera of a batch: a Camera (or a projection) and its images."""
    camera: object
    targets: RenderTargets


class RawSplats:
    """io::RawSplats (modules/io/include/lrt/io/RawSplats.h:15-80): a splat file's
    records, one float a property in the file's order, held by the library on the
    host, and their encoding (a dict of lrt_splat_encoding's fields)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def count(self) -> int:
        return int(lib.lrt_raw_splats_count(self._h))

    @property
    def encoding(self) -> dict:
        e = _abi.SplatEncodingC()
        lib.lrt_raw_splats_encoding(self._h, C.byref(e))
        return {name: int(getattr(e, name)) for name in _abi.ENCODING_FIELDS}

    @property
    def records(self):
        """A copy of the records as a float32 [count, floats_per_record] array."""
        import numpy as np
        n, w = self.count, self.encoding["floats_per_record"]
        ptr = lib.lrt_raw_splats_records(self._h)
        return np.ctypeslib.as_array(ptr, shape=(n * w,)).reshape(n, w).copy() if n * w else np.zeros((n, w), np.float32)

    def close(self):
        if self._h:
            lib.lrt_raw_splats_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_splats(path: str) -> RawSplats:
    """io::readSplats (modules/io/src/SplatReaders.cpp:32-195): a 3DGS PLY or a
    .splat file read on the host; raises LrtError with the reference's codes."""
    h = C.c_void_p()
    code = lib.lrt_splats_read(os.fsencode(path), C.byref(h))
    if code != _abi.LRT_OK:
        raise LrtError(code, lib.lrt_last_error(None).decode())
    return RawSplats(h)


def raw_splats(records, encoding: dict) -> RawSplats:
    """Raw records from memory: records float32 [count, floats_per_record]; encoding
    the lrt_splat_encoding fields (missing ones 0)."""
    import numpy as np
    rec = np.ascontiguousarray(records, dtype=np.float32)
    e = _abi.SplatEncodingC(**{k: int(v) for k, v in encoding.items()})
    if rec.ndim != 2 or rec.shape[1] != e.floats_per_record:
        raise LrtError(3, "records must be [count, floats_per_record]")
    h = C.c_void_p()
    code = lib.lrt_raw_splats_create(rec.ctypes.data_as(C.c_void_p), rec.shape[0], C.byref(e), C.byref(h))
    if code != _abi.LRT_OK:
        raise LrtError(code, lib.lrt_last_error(None).decode())
    return RawSplats(h)


def write_exr(path: str, colour, depth=None, half: bool = True):
    """io::writeExr (modules/io/src/Exr.cpp:130-157): colour float32 [H, W, 4] and
    depth float32 [H, W] (numpy or torch, bottom row first, as RenderTargets hold
    them) into a ZIP-compressed scanline EXR, R G B A half (or float) and Z float."""
    c = np.ascontiguousarray(colour.cpu().numpy() if hasattr(colour, "cpu") else colour, dtype=np.float32)
    if c.ndim != 3 or c.shape[2] != 4:
        raise LrtError(3, "colour must be [height, width, 4]")
    h, w = c.shape[:2]
    d = None
    if depth is not None:
        d = np.ascontiguousarray(depth.cpu().numpy() if hasattr(depth, "cpu") else depth, dtype=np.float32)
        if d.shape != (h, w):
            raise LrtError(3, "depth must be [height, width]")
    code = lib.lrt_write_exr(os.fsencode(path), w, h, c.ctypes.data_as(C.c_void_p),
                             d.ctypes.data_as(C.c_void_p) if d is not None else None, 1 if half else 0)
    if code != _abi.LRT_OK:
        raise LrtError(code, lib.lrt_last_error(None).decode())


@dataclass
class CloudInfo:
    """What a decode made (scene::GpuSplats count/declared/bounds, GpuClouds.h:41-76)."""
    declared: int
    count: int
    rest_per_colour: int
    sh_words: int
    bounds_min: tuple
    bounds_max: tuple
    upload_ms: float = 0.0   # GPU time of the records' copies to the device
    decode_ms: float = 0.0   # GPU time of validate + scan + decode


def _instances_c(instances: Sequence[SplatInstance]):
    arr = (_abi.SplatInstanceC * max(1, len(instances)))()
    for k, inst in enumerate(instances):
        if inst.splats is not None:
            arr[k].splats = C.pointer(inst.splats.c)
        arr[k].object_to_world[:] = list(np.asarray(inst.object_to_world, dtype=np.float64).reshape(16))
        arr[k].edit = inst.edit.to_c()
    return arr


class TileRasterizer:
    """render::TileRasterizer (TileRasterizer.h:136-193) on one HIP stream."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self._ctx = C.c_void_p()
        self.device = device
        s = stream if stream is not None else torch.cuda.current_stream(device)
        code = lib.lrt_context_create(device, C.c_void_p(s.cuda_stream), C.byref(self._ctx))
        if code != _abi.LRT_OK:
            raise LrtError(code, "lrt_context_create: " + lib.lrt_last_error(None).decode())
        self._stream = s

    @staticmethod
    def create(device: int = 0, stream=None) -> "TileRasterizer":
        return TileRasterizer(device, stream)

    def close(self):
        if self._ctx:
            lib.lrt_context_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code: int):
        if code != _abi.LRT_OK:
            raise LrtError(code, lib.lrt_last_error(self._ctx).decode())

    def decode_cloud(self, raw: RawSplats, max_degree: int = 3):
        """CloudLoader::upload (GpuClouds.cpp:151-227): the records validated and
        packed on the GPU, kept ones in file order. Returns (GpuSplats, CloudInfo);
        the splats' tensors are views of buffers sized for the declared count."""
        import torch
        n = raw.count
        words = int(lib.lrt_cloud_sh_words(raw.handle, max_degree))
        dev = torch.device("cuda", self.device)
        positions = torch.empty((n, 4), dtype=torch.float32, device=dev)
        shape = torch.empty((n, 4), dtype=torch.int32, device=dev)
        # one word a splat means no rest coefficient is kept: no harmonics buffer
        sh = torch.empty((n, words), dtype=torch.int32, device=dev) if words > 1 else None
        info = _abi.CloudInfoC()
        self._check(lib.lrt_cloud_decode(self._ctx, raw.handle, max_degree, C.c_void_p(positions.data_ptr()),
                                         C.c_void_p(shape.data_ptr()),
                                         C.c_void_p(sh.data_ptr()) if sh is not None else None, C.byref(info)))
        k = int(info.count)
        rest = int(info.rest_per_colour)
        cloud = GpuSplats(positions[:k], shape[:k], sh[:k] if rest else None, rest)
        return cloud, CloudInfo(int(info.declared), k, rest, int(info.sh_words), tuple(info.bounds_min),
                                tuple(info.bounds_max), float(info.upload_ms), float(info.decode_ms))

    def render(self, camera_or_projection, instances: Sequence[SplatInstance], settings: RenderSettings,
               targets: RenderTargets, points: Sequence[PointInstance] = (),
               under: Optional[RenderTargets] = None) -> FrameStats:
        """TileRasterizer::render (TileRasterizer.cpp:188-484). Point instances are
        drawn as opaque discs in the splats' depth order; `under` (the frame's
        size) lies under the splats and ends each pixel's walk at its depth.
        Either selects the composite blend (splat_blend.slang:116-120)."""
        proj = camera_or_projection
        if isinstance(proj, Camera):
            proj = projection_for(proj, settings.width, settings.height)
        arr = _instances_c(instances)
        t = _abi.RenderTargetsC(targets.width, targets.height, targets.colour.data_ptr(),
                                targets.depth.data_ptr() if targets.depth is not None else None)
        st = _abi.FrameStatsC()
        if not points and under is None:
            self._check(lib.lrt_raster_render(self._ctx, C.byref(proj), arr, len(instances),
                                              C.byref(settings.to_c()), C.byref(t), C.byref(st)))
        else:
            parr = (_abi.PointInstanceC * max(1, len(points)))()
            for k, inst in enumerate(points):
                if inst.points is not None:
                    parr[k].points = C.pointer(inst.points.c)
                parr[k].object_to_world[:] = list(np.asarray(inst.object_to_world, dtype=np.float64).reshape(16))
                parr[k].style = inst.style.to_c()
            u = None
            if under is not None:
                if under.depth is None:
                    raise LrtError(3, "an under-layer needs both colour and depth")
                u = _abi.RenderTargetsC(under.width, under.height, under.colour.data_ptr(), under.depth.data_ptr())
            self._check(lib.lrt_raster_render_composite(self._ctx, C.byref(proj), arr, len(instances), parr,
                                                        len(points), C.byref(u) if u is not None else None,
                                                        C.byref(settings.to_c()), C.byref(t), C.byref(st)))
        return FrameStats(**{name: getattr(st, name) for name, _ in _abi.FrameStatsC._fields_})

    def reference_render(self, camera_or_projection, instances: Sequence[SplatInstance], settings: RenderSettings,
                         keyed: bool = False):
        """The brute force in FP64 on the GPU: (colour [H, W, 4], depth [H, W]) float64 CUDA tensors."""
        import torch
        proj = camera_or_projection
        if isinstance(proj, Camera):
            proj = projection_for(proj, settings.width, settings.height)
        colour = torch.empty((settings.height, settings.width, 4), dtype=torch.float64, device="cuda")
        depth = torch.empty((settings.height, settings.width), dtype=torch.float64, device="cuda")
        self._check(lib.lrt_reference_render(self._ctx, C.byref(proj), _instances_c(instances), len(instances),
                                             C.byref(settings.to_c()), colour.data_ptr(), depth.data_ptr(),
                                             1 if keyed else 0))
        return colour, depth

    def reference_render_pixels(self, camera_or_projection, instances: Sequence[SplatInstance],
                                settings: RenderSettings, pixels, keyed: bool = False):
        """The FP64 brute force at the listed pixels only (pixels: int32/uint32 CUDA
        tensor of y * width + x): (colour [n, 4], depth [n]) float64 CUDA tensors."""
        import torch
        proj = camera_or_projection
        if isinstance(proj, Camera):
            proj = projection_for(proj, settings.width, settings.height)
        n = int(pixels.numel())
        colour = torch.empty((max(n, 1), 4), dtype=torch.float64, device="cuda")
        depth = torch.empty((max(n, 1),), dtype=torch.float64, device="cuda")
        self._check(lib.lrt_reference_render_pixels(self._ctx, C.byref(proj), _instances_c(instances), len(instances),
                                                    C.byref(settings.to_c()), pixels.data_ptr(), n,
                                                    colour.data_ptr(), depth.data_ptr(), 1 if keyed else 0))
        return colour[:n], depth[:n]

    def _views_c(self, views: Sequence[View]):
        arr = (_abi.ViewC * max(1, len(views)))()
        for k, v in enumerate(views):
            t = v.targets
            proj = v.camera if isinstance(v.camera, _abi.ProjectionC) else projection_for(v.camera, t.width, t.height)
            arr[k].projection = proj
            arr[k].targets = _abi.RenderTargetsC(t.width, t.height, t.colour.data_ptr(),
                                                 t.depth.data_ptr() if t.depth is not None else None)
        return arr

    def render_views(self, views: Sequence[View], instances: Sequence[SplatInstance],
                     settings: RenderSettings) -> FrameStats:
        """Every view of the instances in one launch sequence (each view's images as render() draws them)."""
        st = _abi.FrameStatsC()
        self._check(lib.lrt_raster_render_views(self._ctx, self._views_c(views), len(views), _instances_c(instances),
                                                len(instances), C.byref(settings.to_c()), C.byref(st)))
        return FrameStats(**{name: getattr(st, name) for name, _ in _abi.FrameStatsC._fields_})

    def autotune_views(self, views: Sequence[View], instances: Sequence[SplatInstance],
                       settings: RenderSettings) -> tuple:
        """The fastest launch choices for this batch: (settings with them, its batch ms)."""
        import dataclasses
        c = settings.to_c()
        ms = C.c_float()
        self._check(lib.lrt_autotune_views(self._ctx, self._views_c(views), len(views), _instances_c(instances),
                                           len(instances), C.byref(c), C.byref(ms)))
        return dataclasses.replace(settings, blend_batch=c.blend_batch, persistent_blend=bool(c.persistent_blend),
                                   use_graph=bool(c.use_graph), tile_size=c.tile_size,
                                   radix_bits=c.radix_bits), float(ms.value)

    def render_diff(self, camera_or_projection, splats: DiffSplats, settings: RenderSettings,
                    targets: RenderTargets, object_to_world=IDENTITY) -> FrameStats:
        """A frame of one FP32-parameter cloud, kept for backward()."""
        proj = camera_or_projection
        if isinstance(proj, Camera):
            proj = projection_for(proj, settings.width, settings.height)
        t = _abi.RenderTargetsC(targets.width, targets.height, targets.colour.data_ptr(),
                                targets.depth.data_ptr() if targets.depth is not None else None)
        st = _abi.FrameStatsC()
        self._diff_serial = getattr(self, "_diff_serial", 0) + 1
        # The backward reads the parameters (and the 3D filter) again: their
        # tensors stay referenced here until the next differentiable frame, so
        # a caller's temporary is not freed and its memory reused in between.
        self._diff_inputs = splats
        self._check(lib.lrt_raster_render_diff(self._ctx, C.byref(proj), C.byref(splats.c), _mat(object_to_world),
                                               C.byref(settings.to_c()), C.byref(t), C.byref(st)))
        return FrameStats(**{name: getattr(st, name) for name, _ in _abi.FrameStatsC._fields_})

    def backward(self, grad_colour, grads: dict) -> tuple:
        """Gradients of the last render_diff's parameters from dL/d(colour) [H, W, 4].

        grads: float32 device tensors keyed means, log_scales, quats, opacity_logits, sh,
        shaped as the parameters; overwritten. Returns the GPU ms of the blend's and the
        preprocess's backward.
        """
        import torch
        if grad_colour.dtype != torch.float32 or not grad_colour.is_contiguous() or not grad_colour.is_cuda:
            raise LrtError(3, "grad_colour must be a contiguous float32 CUDA tensor")
        g = _abi.DiffGradsC(grads["means"].data_ptr(), grads["log_scales"].data_ptr(), grads["quats"].data_ptr(),
                            grads["opacity_logits"].data_ptr(), grads["sh"].data_ptr())
        ms = (C.c_float * 2)()
        self._check(lib.lrt_raster_backward(self._ctx, C.c_void_p(grad_colour.data_ptr()), C.byref(g), ms))
        return float(ms[0]), float(ms[1])

    def render_diff_views(self, views: Sequence[View], splats: DiffSplats, settings: RenderSettings,
                          object_to_world=IDENTITY) -> FrameStats:
        """Every view of one FP32-parameter cloud in one launch sequence (each view's images as
        render_diff draws them), kept for backward_views()."""
        st = _abi.FrameStatsC()
        self._diff_serial = getattr(self, "_diff_serial", 0) + 1
        self._diff_inputs = splats   # read again by backward_views (see render_diff)
        self._check(lib.lrt_raster_render_diff_views(self._ctx, self._views_c(views), len(views), C.byref(splats.c),
                                                     _mat(object_to_world), C.byref(settings.to_c()), C.byref(st)))
        return FrameStats(**{name: getattr(st, name) for name, _ in _abi.FrameStatsC._fields_})

    def backward_views(self, grad_colours: Sequence, grads: dict) -> tuple:
        """Gradients of the last render_diff_views' parameters, summed over its views, from
        each view's dL/d(colour) [H_v, W_v, 4] (in view order). grads as backward()'s.
        Returns the GPU ms of the blend's and the preprocess's backward."""
        import torch
        for g in grad_colours:
            if g.dtype != torch.float32 or not g.is_contiguous() or not g.is_cuda:
                raise LrtError(3, "each colour gradient must be a contiguous float32 CUDA tensor")
        ptrs = (C.c_void_p * max(1, len(grad_colours)))(*[g.data_ptr() for g in grad_colours])
        g = _abi.DiffGradsC(grads["means"].data_ptr(), grads["log_scales"].data_ptr(), grads["quats"].data_ptr(),
                            grads["opacity_logits"].data_ptr(), grads["sh"].data_ptr())
        ms = (C.c_float * 2)()
        self._check(lib.lrt_raster_backward_views(self._ctx, ptrs, len(grad_colours), C.byref(g), ms))
        return float(ms[0]), float(ms[1])

    def debug(self, which: int) -> np.ndarray:
        """What the last render left on the device (LRT_DEBUG_*), copied to host."""
        need = C.c_size_t()
        self._check(lib.lrt_raster_debug(self._ctx, which, None, 0, C.byref(need)))
        dtype = {_abi.DEBUG_PROJ: np.float32, _abi.DEBUG_KEYS: np.uint64, _abi.DEBUG_FINAL_T: np.float32,
                 _abi.DEBUG_KERNEL_MS: np.float32}.get(which, np.uint32)
        out = np.empty(need.value // np.dtype(dtype).itemsize, dtype=dtype)
        if need.value:
            self._check(lib.lrt_raster_debug(self._ctx, which, out.ctypes.data, need.value, None))
        return out

    def prefix_sum(self, inp, out, total):
        """PrefixSum::apply (PrefixSum.h:43-44) on torch uint32/int32 device tensors."""
        self._check(lib.lrt_prefix_sum(self._ctx, inp.data_ptr(), out.data_ptr(), total.data_ptr(), inp.numel()))

    def radix_sort(self, keys, values, scratch_keys, scratch_values, key_bits: int, radix_bits: int = 8) -> bool:
        """RadixSort::sort (RadixSort.h:56-57), digits up to radix_bits (8..10) wide; returns True
        when the result is in the scratch pair."""
        flag = C.c_int32()
        if radix_bits == 8:
            self._check(lib.lrt_radix_sort(self._ctx, keys.data_ptr(), values.data_ptr(), scratch_keys.data_ptr(),
                                           scratch_values.data_ptr(), keys.numel(), key_bits, C.byref(flag)))
        else:
            self._check(lib.lrt_radix_sort_digits(self._ctx, keys.data_ptr(), values.data_ptr(),
                                                  scratch_keys.data_ptr(), scratch_values.data_ptr(), keys.numel(),
                                                  key_bits, radix_bits, C.byref(flag)))
        return bool(flag.value)

    def synchronize(self):
        self._check(lib.lrt_context_synchronize(self._ctx))

    def copy_to_host(self, device_ptr: int, nbytes: int, dtype) -> np.ndarray:
        """Device memory the library owns, read after the stream's queued work."""
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        if nbytes:
            self._check(lib.lrt_copy_to_host(self._ctx, C.c_void_p(device_ptr), out.ctypes.data, nbytes))
        return out

    def lod_build(self, splats, settings: Optional["LodBuildSettings"] = None) -> "LodCloud":
        """LodBuilder::build (Lod.cpp:81-283): the cloud's levels of detail."""
        c = (settings or LodBuildSettings()).to_c()
        h = C.c_void_p()
        self._check(lib.lrt_lod_build(self._ctx, C.byref(splats.c), C.byref(c), C.byref(h)))
        return LodCloud(h, self)

    def lod_select(self, camera_or_projection, instances: Sequence["LodInstance"], threshold: float,
                   settings: RenderSettings):
        """CutSelector::select (Lod.cpp:350-547): (splat instances to render, one
        CutStats an instance). The instances' clouds are the context's and are
        rewritten by the next call."""
        proj = camera_or_projection
        if isinstance(proj, Camera):
            proj = projection_for(proj, settings.width, settings.height)
        n = len(instances)
        arr = (_abi.LodInstanceC * max(1, n))()
        for k, inst in enumerate(instances):
            arr[k].cloud = inst.cloud.handle if inst.cloud is not None else None
            arr[k].object_to_world[:] = list(np.asarray(inst.object_to_world, dtype=np.float64).reshape(16))
            arr[k].edit = inst.edit.to_c()
            arr[k].has_threshold = 1 if inst.threshold is not None else 0
            arr[k].threshold = inst.threshold if inst.threshold is not None else 0.0
        out = (_abi.SplatInstanceC * max(1, n))()
        count = C.c_uint32()
        stats = (_abi.CutStatsC * max(1, n))()
        self._check(lib.lrt_lod_select(self._ctx, C.byref(proj), arr, n, threshold, out, C.byref(count), stats))
        drawn = []
        for k in range(count.value):
            view = SplatsView(out[k].splats.contents)
            drawn.append(SplatInstance(view, tuple(out[k].object_to_world), _edit_from_c(out[k].edit)))
        return drawn, [CutStats(stats[k].splats, stats[k].merged, stats[k].available) for k in range(n)]


@dataclass
class LodBuildSettings:
    """LodBuildSettings (modules/lod/include/lrt/lod/Lod.h:78-87) without chunks."""
    max_group_fraction: float = 0.5
    coarsest_level: int = 1

    def to_c(self) -> _abi.LodBuildSettingsC:
        return _abi.LodBuildSettingsC(self.max_group_fraction, self.coarsest_level)


@dataclass
class CutStats:
    """CutStats (Lod.h:108-117) without the streaming needs."""
    splats: int = 0
    merged: int = 0
    available: int = 0


class SplatsView:
    """A packed cloud the library owns (an LOD level, a cut), as a SplatInstance's splats."""

    def __init__(self, c: _abi.Splats):
        self._c = _abi.Splats(c.positions, c.shape, c.sh, c.filter3d, c.count, c.rest_per_colour, c.sh_words)
        self.count, self.rest_per_colour, self.sh_words = c.count, c.rest_per_colour, c.sh_words

    @property
    def c(self) -> _abi.Splats:
        return self._c

    def read(self, rast: "TileRasterizer") -> dict:
        """positions float32 [N, 4], shape uint32 [N, 4], sh uint32 [N, sh_words] or None."""
        n, c = self.count, self._c
        out = dict(positions=rast.copy_to_host(c.positions or 0, n * 16, np.float32).reshape(n, 4),
                   shape=rast.copy_to_host(c.shape or 0, n * 16, np.uint32).reshape(n, 4), sh=None)
        if self.rest_per_colour and c.sh:
            out["sh"] = rast.copy_to_host(c.sh, n * self.sh_words * 4, np.uint32).reshape(n, self.sh_words)
        return out


class LodCloud:
    """A cloud with its levels of detail (LodCloud, Lod.h:51-76), owned by the library."""

    def __init__(self, handle: C.c_void_p, rast: "TileRasterizer"):
        self._h = handle
        self._rast = rast
        info = _abi.LodInfoC()
        if lib.lrt_lod_get_info(handle, C.byref(info)) != _abi.LRT_OK:
            raise LrtError(3, "lrt_lod_get_info")
        k = int(info.levels)
        self.count, self.rest_per_colour, self.sh_words = int(info.count), int(info.rest_per_colour), int(info.sh_words)
        self.levels = [dict(level=int(info.level[i]), first=int(info.first[i]), groups=int(info.groups[i]))
                       for i in range(k)]
        self.merged = int(info.merged)
        self.bounds_lo = tuple(float(v) for v in info.bounds_lo)
        self.extent = float(info.extent)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def splats(self) -> tuple:
        """(the Morton-ordered splats as a SplatsView, each one's finest-level group as numpy)."""
        c = _abi.Splats()
        groups = C.c_void_p()
        if lib.lrt_lod_splats(self._h, C.byref(c), C.byref(groups)) != _abi.LRT_OK:
            raise LrtError(3, "lrt_lod_splats")
        return SplatsView(c), self._rast.copy_to_host(groups.value or 0, self.count * 4, np.uint32)

    def gaussians(self, index: int) -> tuple:
        """(stored level `index`'s merged Gaussians as a SplatsView, their cell codes as numpy)."""
        c = _abi.Splats()
        cells = C.c_void_p()
        if lib.lrt_lod_gaussians(self._h, index, C.byref(c), C.byref(cells)) != _abi.LRT_OK:
            raise LrtError(3, "lrt_lod_gaussians")
        return SplatsView(c), self._rast.copy_to_host(cells.value or 0, int(c.count) * 4, np.uint32)

    def close(self):
        if self._h:
            lib.lrt_lod_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class LodInstance:
    """LodInstance (Lod.h:101-106): threshold None takes the call's."""
    cloud: Optional[LodCloud]
    object_to_world: tuple = IDENTITY
    edit: SplatEdit = field(default_factory=SplatEdit)
    threshold: Optional[float] = None


def _edit_from_c(e: _abi.SplatEdit) -> SplatEdit:
    return SplatEdit(active=bool(e.active), shape=e.shape, mode=e.mode, invert=bool(e.invert),
                     centre=tuple(e.centre), size=tuple(e.size), tint=tuple(e.tint), saturation=e.saturation,
                     brightness=e.brightness, opacity=e.opacity, min_opacity=e.min_opacity, max_scale=e.max_scale)


def frame_params_bytes(projection: _abi.ProjectionC, instance: SplatInstance, settings: RenderSettings,
                       base: int = 0, count: int = 0, rest_per_colour: int = 0) -> bytes:
    """The kernel constant block one instance would be drawn with (host only)."""
    splats = _abi.Splats(None, None, None, None, count, rest_per_colour,
                         1 if rest_per_colour == 0 else (rest_per_colour * 3 + 1) // 2)
    inst = _abi.SplatInstanceC()
    inst.splats = C.pointer(splats)
    inst.object_to_world[:] = list(np.asarray(instance.object_to_world, dtype=np.float64).reshape(16))
    inst.edit = instance.edit.to_c()
    need = C.c_size_t()
    lib.lrt_frame_params(C.byref(projection), C.byref(inst), C.byref(settings.to_c()), base, None, 0, C.byref(need))
    buf = (C.c_char * need.value)()
    code = lib.lrt_frame_params(C.byref(projection), C.byref(inst), C.byref(settings.to_c()), base, buf,
                                need.value, None)
    if code != _abi.LRT_OK:
        raise LrtError(code, "lrt_frame_params")
    return bytes(buf)
