"""The seeded synthetic scenes and cameras of the benchmark workloads.

Scenes are generated and packed by liblrt_hip.so (lrt_synth_scene) into the
GpuSplats layout in host memory. Each splat draws from its own seeded stream,
so a scene of N splats is the first N splats of any larger one.

  plain-fwd-65k-640x480  65,536 splats, means U[-1,1]^3 4 units down -Z of a
            camera at the origin, log-scales N(-4, 0.5), opacity sigmoid(N(0,1)),
            degree 0, 640x480, 60 degree horizontal field of view (the plain cube).
  scene-fwd-3M-1080p  3,000,000 splats in 256 anisotropic clusters (centres U[-10,10]^3,
            members N(centre, diag(U[0.05,1])^2)), log-scales N(-5, 0.8), degree 3
            (band l ~ N(0, 0.3 / l)), 1920x1080, camera on an orbit of radius 15
            about the origin, 60 degrees; Mip-Splatting 3D filter on, 2D filter 0.1.
  plain-fwd-1M-720p  1,000,000 splats of the plain cube, degree 3, 1280x720.
  sort-stress-6M-2160p  6,000,000 splats at 3840x2160: 80% of the plain cube, 20% pulled toward
            the camera (into 0.2-1 of their distance) with log-scales N(-3.6, 0.3),
            which puts about 7e8 (tile, splat) pairs in the frame (the plain-cube
            part alone gives 51 a splat at this size); opacity sigmoid(N(1,1)),
            degree 3.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _abi
from ._abi import LrtError, lib
from .raster import Camera, Lens

HAPERTURE = 24.576


def focal_for_fov(fov_degrees: float) -> float:
    """Focal length (mm, 24.576 mm aperture) of a horizontal field of view."""
    return HAPERTURE / (2.0 * math.tan(math.radians(fov_degrees) / 2.0))


@dataclass
class HostScene:
    positions: np.ndarray          # float32 [N, 4]
    shape: np.ndarray              # uint32 [N, 4]
    sh: Optional[np.ndarray]       # uint32 [N, sh_words] or None at degree 0
    rest_per_colour: int
    filter3d: Optional[np.ndarray] = None   # float32 [N]

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])

    @property
    def sh_words(self) -> int:
        return 1 if self.rest_per_colour == 0 else (self.rest_per_colour * 3 + 1) // 2

    def to_device(self, device="cuda"):
        import torch
        from .raster import GpuSplats
        pos = torch.from_numpy(self.positions).to(device)
        shp = torch.from_numpy(self.shape.view(np.int32)).to(device)
        sh = torch.from_numpy(self.sh.view(np.int32)).to(device) if self.sh is not None else None
        f3 = torch.from_numpy(self.filter3d).to(device) if self.filter3d is not None else None
        return GpuSplats(pos, shp, sh, self.rest_per_colour, f3)


REST = {0: 0, 1: 3, 2: 8, 3: 15}


def synth(kind: int, count: int, seed: int = 0, degree: int = 0, offset=(0.0, 0.0, -4.0),
          log_scale=(-4.0, 0.5), opacity_logit=(0.0, 1.0), sh_sd: float = 0.3, clusters: int = 256,
          box_half: float = 10.0, cluster_sd=(0.05, 1.0), near=(0.2, -1.0, 0.5)) -> HostScene:
    spec = _abi.SceneSpecC()
    spec.kind, spec.degree, spec.seed = kind, degree, seed
    spec.offset[:] = list(offset)
    spec.log_scale_mean, spec.log_scale_sd = log_scale
    spec.opacity_logit_mean, spec.opacity_logit_sd = opacity_logit
    spec.sh_sd, spec.clusters, spec.box_half = sh_sd, clusters, box_half
    spec.cluster_sd_lo, spec.cluster_sd_hi = cluster_sd
    spec.near_fraction, spec.near_log_scale_mean, spec.near_log_scale_sd = near
    rest = REST[degree]
    words = 1 if rest == 0 else (rest * 3 + 1) // 2
    pos = np.empty((count, 4), np.float32)
    shape = np.empty((count, 4), np.uint32)
    sh = np.empty((count, words), np.uint32) if rest else None
    code = lib.lrt_synth_scene(C.byref(spec), count, pos.ctypes.data, shape.ctypes.data,
                               sh.ctypes.data if sh is not None else None)
    if code != _abi.LRT_OK:
        raise LrtError(code, "lrt_synth_scene")
    return HostScene(pos, shape, sh, rest)


def pack_splats(fields: np.ndarray, rest_per_colour: int) -> HostScene:
    """Packs rows of (x y z, opacity, 3 log-scales, rotation xyzw, base rgb, rest rgb per basis)."""
    fields = np.ascontiguousarray(fields, dtype=np.float32)
    n = fields.shape[0]
    words = 1 if rest_per_colour == 0 else (rest_per_colour * 3 + 1) // 2
    pos = np.empty((n, 4), np.float32)
    shape = np.empty((n, 4), np.uint32)
    sh = np.empty((n, words), np.uint32) if rest_per_colour else None
    code = lib.lrt_pack_splats(fields.ctypes.data, n, rest_per_colour, pos.ctypes.data, shape.ctypes.data,
                               sh.ctypes.data if sh is not None else None)
    if code != _abi.LRT_OK:
        raise LrtError(code, "lrt_pack_splats")
    return HostScene(pos, shape, sh, rest_per_colour)


def orbit_camera(radius: float, angle_degrees: float, height: float = 0.0, fov: float = 60.0) -> Camera:
    a = math.radians(angle_degrees)
    cam = Camera.looking_at((radius * math.sin(a), height, radius * math.cos(a)), (0.0, 0.0, 0.0))
    cam.lens = Lens(focal=focal_for_fov(fov))
    return cam


def filter3d_for(positions: np.ndarray, cameras, width: int, height: int, scale: float = 0.2) -> np.ndarray:
    """Mip-Splatting's 3D filter size per splat: sqrt(scale) * min over the cameras
    that see it (view depth > 0.2, centre within 1.15x the frame) of depth / focal
    in pixels; splats no camera sees take the largest value found."""
    from .raster import projection_for
    pts = np.asarray(positions[:, :3], dtype=np.float64)
    best = np.full(pts.shape[0], np.inf)
    for cam in cameras:
        p = projection_for(cam, width, height)
        m = np.array(p.world_to_view[:], dtype=np.float64).reshape(4, 4)
        v = pts @ m[:3, :3].T + m[:3, 3]
        z = v[:, 2]
        ok = z > 0.2
        zs = np.where(ok, z, 1.0)
        x = p.focal_x * v[:, 0] / zs + p.centre_x
        y = p.focal_y * v[:, 1] / zs + p.centre_y
        ok &= (x >= -0.15 * width) & (x <= 1.15 * width) & (y >= -0.15 * height) & (y <= 1.15 * height)
        best = np.where(ok, np.minimum(best, z / p.focal_x), best)
    finite = np.isfinite(best)
    fill = best[finite].max() if finite.any() else 1.0
    best = np.where(finite, best, fill)
    return (best * math.sqrt(scale)).astype(np.float32)


SH_C0 = 0.28209479177387814


def diff_params(count: int, seed: int = 0, degree: int = 3, offset=(0.0, 0.0, -4.0), log_scale=(-4.0, 0.5),
                opacity_logit=(0.0, 1.0), sh_sd: float = 0.3) -> dict:
    """FP32 parameters of a plain cube (the differentiable frame's
    input): means U[-1,1]^3 + offset, log-scales N(mean, sd), quaternions of a
    4-D normal (uniform rotations), opacity logits N(mean, sd), dc from a base
    colour U[0,1] (base = 0.5 + C0 dc), rest band l ~ N(0, sh_sd / l)."""
    rng = np.random.default_rng(seed)
    rest = REST[degree]
    means = (rng.uniform(-1.0, 1.0, (count, 3)) + np.asarray(offset)).astype(np.float32)
    log_scales = rng.normal(log_scale[0], log_scale[1], (count, 3)).astype(np.float32)
    quats = rng.normal(0.0, 1.0, (count, 4)).astype(np.float32)
    logits = rng.normal(opacity_logit[0], opacity_logit[1], count).astype(np.float32)
    sh = np.zeros((count, 1 + rest, 3), np.float32)
    sh[:, 0, :] = (rng.uniform(0.0, 1.0, (count, 3)) - 0.5) / SH_C0
    band = [1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3]
    for k in range(rest):
        sh[:, 1 + k, :] = rng.normal(0.0, sh_sd / band[k], (count, 3))
    return dict(means=means, log_scales=log_scales, quats=quats, opacity_logits=logits, sh=sh)


def diff_params_clustered(count: int, seed: int = 0, degree: int = 3, clusters: int = 256, box_half: float = 10.0,
                          cluster_sd=(0.05, 1.0), log_scale=(-5.0, 0.8), opacity_logit=(0.0, 1.0),
                          sh_sd: float = 0.3) -> dict:
    """FP32 parameters of the clustered scene (the batched fwd+bwd input): cluster
    centres U[-box_half, box_half]^3, members N(centre, diag(U[cluster_sd])^2),
    log-scales N(log_scale), quaternions of a 4-D normal, opacity logits
    N(opacity_logit), dc from a base colour U[0,1], rest band l ~ N(0, sh_sd / l)."""
    rng = np.random.default_rng(seed)
    rest = REST[degree]
    centres = rng.uniform(-box_half, box_half, (clusters, 3))
    sds = rng.uniform(cluster_sd[0], cluster_sd[1], (clusters, 3))
    member = rng.integers(0, clusters, count)
    means = (centres[member] + rng.normal(0.0, 1.0, (count, 3)) * sds[member]).astype(np.float32)
    log_scales = rng.normal(log_scale[0], log_scale[1], (count, 3)).astype(np.float32)
    quats = rng.normal(0.0, 1.0, (count, 4)).astype(np.float32)
    logits = rng.normal(opacity_logit[0], opacity_logit[1], count).astype(np.float32)
    sh = np.zeros((count, 1 + rest, 3), np.float32)
    sh[:, 0, :] = (rng.uniform(0.0, 1.0, (count, 3)) - 0.5) / SH_C0
    band = [1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3]
    for k in range(rest):
        sh[:, 1 + k, :] = rng.normal(0.0, sh_sd / band[k], (count, 3))
    return dict(means=means, log_scales=log_scales, quats=quats, opacity_logits=logits, sh=sh)


VIEW_SIZES = ((256, 256), (512, 384), (960, 540), (1280, 720), (1920, 1080))


def sphere_views(count: int = 64, radius: float = 12.0, seed: int = 4, sizes=VIEW_SIZES, fov: float = 60.0):
    """`count` cameras on a sphere of `radius` about the origin (a Fibonacci
    lattice, each looking at the centre), each view's size drawn from `sizes`
    by a seeded generator: a list of (Camera, (width, height))."""
    rng = np.random.default_rng(seed)
    golden = math.pi * (3.0 - math.sqrt(5.0))
    out = []
    for k in range(count):
        y = 1.0 - 2.0 * (k + 0.5) / count
        r = math.sqrt(1.0 - y * y)
        eye = (radius * r * math.cos(golden * k), radius * y, radius * r * math.sin(golden * k))
        up = (0.0, 0.0, 1.0) if abs(y) > 0.95 else (0.0, 1.0, 0.0)
        cam = Camera.looking_at(eye, (0.0, 0.0, 0.0), up)
        cam.lens = Lens(focal=focal_for_fov(fov))
        out.append((cam, tuple(sizes[int(rng.integers(0, len(sizes)))])))
    return out


def perturbed(params: dict, seed: int = 1, means_sd=0.01, logit_sd=0.2, sh_sd=0.05) -> dict:
    """The target scene of the fwd+bwd benchmark: means + N(0, 0.01), logits + N(0, 0.2), SH + N(0, 0.05)."""
    rng = np.random.default_rng(seed)
    out = {k: v.copy() for k, v in params.items()}
    out["means"] += rng.normal(0.0, means_sd, out["means"].shape).astype(np.float32)
    out["opacity_logits"] += rng.normal(0.0, logit_sd, out["opacity_logits"].shape).astype(np.float32)
    out["sh"] += rng.normal(0.0, sh_sd, out["sh"].shape).astype(np.float32)
    return out


@dataclass
class Workload:
    name: str
    scene: HostScene
    camera: Camera
    width: int
    height: int
    dilation: float = 0.3
    filter3d: bool = False


WORKLOADS = ("plain-fwd-65k-640x480", "scene-fwd-3M-1080p", "plain-fwd-1M-720p", "sort-stress-6M-2160p")


def workload(name: str, count: Optional[int] = None, seed: int = 0, view: float = 0.0) -> Workload:
    """The benchmark workload `name` (one of WORKLOADS) as a scene, a camera and a frame size;
    `count` draws the first `count` splats of its scene instead."""
    if name == "plain-fwd-65k-640x480":
        scene = synth(0, count or 65536, seed, degree=0, offset=(0.0, 0.0, -4.0), log_scale=(-4.0, 0.5))
        cam = Camera.looking_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
        cam.lens = Lens(focal=focal_for_fov(60.0))
        return Workload(name, scene, cam, 640, 480)
    if name == "scene-fwd-3M-1080p":
        scene = synth(1, count or 3_000_000, seed, degree=3, log_scale=(-5.0, 0.8), clusters=256, box_half=10.0)
        cam = orbit_camera(15.0, view)
        cams = [orbit_camera(15.0, 360.0 * k / 16.0) for k in range(16)]
        scene.filter3d = filter3d_for(scene.positions, cams, 1920, 1080)
        return Workload(name, scene, cam, 1920, 1080, dilation=0.1, filter3d=True)
    if name == "plain-fwd-1M-720p":
        scene = synth(0, count or 1_000_000, seed, degree=3, offset=(0.0, 0.0, -4.0), log_scale=(-4.0, 0.5))
        cam = Camera.looking_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
        cam.lens = Lens(focal=focal_for_fov(60.0))
        return Workload(name, scene, cam, 1280, 720)
    if name == "sort-stress-6M-2160p":
        scene = synth(2, count or 6_000_000, seed, degree=3, offset=(0.0, 0.0, -4.0), log_scale=(-4.0, 0.5),
                      opacity_logit=(1.0, 1.0), near=(0.2, -3.6, 0.3))
        cam = Camera.looking_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
        cam.lens = Lens(focal=focal_for_fov(60.0))
        return Workload(name, scene, cam, 3840, 2160)
    raise ValueError(f"no workload {name!r}; the workloads are {WORKLOADS}")
