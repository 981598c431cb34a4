"""The C ABI without a GPU: the library loads, exports every function the
headers under include/ declare, and fails as the reference's ErrorCode says
where nothing can run."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from lucabrtrender_amd import _abi, raster

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for header in ("lrt_hip.h", "lrt_synth.h"):
        text = open(os.path.join(ROOT, "include", header)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(lrt_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_function_is_exported_and_bound():
    declared = _declared()
    assert len(declared) >= 17
    lib = C.CDLL(_abi.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_abi.SIGNATURES), declared ^ set(_abi.SIGNATURES)


def test_defaults_are_the_reference_defaults():
    s = _abi.RenderSettingsC()
    _abi.lib.lrt_render_settings_default(C.byref(s))
    assert (s.width, s.height, s.antialias, s.max_sh_degree, s.linearise) == (1920, 1080, 1, 3, 1)
    assert list(s.background) == [0.0, 0.0, 0.0, 0.0] and s.depth_mode == 0
    assert abs(s.depth_threshold - 0.5) < 1e-7 and abs(s.dilation - 0.3) < 1e-7 and s.filter3d == 0
    lens = _abi.Lens()
    _abi.lib.lrt_lens_default(C.byref(lens))
    assert (lens.focal, lens.haperture, lens.vaperture, lens.near_z, lens.far_z) == (50.0, 24.576, 18.672, 0.1,
                                                                                     10000.0)
    e = _abi.SplatEdit()
    _abi.lib.lrt_splat_edit_default(C.byref(e))
    assert (e.active, e.mode, e.opacity, e.saturation) == (0, 2, 1.0, 1.0)


def test_a_context_needs_a_device():
    ctx = C.c_void_p()
    code = _abi.lib.lrt_context_create(0, None, C.byref(ctx))
    if code == _abi.LRT_OK:          # a machine with a GPU
        _abi.lib.lrt_context_destroy(ctx)
        return
    assert code == 6 and not ctx.value                          # DeviceFailure
    assert _abi.lib.lrt_last_error(None).decode()


def test_calls_without_a_context_are_invalid_arguments():
    lib = _abi.lib
    p = _abi.ProjectionC()
    s = _abi.RenderSettingsC()
    t = _abi.RenderTargetsC()
    assert lib.lrt_raster_render(None, C.byref(p), None, 0, C.byref(s), C.byref(t), None) == 3
    assert lib.lrt_raster_debug(None, 0, None, 0, None) == 3
    assert lib.lrt_prefix_sum(None, None, None, None, 0) == 3
    assert lib.lrt_radix_sort(None, None, None, None, None, 0, 8, None) == 3
    assert lib.lrt_context_synchronize(None) == 3
    assert lib.lrt_projection_for(None, 1, 1, C.byref(p)) == 3
    assert lib.lrt_raster_render_views(None, None, 0, None, 0, C.byref(s), None) == 3
    assert lib.lrt_autotune_views(None, None, 0, None, 0, C.byref(s), None) == 3
    assert lib.lrt_reference_render(None, C.byref(p), None, 0, C.byref(s), None, None, 0) == 3
    d = _abi.DiffSplatsC()
    assert lib.lrt_raster_render_diff(None, C.byref(p), C.byref(d), None, C.byref(s), C.byref(t), None) == 3
    assert lib.lrt_raster_backward(None, None, C.byref(_abi.DiffGradsC()), None) == 3


def test_the_static_prefix_sum_refuses_bad_arguments_before_any_device_call():
    lib = _abi.lib
    a, b, t = C.c_void_p(4096), C.c_void_p(8192), C.c_void_p(12288)   # never dereferenced: refused first
    assert lib.lrt_prefix_sum_static(None, b, t, 16, None) == 3
    assert lib.lrt_prefix_sum_static(a, a, t, 16, None) == 3
    assert "differ" in lib.lrt_last_error(None).decode()
    assert lib.lrt_prefix_sum_static(a, b, t, (1 << 28) + 1, None) == 3
    msg = lib.lrt_last_error(None).decode()
    assert "exceeds" in msg and str(1 << 28) in msg


def test_frame_params_need_no_device():
    cam = raster.Camera.looking_at((1.0, 2.0, 3.0), (0.0, 0.0, 0.0))
    p = raster.projection_for(cam, 640, 480)
    blob = raster.frame_params_bytes(p, raster.SplatInstance(None), raster.RenderSettings(width=640, height=480))
    words = np.frombuffer(blob, np.uint32)
    assert tuple(words[:4]) == (640, 480, 40, 30)


def test_synthetic_scenes_are_seeded_and_prefix_stable():
    from lucabrtrender_amd import scenes
    a = scenes.synth(1, 1000, seed=3, degree=3, log_scale=(-5.0, 0.8))
    b = scenes.synth(1, 400, seed=3, degree=3, log_scale=(-5.0, 0.8))
    c = scenes.synth(1, 400, seed=4, degree=3, log_scale=(-5.0, 0.8))
    assert np.array_equal(a.positions[:400], b.positions) and np.array_equal(a.sh[:400], b.sh)
    assert not np.array_equal(b.positions, c.positions)
    assert a.sh.shape == (1000, 23)


# Each ctypes mirror and the header's struct it mirrors.
_MIRRORS = {
    "Lens": "lrt_lens", "CameraC": "lrt_camera", "ProjectionC": "lrt_projection", "Splats": "lrt_splats",
    "SplatEdit": "lrt_splat_edit", "SplatInstanceC": "lrt_splat_instance", "PointsC": "lrt_points",
    "PointStyleC": "lrt_point_style", "PointInstanceC": "lrt_point_instance",
    "RenderSettingsC": "lrt_render_settings", "RenderTargetsC": "lrt_render_targets", "ViewC": "lrt_view",
    "FrameStatsC": "lrt_frame_stats", "DiffSplatsC": "lrt_diff_splats", "DiffGradsC": "lrt_diff_grads",
    "LodBuildSettingsC": "lrt_lod_build_settings", "LodInfoC": "lrt_lod_info",
    "LodInstanceC": "lrt_lod_instance", "CutStatsC": "lrt_cut_stats", "SplatEncodingC": "lrt_splat_encoding",
    "CloudInfoC": "lrt_cloud_info",
}


def test_every_ctypes_mirror_has_the_header_s_layout(tmp_path):
    """Size and every field's offset of each struct, as gcc lays the header
    out, against the ctypes mirror the bindings pass."""
    import shutil
    import subprocess
    from lucabrtrender_amd import _abi
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "lrt_hip.h"', "int main(void) {"]
    expected = []
    for py, c in _MIRRORS.items():
        cls = getattr(_abi, py)
        lines.append(f'printf("%zu\\n", sizeof({c}));')
        expected.append(C.sizeof(cls))
        for name, *_ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({c}, {name}));')
            expected.append(getattr(cls, name).offset)
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    include = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([gcc, "-std=c11", "-I", include, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expected
