"""The C-ABI library: loads, exports every symbol include/mm.h declares, and rejects bad arguments
before touching the device (these cases never launch a kernel, so they run without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "mm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(mm):
    lib = mm.lib()
    names = _declared_symbols()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(mm._SIGS), "python bindings must cover exactly the header"


def test_version_and_errors(mm):
    lib = mm.lib()
    assert lib.mm_abi_version() == 1
    assert lib.mm_strerror(0) == b"ok"
    assert lib.mm_strerror(mm.MM_E_PITCH) == b"row pitch smaller than width"


def test_texture_bytes(mm):
    assert mm.texture_bytes(128, 128, 128, mm.FMT_RGBA8, mm.LAYOUT_LINEAR) == 128 ** 3 * 4
    assert mm.texture_bytes(128, 128, 128, mm.FMT_RGBA8, mm.LAYOUT_APRON) == 129 ** 3 * 4
    assert mm.texture_bytes(32, 32, 32, mm.FMT_RGBA8, mm.LAYOUT_APRON) == 33 ** 3 * 4
    assert mm.texture_bytes(32, 32, 32, 4, mm.LAYOUT_APRON) == 0  # format codes 3 and 4 are not assigned
    assert mm.texture_bytes(512, 512, 1, mm.FMT_RG8, mm.LAYOUT_APRON) == 513 * 513 * 2
    assert mm.texture_bytes(128, 128, 1, mm.FMT_FP16X4, mm.LAYOUT_APRON) == 129 * 129 * 8
    assert mm.texture_bytes(8, 8, 1, 99, mm.LAYOUT_LINEAR) == 0
    assert mm.texture_bytes(8, 8, 1, mm.FMT_RGBA8, 7) == 0


def test_presets(mm):
    r = mm.march_params("reference")
    assert ctypes.sizeof(r) == r.size
    # compute-clouds.comp constants
    assert (r.inner_diameter, r.outer_diameter, r.center_y, r.thickness, r.proj_radius) == (
        2e6, 2.04e6, -995000.0, 20000.0, 1e6)
    assert (r.max_steps, r.miss_limit, r.subtexel_bits, r.flags) == (100, 10, 8, 0)
    s = mm.march_params("scenario")
    assert s.thickness == 2500.0 and s.center_y == -6371000.0 and s.proj_radius == 6372500.0
    assert s.inner_diameter == 2 * 6372500.0 and s.outer_diameter == 2 * 6375000.0
    assert (s.max_coarse_steps, s.max_fine_steps, s.max_steps) == (128, 256, -1)
    assert s.flags == mm.MARCH_ISECT_FP64 | mm.MARCH_BOUNDED


class _FakeTex:
    """Texture descriptors with a dummy address: only used where validation must fail first."""

    def __init__(self, mm, w, h, d, fmt, layout, nbytes=None):
        n = mm.texture_bytes(w, h, d, fmt, layout) if nbytes is None else nbytes
        self.desc = mm.Texture(0x1000, n, w, h, d, fmt, layout, 0, (ctypes.c_float * 4)(1, 1, 1, 1))


def _noise(mm, **over):
    t = {"weather": _FakeTex(mm, 512, 512, 1, mm.FMT_RGBA8, 0), "curl": _FakeTex(mm, 128, 128, 1, mm.FMT_RGBA8, 0),
         "low_res": _FakeTex(mm, 128, 128, 128, mm.FMT_RGBA8, 0), "hi_res": _FakeTex(mm, 32, 32, 32, mm.FMT_RGBA8, 0)}
    t.update(over)
    return mm.NoiseSetC(t["weather"].desc, t["curl"].desc, t["low_res"].desc, t["hi_res"].desc)


def _march(mm, noise=None, width=64, height=64, pitch=64, phase=0, params=None, cam=None, sun=None, sky=None,
           out=0x2000):
    c, s, k = mm.build_ubos(mm.scenario_scene())
    cam, sun, sky = cam or c, sun or s, sky or k
    noise = noise or _noise(mm)
    pp = ctypes.byref(params) if params is not None else None
    return mm.lib().mm_march_clouds(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky), ctypes.byref(noise),
                                    None, out, width, height, pitch, phase, pp, None)


@pytest.fixture
def cpu_only():
    import torch

    if torch.cuda.is_available():
        pytest.skip("validation-with-dummy-pointers cases run on the CPU only; the GPU suite covers them")


def test_march_rejects_bad_arguments(mm, cpu_only):
    lib = mm.lib()
    assert lib.mm_march_clouds(None, None, None, None, None, None, 64, 64, 64, 0, None, None) == mm.MM_E_NULL
    assert _march(mm, out=None) == mm.MM_E_NULL
    assert _march(mm, width=0) == mm.MM_E_DIM
    assert _march(mm, height=-4) == mm.MM_E_DIM
    assert _march(mm, pitch=63) == mm.MM_E_PITCH
    assert _march(mm, phase=16) == mm.MM_E_PHASE
    assert _march(mm, phase=-3) == mm.MM_E_PHASE
    p = mm.march_params("reference")
    p.size = 4
    assert _march(mm, params=p) == mm.MM_E_VERSION
    p = mm.march_params("reference", subtexel_bits=4)
    assert _march(mm, params=p) == mm.MM_E_PARAM
    p = mm.march_params("reference", thickness=float("nan"))
    assert _march(mm, params=p) == mm.MM_E_PARAM
    p = mm.march_params("reference", fine_step_mul=1.5)
    assert _march(mm, params=p) == mm.MM_E_PARAM
    # NaN uniforms and a zero sun direction
    c, s, k = mm.build_ubos(mm.scenario_scene())
    c.view[3] = float("nan")
    assert _march(mm, cam=c) == mm.MM_E_PARAM
    c, s, k = mm.build_ubos(mm.scenario_scene())
    for i in range(4, 8):
        s.directionBasis[i] = 0.0
    assert _march(mm, sun=s) == mm.MM_E_PARAM
    # phase carried in sun.color.a must be in [0, 16)
    c, s, k = mm.build_ubos(mm.scenario_scene())
    s.color[3] = 16.0
    assert _march(mm, sun=s, phase=mm.PHASE_FROM_UBO) == mm.MM_E_PHASE
    s.color[3] = -1.0
    assert _march(mm, sun=s, phase=mm.PHASE_FROM_UBO) == mm.MM_E_PHASE


def test_march_rejects_bad_textures(mm, cpu_only):
    # unknown format / layout, non power-of-two, non-cubic volume, allocation too small, mixed set
    assert _march(mm, _noise(mm, weather=_FakeTex(mm, 512, 512, 1, 42, 0, nbytes=1 << 20))) == mm.MM_E_FORMAT
    assert _march(mm, _noise(mm, weather=_FakeTex(mm, 500, 500, 1, mm.FMT_RGBA8, 0))) == mm.MM_E_DIM
    assert _march(mm, _noise(mm, low_res=_FakeTex(mm, 128, 128, 64, mm.FMT_RGBA8, 0))) == mm.MM_E_DIM
    assert _march(mm, _noise(mm, hi_res=_FakeTex(mm, 32, 32, 32, mm.FMT_RGBA8, 0, nbytes=100))) == mm.MM_E_SIZE
    assert _march(mm, _noise(mm, hi_res=_FakeTex(mm, 32, 32, 32, mm.FMT_RGBA8, mm.LAYOUT_APRON))) == mm.MM_E_FORMAT
    assert _march(mm, _noise(mm, curl=_FakeTex(mm, 128, 128, 1, mm.FMT_FP8X4, 0))) == mm.MM_E_FORMAT
    # a volume of 4 GiB or more (1024^3 RGBA8 = exactly 4 GiB): the kernels use 32-bit byte offsets
    assert _march(mm, _noise(mm, low_res=_FakeTex(mm, 1024, 1024, 1024, mm.FMT_RGBA8, 0))) == mm.MM_E_DIM
    n = _noise(mm)
    n.weather.data = None
    assert _march(mm, n) == mm.MM_E_NULL


def test_generate_rejects_bad_arguments(mm, cpu_only):
    lib = mm.lib()
    p = mm.noise_params()
    n = _noise(mm)
    assert lib.mm_generate_noise(None, ctypes.byref(n), None, None) == mm.MM_E_NULL
    bad = mm.noise_params()
    bad.size = 3
    assert lib.mm_generate_noise(ctypes.byref(bad), ctypes.byref(n), None, None) == mm.MM_E_VERSION
    bad = mm.noise_params(coverage=1.5)
    assert lib.mm_generate_noise(ctypes.byref(bad), ctypes.byref(n), None, None) == mm.MM_E_PARAM
    bad = mm.noise_params(what=64)
    assert lib.mm_generate_noise(ctypes.byref(bad), ctypes.byref(n), None, None) == mm.MM_E_PARAM
    n2 = _noise(mm, curl=_FakeTex(mm, 1024, 1024, 1, mm.FMT_RGBA8, 0))
    assert lib.mm_generate_noise(ctypes.byref(p), ctypes.byref(n2), None, None) == mm.MM_E_DIM
    n3 = _noise(mm, low_res=_FakeTex(mm, 128, 128, 128, mm.FMT_FP16X4, 0))
    assert lib.mm_generate_noise(ctypes.byref(p), ctypes.byref(n3), None, None) == mm.MM_E_FORMAT
    src = _FakeTex(mm, 32, 32, 32, mm.FMT_RGBA8, 0)
    dst = _FakeTex(mm, 16, 16, 16, mm.FMT_RGBA8, mm.LAYOUT_APRON)
    assert lib.mm_pack_texture(3, ctypes.byref(src.desc), ctypes.byref(dst.desc), None) == mm.MM_E_DIM
    assert lib.mm_pack_texture(9, ctypes.byref(src.desc), ctypes.byref(dst.desc), None) == mm.MM_E_PARAM
