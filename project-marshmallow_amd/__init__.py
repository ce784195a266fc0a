"""Host-side mirror of the reference's cloud compute operator over the C ABI of libmm.so.

The reference binds compute-clouds.comp through ``ComputeShader`` (SkyEngine/SkyEngine/Shader.h:286-377):
``updateUniformBuffers(cam, camPrev, sky, sun)`` copies the uniform blocks (Shader.cpp:967-992) and
``bindShader`` + ``vkCmdDispatch`` record one 1/16-pixel march into a ping-ponged RGBA32F storage
image (Shader.h:358-376, VulkanApplication.cpp:1061-1071). :class:`ComputeShader` below keeps those
names and that behaviour on top of ``mm_march_clouds``; the uniform producers ``SkyManager`` /
``Camera`` (SkyManager.cpp, camera.cpp) are reached through ``mm_build_ubos``.

Device memory and streams come from PyTorch (plumbing only); every computation runs in libmm.so.
Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MM_LIBRARY", os.path.join(_HERE, "libmm.so"))

# ---------------------------------------------------------------- constants (include/mm.h)
MM_OK = 0
ERRORS = {
    -1: "MM_E_NULL", -2: "MM_E_DIM", -3: "MM_E_PHASE", -4: "MM_E_PITCH", -5: "MM_E_FORMAT",
    -6: "MM_E_PARAM", -7: "MM_E_DEVICE", -8: "MM_E_SIZE", -9: "MM_E_VERSION", -10: "MM_E_IO",
}
(MM_E_NULL, MM_E_DIM, MM_E_PHASE, MM_E_PITCH, MM_E_FORMAT, MM_E_PARAM, MM_E_DEVICE, MM_E_SIZE, MM_E_VERSION,
 MM_E_IO) = range(-1, -11, -1)

FMT_RGBA8, FMT_RG8, FMT_FP8X4, FMT_FP16X4 = 1, 2, 5, 6
LAYOUT_LINEAR, LAYOUT_APRON = 0, 1
GEN_WEATHER, GEN_CURL, GEN_LOW_RES, GEN_HI_RES, GEN_ALL = 1, 2, 4, 8, 15
STATIC_CURL_MAX = 256  # MM_STATIC_CURL_MAX: largest curl map of mm_generate_noise_noalloc
MARCH_ISECT_FP64, MARCH_BOUNDED, MARCH_NO_BATCH = 1, 2, 4
CONE_PLAIN, CONE_SURROGATE = 0, 1
PHASE_ALL, PHASE_FROM_UBO = -1, -2
TEXEL_BYTES = {FMT_RGBA8: 4, FMT_RG8: 2, FMT_FP8X4: 4, FMT_FP16X4: 8}

# Texel format sets the march kernel runs on: (weather, curl, low-res, hi-res) formats + layout.
FORMAT_SETS = {
    "ref": ((FMT_RGBA8, FMT_RGBA8, FMT_RGBA8, FMT_RGBA8), LAYOUT_LINEAR),
    "fast": ((FMT_RG8, FMT_RGBA8, FMT_RGBA8, FMT_RGBA8), LAYOUT_APRON),
    "prec": ((FMT_RG8, FMT_FP16X4, FMT_FP8X4, FMT_FP16X4), LAYOUT_APRON),
}


class MMError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {ERRORS.get(code, code)} ({code})")
        self.code = code


# ---------------------------------------------------------------- ABI structs
class CameraUBO(ctypes.Structure):
    _fields_ = [("view", ctypes.c_float * 16), ("proj", ctypes.c_float * 16),
                ("cameraPosition", ctypes.c_float * 4), ("cameraParams", ctypes.c_float * 4)]


class SunUBO(ctypes.Structure):
    _fields_ = [("location", ctypes.c_float * 4), ("direction", ctypes.c_float * 4), ("color", ctypes.c_float * 4),
                ("directionBasis", ctypes.c_float * 16), ("intensity", ctypes.c_float), ("_pad", ctypes.c_float * 3)]


class SkyUBO(ctypes.Structure):
    _fields_ = [("betaR", ctypes.c_float * 4), ("betaV", ctypes.c_float * 4), ("wind", ctypes.c_float * 4),
                ("mie_directional", ctypes.c_float), ("_pad", ctypes.c_float * 3)]


class Texture(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("bytes", ctypes.c_size_t), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("depth", ctypes.c_int32), ("format", ctypes.c_int32),
                ("layout", ctypes.c_int32), ("_pad", ctypes.c_int32), ("scale", ctypes.c_float * 4)]


class NoiseSetC(ctypes.Structure):
    _fields_ = [("weather", Texture), ("curl", Texture), ("low_res", Texture), ("hi_res", Texture)]


class NoiseParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("what", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("weather_seed", ctypes.c_uint32), ("coverage", ctypes.c_float), ("_reserved", ctypes.c_int32 * 7)]


class MarchParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("inner_diameter", ctypes.c_float), ("outer_diameter", ctypes.c_float), ("center_y", ctypes.c_float),
                ("thickness", ctypes.c_float), ("proj_radius", ctypes.c_float),
                ("coarse_step_frac", ctypes.c_float), ("fine_step_mul", ctypes.c_float), ("miss_limit", ctypes.c_int32),
                ("opacity_cutoff", ctypes.c_float), ("max_steps", ctypes.c_int32),
                ("max_coarse_steps", ctypes.c_int32), ("max_fine_steps", ctypes.c_int32),
                ("wind_strength", ctypes.c_float), ("weather_scale", ctypes.c_float), ("low_scale", ctypes.c_float),
                ("curl_scale", ctypes.c_float), ("hi_scale", ctypes.c_float), ("subtexel_bits", ctypes.c_int32),
                ("cone_mode", ctypes.c_int32), ("surrogate_threshold", ctypes.c_float),
                ("surrogate", ctypes.c_void_p), ("tile_width", ctypes.c_int32), ("lanes_per_ray", ctypes.c_int32),
                ("_reserved", ctypes.c_int32 * 6)]


class Scene(ctypes.Structure):
    _fields_ = [("camera_pos", ctypes.c_float * 3), ("pitch", ctypes.c_float), ("yaw", ctypes.c_float),
                ("fov_deg", ctypes.c_float), ("aspect", ctypes.c_float), ("sun_elevation", ctypes.c_float),
                ("sun_azimuth", ctypes.c_float), ("time", ctypes.c_float), ("wind", ctypes.c_float * 3),
                ("rayleigh", ctypes.c_float), ("mie", ctypes.c_float), ("mie_directional", ctypes.c_float),
                ("turbidity", ctypes.c_float), ("phase", ctypes.c_int32)]


class ShadowParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("inner_diameter", ctypes.c_float),
                ("center_y", ctypes.c_float), ("thickness", ctypes.c_float), ("proj_radius", ctypes.c_float),
                ("step", ctypes.c_float), ("steps", ctypes.c_int32), ("origin_scale", ctypes.c_float),
                ("wind_strength", ctypes.c_float), ("weather_scale", ctypes.c_float), ("low_scale", ctypes.c_float),
                ("subtexel_bits", ctypes.c_int32), ("_reserved", ctypes.c_int32 * 7)]


class PostParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("subtexel_bits", ctypes.c_int32),
                ("god_ray_samples", ctypes.c_int32), ("god_ray_density", ctypes.c_float),
                ("god_ray_decay", ctypes.c_float), ("god_ray_exposure", ctypes.c_float), ("blur_scale", ctypes.c_float),
                ("blur_gain", ctypes.c_float), ("exposure", ctypes.c_float), ("inv_gamma", ctypes.c_float),
                ("white_point", ctypes.c_float), ("vignette", ctypes.c_float * 3), ("_reserved", ctypes.c_int32 * 7)]


class View(ctypes.Structure):
    _fields_ = [("camera", CameraUBO), ("sun", SunUBO), ("sky", SkyUBO), ("weather", Texture),
                ("out_rgba32f", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("row_pitch_px", ctypes.c_int32), ("phase", ctypes.c_int32)]


# Every symbol include/mm.h declares, with its ctypes signature.
_P = ctypes.POINTER
_SIGS = {
    "mm_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "mm_abi_version": (ctypes.c_int, []),
    "mm_texture_bytes": (ctypes.c_size_t, [ctypes.c_int32] * 5),
    "mm_noise_params_default": (None, [_P(NoiseParams)]),
    "mm_noise_workspace_bytes": (ctypes.c_size_t, [_P(NoiseSetC)]),
    "mm_generate_noise": (ctypes.c_int, [_P(NoiseParams), _P(NoiseSetC), ctypes.c_void_p, ctypes.c_void_p]),
    "mm_generate_noise_noalloc": (ctypes.c_int, [_P(NoiseParams), _P(NoiseSetC), ctypes.c_void_p]),
    "mm_pack_texture": (ctypes.c_int, [ctypes.c_int32, _P(Texture), _P(Texture), ctypes.c_void_p]),
    "mm_march_params_reference": (None, [_P(MarchParams)]),
    "mm_march_params_scenario": (None, [_P(MarchParams)]),
    "mm_march_clouds": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(NoiseSetC), _P(Texture),
                                       ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       _P(MarchParams), ctypes.c_void_p]),
    "mm_march_clouds_counted": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(NoiseSetC), _P(Texture),
                                               ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, _P(MarchParams), ctypes.c_void_p, ctypes.c_void_p]),
    "mm_march_clouds_counted_ex": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(NoiseSetC), _P(Texture),
                                                  ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.c_int32, _P(MarchParams), ctypes.c_void_p, ctypes.c_void_p]),
    "mm_march_clouds_state": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(NoiseSetC), _P(Texture),
                                             ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, _P(MarchParams), ctypes.c_void_p, ctypes.c_void_p]),
    "mm_reproject": (ctypes.c_int, [_P(CameraUBO), _P(CameraUBO), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_int32, _P(MarchParams), ctypes.c_void_p]),
    "mm_render_frame": (ctypes.c_int, [_P(CameraUBO), _P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(NoiseSetC),
                                       _P(Texture), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, _P(MarchParams), ctypes.c_void_p]),
    "mm_shadow_params_reference": (None, [_P(ShadowParams)]),
    "mm_cloud_shadow": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(Texture), _P(Texture),
                                       ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, _P(ShadowParams),
                                       ctypes.c_void_p]),
    "mm_post_params_reference": (None, [_P(PostParams)]),
    "mm_post_process": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       _P(PostParams), ctypes.c_void_p]),
    "mm_image_info": (ctypes.c_int, [ctypes.c_char_p, _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "mm_load_image_rgba8": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t, _P(ctypes.c_int32),
                                           _P(ctypes.c_int32)]),
    "mm_load_volume_rgba8": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t,
                                            _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "mm_view_plan_create": (ctypes.c_int, [_P(View), ctypes.c_int32, _P(NoiseSetC), _P(MarchParams),
                                           _P(ctypes.c_void_p), ctypes.c_void_p]),
    "mm_view_plan_update": (ctypes.c_int, [ctypes.c_void_p, _P(View), ctypes.c_int32, _P(MarchParams),
                                           ctypes.c_void_p]),
    "mm_view_plan_info": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "mm_march_views": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mm_view_plan_destroy": (None, [ctypes.c_void_p]),
    "mm_march_schedule_create": (ctypes.c_int, [ctypes.c_int32, _P(ctypes.c_void_p)]),
    "mm_march_schedule_destroy": (None, [ctypes.c_void_p]),
    "mm_march_clouds_scheduled": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(NoiseSetC), _P(Texture),
                                                 ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, _P(MarchParams), ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p]),
    "mm_surrogate_bytes": (ctypes.c_size_t, []),
    "mm_surrogate_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "mm_surrogate_fit_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "mm_surrogate_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_float, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "mm_surrogate_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mm_surrogate_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    "mm_surrogate_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "mm_march_clouds_collect": (ctypes.c_int, [_P(CameraUBO), _P(SunUBO), _P(SkyUBO), _P(NoiseSetC), _P(Texture),
                                               ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, _P(MarchParams), ctypes.c_void_p, ctypes.c_uint32,
                                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    "mm_scene_default": (None, [_P(Scene)]),
    "mm_scene_scenario": (None, [_P(Scene)] + [ctypes.c_float] * 6),
    "mm_build_ubos": (ctypes.c_int, [_P(Scene), _P(CameraUBO), _P(SunUBO), _P(SkyUBO)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libmm.so (built in-tree by ``make`` in this directory). Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmm.so not built: {LIB_PATH} (run __graft_entry__.build())")
        try:
            # One HIP runtime per process: PyTorch-ROCm bundles its own libamdhip64.so.7; loading it
            # first lets libmm.so bind to that same runtime (streams and allocations are shared).
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(code: int, what: str) -> None:
    if code != MM_OK:
        raise MMError(code, what)


# ---------------------------------------------------------------- presets / builders (host only)
def march_params(preset: str = "reference", **overrides) -> MarchParams:
    p = MarchParams()
    if preset == "reference":
        lib().mm_march_params_reference(ctypes.byref(p))
    elif preset == "scenario":
        lib().mm_march_params_scenario(ctypes.byref(p))
    else:
        raise ValueError(preset)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def noise_params(seed: int = 1, weather_seed: int = 1, coverage: float = 0.5, what: int = GEN_ALL) -> NoiseParams:
    p = NoiseParams()
    lib().mm_noise_params_default(ctypes.byref(p))
    p.seed, p.weather_seed, p.coverage, p.what = seed, weather_seed, coverage, what
    return p


def scenario_scene(altitude=200.0, pitch_deg=20.0, fov_deg=60.0, aspect=1.0, sun_elevation_deg=25.0, time=0.0,
                   phase=0) -> Scene:
    s = Scene()
    lib().mm_scene_scenario(ctypes.byref(s), altitude, pitch_deg, fov_deg, aspect, sun_elevation_deg, time)
    s.phase = phase
    return s


def build_ubos(scene: Scene):
    cam, sun, sky = CameraUBO(), SunUBO(), SkyUBO()
    check(lib().mm_build_ubos(ctypes.byref(scene), ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky)),
          "mm_build_ubos")
    return cam, sun, sky


def texture_bytes(w: int, h: int, d: int, fmt: int, layout: int) -> int:
    return int(lib().mm_texture_bytes(w, h, d, fmt, layout))


# ---------------------------------------------------------------- device textures
@dataclass
class DeviceTexture:
    tensor: object          # torch.uint8 device tensor holding the texels
    desc: Texture


def device_texture(w: int, h: int, d: int, fmt: int, layout: int, device="cuda", scale=(1.0, 1.0, 1.0, 1.0),
                   fill: int | None = None) -> DeviceTexture:
    import torch

    n = texture_bytes(w, h, d, fmt, layout)
    if n == 0:
        raise MMError(MM_E_FORMAT, "texture_bytes")
    t = torch.empty(n, dtype=torch.uint8, device=device)
    if fill is not None:
        t.fill_(fill)
    desc = Texture(t.data_ptr(), n, w, h, d, fmt, layout, 0, (ctypes.c_float * 4)(*scale))
    return DeviceTexture(t, desc)


class NoiseSet:
    """The four march textures (bindings 4, 6, 7, 8 of compute-clouds.comp) in one format set."""

    def __init__(self, fset: str = "fast", weather_dim=512, curl_dim=128, low_dim=128, hi_dim=32, device="cuda",
                 fp8_scale=1.0 / 448.0):
        fmts, layout = FORMAT_SETS[fset]
        self.fset = fset
        low_scale = (fp8_scale,) * 4 if fmts[2] == FMT_FP8X4 else (1.0,) * 4
        self.weather = device_texture(weather_dim, weather_dim, 1, fmts[0], layout, device)
        self.curl = device_texture(curl_dim, curl_dim, 1, fmts[1], layout, device)
        self.low_res = device_texture(low_dim, low_dim, low_dim, fmts[2], layout, device, scale=low_scale)
        self.hi_res = device_texture(hi_dim, hi_dim, hi_dim, fmts[3], layout, device)
        self.c = NoiseSetC(self.weather.desc, self.curl.desc, self.low_res.desc, self.hi_res.desc)
        self.workspace = None

    def generate(self, params: NoiseParams, stream: int = 0, workspace=True) -> None:
        import torch

        ws = None
        if workspace:
            if self.workspace is None:
                nbytes = int(lib().mm_noise_workspace_bytes(ctypes.byref(self.c)))
                self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.weather.tensor.device)
            ws = self.workspace.data_ptr()
        check(lib().mm_generate_noise(ctypes.byref(params), ctypes.byref(self.c), ws, stream), "mm_generate_noise")


def current_stream() -> int:
    import torch

    return torch.cuda.current_stream().cuda_stream


class MarchSchedule:
    """Per-tile cost history for mm_march_clouds_scheduled (longest tiles first from the previous
    launch of the same shape); holds device buffers for up to ``max_tiles`` tiles."""

    def __init__(self, max_tiles: int = 1 << 18):
        h = ctypes.c_void_p()
        check(lib().mm_march_schedule_create(max_tiles, ctypes.byref(h)), "mm_march_schedule_create")
        self.ptr = h.value

    def close(self) -> None:
        if self.ptr:
            lib().mm_march_schedule_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def march_clouds(cam: CameraUBO, sun: SunUBO, sky: SkyUBO, noise: NoiseSet, out, width: int, height: int,
                 phase: int = PHASE_ALL, params: MarchParams | None = None, night: DeviceTexture | None = None,
                 pitch: int | None = None, stream: int | None = None, counts=None,
                 schedule: MarchSchedule | None = None) -> None:
    """One cloud dispatch into ``out`` (a float32 device tensor of >= height*pitch*4 elements). ``counts``
    (int64 device tensor of 4 or 6 elements) accumulates the work counters of mm_march_clouds_counted
    (4) or mm_march_clouds_counted_ex (6: also the weather-map and low-res reads). ``schedule``: march
    through mm_march_clouds_scheduled (4 counters at most)."""
    pitch = width if pitch is None else pitch
    stream = current_stream() if stream is None else stream
    pp = ctypes.byref(params) if params is not None else None
    nt = ctypes.byref(night.desc) if night is not None else None
    if schedule is not None:
        if counts is not None and counts.numel() != 4:
            raise ValueError("a scheduled march takes 4 counters")
        code = lib().mm_march_clouds_scheduled(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky),
                                               ctypes.byref(noise.c), nt, out.data_ptr(), width, height, pitch, phase,
                                               pp, schedule.ptr, counts.data_ptr() if counts is not None else None,
                                               stream)
        check(code, "mm_march_clouds_scheduled")
        return
    if counts is not None and counts.numel() >= 6:
        code = lib().mm_march_clouds_counted_ex(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky),
                                                ctypes.byref(noise.c), nt, out.data_ptr(), width, height, pitch,
                                                phase, pp, counts.data_ptr(), stream)
        check(code, "mm_march_clouds_counted_ex")
    elif counts is None:
        code = lib().mm_march_clouds(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky), ctypes.byref(noise.c), nt,
                                     out.data_ptr(), width, height, pitch, phase, pp, stream)
        check(code, "mm_march_clouds")
    else:
        code = lib().mm_march_clouds_counted(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky),
                                             ctypes.byref(noise.c), nt, out.data_ptr(), width, height, pitch, phase,
                                             pp, counts.data_ptr(), stream)
        check(code, "mm_march_clouds_counted")


def march_clouds_state(cam: CameraUBO, sun: SunUBO, sky: SkyUBO, noise: NoiseSet, out, state, width: int,
                       height: int, phase: int = PHASE_ALL, params: MarchParams | None = None,
                       pitch: int | None = None, stream: int | None = None) -> None:
    """The march with each pixel's (final opacity accumulator, light-energy accumulator) written into
    ``state`` (float32 device tensor of >= height*pitch*2 elements) beside its colour in ``out``."""
    pitch = width if pitch is None else pitch
    stream = current_stream() if stream is None else stream
    pp = ctypes.byref(params) if params is not None else None
    code = lib().mm_march_clouds_state(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky), ctypes.byref(noise.c),
                                       None, out.data_ptr(), width, height, pitch, phase, pp, state.data_ptr(), stream)
    check(code, "mm_march_clouds_state")


class ComputeShader:
    """Mirror of the reference's ComputeShader operator (Shader.h:286-377) for compute-clouds.comp.

    Holds the two ping-ponged RGBA32F storage images (storageBufferSetA/B) and the uniform copies;
    ``bind_shader`` swaps the images like Shader.h:358-376, ``dispatch`` records the march of the
    pixel phase carried in ``sun.color.a`` (comp:292) into the current target image.
    """

    def __init__(self, width: int, height: int, noise: NoiseSet, night: DeviceTexture | None = None,
                 params: MarchParams | None = None, device="cuda"):
        import torch

        self.width, self.height = width, height
        self.noise, self.night, self.params = noise, night, params
        self.storage_a = torch.zeros(height * width * 4, dtype=torch.float32, device=device)
        self.storage_b = torch.zeros(height * width * 4, dtype=torch.float32, device=device)
        self.swapped_buffers = False
        self.cam, self.cam_prev, self.sky, self.sun = CameraUBO(), CameraUBO(), SkyUBO(), SunUBO()

    def update_uniform_buffers(self, cam: CameraUBO, cam_prev: CameraUBO, sky: SkyUBO, sun: SunUBO) -> None:
        # Shader.cpp:967-992 memcpy's each block into its host-coherent UBO.
        ctypes.memmove(ctypes.byref(self.cam), ctypes.byref(cam), ctypes.sizeof(CameraUBO))
        ctypes.memmove(ctypes.byref(self.cam_prev), ctypes.byref(cam_prev), ctypes.sizeof(CameraUBO))
        ctypes.memmove(ctypes.byref(self.sky), ctypes.byref(sky), ctypes.sizeof(SkyUBO))
        ctypes.memmove(ctypes.byref(self.sun), ctypes.byref(sun), ctypes.sizeof(SunUBO))

    def bind_shader(self):
        target = self.storage_b if self.swapped_buffers else self.storage_a
        self.swapped_buffers = not self.swapped_buffers
        return target

    def dispatch(self, stream: int | None = None):
        target = self.bind_shader()
        march_clouds(self.cam, self.sun, self.sky, self.noise, target, self.width, self.height,
                     phase=PHASE_FROM_UBO, params=self.params, night=self.night, stream=stream)
        return target


def reproject(cam: CameraUBO, cam_prev: CameraUBO, source, target, width: int, height: int,
              params: MarchParams | None = None, pitch: int | None = None, stream: int | None = None) -> None:
    """reproject.comp's pass: target (float32 device tensor) from the previous frame source."""
    pitch = width if pitch is None else pitch
    stream = current_stream() if stream is None else stream
    pp = ctypes.byref(params) if params is not None else None
    check(lib().mm_reproject(ctypes.byref(cam), ctypes.byref(cam_prev), source.data_ptr(), target.data_ptr(), width,
                             height, pitch, pp, stream), "mm_reproject")


def shadow_params(**overrides) -> ShadowParams:
    p = ShadowParams()
    lib().mm_shadow_params_reference(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def cloud_shadow(cam: CameraUBO, sun: SunUBO, sky: SkyUBO, noise: NoiseSet, points, out,
                 params: ShadowParams | None = None, stream: int | None = None) -> None:
    """model.frag's shadow march for device points [n, 3] (float32) into out [n]."""
    stream = current_stream() if stream is None else stream
    pp = ctypes.byref(params) if params is not None else None
    check(lib().mm_cloud_shadow(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky), ctypes.byref(noise.weather.desc),
                                ctypes.byref(noise.low_res.desc), points.data_ptr(), int(points.shape[0]),
                                out.data_ptr(), pp, stream), "mm_cloud_shadow")


def post_params(**overrides) -> PostParams:
    p = PostParams()
    lib().mm_post_params_reference(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def post_process(cam: CameraUBO, sun: SunUBO, image, width: int, height: int, god_alpha, hdr=None, rgba8=None,
                 pitch: int | None = None, params: PostParams | None = None, stream: int | None = None) -> None:
    """God rays -> radial blur -> tonemap of a device RGBA32F cloud image (mm_post_process).
    god_alpha: device float32 scratch [2 * pitch * height] (plane 1 = framebuffer 1's alpha);
    hdr: optional float32 [pitch * height * 4]; rgba8: optional uint8 [pitch * height * 4]."""
    stream = current_stream() if stream is None else stream
    pitch = width if pitch is None else pitch
    pp = ctypes.byref(params) if params is not None else None
    check(lib().mm_post_process(ctypes.byref(cam), ctypes.byref(sun), image.data_ptr(), width, height, pitch,
                                god_alpha.data_ptr(), hdr.data_ptr() if hdr is not None else None,
                                rgba8.data_ptr() if rgba8 is not None else None, pp, stream), "mm_post_process")


def load_image(path: str):
    """Decode a TGA or PNG file to an (h, w, 4) uint8 array (mm_load_image_rgba8: stb_image's
    STBI_rgb_alpha semantics, rows top first)."""
    import numpy as np

    w, h = ctypes.c_int32(), ctypes.c_int32()
    check(lib().mm_image_info(path.encode(), ctypes.byref(w), ctypes.byref(h)), "mm_image_info")
    a = np.zeros((h.value, w.value, 4), np.uint8)
    check(lib().mm_load_image_rgba8(path.encode(), a.ctypes.data, a.nbytes, ctypes.byref(w), ctypes.byref(h)),
          "mm_load_image_rgba8")
    return a


def load_volume(prefix: str, depth: int):
    """Decode the TGA slices "<prefix>(<i>).tga", i < depth, to a (depth, h, w, 4) uint8 array
    (mm_load_volume_rgba8, the reference's Texture3D::initFromFile layout)."""
    import numpy as np

    w, h = ctypes.c_int32(), ctypes.c_int32()
    check(lib().mm_image_info(f"{prefix}(0).tga".encode(), ctypes.byref(w), ctypes.byref(h)), "mm_image_info")
    a = np.zeros((depth, h.value, w.value, 4), np.uint8)
    check(lib().mm_load_volume_rgba8(prefix.encode(), depth, a.ctypes.data, a.nbytes, ctypes.byref(w),
                                     ctypes.byref(h)), "mm_load_volume_rgba8")
    return a


def upload_rgba8(texture: "DeviceTexture", rgba, kind: int, stream: int | None = None) -> None:
    """Host RGBA8 (h, w, 4) or (d, h, w, 4) into a device texture of any format: copy into an RGBA8
    LINEAR staging texture, then mm_pack_texture (kind 0 weather, 1 curl, 2 low-res, 3 hi-res)."""
    import torch

    stream = current_stream() if stream is None else stream
    d = texture.desc
    stage = device_texture(d.width, d.height, d.depth, FMT_RGBA8, LAYOUT_LINEAR, "cuda")
    stage.tensor.copy_(torch.from_numpy(rgba.reshape(-1)))
    check(lib().mm_pack_texture(kind, ctypes.byref(stage.desc), ctypes.byref(d), stream), "mm_pack_texture")
    torch.cuda.synchronize()


def previous_camera(cam: CameraUBO) -> CameraUBO:
    """The next frame's UniformCameraObjectPrev: this frame's view, proj and position
    (VulkanApplication.cpp:354-358)."""
    p = CameraUBO()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(cam), ctypes.sizeof(CameraUBO))
    return p


class CloudFrames:
    """The reference's per-frame compute work (VulkanApplication.cpp:1046-1071, 372-388): the pixel
    phase advances by one each frame (sun.color.a = (a + 1) % 16), the previous frame is reprojected
    into the current target at full resolution, then 1/16 of the pixels are marched into it; the
    two RGBA32F images swap every frame."""

    def __init__(self, width: int, height: int, noise: NoiseSet, params: MarchParams | None = None,
                 night: DeviceTexture | None = None, device="cuda"):
        import torch

        self.width, self.height, self.noise, self.params, self.night = width, height, noise, params, night
        self.images = [torch.zeros(width * height * 4, dtype=torch.float32, device=device) for _ in range(2)]
        self.cur = 0
        self.phase = -1
        self.cam_prev = None

    def frame(self, cam: CameraUBO, sun: SunUBO, sky: SkyUBO, stream: int | None = None):
        """Render one frame; returns the target image (float32 tensor, height*width*4)."""
        self.phase = (self.phase + 1) % 16
        sun.color[3] = float(self.phase)
        prev = self.cam_prev if self.cam_prev is not None else previous_camera(cam)
        src, dst = self.images[self.cur], self.images[1 - self.cur]
        stream = current_stream() if stream is None else stream
        pp = ctypes.byref(self.params) if self.params is not None else None
        nt = ctypes.byref(self.night.desc) if self.night is not None else None
        check(lib().mm_render_frame(ctypes.byref(cam), ctypes.byref(prev), ctypes.byref(sun), ctypes.byref(sky),
                                    ctypes.byref(self.noise.c), nt, src.data_ptr(), dst.data_ptr(), self.width,
                                    self.height, self.width, pp, stream), "mm_render_frame")
        self.cur = 1 - self.cur
        self.cam_prev = previous_camera(cam)
        return dst


# ---------------------------------------------------------------- learned light-cone surrogate
SURROGATE_FEATURES = 16
SURROGATE_ROW = 17
# parameter layout of the FP32 master copy (offsets in floats): W1[16][64], b1, W2[64][64], b2,
# W3[64][64], b3, W4[64][2], b4[2]
SURROGATE_PARAMS = 9538


class Surrogate:
    """One device model block (parameters, Adam state, feature statistics, BF16 fragments)."""

    def __init__(self, seed: int = 1, device="cuda", stream: int | None = None):
        import torch

        self.block = torch.zeros(int(lib().mm_surrogate_bytes()), dtype=torch.uint8, device=device)
        stream = current_stream() if stream is None else stream
        check(lib().mm_surrogate_init(self.ptr, seed, stream), "mm_surrogate_init")

    @property
    def ptr(self) -> int:
        return self.block.data_ptr()

    def params(self):
        """View of the FP32 master parameters (float32 tensor of SURROGATE_PARAMS)."""
        return self.block[: SURROGATE_PARAMS * 4].view(dtype=__import__("torch").float32)

    def fit_stats(self, samples, n: int | None = None, stream: int | None = None) -> None:
        n = samples.shape[0] if n is None else n
        stream = current_stream() if stream is None else stream
        check(lib().mm_surrogate_fit_stats(self.ptr, samples.data_ptr(), n, stream), "mm_surrogate_fit_stats")

    def train(self, samples, steps: int, batch: int = 8192, lr: float = 1e-3, seed: int = 1, n: int | None = None,
              loss_out=None, stream: int | None = None) -> None:
        n = samples.shape[0] if n is None else n
        stream = current_stream() if stream is None else stream
        lp = loss_out.data_ptr() if loss_out is not None else None
        check(lib().mm_surrogate_train(self.ptr, samples.data_ptr(), n, batch, steps, lr, seed, lp, stream),
              "mm_surrogate_train")

    def load(self, params=None, stats=None, stream: int | None = None) -> None:
        """Set FP32 parameters (float32 device tensor [SURROGATE_PARAMS]) and/or stats ([32])."""
        stream = current_stream() if stream is None else stream
        pp = params.data_ptr() if params is not None else None
        sp = stats.data_ptr() if stats is not None else None
        check(lib().mm_surrogate_load(self.ptr, pp, sp, stream), "mm_surrogate_load")

    def gradient(self, samples, grad_out, loss_out, n: int | None = None, stream: int | None = None) -> None:
        n = samples.shape[0] if n is None else n
        stream = current_stream() if stream is None else stream
        check(lib().mm_surrogate_gradient(self.ptr, samples.data_ptr(), n, grad_out.data_ptr(), loss_out.data_ptr(),
                                          stream), "mm_surrogate_gradient")

    def stats(self):
        """Feature mean[16] and inverse std[16] (float32 tensor view)."""
        import torch

        off = 4 * ((SURROGATE_PARAMS * 4 + 255) // 256 * 256)
        return self.block[off: off + 32 * 4].view(dtype=torch.float32)

    def predict(self, samples, out, n: int | None = None, stream: int | None = None) -> None:
        n = samples.shape[0] if n is None else n
        stream = current_stream() if stream is None else stream
        check(lib().mm_surrogate_predict(self.ptr, samples.data_ptr(), n, out.data_ptr(), stream),
              "mm_surrogate_predict")


def march_collect(cam: CameraUBO, sun: SunUBO, sky: SkyUBO, noise: NoiseSet, out, width: int, height: int,
                  samples, count, keep_one_in: int = 1, seed: int = 1, phase: int = PHASE_ALL,
                  params: MarchParams | None = None, stream: int | None = None) -> None:
    """Plain march that also records surrogate training rows into ``samples`` (float32 [cap, 17])."""
    stream = current_stream() if stream is None else stream
    pp = ctypes.byref(params) if params is not None else None
    code = lib().mm_march_clouds_collect(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky),
                                         ctypes.byref(noise.c), None, out.data_ptr(), width, height, width, phase, pp,
                                         samples.data_ptr(), samples.shape[0], keep_one_in, seed, count.data_ptr(),
                                         stream)
    check(code, "mm_march_clouds_collect")


# ---------------------------------------------------------------- many views in one launch
class ViewPlan:
    """Per-view constants in device memory for one persistent multi-view launch (mm_view_plan)."""

    def __init__(self, views, shared: NoiseSet, params: MarchParams | None = None, stream: int | None = None):
        self.views = (View * len(views))(*views)
        self.n = len(views)
        self.shared = shared            # keeps the shared textures alive
        stream = current_stream() if stream is None else stream
        h = ctypes.c_void_p()
        pp = ctypes.byref(params) if params is not None else None
        check(lib().mm_view_plan_create(self.views, self.n, ctypes.byref(shared.c), pp, ctypes.byref(h), stream),
              "mm_view_plan_create")
        self.handle = h

    def update(self, views, params: MarchParams | None = None, stream: int | None = None) -> None:
        arr = (View * len(views))(*views)
        stream = current_stream() if stream is None else stream
        pp = ctypes.byref(params) if params is not None else None
        check(lib().mm_view_plan_update(self.handle, arr, len(views), pp, stream), "mm_view_plan_update")
        self.views = arr

    def info(self):
        t, g = ctypes.c_int32(), ctypes.c_int32()
        check(lib().mm_view_plan_info(self.handle, ctypes.byref(t), ctypes.byref(g)), "mm_view_plan_info")
        return t.value, g.value

    def launch(self, counts=None, stream: int | None = None) -> None:
        stream = current_stream() if stream is None else stream
        check(lib().mm_march_views(self.handle, counts.data_ptr() if counts is not None else None, stream),
              "mm_march_views")

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().mm_view_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_view(cam: CameraUBO, sun: SunUBO, sky: SkyUBO, weather: DeviceTexture, out, width: int, height: int,
              phase: int = PHASE_ALL, pitch: int | None = None) -> View:
    return View(cam, sun, sky, weather.desc, out.data_ptr(), width, height, width if pitch is None else pitch, phase)
