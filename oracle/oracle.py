"""ctypes loader of liboracle.so, the CPU restatement of the cloud path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, always as the checker, never as the thing measured for the GPU path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"] + (["-B"] if force else []), check=True)
    return LIB_PATH


class OrcTextures(ctypes.Structure):
    _fields_ = [("weather", ctypes.c_void_p), ("weather_w", ctypes.c_int), ("weather_h", ctypes.c_int),
                ("curl", ctypes.c_void_p), ("curl_w", ctypes.c_int), ("curl_h", ctypes.c_int),
                ("low", ctypes.c_void_p), ("low_dim", ctypes.c_int),
                ("hi", ctypes.c_void_p), ("hi_dim", ctypes.c_int),
                ("night", ctypes.c_void_p), ("night_w", ctypes.c_int), ("night_h", ctypes.c_int)]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    def __init__(self, path: str | None = None):
        path = path or build()
        L = ctypes.CDLL(path)
        f, d, i, u32, vp = ctypes.c_float, ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p
        sigs = {
            "orc_hash_noise": (f, [f, f, f, i]),
            "orc_perlin": (f, [f, f, f, f, i]),
            "orc_fbm": (f, [f, f, f, f, i, i]),
            "orc_curl_fbm": (None, [f, f, f, i, i, vp]),
            "orc_curl_generate": (None, [i, i, vp, vp, vp]),
            "orc_sin_poly": (d, [d]),
            "orc_set_curl_mutation": (None, [i]),
            "orc_hash3": (u32, [u32, u32, u32, u32, vp]),
            "orc_worley": (f, [f, f, f, i, u32]),
            "orc_perlin_tiled": (f, [f, f, f, i, u32]),
            "orc_worley_fbm": (f, [f, f, f, i, u32]),
            "orc_perlin_fbm": (f, [f, f, f, i, u32]),
            "orc_value2": (f, [f, f, i, u32]),
            "orc_value_fbm": (f, [f, f, i, u32]),
            "orc_set_noise_mutation": (None, [i]),
            "orc_low_res_texel": (None, [i, i, i, i, u32, vp]),
            "orc_hi_res_texel": (None, [i, i, i, i, u32, vp]),
            "orc_weather_texel": (None, [i, i, i, u32, f, vp]),
            "orc_generate_low_res": (None, [i, u32, vp]),
            "orc_generate_hi_res": (None, [i, u32, vp]),
            "orc_generate_weather": (None, [i, u32, f, vp]),
            "orc_quantize_unorm8": (ctypes.c_uint8, [f]),
            "orc_generate_low_res_f32": (None, [i, u32, vp]),
            "orc_generate_hi_res_f32": (None, [i, u32, vp]),
            "orc_sample2d": (None, [vp, i, i, f, f, i, vp]),
            "orc_sample3d": (None, [vp, i, f, f, f, i, vp]),
            "orc_march_f32": (i, [vp, vp, vp, vp, vp, i, i, i, i, i, vp, vp, i]),
            "orc_march_f64": (i, [vp, vp, vp, vp, vp, i, i, i, i, i, vp, vp, i]),
            "orc_march_pixels_f32": (i, [vp, vp, vp, vp, vp, i, i, vp, i, vp, i]),
            "orc_march_pixels_f64": (i, [vp, vp, vp, vp, vp, i, i, vp, i, vp, i]),
            "orc_max_threads": (i, []),
            "orc_march_state_f32": (i, [vp, vp, vp, vp, vp, i, i, vp, vp, i]),
            "orc_march_state_f64": (i, [vp, vp, vp, vp, vp, i, i, vp, vp, i]),
            "orc_march_pixels_state_f64": (i, [vp, vp, vp, vp, vp, i, i, vp, i, vp, vp, i]),
            "orc_march_trace_f32": (i, [vp, vp, vp, vp, vp, i, i, i, i, vp, i, vp, vp]),
            "orc_hg_phase": (f, [f, f]),
            "orc_hg_phase_f64": (d, [d, d]),
            "orc_rayleigh_phase": (f, [f]),
            "orc_layer_density": (f, [f, f]),
            "orc_height_bias_coverage": (f, [f, f]),
            "orc_remap_clamped": (f, [f, f, f, f, f]),
            "orc_pow_rt": (f, [f, f]),
            "orc_ray_sphere_f32": (f, [vp, vp, vp, f, vp]),
            "orc_ray_sphere_f64": (d, [vp, vp, vp, d, vp]),
            "orc_reproject": (i, [vp, vp, vp, vp, i, i, i, vp, i, i]),
            "orc_cloud_shadow": (i, [vp, vp, vp, vp, i, i, vp, i, vp, i, vp, vp]),
            "orc_shadow_layer_density": (f, [f, f]),
            "orc_post_sun_ndc": (None, [vp, vp, vp]),
            "orc_post_uc2": (f, [f]),
            "orc_post_pow": (f, [f, f]),
            "orc_post_process": (i, [vp, vp, vp, i, i, i, vp, vp, vp, vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        self.L = L

    # ---- generators
    def curl(self, dim: int = 128, sin_kind: int = 0):
        rgba = np.zeros((dim, dim, 4), np.uint8)
        raw = np.zeros((dim, dim, 3), np.float32)
        mm = np.zeros(6, np.float32)
        self.L.orc_curl_generate(dim, sin_kind, _ptr(rgba), _ptr(raw), _ptr(mm))
        return rgba, raw, mm

    def low_res(self, dim: int = 128, seed: int = 1) -> np.ndarray:
        a = np.zeros((dim, dim, dim, 4), np.uint8)
        self.L.orc_generate_low_res(dim, seed, _ptr(a))
        return a

    def hi_res(self, dim: int = 32, seed: int = 1) -> np.ndarray:
        a = np.zeros((dim, dim, dim, 4), np.uint8)
        self.L.orc_generate_hi_res(dim, seed, _ptr(a))
        return a

    def low_res_f32(self, dim: int = 128, seed: int = 1) -> np.ndarray:
        a = np.zeros((dim, dim, dim, 4), np.float32)
        self.L.orc_generate_low_res_f32(dim, seed, _ptr(a))
        return a

    def hi_res_f32(self, dim: int = 32, seed: int = 1) -> np.ndarray:
        a = np.zeros((dim, dim, dim, 4), np.float32)
        self.L.orc_generate_hi_res_f32(dim, seed, _ptr(a))
        return a

    def weather(self, dim: int = 512, seed: int = 1, coverage: float = 0.5) -> np.ndarray:
        a = np.zeros((dim, dim, 4), np.uint8)
        self.L.orc_generate_weather(dim, seed, coverage, _ptr(a))
        return a

    # ---- march
    @staticmethod
    def textures(weather, curl, low, hi, night=None):
        for a in (weather, curl, low, hi) + ((night,) if night is not None else ()):
            assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"] and a.shape[-1] == 4
        t = OrcTextures(_ptr(weather).value, weather.shape[1], weather.shape[0],
                        _ptr(curl).value, curl.shape[1], curl.shape[0],
                        _ptr(low).value, low.shape[0], _ptr(hi).value, hi.shape[0],
                        _ptr(night).value if night is not None else None,
                        night.shape[1] if night is not None else 0, night.shape[0] if night is not None else 0)
        t._keep = (weather, curl, low, hi, night)
        return t

    def march(self, cam, sun, sky, tex: OrcTextures, params, width: int, height: int, phase: int = -1,
              rows=None, f64: bool = False, threads: int = 0, counts: bool = False):
        y0, y1 = (0, height) if rows is None else rows
        out = np.zeros((height, width, 4), np.float64 if f64 else np.float32)
        cnt = np.zeros((height, width, 2), np.int32) if counts else None
        fn = self.L.orc_march_f64 if f64 else self.L.orc_march_f32
        r = fn(ctypes.addressof(cam), ctypes.addressof(sun), ctypes.addressof(sky), ctypes.addressof(tex),
               ctypes.addressof(params), width, height, phase, y0, y1, _ptr(out),
               _ptr(cnt) if counts else None, threads)
        if r != 0:
            raise RuntimeError(f"oracle march failed: {r}")
        return (out, cnt) if counts else out

    def march_state(self, cam, sun, sky, tex: OrcTextures, params, width: int, height: int, f64: bool = False,
                    xy=None, threads: int = 0):
        """RGBA and the two transmittance forms per pixel (opacity accumulator after compose, light-energy
        accumulator), for every pixel or (FP64 only) the listed (x, y) pixels."""
        dt = np.float64 if f64 else np.float32
        n = height * width if xy is None else len(xy)
        out = np.zeros((n, 4), dt)
        state = np.zeros((n, 2), dt)
        a = (ctypes.addressof(cam), ctypes.addressof(sun), ctypes.addressof(sky), ctypes.addressof(tex),
             ctypes.addressof(params), width, height)
        if xy is None:
            fn = self.L.orc_march_state_f64 if f64 else self.L.orc_march_state_f32
            r = fn(*a, _ptr(out), _ptr(state), threads)
        else:
            assert f64, "pixel lists: FP64 only"
            xy = np.ascontiguousarray(xy, dtype=np.int32)
            r = self.L.orc_march_pixels_state_f64(*a, _ptr(xy), len(xy), _ptr(out), _ptr(state), threads)
        if r != 0:
            raise RuntimeError(f"oracle march failed: {r}")
        return out, state

    def reproject(self, cam, cam_prev, src: np.ndarray, params, width: int, height: int) -> np.ndarray:
        """reproject.comp restated: the target image (height, width, 4) float32 from src."""
        src = np.ascontiguousarray(src, dtype=np.float32).reshape(height, width, 4)
        out = np.zeros((height, width, 4), np.float32)
        r = self.L.orc_reproject(ctypes.addressof(cam), ctypes.addressof(cam_prev), _ptr(src), _ptr(out), width,
                                 height, width, ctypes.addressof(params), 0, height)
        if r != 0:
            raise RuntimeError(f"oracle reproject failed: {r}")
        return out

    def post_process(self, cam, sun, image: np.ndarray, params):
        """The post chain restated on an (h, w, 4) float32 image: (god-ray alpha (h, w), radial-blur
        RGBA (h, w, 4), tonemapped RGBA8 (h, w, 4))."""
        img = np.ascontiguousarray(image, np.float32)
        h, w = img.shape[:2]
        alpha = np.zeros((h, w), np.float32)
        hdr = np.zeros((h, w, 4), np.float32)
        out = np.zeros((h, w, 4), np.uint8)
        r = self.L.orc_post_process(ctypes.addressof(cam), ctypes.addressof(sun), _ptr(img), w, h, w, _ptr(alpha),
                                    _ptr(hdr), _ptr(out), ctypes.addressof(params))
        if r != 0:
            raise RuntimeError(f"oracle post chain failed: {r}")
        return alpha, hdr, out

    def post_sun_ndc(self, cam, sun):
        xy = np.zeros(2, np.float32)
        self.L.orc_post_sun_ndc(ctypes.addressof(cam), ctypes.addressof(sun), _ptr(xy))
        return xy

    def cloud_shadow(self, cam, sun, sky, weather: np.ndarray, low: np.ndarray, points: np.ndarray, params):
        """model.frag's shadow march restated: RGBA8 weather (h, w, 4) and low-res (n, n, n, 4)."""
        weather = np.ascontiguousarray(weather, np.uint8)
        low = np.ascontiguousarray(low, np.uint8)
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.zeros(len(pts), np.float32)
        r = self.L.orc_cloud_shadow(ctypes.addressof(cam), ctypes.addressof(sun), ctypes.addressof(sky), _ptr(weather),
                                    weather.shape[1], weather.shape[0], _ptr(low), low.shape[0], _ptr(pts), len(pts),
                                    _ptr(out), ctypes.addressof(params))
        if r != 0:
            raise RuntimeError(f"oracle cloud shadow failed: {r}")
        return out

    def march_pixels(self, cam, sun, sky, tex: OrcTextures, params, width: int, height: int, xy: np.ndarray,
                     f64: bool = False, threads: int = 0):
        xy = np.ascontiguousarray(xy, dtype=np.int32)
        out = np.zeros((len(xy), 4), np.float64 if f64 else np.float32)
        fn = self.L.orc_march_pixels_f64 if f64 else self.L.orc_march_pixels_f32
        r = fn(ctypes.addressof(cam), ctypes.addressof(sun), ctypes.addressof(sky), ctypes.addressof(tex),
               ctypes.addressof(params), width, height, _ptr(xy), len(xy), _ptr(out), threads)
        if r != 0:
            raise RuntimeError(f"oracle march failed: {r}")
        return out


def tonemap(rgb: np.ndarray) -> np.ndarray:
    """Uncharted-2 tonemap of tonemap.frag:11-28 (exposure 0.7, white 50.2, gamma 1/2.2), no vignette.

    Used to normalise HDR march output before per-channel comparison; negative inputs are clamped
    to 0 first (the reference would produce NaN from pow of a negative).
    """
    x = np.maximum(np.asarray(rgb, np.float64), 0.0)

    def uc2(v):
        return ((v * (0.15 * v + 0.1 * 0.5) + 0.2 * 0.02) / (v * (0.15 * v + 0.5) + 0.2 * 0.3)) - 0.02 / 0.3

    c = uc2(0.7 * x) / uc2(50.2)
    return np.power(np.maximum(c, 0.0), 1.0 / 2.2)


def parity_stats(gpu: np.ndarray, ref: np.ndarray) -> dict:
    """Per-channel error after the tonemap (RGB) and raw (alpha), in units of 1/255, and the relative
    error of the raw HDR RGB (|gpu - ref| / max(|ref|, 1e-6), per channel)."""
    g = np.concatenate([tonemap(gpu[..., :3]), np.asarray(gpu[..., 3:4], np.float64)], -1)
    r = np.concatenate([tonemap(ref[..., :3]), np.asarray(ref[..., 3:4], np.float64)], -1)
    err = np.abs(g - r) * 255.0
    per_px = err.reshape(-1, 4).max(1)
    hg, hr = np.asarray(gpu[..., :3], np.float64), np.asarray(ref[..., :3], np.float64)
    rel = np.abs(hg - hr) / np.maximum(np.abs(hr), 1e-6)
    return {
        "frac_within_1": float((per_px <= 1.0).mean()),
        "max": float(per_px.max()) if per_px.size else 0.0,
        "mean": float(err.mean()) if err.size else 0.0,
        "hdr_rel_mean": float(rel.mean()) if rel.size else 0.0,
        "hdr_rel_max": float(rel.max()) if rel.size else 0.0,
        "n": int(per_px.size),
        "nan": int(np.isnan(gpu).sum()),
    }
