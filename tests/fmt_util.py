"""numpy restatement of the texel formats and layouts of include/mm.h, used to check GPU-written textures.

Written independently of the HIP code: FP8 e4m3 encoding goes through a table of all representable
values (nearest, ties to even code) rather than bit manipulation.
"""
import numpy as np

FMT_RGBA8, FMT_RG8, FMT_FP8X4, FMT_FP16X4 = 1, 2, 5, 6
LAYOUT_LINEAR, LAYOUT_APRON = 0, 1


def apron(a: np.ndarray, spatial_dims: int) -> np.ndarray:
    """Append one texel per spatial axis that repeats texel 0 (REPEAT wrap made explicit)."""
    for ax in range(spatial_dims):
        first = np.take(a, [0], axis=ax)
        a = np.concatenate([a, first], axis=ax)
    return a


def e4m3_table():
    codes = np.arange(0x7F)  # 0x7F is NaN in OCP e4m3
    e = codes >> 3
    m = codes & 7
    vals = np.where(e == 0, m / 8.0 * 2.0 ** -6, (1 + m / 8.0) * 2.0 ** (e.astype(np.float64) - 7))
    return vals


_E4M3 = e4m3_table()


def fp8_e4m3_encode(v: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even OCP e4m3 of non-negative values, saturating at 448."""
    v = np.asarray(v, np.float64)
    v = np.clip(v, 0.0, 448.0)
    hi = np.searchsorted(_E4M3, v, side="left").clip(0, len(_E4M3) - 1)
    lo = (hi - 1).clip(0)
    dlo = np.abs(v - _E4M3[lo])
    dhi = np.abs(_E4M3[hi] - v)
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (hi % 2 == 0))
    return np.where(pick_hi, hi, lo).astype(np.uint8)


def fp8_e4m3_decode(c: np.ndarray) -> np.ndarray:
    return _E4M3[np.asarray(c, np.int64)]


def pack_weather(rgba: np.ndarray, fmt: int, layout: int) -> bytes:
    a = rgba if fmt == FMT_RGBA8 else np.stack([rgba[..., 0], rgba[..., 2]], -1)
    if layout == LAYOUT_APRON:
        a = apron(a, 2)
    return np.ascontiguousarray(a).tobytes()


def pack_curl(rgba: np.ndarray, layout: int) -> bytes:
    a = apron(rgba, 2) if layout == LAYOUT_APRON else rgba
    return np.ascontiguousarray(a).tobytes()


def pack_low(rgba: np.ndarray, fmt: int, layout: int, f32: np.ndarray | None = None, inv_scale=None) -> bytes:
    if fmt == FMT_RGBA8:
        a = rgba
    elif fmt == FMT_FP8X4:
        a = fp8_e4m3_encode(np.asarray(f32, np.float64) * np.asarray(inv_scale, np.float64))
    else:
        raise ValueError(fmt)
    if layout == LAYOUT_APRON:
        a = apron(a, 3)
    return np.ascontiguousarray(a).tobytes()


def pack_hi(rgba: np.ndarray, fmt: int, layout: int, f32: np.ndarray | None = None) -> bytes:
    if fmt == FMT_RGBA8:
        a = rgba
    elif fmt == FMT_FP16X4:
        h = np.zeros(f32.shape, np.float16)
        h[..., :3] = f32[..., :3].astype(np.float16)
        a = h
    else:
        raise ValueError(fmt)
    if layout == LAYOUT_APRON:
        a = apron(a, 3)
    return np.ascontiguousarray(a).tobytes()
