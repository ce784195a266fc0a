"""Closed-form and brute-force pins of the march restatement's pieces (oracle/march.cpp).

Each expected value is computed independently of the oracle's code path: hand-derived known answers
for the shader's phase functions and layer profiles (compute-clouds.comp:73-86, 193-208), a
weighted-sum trilinear filter for the sampler (Texture.cpp:29-52, 315-338), and high-precision
decimal arithmetic for the ray/shell intersection (comp:147-177).
"""
import ctypes
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import _ptr


@pytest.mark.parametrize("g", [0.0, 0.3, 0.6, 0.8, 0.89])
def test_hg_phase_closed_forms(orc, g):
    # hg(1, g) = (1 + g) / (4 pi (1 - g)^2), hg(-1, g) = (1 - g) / (4 pi (1 + g)^2), hg(c, 0) = 1/(4 pi)
    fwd, back = (1 + g) / (4 * math.pi * (1 - g) ** 2), (1 - g) / (4 * math.pi * (1 + g) ** 2)
    assert orc.L.orc_hg_phase_f64(1.0, g) == pytest.approx(fwd, rel=1e-12)
    assert orc.L.orc_hg_phase_f64(-1.0, g) == pytest.approx(back, rel=1e-12)
    # FP32: 1 - 2g + g^2 cancels near g = 1 (relative error up to ~1e-5 at g = 0.89)
    assert orc.L.orc_hg_phase(1.0, g) == pytest.approx(fwd, rel=1e-5)
    assert orc.L.orc_hg_phase(-1.0, g) == pytest.approx(back, rel=1e-5)
    assert orc.L.orc_hg_phase_f64(0.3, 0.0) == pytest.approx(1 / (4 * math.pi), rel=1e-12)


def test_rayleigh_phase(orc):
    for c in (-1.0, 0.0, 0.5, 1.0):
        assert orc.L.orc_rayleigh_phase(c) == pytest.approx(3 / (16 * math.pi) * (1 + c * c), rel=1e-6)


@pytest.mark.parametrize("h,t,expect", [
    # type 0 -> stratus profile max(0, h/0.1 * (1 - (h-0.2)/0.1))
    (0.0, 0.0, 0.0), (0.05, 0.0, 1.25), (0.1, 0.0, 2.0), (0.2, 0.0, 2.0), (0.3, 0.0, 0.0), (0.7, 0.0, 0.0),
    # type 1 -> cumulus profile max(0, h/0.2 * (1 - (h-0.7)/0.2))
    (0.1, 1.0, 2.0), (0.5, 1.0, 5.0), (0.7, 1.0, 3.5), (0.9, 1.0, 0.0),
    # type 0.5 -> stratocumulus profile max(0, h/0.2 * (1 - (h-0.2)/0.5))
    (0.2, 0.5, 1.0), (0.45, 0.5, 1.125), (0.7, 0.5, 0.0),
    # clamped height
    (1.5, 1.0, 0.0), (-0.2, 0.5, 0.0),
])
def test_cloud_layer_density_known_answers(orc, h, t, expect):
    assert orc.L.orc_layer_density(h, t) == pytest.approx(expect, abs=2e-6)


def test_cloud_layer_density_vanishes_outside_the_profiles(orc):
    # the march skips its texture reads where every profile is 0 (h = 0 and h >= 0.9, the cumulus
    # top), which is only exact if the layer density is 0 there for every cloud type
    hs = np.arange(np.float32(0.9).view(np.int32), np.float32(1.0).view(np.int32) + 1, 4099).view(np.float32)
    for t in np.linspace(0.0, 1.0, 33, dtype=np.float32):
        assert orc.L.orc_layer_density(0.0, float(t)) == 0.0
        assert max(orc.L.orc_layer_density(float(h), float(t)) for h in hs) == 0.0
    # and not beyond: just below the cumulus top a cumulus-type cloud is still dense
    assert orc.L.orc_layer_density(0.89, 1.0) > 1e-3


@pytest.mark.parametrize("cov,height,exp", [(0.5, 0.5, 1.0), (0.5, 0.7, 1.0), (0.5, 0.75, 0.9), (0.5, 0.8, 0.8),
                                            (0.3, 0.85, 0.8), (0.0, 0.9, 0.8)])
def test_height_bias_coverage(orc, cov, height, exp):
    # pow(coverage, clamp(remap(height, .7, .8, 1, .8), .8, 1)) (comp:206-208)
    assert orc.L.orc_height_bias_coverage(cov, height) == pytest.approx(cov ** exp, rel=1e-5, abs=1e-7)


def test_remap_clamped(orc):
    assert orc.L.orc_remap_clamped(0.65, 0.3, 1.0, 0.0, 1.0) == pytest.approx(0.5, abs=1e-6)
    assert orc.L.orc_remap_clamped(0.1, 0.3, 1.0, 0.0, 1.0) == 0.0
    assert orc.L.orc_remap_clamped(2.0, 0.3, 1.0, 0.0, 1.0) == 1.0
    # a zero-width range (erosion == 1) yields 0/0 or -inf, which clamp maps into [0, 1]
    v = orc.L.orc_remap_clamped(1.0, 1.0, 1.0, 0.0, 1.0)
    assert 0.0 <= v <= 1.0


def _trilinear_ref(tex, u, v, w, bits):
    """Weighted-sum trilinear filter with REPEAT wrap, written independently of the lerp chain."""
    n = tex.shape[0]
    out = np.zeros(4)
    idx, frac = [], []
    for c in (u, v, w):
        x = c * n - 0.5
        if bits:
            x = math.floor(x * 256 + 0.5) / 256
        i = math.floor(x)
        idx.append(i)
        frac.append(x - i)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wgt = ((frac[0] if dx else 1 - frac[0]) * (frac[1] if dy else 1 - frac[1]) *
                       (frac[2] if dz else 1 - frac[2]))
                out += wgt * tex[(idx[2] + dz) % n, (idx[1] + dy) % n, (idx[0] + dx) % n].astype(np.float64)
    return out / 255.0


@pytest.mark.parametrize("n", [2, 3, 4])
@pytest.mark.parametrize("bits", [0, 8])
def test_sampler_3d_brute_force(orc, n, bits):
    rng = np.random.default_rng(n * 10 + bits)
    tex = rng.integers(0, 256, (n, n, n, 4), dtype=np.uint8)
    coords = list(rng.uniform(-2.0, 3.0, (60, 3))) + [(0, 0, 0), (1, 1, 1), (-1e-7, 0.5, 0.5), (0.5, 1.0, -1e-7),
                                                      (1.0 / (2 * n), 0.25, 0.75)]
    out = np.zeros(4, np.float32)
    for u, v, w in coords:
        u, v, w = (float(np.float32(c)) for c in (u, v, w))
        orc.L.orc_sample3d(_ptr(tex), n, u, v, w, bits, _ptr(out))
        np.testing.assert_allclose(out, _trilinear_ref(tex, u, v, w, bits), atol=3e-6)


def test_sampler_2d_texel_centres_return_texels(orc):
    rng = np.random.default_rng(5)
    tex = rng.integers(0, 256, (4, 8, 4), dtype=np.uint8)  # h=4, w=8
    out = np.zeros(4, np.float32)
    for y in range(4):
        for x in range(8):
            orc.L.orc_sample2d(_ptr(tex), 8, 4, (x + 0.5) / 8, (y + 0.5) / 4, 8, _ptr(out))
            np.testing.assert_allclose(out * 255, tex[y, x], atol=1e-4)
    # REPEAT: u = 0 is halfway between the last and the first texel of the row
    orc.L.orc_sample2d(_ptr(tex), 8, 4, 0.0, 0.5 / 4, 0, _ptr(out))
    np.testing.assert_allclose(out * 255, (tex[0, 0].astype(float) + tex[0, 7]) / 2, atol=1e-4)


def _isect(orc, ro, rd, c, w, f64):
    v = ctypes.c_int(0)
    if f64:
        a = [np.array(x, np.float64) for x in (ro, rd, c)]
        t = orc.L.orc_ray_sphere_f64(_ptr(a[0]), _ptr(a[1]), _ptr(a[2]), w, ctypes.byref(v))
    else:
        a = [np.array(x, np.float32) for x in (ro, rd, c)]
        t = orc.L.orc_ray_sphere_f32(_ptr(a[0]), _ptr(a[1]), _ptr(a[2]), w, ctypes.byref(v))
    return t, bool(v.value)


def _quirk_t_exact(ro, rd, c, w):
    """comp:147-177 in 50-digit decimal: ro' = (ro - c)/w, nearest non-negative root, t = |p_world - ro'|."""
    getcontext().prec = 50
    D = [Decimal(float(x)) for x in ro], [Decimal(float(x)) for x in rd], [Decimal(float(x)) for x in c]
    W = Decimal(float(w))
    ro_, rd_, c_ = D
    rl = [(a - b) / W for a, b in zip(ro_, c_)]
    A = sum(x * x for x in rd_)
    B = 2 * sum(x * y for x, y in zip(rd_, rl))
    C = sum(x * x for x in rl) - Decimal("0.25")
    disc = B * B - 4 * A * C
    assert disc >= 0
    s = disc.sqrt()
    t = (-s - B) / A / 2
    if t < 0:
        t = (s - B) / A / 2
    assert t >= 0
    p = [(r + d * t) * W + cc for r, d, cc in zip(rl, rd_, c_)]
    return float(sum((a - b) ** 2 for a, b in zip(p, rl)).sqrt())


def test_ray_sphere_textbook_and_quirk(orc):
    # unit-radius sphere (diameter 2) at the origin: t is the textbook distance 1
    t, ok = _isect(orc, (0, 0, 0), (0, 1, 0), (0, 0, 0), 2.0, True)
    assert ok and t == pytest.approx(1.0, abs=1e-12)
    # off-centre origin: the reference returns |p_world - ro_local| (comp:173), not the ray distance
    t, ok = _isect(orc, (0, 0, 0), (0, 1, 0), (0, -10, 0), 22.0, True)
    assert ok and t == pytest.approx(1.0 - 10.0 / 22.0, abs=1e-12)
    # miss: ray pointing away from a sphere in front of the origin
    _, ok = _isect(orc, (0, 0, 0), (0, 1, 0), (0, -10, 0), 2.0, True)
    assert not ok


def test_ray_sphere_fp32_cancellation_at_earth_radius(orc):
    # Scenario shell: Earth radius 6.371e6 m, inner shell at +1500 m, camera at 200 m. FP64 matches the
    # exact value of the same formula; FP32 loses metres to the cancellation in C = |ro'|^2 - 0.25.
    R = 6371000.0
    worst32 = 0.0
    for elev in (2.0, 5.0, 10.0, 30.0, 90.0):
        th = math.radians(elev)
        rd = (0.0, math.sin(th), -math.cos(th))
        rd = tuple(float(np.float32(x)) for x in rd)
        args = ((0.0, 200.0, 0.0), rd, (0.0, -R, 0.0), 2 * (R + 1500.0))
        exact = _quirk_t_exact(*args)
        t64, ok64 = _isect(orc, *args, True)
        t32, ok32 = _isect(orc, *args, False)
        assert ok64 and ok32
        assert abs(t64 - exact) < 1e-6 * max(1.0, exact)
        worst32 = max(worst32, abs(t32 - exact))
    assert worst32 > 0.05  # metres: the FP32 formula is visibly off at Earth radius


def test_pow_rt_against_the_library_pow(orc):
    """The coverage exponent's pow (comp:206-208) is evaluated by a written-out exp2(y log2 x) that the
    kernel repeats operation for operation. Against float64 pow: within (1 + |y log2 x|) float ulps
    (GLSL allows 3 + 2 |y log2 x| for exp2 alone), exact for pow(x, 1) and for powers of two whose
    result is a power of two, and 0 for x <= 0."""
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(1e-6, 0.9, 4000), 10 ** rng.uniform(-30, 30, 4000)]).astype(np.float32)
    y = np.concatenate([rng.uniform(0.8, 1.0, 4000), rng.uniform(-3, 3, 4000)]).astype(np.float32)
    z = np.abs(y.astype(np.float64) * np.log2(x.astype(np.float64)))
    keep = z < 120  # results inside the float range
    x, y, z = x[keep], y[keep], z[keep]
    got = np.array([orc.L.orc_pow_rt(float(a), float(b)) for a, b in zip(x, y)], np.float64)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    ulps = np.abs(got - ref) / ref / 2.0 ** -23
    assert (ulps <= 1.0 + z).all(), float((ulps / (1.0 + z)).max())
    for v in (1e-7, 0.3, 0.9, 123.0):
        assert orc.L.orc_pow_rt(v, 1.0) == np.float32(v)
    for k in range(-20, 21):
        for yy in (0.5, 0.25, 2.0, -1.0):
            if (k * yy) == int(k * yy):
                assert orc.L.orc_pow_rt(2.0 ** k, yy) == 2.0 ** (k * yy)
    assert orc.L.orc_pow_rt(0.0, 0.9) == 0.0 and orc.L.orc_pow_rt(-0.5, 0.9) == 0.0
