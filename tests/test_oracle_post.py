"""The post-chain oracle (god-ray.frag, radialBlur.frag, tonemap.frag) pinned by closed forms.

Synthetic uniforms put the sun where a case needs it: with proj = view = identity the sun's NDC
position is sun.location.xy / sun.location.w.
"""
import numpy as np


def _ubos(mm, sx=0.2, sy=-0.3, w=1.0, dir_y=0.5, color=(1.0, 0.9, 0.8), intensity=2.0):
    cam, sun = mm.CameraUBO(), mm.SunUBO()
    for i in range(4):
        cam.view[5 * i] = 1.0
        cam.proj[5 * i] = 1.0
    sun.location[0], sun.location[1], sun.location[2], sun.location[3] = sx * w, sy * w, 0.0, w
    sun.direction[1] = dir_y
    sun.color[0], sun.color[1], sun.color[2] = color
    sun.intensity = intensity
    return cam, sun


def _uc2(x):
    x = np.asarray(x, np.float64)
    return ((x * (0.15 * x + 0.05) + 0.004) / (x * (0.15 * x + 0.5) + 0.06)) - 0.02 / 0.3


def test_sun_position_is_proj_view_location_over_w(mm, orc):
    cam, sun = _ubos(mm, sx=0.25, sy=-0.5, w=2.0)
    assert np.allclose(orc.post_sun_ndc(cam, sun), [0.25, -0.5])
    # a real camera: the projected direction of the sun agrees with a float64 evaluation
    cam, sun, _ = mm.build_ubos(mm.scenario_scene(aspect=16 / 9))
    P = np.array(cam.proj, np.float64).reshape(4, 4).T
    V = np.array(cam.view, np.float64).reshape(4, 4).T
    s = P @ V @ np.array(sun.location, np.float64)
    assert np.allclose(orc.post_sun_ndc(cam, sun), s[:2] / s[3], atol=1e-5)


def test_uniform_image_closed_forms(mm, orc):
    H, W = 24, 40
    a, rgb = 0.6, np.array([0.3, 1.5, 4.0], np.float32)
    img = np.empty((H, W, 4), np.float32)
    img[..., :3] = rgb
    img[..., 3] = a
    cam, sun = _ubos(mm)
    p = mm.post_params()
    alpha, hdr, out = orc.post_process(cam, sun, img, p)
    # god rays: 0.9 * (0.5 a + sum_i 0.5 a / 8 * 0.99^i); every tap sees a (clamp to edge)
    want = 0.9 * (0.5 * a + sum(0.5 * a / 8 * 0.99 ** i for i in range(8)))
    assert np.allclose(alpha, want, rtol=1e-6)
    # radial blur: mean of 10 taps of that alpha times 1.1, then sun.color * intensity * it + 0.5 rgb
    gc = np.array([1.0, 0.9, 0.8]) * 2.0
    assert np.allclose(hdr[..., :3], gc * (1.1 * want) + 0.5 * rgb, rtol=1e-5)
    assert (hdr[..., 3] == 1.0).all()
    # tonemap with the vignette, against a float64 evaluation (1 LSB for the float rounding)
    y, x = np.mgrid[0:H, 0:W]
    u, v = (x + 0.5) / W, (y + 0.5) / H
    vig = (u - 0.5) ** 2 + (v - 0.5) ** 2
    t = _uc2(0.7 * hdr[..., :3].astype(np.float64)) / _uc2(50.2)
    t = t ** (1 / 2.2)
    t = t * (1 - vig[..., None]) + np.array([0.1, 0.05, 0.13]) * vig[..., None]
    q = np.floor(np.clip(t, 0, 1) * 255 + 0.5)
    assert np.abs(out[..., :3].astype(int) - q).max() <= 1
    assert (out[..., 3] == 255).all()


def test_tonemap_white_point_maps_to_one(mm, orc):
    # exposure * x = white point: the curve times 1 / curve(white) is 1 and pow keeps it; the centre
    # pixel of an odd-sized image has vignette weight 0, so it is 255, and a corner pixel is pulled
    # toward the vignette colour by its weight (u - 0.5)^2 + (v - 0.5)^2
    H, W = 9, 9
    cam, sun = _ubos(mm, dir_y=-1.0)  # sun below the horizon: rgb passes through the radial blur
    img = np.full((H, W, 4), 50.2 / 0.7, np.float32)
    p = mm.post_params()
    _, hdr, out = orc.post_process(cam, sun, img, p)
    assert np.allclose(hdr[..., :3], 50.2 / 0.7, rtol=1e-6)
    assert (out[4, 4, :3] >= 254).all()
    vig = 2 * (0.5 / 9 - 0.5) ** 2
    want = np.floor(((1 - vig) + np.array([0.1, 0.05, 0.13]) * vig) * 255 + 0.5)
    assert np.abs(out[0, 0, :3].astype(int) - want).max() <= 1


def test_sun_below_horizon_passes_colour_through(mm, orc):
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 3, (17, 29, 4)).astype(np.float32)
    cam, sun = _ubos(mm, dir_y=-0.1)
    alpha, hdr, _ = orc.post_process(cam, sun, img, mm.post_params())
    assert (alpha == 1.0).all()                    # god-ray.frag:55-57
    assert np.array_equal(hdr[..., :3], img[..., :3])  # radialBlur.frag:44-46 (centre taps are exact)


def test_god_rays_step_toward_the_sun(mm, orc):
    # alpha 1 on the left half only, sun at the left edge: a right-half pixel gathers alpha only if
    # its taps walk toward the sun (a sign slip walks them right, into alpha 0)
    H, W = 8, 64
    img = np.zeros((H, W, 4), np.float32)
    img[:, : W // 2, 3] = 1.0
    cam, sun = _ubos(mm, sx=-1.0, sy=0.0)
    alpha, _, _ = orc.post_process(cam, sun, img, mm.post_params())
    right = alpha[:, W // 2 + 2 : W // 2 + 8]
    assert (right > 0.05).all()
    cam, sun = _ubos(mm, sx=1.0, sy=0.0)  # sun on the right: the same pixels stay dark
    alpha, _, _ = orc.post_process(cam, sun, img, mm.post_params())
    assert (alpha[:, W // 2 + 2 : W // 2 + 8] == 0.0).all()


def test_radial_blur_taps_are_symmetric_about_the_pixel(mm, orc):
    # an alpha ramp along the sun direction: symmetric taps average to the centre value
    H, W = 16, 256
    img = np.zeros((H, W, 4), np.float32)
    img[..., 3] = np.linspace(0.0, 1.0, W, dtype=np.float32)[None, :]
    cam, sun = _ubos(mm, sx=3.0, sy=0.0, intensity=1.0, color=(1.0, 1.0, 1.0))
    p = mm.post_params(god_ray_density=0.0)  # god-ray taps all at the pixel: alpha = 0.9 * a * (0.5 + ...)
    alpha, hdr, _ = orc.post_process(cam, sun, img, p)
    acc = hdr[..., 0] / 1.1  # rgb = 0 here, so hdr.r = 1.1 * mean(alpha taps)
    inner = slice(100, 210)  # the widest tap (0.06 * dist * W px) stays inside the image
    assert np.allclose(acc[:, inner], alpha[:, inner], atol=2e-3)


def test_series_pow_against_float64(orc):
    # pow as the tonemap evaluates it (exp2(y log2 x) in float series) against float64 pow
    xs = np.concatenate([np.geomspace(1e-30, 1e4, 4000), [1.0, 0.5, 2.0, 1e-40]]).astype(np.float32)
    for y in (1 / 2.2, 2.2, 0.5, 3.0):
        got = np.array([orc.L.orc_post_pow(float(x), y) for x in xs], np.float64)
        want = np.power(xs.astype(np.float64), np.float32(y).astype(np.float64))
        ok = want > 1e-37  # normal results
        # float rounding of t = y log2 x moves 2^t by ln2 |t| 2^-24 relative: the bound grows with |t|
        t = np.abs(y * np.log2(xs.astype(np.float64)))
        bound = (1e-6 + 2.0 * t * np.log(2.0) * 2.0 ** -24) * want
        assert np.all(np.abs(got[ok] - want[ok]) <= bound[ok]), y
        if y == 1 / 2.2:  # the tonemap's gamma on the values it sees
            tone = ok & (xs >= 1e-4) & (xs <= 4.0)
            assert np.all(np.abs(got[tone] - want[tone]) <= 1e-6 * want[tone])
    assert orc.L.orc_post_pow(0.0, 0.45) == 0.0
    assert np.isnan(orc.L.orc_post_pow(-1.0, 0.45))
    assert orc.L.orc_post_pow(float("inf"), 0.45) == float("inf")
    assert abs(orc.L.orc_post_pow(1.0, 0.45) - 1.0) <= 6e-8


def test_parity_statistics_closed_forms():
    """The statistics the march parity tests assert on: tonemapped RGB and raw alpha errors in
    1/255 units, and the raw HDR relative error, on perturbations whose values are known."""
    from oracle import parity_stats, tonemap

    ref = np.zeros((4, 5, 4), np.float32)
    ref[..., :3] = 2.0
    ref[..., 3] = 0.5
    st = parity_stats(ref.copy(), ref)
    assert st["max"] == 0.0 and st["frac_within_1"] == 1.0 and st["hdr_rel_max"] == 0.0 and st["n"] == 20
    g = ref.copy()
    g[0, 0, 3] += 3.0 / 255.0                      # alpha is compared raw: 3/255 on one pixel
    g[1, 1, 0] *= 1.5                              # HDR red 50 % high on another
    st = parity_stats(g, ref)
    tm = float((tonemap(np.array([3.0])) - tonemap(np.array([2.0])))[0]) * 255.0
    assert abs(st["max"] - max(3.0, tm)) < 1e-4
    assert st["frac_within_1"] == (20 - 1 - (tm > 1.0)) / 20
    assert abs(st["hdr_rel_max"] - 0.5) < 1e-7
    assert abs(st["hdr_rel_mean"] - 0.5 / 60) < 1e-9
    g[2, 2, 1] = np.nan
    assert parity_stats(g, ref)["nan"] == 1
