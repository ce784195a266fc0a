"""Pins the cloud-shadow oracle (model.frag restated): the mesh pass's own layer profile (hand
values), empty and saturated volumes, and the max-accumulation bound."""
import numpy as np
import pytest


def test_layer_profile_known_values(orc):
    f = orc.L.orc_shadow_layer_density
    # remap is unclamped: stratus(h) = max(0, remap(h, 0, .1, 0, 1) * remap(h, .2, .3, 1, 0)),
    # strato(h) = max(0, remap(h, 0, .2, 0, 1) * remap(h, .2, .7, 1, 0))
    # type 0 -> d1 = stratus, weight 0 on d2
    assert f(0.05, 0.0) == pytest.approx(0.5 * 2.5, abs=1e-5)
    assert f(0.15, 0.0) == pytest.approx(1.5 * 1.5, abs=1e-5)
    # type 1 -> d2 = mix(strato, stratus, 1) = stratus (model.frag's d2; compute-clouds.comp has cumulus)
    assert f(0.25, 1.0) == pytest.approx(2.5 * 0.5, abs=1e-5)
    assert f(0.5, 1.0) == 0.0
    # type 0.5 -> d1 = d2 = strato
    assert f(0.1, 0.5) == pytest.approx(0.5 * 1.2, abs=1e-5)
    assert f(0.45, 0.5) == pytest.approx(2.25 * 0.5, abs=1e-5)


def _scene(mm):
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=1.0, sun_elevation_deg=40.0))
    return cam, sun, sky


def _points(n=400, seed=1):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 3), np.float32)
    p[:, 0] = rng.uniform(-3000, 3000, n)
    p[:, 2] = rng.uniform(-3000, 3000, n)
    p[:, 1] = rng.uniform(0, 30, n)
    return p


def test_empty_and_saturated_volumes(mm, orc):
    cam, sun, sky = _scene(mm)
    prm = mm.shadow_params()
    weather = np.zeros((64, 64, 4), np.uint8)
    weather[..., 0] = 128
    weather[..., 2] = 0       # type 0: stratus band
    pts = _points()
    empty = orc.cloud_shadow(cam, sun, sky, weather, np.zeros((16, 16, 16, 4), np.uint8), pts, prm)
    assert (empty == 0).all()
    full = np.full((16, 16, 16, 4), 255, np.uint8)
    full[..., 1:] = 0         # no erosion
    sat = orc.cloud_shadow(cam, sun, sky, weather, full, pts, prm)
    assert (sat >= 0).all() and (sat <= 1).all()
    # the shadow is the max over the steps of layer(h) * 1 (full shape, no erosion): 1 where any
    # step lands in the stratus band with layer >= 0.99
    assert (np.isin(sat, [0.0, 1.0]) | ((sat > 0) & (sat < 1))).all()


def test_random_volume_bounds(mm, orc):
    cam, sun, sky = _scene(mm)
    prm = mm.shadow_params()
    w = orc.weather(128, 3, 0.6)
    low = orc.low_res(32, 2)
    s = orc.cloud_shadow(cam, sun, sky, w, low, _points(2000, 2), prm)
    assert np.isfinite(s).all() and (s >= 0).all() and (s <= 1).all()
    # fewer steps can only lower the maximum
    prm2 = mm.shadow_params(steps=3)
    s3 = orc.cloud_shadow(cam, sun, sky, w, low, _points(2000, 2), prm2)
    assert (s3 <= s + 1e-7).all()
