"""Pins the reprojection oracle (reproject.comp restated): identity motion, the geometry of the
centre tap against an independent float64 projection, and a camera yaw that shifts the image."""
import ctypes

import numpy as np
import pytest


def _cam(mm, yaw_deg=0.0, altitude=200.0, W=96, H=64):
    s = mm.scenario_scene(altitude=altitude, aspect=W / H)
    s.yaw = s.yaw + np.deg2rad(yaw_deg)
    return mm.build_ubos(s)[0]


def test_identity_motion_copies_the_image(mm, orc):
    W, H = 96, 64
    cam = _cam(mm, W=W, H=H)
    rng = np.random.default_rng(1)
    src = rng.uniform(0, 4, (H, W, 4)).astype(np.float32)
    for preset in ("reference", "scenario"):
        out = orc.reproject(cam, cam, src, mm.march_params(preset), W, H)
        assert np.array_equal(out[..., 3], src[..., 3])
        # ten equal taps summed and divided by ten: within a few ulp of the source
        assert np.allclose(out[..., :3], src[..., :3], rtol=4e-7, atol=0)


def _expected_centre(mm, cam, prev, prm, W, H):
    """Independent float64 restatement of the centre-tap coordinate (numpy, vectorised)."""
    V = np.array(cam.view, np.float64).reshape(4, 4)       # column-major: V[col, row]
    P = np.array(prev.view, np.float64).reshape(4, 4)
    right, up, look = V[:3, 0], V[:3, 1], V[:3, 2]
    pos = np.array(cam.cameraPosition[:3], np.float64)
    aspect, tanh = cam.cameraParams[0], cam.cameraParams[1]
    ys, xs = np.mgrid[0:H, 0:W]
    sx, sy = xs / W * 2 - 1, ys / H * 2 - 1
    d = (-look)[None, None] + (aspect * sx * tanh)[..., None] * right - (sy * tanh)[..., None] * up
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    c = pos.copy()
    c[1] = prm.center_y
    r = prm.inner_diameter / 2
    oc = pos - c
    b = (d * oc).sum(-1)
    disc = b * b - ((oc * oc).sum() - r * r)
    t = -b + np.sqrt(disc)  # camera inside the sphere: the far root
    hit = pos + d * t[..., None]
    q = hit @ P[:3, :3] + P[3, :3]
    u = q[..., 0] / -q[..., 2] / tanh / aspect * 0.5 + 0.5
    v = -q[..., 1] / -q[..., 2] / tanh * 0.5 + 0.5
    return u * W, v * H


@pytest.mark.parametrize("preset", ["reference", "scenario"])
def test_centre_tap_geometry(mm, orc, preset):
    W, H = 96, 64
    prm = mm.march_params(preset)
    cam = _cam(mm, yaw_deg=1.5, W=W, H=H)
    prev = _cam(mm, yaw_deg=0.0, W=W, H=H)
    ids = np.arange(W * H, dtype=np.float32).reshape(H, W)
    src = np.zeros((H, W, 4), np.float32)
    src[..., 3] = ids
    out = orc.reproject(cam, prev, src, prm, W, H)
    got = out[..., 3].astype(np.int64)
    gx, gy = got % W, got // W
    ex, ey = _expected_centre(mm, cam, prev, prm, W, H)
    ex = np.clip(np.floor(ex + 0.5), 0, W - 1)
    ey = np.clip(np.floor(ey + 0.5), 0, H - 1)
    # the float32 shader path rounds to the same texel except at rounding ties
    assert np.mean((gx == ex) & (gy == ey)) > 0.995
    assert np.max(np.abs(gx - ex)) <= 1 and np.max(np.abs(gy - ey)) <= 1


def test_yaw_shifts_the_image(mm, orc):
    W, H = 128, 64
    prm = mm.march_params("reference")
    fov_px = 2 * np.degrees(np.arctan(_cam(mm, W=W, H=H).cameraParams[0] * _cam(mm, W=W, H=H).cameraParams[1])) / W
    cam = _cam(mm, yaw_deg=10 * fov_px, W=W, H=H)
    prev = _cam(mm, W=W, H=H)
    src = np.zeros((H, W, 4), np.float32)
    src[..., 0] = np.arange(W, dtype=np.float32)[None, :]
    out = orc.reproject(cam, prev, src, prm, W, H)
    shift = out[H // 2, 20:W - 20, 0] - np.arange(20, W - 20)
    # a ~10-pixel horizontal shift (sign set by the yaw direction) in every column; it grows toward
    # the edges (perspective) and the 10-tap motion blur averages the gradient along the shift
    assert 8.0 <= np.abs(np.median(shift)) <= 12.0
    assert (np.sign(shift) == np.sign(np.median(shift))).all() and np.ptp(shift) <= 6.5
