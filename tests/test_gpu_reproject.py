"""Temporal reprojection on the GPU against the oracle (bit-exact), and the per-frame schedule
(reproject, then march one pixel phase) converging to the full-frame march for a still camera."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(mm, W, H, yaw_deg=0.0, altitude=200.0, pitch=20.0):
    s = mm.scenario_scene(altitude=altitude, pitch_deg=pitch, aspect=W / H)
    s.yaw = s.yaw + np.deg2rad(yaw_deg)
    return mm.build_ubos(s)[0]


@pytest.mark.parametrize("preset", ["reference", "scenario"])
@pytest.mark.parametrize("W,H", [(96, 64), (1920, 1080), (250, 131)])
def test_reproject_matches_oracle(mm, gpu, orc, preset, W, H):
    import torch

    prm = mm.march_params(preset)
    rng = np.random.default_rng(W + H)
    src = rng.uniform(0, 3, (H, W, 4)).astype(np.float32)
    s = torch.from_numpy(src.reshape(-1)).cuda()
    for yaw, alt, pitch in ((0.0, 200.0, 20.0), (2.0, 260.0, 20.5), (-7.0, 200.0, 23.0)):
        cam = _cam(mm, W, H, yaw, alt, pitch)
        prev = _cam(mm, W, H)
        d = torch.full_like(s, -1.0)
        mm.reproject(cam, prev, s, d, W, H, params=prm)
        torch.cuda.synchronize()
        got = d.cpu().numpy().reshape(H, W, 4)
        ref = orc.reproject(cam, prev, src, prm, W, H)
        assert np.array_equal(got, ref), (preset, W, H, yaw, np.argwhere(got != ref)[:5])


def test_reproject_rejects_bad_arguments(mm, gpu):
    import torch

    W, H = 32, 16
    cam = _cam(mm, W, H)
    a = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    with pytest.raises(mm.MMError) as e:
        mm.reproject(cam, cam, a, a, W, H)
    assert e.value.code == mm.MM_E_PARAM
    b = torch.zeros_like(a)
    with pytest.raises(mm.MMError) as e:
        mm.reproject(cam, cam, a, b, W, H, pitch=W - 1)
    assert e.value.code == mm.MM_E_PITCH


def test_frames_of_a_still_camera_converge_to_the_full_march(mm, gpu):
    """16 frames (phases 0..15) with a still camera: every pixel was marched once and then carried
    by reprojections that copy it (ten equal taps / 10), so the image equals the full-frame march
    up to a few ulp."""
    import torch

    W, H = 320, 180
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=1, weather_seed=2, coverage=0.5), mm.current_stream())
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    prm = mm.march_params("scenario")
    frames = mm.CloudFrames(W, H, ns, params=prm)
    for _ in range(16):
        img = frames.frame(cam, sun, sky)
    full = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    mm.march_clouds(cam, sun, sky, ns, full, W, H, params=prm)
    torch.cuda.synchronize()
    a, b = img.cpu().numpy(), full.cpu().numpy()
    assert np.allclose(a, b, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("phase", [0, 6, 15])
def test_render_frame_equals_reprojection_then_march(mm, gpu, phase):
    """mm_render_frame runs the reprojection (minus the phase's pixels) beside the phase march on a
    second stream: its image equals mm_reproject followed by mm_march_clouds of that phase, in the
    reference's order, bit for bit, for a moving camera and a ragged size."""
    import ctypes

    import torch

    W, H = 203, 117
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=3, weather_seed=4, coverage=0.6), mm.current_stream())
    prm = mm.march_params("scenario")
    cam0, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H, pitch_deg=18.0))
    cam1, _, _ = mm.build_ubos(mm.scenario_scene(aspect=W / H, pitch_deg=20.0))
    prev = mm.previous_camera(cam0)
    src = torch.rand(W * H * 4, dtype=torch.float32, device="cuda")
    sun.color[3] = float(phase)
    a = torch.full_like(src, -5.0)
    mm.check(mm.lib().mm_render_frame(ctypes.byref(cam1), ctypes.byref(prev), ctypes.byref(sun), ctypes.byref(sky),
                                      ctypes.byref(ns.c), None, src.data_ptr(), a.data_ptr(), W, H, W,
                                      ctypes.byref(prm), mm.current_stream()), "mm_render_frame")
    b = torch.full_like(src, -5.0)
    mm.reproject(cam1, prev, src, b, W, H, params=prm)
    mm.march_clouds(cam1, sun, sky, ns, b, W, H, phase=phase, params=prm)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert not (a == -5.0).any()
