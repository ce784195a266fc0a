"""Caller side of the boundary: std140 layout of the uniform blocks and the scene -> UBO builder
(restating VulkanApplication.cpp:351-395, camera.cpp:27-39/218-238, SkyManager.cpp:15-70)."""
import ctypes
import math

import numpy as np
import pytest

DEG2RAD = np.float32(0.01745)  # camera.h:15


def test_std140_sizes_and_offsets(mm):
    # compute-clouds.comp:12-41 std140 blocks; host structs Shader.h:24-28, SkyManager.h:8-14, 28-36
    assert ctypes.sizeof(mm.CameraUBO) == 160
    assert mm.CameraUBO.view.offset == 0 and mm.CameraUBO.proj.offset == 64
    assert mm.CameraUBO.cameraPosition.offset == 128 and mm.CameraUBO.cameraParams.offset == 144
    assert ctypes.sizeof(mm.SunUBO) == 128
    assert (mm.SunUBO.location.offset, mm.SunUBO.direction.offset, mm.SunUBO.color.offset) == (0, 16, 32)
    assert mm.SunUBO.directionBasis.offset == 48 and mm.SunUBO.intensity.offset == 112
    assert ctypes.sizeof(mm.SkyUBO) == 64
    assert (mm.SkyUBO.betaR.offset, mm.SkyUBO.betaV.offset, mm.SkyUBO.wind.offset) == (0, 16, 32)
    assert mm.SkyUBO.mie_directional.offset == 48


@pytest.mark.parametrize("elev", [2.0, 5.0, 25.0, 60.0, 70.0])
def test_sun_closed_forms(mm, elev):
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(sun_elevation_deg=elev, phase=7))
    d = np.array(sun.direction[:3], np.float64)
    # SkyManager parametrisation with azimuth 0.25: dir = (0, sin e, -cos e)
    np.testing.assert_allclose(d, [0.0, math.sin(math.radians(elev)), -math.cos(math.radians(elev))], atol=2e-6)
    assert sun.direction[3] == 0.0
    np.testing.assert_allclose(sun.location[:3], 400000.0 * d, rtol=1e-6, atol=0.1)  # SUN_DISTANCE
    # intensity = 1000 max(0, 1 - e^(-(cutoff - acos(dir.y)) / 1.5)) (SkyManager.cpp:28-34)
    expect = 1000.0 * max(0.0, 1.0 - math.exp(-((1.6110731557 - math.acos(d[1])) / 1.5)))
    assert sun.intensity == pytest.approx(expect, rel=2e-5)
    # colour: sunset blend t = clamp(13 dir.y) (SkyManager.cpp:15-26); white once dir.y >= 1/13
    t = min(1.0, max(0.0, 13 * d[1]))
    np.testing.assert_allclose(sun.color[:3], [(1 - t) * 2 + t, (1 - t) * 0.33922 + t, (1 - t) * 0.0431 + t], atol=2e-6)
    assert sun.color[3] == 7.0  # pixel phase rides in sun.color.a (VulkanApplication.cpp:384)
    # directionBasis: column 1 = sun direction, columns orthonormal
    B = np.array(sun.directionBasis, np.float64).reshape(4, 4)  # B[col] = column
    np.testing.assert_allclose(B[1, :3], d, atol=1e-7)
    M = B[:3, :3]
    np.testing.assert_allclose(M @ M.T, np.eye(3), atol=2e-6)
    # sky: betaR = RAYLEIGH_TOTAL * (rayleigh - 1 + sunFade) with sunFade = 1 above the horizon
    np.testing.assert_allclose(sky.betaR[:3], 2.0 * np.array([5.804542996261093e-6, 1.3562911419845635e-5,
                                                              3.0265902468824876e-5]), rtol=1e-6)
    # betaV = 0.434 * (0.2 * turbidity * 1e-17) * MIE_CONST * mie
    np.testing.assert_allclose(sky.betaV[:3], 0.434 * (0.2 * 10 * 1e-17) * np.array(
        [1.839991851443397, 2.779802391966052, 4.079047954386109]) * 0.005, rtol=1e-5)
    assert sky.mie_directional == pytest.approx(0.8) and list(sky.wind) == pytest.approx([1, 0.05, 1, 0])


def test_night_sun(mm):
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(sun_elevation_deg=-10.0))
    assert sun.direction[1] < 0
    assert sun.intensity == 2.0                       # SkyManager.cpp:31-33
    assert list(sun.color[:3]) == pytest.approx([0.8, 0.9, 1.0])
    B = np.array(sun.directionBasis, np.float64).reshape(4, 4)
    np.testing.assert_allclose(B[1, :3], -np.array(sun.direction[:3]), atol=1e-7)  # basis *= -1 (:67-69)
    assert B[3, 3] == -1.0
    np.testing.assert_allclose(sun.location[:3], -400000.0 * np.array(sun.direction[:3]), rtol=1e-6)


def test_camera_view_and_params(mm):
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(altitude=200.0, pitch_deg=20.0, fov_deg=60.0, aspect=16 / 9))
    V = np.array(cam.view, np.float64).reshape(4, 4)  # V[col][row]
    right, up, fwd = V[:3, 0], V[:3, 1], V[:3, 2]     # rows 0, 1, 2 of the view matrix
    R = np.stack([right, up, fwd])
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=2e-6)
    pitch = float(np.float32(20.0) * DEG2RAD)
    # the shader looks along -forward (comp:312, 318): up by `pitch`, towards -z
    np.testing.assert_allclose(-fwd, [0.0, math.sin(pitch), -math.cos(pitch)], atol=2e-6)
    np.testing.assert_allclose(V[3, :3], -(R @ np.array([0, 200.0, 0])), atol=1e-3)  # view * translate(-pos)
    assert list(cam.cameraPosition) == [0.0, 200.0, 0.0, 1.0]
    assert cam.cameraParams[0] == pytest.approx(16 / 9)
    assert cam.cameraParams[1] == pytest.approx(math.tan(0.5 * 0.01745 * 60.0), rel=1e-6)  # getHTanFov
    assert cam.proj[5] < 0  # proj[1][1] *= -1 (VulkanApplication.cpp:362)


def test_builder_rejects_bad_input(mm):
    s = mm.scenario_scene()
    s.camera_pos[1] = float("nan")
    with pytest.raises(mm.MMError) as e:
        mm.build_ubos(s)
    assert e.value.code == mm.MM_E_PARAM
    s = mm.scenario_scene(phase=16)
    with pytest.raises(mm.MMError) as e:
        mm.build_ubos(s)
    assert e.value.code == mm.MM_E_PHASE
    s = mm.scenario_scene(fov_deg=0.0)
    with pytest.raises(mm.MMError):
        mm.build_ubos(s)
    assert mm.lib().mm_build_ubos(None, None, None, None) == mm.MM_E_NULL
