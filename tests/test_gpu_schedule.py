"""Tile schedule (mm_march_clouds_scheduled): the launch after a recorded one marches its tiles
longest first. The order only decides which tiles start first, so every scheduled launch must give
the plain launch's image and work counters bit for bit -- with the output pre-filled, a tile left out
or marched twice in place of another would show."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(mm, W, H, sun_deg=25.0, t=0.0):
    return mm.build_ubos(mm.scenario_scene(aspect=W / H, sun_elevation_deg=sun_deg, time=t))


def _noise(mm, seed=1, coverage=0.5):
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=seed, weather_seed=seed, coverage=coverage), mm.current_stream())
    return ns


def _march(mm, ns, ubos, W, H, prm, phase=-1, schedule=None):
    import torch

    out = torch.full((W * H * 4,), -1.0, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(4, dtype=torch.int64, device="cuda")
    mm.march_clouds(*ubos, ns, out, W, H, phase=phase, params=prm, counts=cnt, schedule=schedule)
    torch.cuda.synchronize()
    return out.cpu().numpy(), cnt.cpu().numpy()


@pytest.mark.parametrize("case", ["frame_default", "frame_2lanes", "frame_1lane_8x8", "phase_8lanes",
                                  "surrogate_threshold0", "reference_preset", "no_params"])
def test_scheduled_launches_equal_the_plain_march(mm, gpu, case):
    """Scheduled full frames default to one lane per ray in 16 x 4 tiles (the plain call to two lanes
    per ray in 8 x 4); every shape gives the same image."""
    W, H = 480, 270
    ns = _noise(mm)
    phase = -1
    if case == "frame_default":
        prm = mm.march_params("scenario")
    elif case == "frame_2lanes":
        prm = mm.march_params("scenario", lanes_per_ray=2)
    elif case == "frame_1lane_8x8":
        prm = mm.march_params("scenario", lanes_per_ray=1, tile_width=8)
    elif case == "phase_8lanes":
        prm = mm.march_params("scenario")
        phase = 6
    elif case == "surrogate_threshold0":
        model = mm.Surrogate(seed=3)
        prm = mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=0.0, surrogate=model.ptr)
    elif case == "reference_preset":
        prm = mm.march_params("reference")
    else:
        prm = None  # NULL params: the reference preset, scheduled in 16 x 4 tiles
    sched = mm.MarchSchedule(1 << 16)
    for rep, (sun_deg, t) in enumerate([(25.0, 0.0), (25.0, 0.0), (12.0, 300.0), (12.0, 300.0)]):
        ubos = _scene(mm, W, H, sun_deg, t)
        ref, cref = _march(mm, ns, ubos, W, H, prm, phase)
        got, cgot = _march(mm, ns, ubos, W, H, prm, phase, schedule=sched)
        assert np.array_equal(got, ref), (case, rep)
        assert (cgot == cref).all(), (case, rep, cgot, cref)
    sched.close()


def test_scheduled_1080p_frames_equal_the_plain_march(mm, gpu):
    """The headline shape: 32 400 tiles, so every order launch runs several workgroups."""
    W, H = 1920, 1080
    ns = _noise(mm)
    prm = mm.march_params("scenario")
    sched = mm.MarchSchedule(1 << 16)
    for rep, (sun_deg, t) in enumerate([(25.0, 0.0), (25.0, 0.0), (25.0, 0.0), (40.0, 120.0)]):
        ubos = _scene(mm, W, H, sun_deg, t)
        ref, cref = _march(mm, ns, ubos, W, H, prm)
        got, cgot = _march(mm, ns, ubos, W, H, prm, schedule=sched)
        assert np.array_equal(got, ref), rep
        assert (cgot == cref).all(), (rep, cgot, cref)
    sched.close()


def test_schedule_follows_shape_changes(mm, gpu):
    """A launch of another shape marches in the default order and starts a new record; going back
    to the first shape then uses that shape's costs again after one launch."""
    ns = _noise(mm, seed=2, coverage=0.7)
    prm = mm.march_params("scenario")
    sched = mm.MarchSchedule(1 << 16)
    for W, H in [(320, 180), (320, 180), (256, 256), (256, 256), (320, 180), (320, 180)]:
        ubos = _scene(mm, W, H)
        ref, _ = _march(mm, ns, ubos, W, H, prm)
        got, _ = _march(mm, ns, ubos, W, H, prm, schedule=sched)
        assert np.array_equal(got, ref), (W, H)
    sched.close()


def test_schedule_capacity_and_arguments(mm, gpu):
    import torch

    ns = _noise(mm)
    prm = mm.march_params("scenario")
    W, H = 320, 180  # 20 x 45 = 900 tiles of 16 x 4 rays (the scheduled default shape)
    small = mm.MarchSchedule(500)
    out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    with pytest.raises(mm.MMError) as err:
        mm.march_clouds(*_scene(mm, W, H), ns, out, W, H, params=prm, schedule=small)
    assert err.value.code == mm.MM_E_SIZE
    small.close()
    import ctypes

    h = ctypes.c_void_p()
    assert mm.lib().mm_march_schedule_create(0, ctypes.byref(h)) == mm.MM_E_PARAM
    assert mm.lib().mm_march_schedule_create(16, None) == mm.MM_E_NULL
    cam, sun, sky = _scene(mm, W, H)
    assert mm.lib().mm_march_clouds_scheduled(ctypes.byref(cam), ctypes.byref(sun), ctypes.byref(sky),
                                              ctypes.byref(ns.c), None, out.data_ptr(), W, H, W, -1,
                                              ctypes.byref(prm), None, None, mm.current_stream()) == mm.MM_E_NULL
