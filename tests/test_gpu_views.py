"""Many views in one launch: each view's image equals its own single-view march, the
launch replays correctly from a HIP graph, and neither the tile shape nor the lanes per ray change
the image."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vs(mm, gpu):
    from marshmallow_amd import multiview

    specs = multiview.view_specs(n=6, seed=3, lo=(64, 48), hi=(300, 170))
    specs.append(multiview.ViewSpec(37, 23, 2500.0, 12.0, 0.6, 77))   # ragged, camera inside the shell
    return multiview.ViewSet(mm, specs)


def _single(mm, vs, prm):
    import torch

    outs = [torch.zeros_like(o) for o in vs.outs]
    vs.march_one_by_one(prm, outs)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("lanes,tile_width", [(1, 8), (1, 16), (0, 32), (2, 8), (4, 8), (8, 8)])
def test_views_equal_single_view_marches(mm, vs, lanes, tile_width):
    """lanes 1 (and 0, the default): 64-ray tiles tile_width wide; otherwise the ray-parallel kernel's
    8 x (8 / lanes) tiles."""
    import torch

    prm = mm.march_params("scenario", tile_width=tile_width, lanes_per_ray=lanes)
    ref = _single(mm, vs, mm.march_params("scenario"))
    for o in vs.outs:
        o.zero_()
    plan = mm.ViewPlan(vs.views(), vs.shared, prm)
    total, grid = plan.info()
    tw, th = (tile_width, 64 // tile_width) if lanes <= 1 else (8, 8 // lanes)
    assert total == sum(-(-s.width // tw) * -(-s.height // th) for s in vs.specs)
    assert grid == total
    cnt = torch.zeros(4, dtype=torch.int64, device="cuda")
    plan.launch(cnt)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(vs.outs, ref)):
        assert torch.equal(a, b), (i, vs.specs[i])
    assert int(cnt[0]) > 0 and int(cnt[2]) == int(cnt[3])
    plan.close()


def test_views_phase_and_update(mm, vs):
    """A phase plan writes only its phase's pixels. An update that would change the launch shape
    (phase 5 -> every pixel: more tiles) or the kernel (bounded-step checks -> the general ones) is
    refused, and a same-shape update (new suns, same phase) takes effect in a HIP graph captured
    before it."""
    import torch

    from marshmallow_amd import multiview

    prm = mm.march_params("scenario")
    plan = mm.ViewPlan(vs.views(phase=5), vs.shared, prm)
    for o in vs.outs:
        o.fill_(-1.0)
    plan.launch()
    torch.cuda.synchronize()
    for i, s in enumerate(vs.specs):
        img = vs.outs[i].view(s.height, s.width, 4)
        m = torch.zeros(s.height, s.width, dtype=torch.bool, device="cuda")
        m[1::4, 1::4] = True  # phase 5 -> (x % 4, y % 4) = (1, 1)
        assert (img[~m] == -1.0).all()
        if m.any():
            assert (img[m][:, 3] >= 0).all()
    t5, _ = plan.info()
    with pytest.raises(mm.MMError) as err:
        plan.update(vs.views(), prm)
    assert err.value.code == mm.MM_E_PARAM
    assert plan.info()[0] == t5
    # params whose step checks need the general loop (a counted-step limit) select another kernel
    general = mm.march_params("scenario", max_steps=1000)
    with pytest.raises(mm.MMError) as err:
        plan.update(vs.views(phase=5), general)
    assert err.value.code == mm.MM_E_PARAM

    g = multiview.capture(plan)
    ubos = [mm.build_ubos(mm.scenario_scene(altitude=s.altitude, aspect=s.width / s.height,
                                            sun_elevation_deg=s.sun_deg + 7.0)) for s in vs.specs]
    plan.update([mm.make_view(*ubos[i], vs.weathers[i], vs.outs[i], s.width, s.height, phase=5)
                 for i, s in enumerate(vs.specs)], prm)
    for o in vs.outs:
        o.fill_(-1.0)
    g.replay()
    torch.cuda.synchronize()
    for i, s in enumerate(vs.specs):
        ref = torch.full_like(vs.outs[i], -1.0)
        ns = multiview._Bound(mm.NoiseSetC(vs.weathers[i].desc, vs.shared.curl.desc, vs.shared.low_res.desc,
                                           vs.shared.hi_res.desc))
        mm.march_clouds(*ubos[i], ns, ref, s.width, s.height, phase=5, params=prm)
        torch.cuda.synchronize()
        assert torch.equal(vs.outs[i], ref), (i, s)
    plan.close()


def test_views_graph_replay(mm, vs):
    """The launch captured into a HIP graph: every replay (with eager work and syncs in between)
    reproduces the eager result."""
    import torch

    from marshmallow_amd import multiview

    prm = mm.march_params("scenario", tile_width=16)
    plan = mm.ViewPlan(vs.views(), vs.shared, prm)
    plan.launch()
    torch.cuda.synchronize()
    eager = [o.clone() for o in vs.outs]
    g = multiview.capture(plan)
    for rep in range(3):
        for o in vs.outs:
            o.zero_()
        vs.march_one_by_one(mm.march_params("scenario"), [torch.empty_like(o) for o in vs.outs])  # eager work
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(vs.outs, eager)):
            assert torch.equal(a, b), (rep, i)
    plan.close()


def test_views_surrogate_threshold_zero(mm, vs):
    import torch

    ref = _single(mm, vs, mm.march_params("scenario"))
    model = mm.Surrogate(seed=2)
    prm = mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=0.0, surrogate=model.ptr)
    plan = mm.ViewPlan(vs.views(), vs.shared, prm)
    plan.launch()
    torch.cuda.synchronize()
    for a, b in zip(vs.outs, ref):
        assert torch.equal(a, b)
    plan.close()


def test_tile_shape_and_lanes_per_ray_do_not_change_the_image(mm, gpu):
    import torch

    W, H = 200, 120
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=1, weather_seed=3, coverage=0.6), mm.current_stream())
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    variants = [(tw, flags, 1) for tw in (8, 16, 32) for flags in (0, mm.MARCH_NO_BATCH)]
    variants += [(8, 0, 2), (8, 0, 4), (8, 0, 8), (8, mm.MARCH_BOUNDED, 1), (8, mm.MARCH_BOUNDED, 4)]
    for phase in (mm.PHASE_ALL, 5):
        outs, cnts = [], []
        for tw, flags, lanes in variants:
            prm = mm.march_params("scenario", tile_width=tw, lanes_per_ray=lanes)
            prm.flags |= flags
            o = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
            c = torch.zeros(4, dtype=torch.int64, device="cuda")
            mm.march_clouds(cam, sun, sky, ns, o, W, H, phase=phase, params=prm, counts=c)
            outs.append(o)
            cnts.append(c)
        torch.cuda.synchronize()
        ref_o, ref_c, bnd_o, bnd_c = outs[0], cnts[0], outs[-2], cnts[-2]
        for (tw, flags, lanes), o, c in zip(variants, outs, cnts):
            want_o, want_c = (bnd_o, bnd_c) if flags & mm.MARCH_BOUNDED else (ref_o, ref_c)
            assert torch.equal(o, want_o), (phase, tw, flags, lanes)
            assert torch.equal(c, want_c), (phase, tw, flags, lanes)  # same samples, same cone jobs
