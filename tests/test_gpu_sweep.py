"""The plain FP32 march over the scene sweep of the 1080p workload (coverage 0.3 / 0.5 / 0.8, sun
5 / 25 / 60 deg, wind time 0 and 600 s, five seeded weather maps), at a small resolution, against the
oracle on the same textures, in the reference and the packed format sets, with the oracle's exact
sample counts."""
import itertools

import numpy as np
import pytest

from test_gpu_march import check_parity, gpu_march, upload

pytestmark = pytest.mark.gpu

CASES = [(cov, sun, t, seed) for (cov, sun), (t, seed) in
         zip(itertools.product((0.3, 0.5, 0.8), (5.0, 25.0, 60.0)),
             itertools.cycle([(0.0, 1), (600.0, 2), (0.0, 3), (600.0, 4), (300.0, 5)]))]


@pytest.fixture(scope="module")
def shared(orc):
    return {"curl": orc.curl(128)[0], "low": orc.low_res(128, 1), "hi": orc.hi_res(32, 1)}


@pytest.mark.parametrize("cov,sun,t,seed", CASES)
def test_scene_sweep_parity(mm, gpu, orc, shared, cov, sun, t, seed):
    W, H = 320, 180
    tex = dict(shared, weather=orc.weather(512, seed, cov))
    cam, sunu, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H, sun_elevation_deg=sun, time=t))
    prm = mm.march_params("scenario")
    r, rc = orc.march(cam, sunu, sky, orc.textures(tex["weather"], tex["curl"], tex["low"], tex["hi"]), prm, W, H,
                      counts=True)
    for fset in ("ref", "fast"):
        g, c = gpu_march(mm, upload(mm, tex, fset), cam, sunu, sky, W, H, params=prm, counts=True)
        check_parity(g, r, f"cov {cov} sun {sun} t {t} seed {seed} {fset}")
        assert int(c[0]) == int(rc[..., 0].sum()) and int(c[1]) == int(rc[..., 1].sum())
    assert np.isfinite(r).all()
