"""The learned light cone and the many-view launch at their workloads.

Config 3: the surrogate trained by the library's own kernels to the training spec (2e7 recorded cone
jobs over 64 seeded weather maps, sun uniform in [2, 70] deg, coverage uniform in [0.2, 0.9]; Adam,
batch 8192), validated at 1920x1080 on 8 held-out weather seeds against the plain cone march
(compute-clouds.comp:440-464), which is also its fallback.

Config 4: 64 seeded views of log-uniform size from 256x256 to 1920x1080 (camera altitude, sun and
coverage drawn per view, one weather map per view, shared volumes), marched by one launch captured
into a HIP graph, with the plain cone and with the config-3 surrogate. The reference issues one
dispatch per view (VulkanApplication.cpp:1061-1071): every view must equal its own single-view march,
and sampled pixels the oracle. The autotuner runs once on the set.
"""
import numpy as np
import pytest

from oracle import parity_stats

pytestmark = pytest.mark.gpu

HELD_OUT = tuple(range(901, 909))
# the fallback ladder: 0 is the plain march (every cone on the plain path), 1e30 replaces every cone
THRESHOLDS = (0.0, 0.0001, 0.0002, 0.0005, 0.001, 0.002, 0.005, 0.01, 0.02, 1e30)


@pytest.fixture(scope="module")
def trained(mm, gpu):
    from marshmallow_amd import training

    cfg = training.TrainConfig()  # the training spec: 64 seeds x 312 500 rows = 2e7 rows, 20 000 Adam steps
    assert len(cfg.train_seeds) == 64 and len(cfg.train_seeds) * cfg.rows_per_view == 20_000_000
    model, info = training.train_surrogate(mm, cfg)
    print("config 3 training", info, flush=True)
    assert info["rows"] >= 0.5 * 20_000_000 and info["loss_last"] < 0.2 * info["loss_first"]
    return model, cfg


def test_config3_held_out_seeds_at_1080p(mm, trained):
    """On each of 8 held-out seeds (sun and coverage drawn like the training views), the largest
    threshold of the ladder that keeps the surrogate-lit 1920x1080 march within 1/255 mean per-channel
    error of the plain march, and the fraction of light cones the network replaces there; on average
    the network must replace most of them."""
    from marshmallow_amd import training

    model, cfg = trained
    rng = np.random.default_rng(77)
    rows = []
    for ws in HELD_OUT:
        sun, cov = float(rng.uniform(2.0, 70.0)), float(rng.uniform(0.2, 0.9))
        best = None
        for thr in THRESHOLDS:
            r = training.evaluate(mm, model, width=1920, height=1080, weather_seed=ws, sun_deg=sun, coverage=cov,
                                  threshold=thr)
            if r["mean_err_255"] <= 1.0:
                best = dict(r, threshold=thr)
        print("held-out", ws, round(sun, 1), round(cov, 2), best, flush=True)
        assert best is not None and best["density_samples"][0] > 0, ws
        rows.append(best)
    replaced = np.mean([r["replaced_frac"] for r in rows])
    print("config 3: mean replaced fraction at <= 1/255", replaced,
          "mean error", np.mean([r["mean_err_255"] for r in rows]), flush=True)
    assert replaced > 0.5


@pytest.fixture(scope="module")
def views64(mm, gpu):
    from marshmallow_amd import multiview

    specs = multiview.view_specs(n=64, seed=1)
    vs = multiview.ViewSet(mm, specs)
    print("config 4:", len(specs), "views,", vs.pixels(), "pixels", flush=True)
    return vs


def _single(mm, vs, prm):
    import torch

    outs = [torch.zeros_like(o) for o in vs.outs]
    vs.march_one_by_one(prm, outs)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("cone", ["plain", "surrogate"])
def test_config4_graph_launch_equals_single_view_marches(mm, views64, trained, cone):
    import torch

    from marshmallow_amd import multiview

    vs = views64
    if cone == "plain":
        prm = mm.march_params("scenario")
    else:
        prm = mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=0.01,
                              surrogate=trained[0].ptr)
    ref = _single(mm, vs, prm)
    plan = mm.ViewPlan(vs.views(), vs.shared, prm)
    g = multiview.capture(plan)
    for rep in range(2):
        for o in vs.outs:
            o.fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(vs.outs, ref)):
            assert torch.equal(a, b), (cone, rep, i, vs.specs[i])
    ms = multiview.time_launch(lambda: g.replay(), reps=5)
    # eager launches after recorded ones march all views' tiles longest first (the tile order over
    # the whole set, 0.5 M tiles): the same images
    for rep in range(2):
        for o in vs.outs:
            o.fill_(-1.0)
        plan.launch()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(vs.outs, ref)):
            assert torch.equal(a, b), (cone, "ordered", rep, i, vs.specs[i])
    cnt = torch.zeros(4, dtype=torch.int64, device="cuda")
    plan.launch(cnt)
    torch.cuda.synchronize()
    print(f"config 4 {cone}: {ms:.2f} ms per graph launch, {int(cnt[0]) / (ms * 1e-3):.3e} density samples/s, "
          f"{int(cnt[2])} cone jobs, {int(cnt[3])} on the plain cone", flush=True)
    plan.close()


def test_config4_sampled_pixels_against_the_oracle(mm, orc, views64):
    """256 seeded pixels of every view against the oracle on that view's own textures (the shared
    volumes and its weather map, all regenerated by the oracle's generators)."""
    import torch

    vs = views64
    prm = mm.march_params("scenario")
    plan = mm.ViewPlan(vs.views(), vs.shared, prm)
    plan.launch()
    torch.cuda.synchronize()
    shared = {"curl": orc.curl(128)[0], "low": orc.low_res(128, 1), "hi": orc.hi_res(32, 1)}
    rng = np.random.default_rng(4)
    got, want = [], []
    for i, s in enumerate(vs.specs):
        tex = orc.textures(orc.weather(512, s.weather_seed, s.coverage), shared["curl"], shared["low"], shared["hi"])
        xy = np.stack([rng.integers(0, s.width, 256), rng.integers(0, s.height, 256)], 1).astype(np.int32)
        cam, sun, sky = vs.ubos[i]
        want.append(orc.march_pixels(cam, sun, sky, tex, prm, s.width, s.height, xy))
        img = vs.outs[i].view(s.height, s.width, 4).cpu().numpy()
        got.append(img[xy[:, 1], xy[:, 0]])
    plan.close()
    st = parity_stats(np.concatenate(got), np.concatenate(want))
    print("config 4 sampled vs oracle", st, flush=True)
    assert st["nan"] == 0 and st["frac_within_1"] >= 0.999 and st["max"] <= 0.01


def test_config4_autotuner(mm, views64):
    from marshmallow_amd import multiview

    best, table = multiview.autotune(mm, views64, reps=3)
    print("config 4 autotune best", best, flush=True)
    assert len(table) == 15 and best["mean_err_255"] <= 1.0
    assert best["ms"] == min(r["ms"] for r in table if r["mean_err_255"] <= 1.0)
