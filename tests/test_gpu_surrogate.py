"""Learned light-cone surrogate on the GPU: the MFMA training kernels' gradient against the CPU
oracle, inference against the oracle, the sample recorder, the plain-cone fallback, and what a
trained network gains over the plain march at the 1/255 error bar."""
import numpy as np
import pytest

import surrogate_mlp as S

pytestmark = pytest.mark.gpu


def _rows(rng, n):
    x = rng.normal(0.0, 1.0, (n, 17)).astype(np.float32)
    x[:, 16] = rng.uniform(0.0, 4.0, n)
    return x


def _model(mm, seed, stats=None):
    import torch

    m = mm.Surrogate(seed=seed)
    flat = S.init_params(seed)
    st = stats if stats is not None else np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    m.load(torch.from_numpy(flat.astype(np.float32)).cuda(), torch.from_numpy(st).cuda())
    return m, flat, st


def _rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def test_gradient_matches_oracle(mm, gpu):
    """Exact mean gradient of the MFMA forward/backward vs the oracle with the same bf16 roundings."""
    import torch

    rng = np.random.default_rng(4)
    stats = np.concatenate([rng.normal(0, 0.2, 16), rng.uniform(0.7, 1.5, 16)]).astype(np.float32)
    m, flat, st = _model(mm, 21, stats)
    for n in (64, 200, 4096):  # whole waves, a ragged last wave, many waves
        rows = _rows(rng, n)
        g = torch.zeros(mm.SURROGATE_PARAMS, dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        m.gradient(torch.from_numpy(rows).cuda(), g, loss)
        torch.cuda.synchronize()
        g_gpu = g.cpu().numpy().astype(np.float64)
        g_orc, L = S.gradient(flat, st, rows, bf16_ops=True)
        g64, _ = S.gradient(flat, st, rows)
        assert abs(loss.item() - L) <= 1e-3 * abs(L), (n, loss.item(), L)
        o = 0
        for name, shape in S.SHAPES:
            k = int(np.prod(shape))
            e = _rel_err(g_gpu[o:o + k], g_orc[o:o + k])
            print(f"n={n} {name}: rel err vs bf16 oracle {e:.2e}, vs f64 {_rel_err(g_gpu[o:o + k], g64[o:o + k]):.2e}",
                  flush=True)
            assert e <= 5e-3, (n, name, e)
            o += k


def test_predict_matches_oracle(mm, gpu):
    import torch

    rng = np.random.default_rng(5)
    m, flat, st = _model(mm, 8)
    rows = _rows(rng, 1000)
    out = torch.zeros(2000, dtype=torch.float32, device="cuda")
    m.predict(torch.from_numpy(rows).cuda(), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(-1, 2)
    # the FP16 inference path (the march's light-cone batches): FP32 MFMA accumulation against float64
    ref = S.predict(flat, st, rows, f16_ops=True)
    scale = np.max(np.abs(ref), 0)
    err = np.max(np.abs(got - ref), 0) / scale
    print("predict rel err", err, flush=True)
    assert (err <= 5e-3).all()


def test_training_fits_a_smooth_target(mm, gpu):
    """Adam on the GPU drives the loss of a smooth synthetic target down by an order of magnitude."""
    import torch

    rng = np.random.default_rng(6)
    n = 1 << 16
    x = rng.uniform(-1, 1, (n, 17)).astype(np.float32)
    x[:, 16] = 1.5 + np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]
    dev = torch.from_numpy(x).cuda()
    m = mm.Surrogate(seed=3)
    m.fit_stats(dev)
    losses = torch.zeros(400, dtype=torch.float32, device="cuda")
    m.train(dev, steps=400, batch=4096, lr=2e-3, seed=1, loss_out=losses)
    torch.cuda.synchronize()
    l = losses.cpu().numpy() / 4096
    print("loss first/last", l[:5].mean(), l[-20:].mean(), flush=True)
    assert np.isfinite(l).all() and l[-20:].mean() < 0.1 * l[:5].mean()


def _scene(mm, W, H, sun=25.0, coverage=0.5, seed=1):
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=1, weather_seed=seed, coverage=coverage), mm.current_stream())
    cam, sunu, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H, sun_elevation_deg=sun))
    return ns, cam, sunu, sky


def test_collect_is_the_plain_march(mm, gpu):
    import torch

    W = H = 256
    ns, cam, sun, sky = _scene(mm, W, H)
    prm = mm.march_params("scenario")
    a = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    mm.march_clouds(cam, sun, sky, ns, a, W, H, params=prm)
    rows = torch.zeros((1 << 21, 17), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    mm.march_collect(cam, sun, sky, ns, b, W, H, rows, cnt, keep_one_in=1, seed=1, params=prm)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    n = int(cnt.item())
    assert 0 < n <= rows.shape[0]
    r = rows[:n].cpu().numpy()
    assert np.isfinite(r).all()
    assert (r[:, 0] >= 0).all() and (r[:, 0] <= 1).all()          # relative height
    assert (r[:, 3:5] >= 0).all() and (r[:, 3:5] <= 1).all()      # coverage, type
    assert np.allclose(r[:, 5], sun.directionBasis[5] / np.linalg.norm(sun.directionBasis[4:7]), atol=1e-6)
    assert (r[:, 6:9] >= 0).all() and (r[:, 6:9] <= 1).all()      # probe densities
    assert (np.abs(r[:, 10:16]) <= 1.0 + 1e-3).all()
    assert (r[:, 16] >= 0).all()                                  # density along the light
    # thinning keeps about 1 in k, and the same rows for the same seed
    cnt2 = torch.zeros(1, dtype=torch.int32, device="cuda")
    rows2 = torch.zeros_like(rows)
    mm.march_collect(cam, sun, sky, ns, b, W, H, rows2, cnt2, keep_one_in=8, seed=1, params=prm)
    torch.cuda.synchronize()
    assert 0.08 * n < int(cnt2.item()) < 0.17 * n


def test_threshold_zero_is_the_plain_march(mm, gpu):
    """Every job whose error bound is >= 0 (all of them) takes the plain cone: bit-identical image."""
    import torch

    W = H = 256
    ns, cam, sun, sky = _scene(mm, W, H)
    m = mm.Surrogate(seed=1)
    prm = mm.march_params("scenario")
    a = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    mm.march_clouds(cam, sun, sky, ns, a, W, H, params=prm)
    ps = mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=0.0, surrogate=m.ptr)
    mm.march_clouds(cam, sun, sky, ns, b, W, H, params=ps)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_surrogate_rejects_prec_set_and_null_model(mm, gpu):
    import torch

    W = H = 64
    out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    ns, cam, sun, sky = _scene(mm, W, H)
    with pytest.raises(mm.MMError) as e:
        mm.march_clouds(cam, sun, sky, ns, out, W, H,
                        params=mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate=None))
    assert e.value.code == mm.MM_E_PARAM
    m = mm.Surrogate(seed=1)
    prec = mm.NoiseSet("prec")
    with pytest.raises(mm.MMError) as e:
        mm.march_clouds(cam, sun, sky, prec, out, W, H,
                        params=mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate=m.ptr))
    assert e.value.code == mm.MM_E_FORMAT


def test_trained_surrogate_gain(mm, gpu):
    """Train on recorded cone jobs of 6 scenes, validate on 2 held-out weather seeds. What the trained
    network gains over the method without it: (a) with every cone replaced, its image error is far
    below an untrained network's; (b) some threshold keeps the surrogate-lit march within 1/255 mean
    per-channel error of the plain march while replacing at least half of the cones."""
    from marshmallow_amd import training

    cfg = training.TrainConfig(width=384, height=216, train_seeds=(11, 12, 13, 14, 15, 16), rows_per_view=1 << 17,
                               steps=1500, batch=8192, lr=2e-3)
    model, info = training.train_surrogate(mm, cfg)
    print("training", info, flush=True)
    untrained = mm.Surrogate(seed=5)
    for seed in (101, 102):
        kw = dict(width=384, height=216, weather_seed=seed, sun_deg=30.0, coverage=0.55)
        every = training.evaluate(mm, model, threshold=1e30, **kw)
        base = training.evaluate(mm, untrained, threshold=1e30, **kw)
        print("held-out", seed, "all replaced: trained", every, "untrained", base, flush=True)
        assert every["replaced_frac"] == 1.0
        assert every["mean_err_255"] < 0.5 * base["mean_err_255"]
        ok = []
        for thr in (0.0005, 0.001, 0.002, 0.005, 0.01, 0.02):
            r = training.evaluate(mm, model, threshold=thr, **kw)
            print("  threshold", thr, r, flush=True)
            if r["mean_err_255"] <= 1.0 and r["replaced_frac"] >= 0.5:
                ok.append(thr)
        assert ok, "no threshold meets 1/255 mean error with half of the cones replaced"


def test_surrogate_lanes_per_ray_do_not_change_the_image(mm, gpu):
    """The surrogate-lit march with one lane per ray (cone-batched kernel) and with two (ray-parallel
    kernel, the default for a full frame): the same jobs reach the network and the fallback, in
    another queue order, so the image and the counters are the same bit for bit; likewise one phase."""
    import torch

    W, H = 200, 120
    ns, cam, sun, sky = _scene(mm, W, H, sun=35.0, coverage=0.7, seed=3)
    m = mm.Surrogate(seed=4)
    for phase in (mm.PHASE_ALL, 6):
        outs, cnts = [], []
        for lanes in (1, 2, 0):
            ps = mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=0.01, surrogate=m.ptr,
                                 lanes_per_ray=lanes)
            o = torch.full((W * H * 4,), -1.0, dtype=torch.float32, device="cuda")
            c = torch.zeros(4, dtype=torch.int64, device="cuda")
            mm.march_clouds(cam, sun, sky, ns, o, W, H, phase=phase, params=ps, counts=c)
            outs.append(o)
            cnts.append(c)
        torch.cuda.synchronize()
        print("phase", phase, "jobs", int(cnts[0][2]), "on the plain cone", int(cnts[0][3]), flush=True)
        assert int(cnts[0][2]) > 0
        for o, c in zip(outs[1:], cnts[1:]):
            assert torch.equal(o, outs[0]) and torch.equal(c, cnts[0]), phase
