"""Post chain (mm_post_process: god rays -> radial blur -> tonemap) on the GPU against the oracle."""
import numpy as np
import pytest

from test_oracle_post import _ubos

pytestmark = pytest.mark.gpu


def _run(mm, cam, sun, img, params, pitch=None):
    import torch

    H, W = img.shape[:2]
    pitch = W if pitch is None else pitch
    src = torch.zeros((H, pitch, 4), dtype=torch.float32)
    src[:, :W] = torch.from_numpy(img)
    src = src.cuda()
    alpha = torch.full((2, H, pitch), -7.0, dtype=torch.float32, device="cuda")
    hdr = torch.full((H, pitch, 4), -7.0, dtype=torch.float32, device="cuda")
    out = torch.zeros((H, pitch, 4), dtype=torch.uint8, device="cuda")
    mm.post_process(cam, sun, src, W, H, alpha, hdr, out, pitch=pitch, params=params)
    torch.cuda.synchronize()
    return alpha[1].cpu().numpy()[:, :W], hdr.cpu().numpy()[:, :W], out.cpu().numpy()[:, :W]


@pytest.mark.parametrize("W,H,pitch", [(160, 90, None), (37, 23, 48), (1920, 1080, None)])
def test_post_matches_oracle_on_random_images(mm, gpu, orc, W, H, pitch):
    rng = np.random.default_rng(W * H)
    img = rng.uniform(0.0, 2.0, (H, W, 4)).astype(np.float32)
    img[..., 3] = rng.uniform(0.0, 1.0, (H, W))
    cam, sun, _ = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    p = mm.post_params()
    ga, gh, go = _run(mm, cam, sun, img, p, pitch)
    oa, oh, oo = orc.post_process(cam, sun, img, p)
    assert np.array_equal(ga, oa)
    assert np.array_equal(gh, oh)
    assert np.array_equal(go, oo)


def test_post_on_a_marched_frame(mm, gpu, orc):
    import torch

    W, H = 320, 180
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=1, weather_seed=1, coverage=0.5), mm.current_stream())
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    mm.march_clouds(cam, sun, sky, ns, img, W, H, params=mm.march_params("scenario"))
    torch.cuda.synchronize()
    img = img.cpu().numpy().reshape(H, W, 4)
    p = mm.post_params()
    ga, gh, go = _run(mm, cam, sun, img, p)
    oa, oh, oo = orc.post_process(cam, sun, img, p)
    assert np.array_equal(ga, oa) and np.array_equal(gh, oh) and np.array_equal(go, oo)
    assert 0 < go[..., :3].mean() < 255


@pytest.mark.parametrize("case", ["below", "sun_left", "few_samples", "coarse_subtexel", "sun_behind"])
def test_post_variants(mm, gpu, orc, case):
    rng = np.random.default_rng(11)
    img = rng.uniform(0.0, 3.0, (64, 96, 4)).astype(np.float32)
    p = mm.post_params()
    cam, sun = _ubos(mm)
    if case == "below":
        cam, sun = _ubos(mm, dir_y=-0.2)
    elif case == "sun_left":
        cam, sun = _ubos(mm, sx=-1.7, sy=0.4)
    elif case == "few_samples":
        p = mm.post_params(god_ray_samples=3, god_ray_decay=0.9)
    elif case == "coarse_subtexel":
        p = mm.post_params(subtexel_bits=1)
    elif case == "sun_behind":
        cam, sun = _ubos(mm, sx=0.3, sy=0.2, w=-1.0)  # w < 0: the projected point flips sides
    ga, gh, go = _run(mm, cam, sun, img, p)
    oa, oh, oo = orc.post_process(cam, sun, img, p)
    assert np.array_equal(ga, oa) and np.array_equal(gh, oh) and np.array_equal(go, oo)


def test_post_optional_outputs_and_errors(mm, gpu):
    import torch

    cam, sun = _ubos(mm)
    img = torch.rand((20, 30, 4), dtype=torch.float32, device="cuda")
    alpha = torch.zeros((2, 20, 30), dtype=torch.float32, device="cuda")
    out = torch.zeros((20, 30, 4), dtype=torch.uint8, device="cuda")
    mm.post_process(cam, sun, img, 30, 20, alpha, None, out)            # no HDR output
    hdr = torch.zeros((20, 30, 4), dtype=torch.float32, device="cuda")
    mm.post_process(cam, sun, img, 30, 20, alpha, hdr, None)            # no RGBA8 output
    torch.cuda.synchronize()
    assert (out[..., 3] == 255).all() and (hdr[..., 3] == 1.0).all()
    with pytest.raises(mm.MMError) as e:
        mm.post_process(cam, sun, img, 30, 20, alpha, hdr, out, params=mm.post_params(subtexel_bits=0))
    assert e.value.code == mm.MM_E_PARAM
    with pytest.raises(mm.MMError) as e:
        mm.post_process(cam, sun, img, 30, 20, alpha, hdr, out, pitch=29)
    assert e.value.code == mm.MM_E_PITCH
