"""Transmittance error against the FP64 oracle, stated per config, in the two forms the reference
keeps: the opacity complement 1 - accum after the horizon fade and 0.999 cap (comp:476-486, 496) and
the light-energy accumulator `transmittance` (comp:464). mm_march_clouds_state writes both per pixel.

Config 0: 256x256, seed-1 textures, every pixel. Config 1: the 1920x1080 frame, every pixel.
Config 2: the 3840x2160 frame of the precision variant (FP8 / FP16 volumes), 4096 seeded pixels.
The FP32 kernel is also held to the FP32 oracle: the opacity accumulator is bit-exact (the density
path rounds like the restatement), the light accumulator within a few float ulps (its exp / pow are
the hardware's)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _textures(mm, orc, fset, seed=1, weather_seed=1, coverage=0.5):
    import torch

    ns = mm.NoiseSet(fset)
    ns.generate(mm.noise_params(seed=seed, weather_seed=weather_seed, coverage=coverage), mm.current_stream())
    ref = mm.NoiseSet("ref")
    ref.generate(mm.noise_params(seed=seed, weather_seed=weather_seed, coverage=coverage), mm.current_stream())
    torch.cuda.synchronize()
    t = {k: getattr(ref, k).tensor.cpu().numpy() for k in ("weather", "curl", "low_res", "hi_res")}
    tex = orc.textures(t["weather"].reshape(512, 512, 4), t["curl"].reshape(128, 128, 4),
                       t["low_res"].reshape(128, 128, 128, 4), t["hi_res"].reshape(32, 32, 32, 4))
    return ns, tex


def _gpu_state(mm, ns, cam, sun, sky, W, H, prm):
    import torch

    out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    st = torch.zeros(W * H * 2, dtype=torch.float32, device="cuda")
    mm.march_clouds_state(cam, sun, sky, ns, out, st, W, H, params=prm)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1, 4), st.cpu().numpy().reshape(-1, 2)


def _errors(g, r):
    """Absolute errors of the opacity complement and of the light accumulator."""
    d_op = np.abs((1.0 - g[:, 0].astype(np.float64)) - (1.0 - r[:, 0]))
    d_tr = np.abs(g[:, 1].astype(np.float64) - r[:, 1])
    return {"opacity_complement": {"mean": float(d_op.mean()), "p999": float(np.quantile(d_op, 0.999)),
                                   "max": float(d_op.max())},
            "light_accumulator": {"mean": float(d_tr.mean()), "p999": float(np.quantile(d_tr, 0.999)),
                                  "max": float(d_tr.max())}}


@pytest.mark.parametrize("config,W,H", [(0, 256, 256), (1, 1920, 1080)])
def test_transmittance_error_vs_fp64(mm, gpu, orc, config, W, H):
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    prm = mm.march_params("scenario")
    ns, tex = _textures(mm, orc, "fast")
    rgba, st = _gpu_state(mm, ns, cam, sun, sky, W, H, prm)
    o32, s32 = orc.march_state(cam, sun, sky, tex, prm, W, H)
    o64, s64 = orc.march_state(cam, sun, sky, tex, prm, W, H, f64=True)
    marched = s32[:, 1] != 1.0
    print(f"config {config}: FP32 kernel vs FP64", _errors(st, s64), flush=True)
    print(f"config {config}: FP32 oracle vs FP64", _errors(s32, s64), flush=True)
    # against the FP32 restatement: the opacity accumulator bit for bit, the light accumulator close
    assert np.array_equal(st[:, 0], s32[:, 0])
    rel = np.abs(st[:, 1] - s32[:, 1]) / np.maximum(np.abs(s32[:, 1]), 1e-3)
    assert rel.max() <= 1e-4, float(rel.max())
    assert marched.mean() > 0.2
    # against FP64 (the FP32 march's own rounding, mostly the single-precision step positions):
    # the bar is that of the colours, applied to both transmittance forms
    e = _errors(st, s64)
    assert e["opacity_complement"]["mean"] <= 1.0 / 255 and e["light_accumulator"]["mean"] <= 1.0 / 255


def test_precision_variant_transmittance_error_vs_fp64(mm, gpu, orc):
    """Config 2 (FP8 low-res volume, FP16 hi-res and curl) on 4096 seeded pixels of the 4K frame,
    next to the FP32 (FAST) march of the same frame."""
    W, H = 3840, 2160
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    prm = mm.march_params("scenario")
    rng = np.random.default_rng(1)
    xy = np.stack([rng.integers(0, W, 4096), rng.integers(0, H, 4096)], 1).astype(np.int32)
    res = {}
    for fset in ("fast", "prec"):
        ns, tex = _textures(mm, orc, fset)
        _, st = _gpu_state(mm, ns, cam, sun, sky, W, H, prm)
        res[fset] = st.reshape(H, W, 2)[xy[:, 1], xy[:, 0]]
    _, s64 = orc.march_state(cam, sun, sky, tex, prm, W, H, f64=True, xy=xy)
    e_fast, e_prec = _errors(res["fast"], s64), _errors(res["prec"], s64)
    print("config 2: FAST vs FP64", e_fast, flush=True)
    print("config 2: PREC vs FP64", e_prec, flush=True)
    assert e_fast["opacity_complement"]["mean"] <= 1.0 / 255
    assert np.isfinite(res["prec"]).all() and e_prec["opacity_complement"]["mean"] <= 16.0 / 255
