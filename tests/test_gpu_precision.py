"""Precision variant (reduced-precision noise storage): FP8 e4m3 low-res volume with per-channel FP32
scale, FP16 hi-res volume and curl map, density/transmittance accumulated in FP32, shell
intersection in FP64. Error is reported against the FP64 oracle on 4096 seeded random pixels of a
3840x2160 frame, next to the FP32 march of the same frame."""
import numpy as np
import pytest

import fmt_util as F
from oracle import parity_stats

pytestmark = pytest.mark.gpu


def _fill(mm, ns_slot, arr_bytes):
    import torch

    ns_slot.tensor.copy_(torch.from_numpy(np.frombuffer(arr_bytes, np.uint8).copy()))


def test_fp8_decode_through_the_march(mm, gpu, orc):
    """Constant low-res volume whose FP8 values are e4m3-exact integers (scale 1/255): the PREC
    kernel must reproduce the RGBA8 kernel bit for bit (same filtered values, same arithmetic)."""
    import torch

    codes_val = np.array([160.0, 176.0, 192.0, 208.0])            # 1.25, 1.375, 1.5, 1.625 x 2^7
    codes = F.fp8_e4m3_encode(codes_val)
    assert (F.fp8_e4m3_decode(codes) == codes_val).all()
    weather = orc.weather(512, 1, 0.5)
    W = H = 96
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=1.0))
    prm = mm.march_params("scenario")
    # RGBA8 twin (reference set): constant low-res bytes, zero hi-res erosion, mid-grey curl
    ref = mm.NoiseSet("ref")
    low8 = np.broadcast_to(codes_val.astype(np.uint8), (128, 128, 128, 4)).copy()
    _fill(mm, ref.low_res, low8.tobytes())
    _fill(mm, ref.weather, weather.tobytes())
    _fill(mm, ref.curl, np.full((128, 128, 4), 0, np.uint8).tobytes())
    _fill(mm, ref.hi_res, np.zeros((32, 32, 32, 4), np.uint8).tobytes())
    prec = mm.NoiseSet("prec", fp8_scale=1.0 / 255.0)
    _fill(mm, prec.low_res, F.apron(np.broadcast_to(codes, (128, 128, 128, 4)).copy(), 3).tobytes())
    _fill(mm, prec.weather, F.pack_weather(weather, F.FMT_RG8, F.LAYOUT_APRON))
    _fill(mm, prec.curl, np.zeros(129 * 129 * 4, np.float16).tobytes())
    _fill(mm, prec.hi_res, np.zeros(33 ** 3 * 4, np.float16).tobytes())
    outs = []
    for ns in (ref, prec):
        out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
        mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy().reshape(H, W, 4))
    # curl: RGBA8 byte 0 decodes to 0.0, FP16 zero decodes to 0.0; hi-res erosion 0 in both
    assert np.array_equal(outs[0], outs[1])
    assert outs[0][..., 3].max() > 0


def test_precision_variant_error_vs_fp64(mm, gpu, orc):
    import torch

    W, H = 3840, 2160
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    prm = mm.march_params("scenario")
    rng = np.random.default_rng(1)
    xy = np.stack([rng.integers(0, W, 4096), rng.integers(0, H, 4096)], 1).astype(np.int32)
    res = {}
    for fset in ("fast", "prec"):
        ns = mm.NoiseSet(fset)
        ns.generate(mm.noise_params(seed=1, weather_seed=1, coverage=0.5), mm.current_stream())
        out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
        mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm)
        torch.cuda.synchronize()
        img = out.cpu().numpy().reshape(H, W, 4)
        assert np.isfinite(img).all()
        res[fset] = img[xy[:, 1], xy[:, 0]]
    ref = mm.NoiseSet("ref")
    ref.generate(mm.noise_params(seed=1, weather_seed=1, coverage=0.5), mm.current_stream())
    torch.cuda.synchronize()
    t = {k: getattr(ref, k).tensor.cpu().numpy() for k in ("weather", "curl", "low_res", "hi_res")}
    tex = orc.textures(t["weather"].reshape(512, 512, 4), t["curl"].reshape(128, 128, 4),
                       t["low_res"].reshape(128, 128, 128, 4), t["hi_res"].reshape(32, 32, 32, 4))
    r64 = orc.march_pixels(cam, sun, sky, tex, prm, W, H, xy, f64=True)
    for fset, g in list(res.items()):
        st = parity_stats(g, r64)
        print(f"{fset} vs FP64: {st}", flush=True)
        res[fset + "_stats"] = st
    # the FP32 march is within the parity bar of the FP64 oracle on almost every pixel
    assert res["fast_stats"]["frac_within_1"] >= 0.995
    # FP8/FP16 storage: bounded degradation (mean error a few 1/255 at most)
    assert res["prec_stats"]["mean"] <= 4.0 and res["prec_stats"]["nan"] == 0
