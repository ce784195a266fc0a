"""mm_generate_noise_noalloc (the allocation-free generator entry point, scratch in a fixed device
array) against mm_generate_noise (library or caller workspace), byte for byte, on two calls in a row,
the second with smaller textures; a curl map beyond the scratch limit is refused before any launch."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _pair(mm, fset, dims, seed, weather_seed, coverage):
    import torch

    wd, cd, ld, hd = dims
    a = mm.NoiseSet(fset, weather_dim=wd, curl_dim=cd, low_dim=ld, hi_dim=hd)
    b = mm.NoiseSet(fset, weather_dim=wd, curl_dim=cd, low_dim=ld, hi_dim=hd)
    for t in (a.weather, a.curl, a.low_res, a.hi_res):
        t.tensor.fill_(0xA5)  # the entry point must write every byte the other one writes
    p = mm.noise_params(seed=seed, weather_seed=weather_seed, coverage=coverage)
    s = mm.current_stream()
    mm.check(mm.lib().mm_generate_noise_noalloc(ctypes.byref(p), ctypes.byref(a.c), s), "mm_generate_noise_noalloc")
    b.generate(p, s, workspace=False)
    torch.cuda.synchronize()
    return a, b


@pytest.mark.parametrize("fset", ["fast", "ref", "prec"])
def test_noalloc_generator_equals_the_allocating_one(mm, gpu, fset):
    import torch

    calls = [((512, mm.STATIC_CURL_MAX, 128, 32), 7, 3, 0.5),   # the largest curl map the scratch holds
             ((128, 64, 32, 16), 11, 5, 0.8)]                     # then smaller textures, other seeds
    for dims, seed, wseed, cov in calls:
        a, b = _pair(mm, fset, dims, seed, wseed, cov)
        for name in ("weather", "curl", "low_res", "hi_res"):
            ta, tb = getattr(a, name).tensor, getattr(b, name).tensor
            assert torch.equal(ta, tb), (fset, dims, name)


def test_noalloc_refuses_a_curl_map_beyond_its_scratch(mm, gpu):
    ns = mm.NoiseSet("fast", weather_dim=64, curl_dim=2 * mm.STATIC_CURL_MAX, low_dim=16, hi_dim=16)
    p = mm.noise_params(seed=1, weather_seed=1, coverage=0.5)
    r = mm.lib().mm_generate_noise_noalloc(ctypes.byref(p), ctypes.byref(ns.c), mm.current_stream())
    assert r == mm.MM_E_SIZE
    p.what = mm.GEN_LOW_RES | mm.GEN_HI_RES | mm.GEN_WEATHER  # without the curl map it needs no scratch
    assert mm.lib().mm_generate_noise_noalloc(ctypes.byref(p), ctypes.byref(ns.c), mm.current_stream()) == 0
