"""GPU noise generators (mm_generate_noise) against the oracle, byte for byte, in every storage format.

The curl map is additionally pinned to the reference's own output (CurlNoiseFBM.tga, via its SHA-256).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import fmt_util as F

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "curl_fbm_fixture.json")


@pytest.fixture(scope="module")
def ref(orc):
    orc.L.orc_set_curl_mutation(0)
    r = {
        "curl": orc.curl(128),
        "low": orc.low_res(128, 7),
        "low_f32": orc.low_res_f32(128, 7),
        "hi": orc.hi_res(32, 7),
        "hi_f32": orc.hi_res_f32(32, 7),
        "weather": orc.weather(512, 3, 0.5),
    }
    return r


def _gen(mm, fset, seed=7, weather_seed=3, coverage=0.5, workspace=True):
    import torch

    ns = mm.NoiseSet(fset)
    ns.generate(mm.noise_params(seed=seed, weather_seed=weather_seed, coverage=coverage), mm.current_stream(),
                workspace=workspace)
    torch.cuda.synchronize()
    return ns


def _bytes(t):
    return t.tensor.cpu().numpy().tobytes()


def test_curl_generator_matches_reference_output(mm, gpu, ref):
    ns = _gen(mm, "ref")
    got = np.frombuffer(_bytes(ns.curl), np.uint8).reshape(128, 128, 4)
    assert (got == ref["curl"][0]).all()
    with open(FIX) as f:
        fx = json.load(f)
    assert hashlib.sha256(np.ascontiguousarray(got[..., :3]).tobytes()).hexdigest() == fx["rgb_sha256"]


def test_generators_rgba8_linear_bit_exact(mm, gpu, ref):
    ns = _gen(mm, "ref")
    assert _bytes(ns.low_res) == ref["low"].tobytes()
    assert _bytes(ns.hi_res) == ref["hi"].tobytes()
    assert _bytes(ns.weather) == ref["weather"].tobytes()


def test_generators_fast_formats_bit_exact(mm, gpu, ref):
    ns = _gen(mm, "fast")
    assert _bytes(ns.weather) == F.pack_weather(ref["weather"], F.FMT_RG8, F.LAYOUT_APRON)
    assert _bytes(ns.curl) == F.pack_curl(ref["curl"][0], F.LAYOUT_APRON)
    assert _bytes(ns.low_res) == F.pack_low(ref["low"], F.FMT_RGBA8, F.LAYOUT_APRON)
    assert _bytes(ns.hi_res) == F.pack_hi(ref["hi"], F.FMT_RGBA8, F.LAYOUT_APRON)


def test_generators_precision_formats(mm, gpu, ref):
    ns = _gen(mm, "prec")
    inv = np.float32(1.0) / np.float32(1.0 / 448.0)
    assert _bytes(ns.low_res) == F.pack_low(None, F.FMT_FP8X4, F.LAYOUT_APRON,
                                            f32=(ref["low_f32"] * inv).astype(np.float32), inv_scale=1.0)
    assert _bytes(ns.hi_res) == F.pack_hi(None, F.FMT_FP16X4, F.LAYOUT_APRON, f32=ref["hi_f32"])
    # curl FP16: the remapped [0, 1] value before 8-bit rounding (ImageUtils.cpp:200-204)
    _, raw, mmx = ref["curl"]
    lo, hi = mmx[0::2].astype(np.float32), mmx[1::2].astype(np.float32)
    m = np.float32(0.0) + ((raw - lo) / (hi - lo) * np.float32(1.0))
    h = np.zeros((128, 128, 4), np.float16)
    h[..., :3] = m.astype(np.float16)
    assert _bytes(ns.curl) == F.apron(h, 2).tobytes()


@pytest.mark.parametrize("dim", [130, 36])
def test_low_res_ragged_dims_bit_exact(mm, gpu, orc, dim):
    # 130: partial bricks in every axis, apron copies, and bricks on both sides of the LDS staging
    # limit in one launch; 36: more lattice cells than texels, every brick hashes in place
    import torch

    want = orc.low_res(dim, 11)
    for fset, fmt, layout in (("ref", F.FMT_RGBA8, F.LAYOUT_LINEAR), ("fast", F.FMT_RGBA8, F.LAYOUT_APRON)):
        ns = mm.NoiseSet(fset, low_dim=dim)
        ns.generate(mm.noise_params(seed=11, what=mm.GEN_LOW_RES), mm.current_stream())
        torch.cuda.synchronize()
        assert _bytes(ns.low_res) == F.pack_low(want, fmt, layout), (fset, dim)


def test_fp8_encoder_matches_table(mm, gpu):
    # every e4m3 code decodes to a value that re-encodes to itself, ties go to the even code
    vals = F.fp8_e4m3_decode(np.arange(0x7F))
    assert (F.fp8_e4m3_encode(vals) == np.arange(0x7F)).all()
    mids = (vals[:-1] + vals[1:]) / 2
    enc = F.fp8_e4m3_encode(mids)
    assert (enc % 2 == 0).all()


def test_generate_without_workspace(mm, gpu, ref):
    ns = _gen(mm, "ref", workspace=False)
    assert (np.frombuffer(_bytes(ns.curl), np.uint8).reshape(128, 128, 4) == ref["curl"][0]).all()


def test_pack_texture_matches_numpy(mm, gpu, ref):
    import ctypes

    import torch

    for kind, arr, fmt, d in [(0, ref["weather"], mm.FMT_RG8, 1), (2, ref["low"], mm.FMT_RGBA8, 128),
                              (3, ref["hi"], mm.FMT_RGBA8, 32)]:
        n = arr.shape[1]
        src = mm.device_texture(n, n, d, mm.FMT_RGBA8, mm.LAYOUT_LINEAR)
        src.tensor.copy_(torch.from_numpy(arr.reshape(-1).view(np.uint8)))
        dst = mm.device_texture(n, n, d, fmt, mm.LAYOUT_APRON)
        mm.check(mm.lib().mm_pack_texture(kind, ctypes.byref(src.desc), ctypes.byref(dst.desc),
                                          mm.current_stream()), "pack")
        torch.cuda.synchronize()
        expect = {0: lambda: F.pack_weather(arr, fmt, 1), 2: lambda: F.pack_low(arr, fmt, 1),
                  3: lambda: F.pack_hi(arr, fmt, 1)}[kind]()
        assert _bytes(dst) == expect


def test_weather_coverage_knob(mm, gpu, orc):
    means = []
    for cov in (0.3, 0.5, 0.8):
        ns = _gen(mm, "ref", coverage=cov, weather_seed=2)
        w = np.frombuffer(_bytes(ns.weather), np.uint8).reshape(512, 512, 4)
        assert (w == orc.weather(512, 2, cov)).all()
        means.append(w[..., 0].mean())
    assert means[0] < means[1] < means[2]


def test_generation_captured_in_a_graph(mm, gpu):
    """mm_generate_noise forks the curl, hi-res and weather generators onto a second stream beside the
    low-res volume and joins them back with events; captured into a HIP graph, every replay (with
    eager work and syncs in between) reproduces the eager textures byte for byte."""
    import torch

    ns = mm.NoiseSet("fast")
    p = mm.noise_params(seed=5, weather_seed=6, coverage=0.4)
    ns.generate(p, mm.current_stream())
    torch.cuda.synchronize()
    slots = ("weather", "curl", "low_res", "hi_res")
    eager = {k: getattr(ns, k).tensor.clone() for k in slots}
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ns.generate(p, mm.current_stream())  # warm-up outside the graph on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        ns.generate(p, mm.current_stream())
    other = mm.NoiseSet("fast")
    for rep in range(3):
        for k in slots:
            getattr(ns, k).tensor.zero_()
        other.generate(mm.noise_params(seed=rep + 9), mm.current_stream())  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in slots:
            assert torch.equal(getattr(ns, k).tensor, eager[k]), (rep, k)
