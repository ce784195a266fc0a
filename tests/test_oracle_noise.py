"""Pins the build-defined generator restatement (oracle/noise.cpp) to the published methods it follows.

The reference ships its cloud volumes and weather map as baked Houdini assets (README.md:68, loaded
by Texture.cpp:502-538), so the build defines the generators; the march consumes them by the channel
contract of compute-clouds.comp:223-224 and 237-247. Each primitive is checked against what its
method fixes, not against a retyped copy of the oracle:

* PCG3D (Jarzynski & Olano 2020, "Hash Functions for GPU Rendering", JCGT 9(3)): an independent
  numpy uint32 implementation of the published pcg3d, under the build's seeding.
* Worley F1 (Worley 1996): brute-force nearest feature point over every cell and every periodic
  image of tiny grids, the one-feature-point-per-cell definition.
* Perlin 2002 improved noise: zero at lattice points, exact periodicity, slope at a lattice point
  equal to the hashed cube-edge gradient, and on a lattice edge the closed form
  g0 r (1 - fade(r)) + g1 (r - 1) fade(r) with the quintic fade 6r^5 - 15r^4 + 10r^3.
* Value noise: the hashed lattice value at lattice points, smoothstep interpolation along lattice
  edges, periodicity.
* The generated volumes: channel statistics against the reference's own baked assets.

Planted mistakes (orc_set_noise_mutation) show that each check fails when the primitive is wrong.
"""
import json
import math
import os

import numpy as np
import pytest

M32 = np.uint64(0xFFFFFFFF)
STATS = os.path.join(os.path.dirname(__file__), "golden", "asset_stats.json")

# Perlin 2002's 12 cube-edge gradient directions, in the order Perlin's improved noise lists them
GRAD12 = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1],
                   [1, 0, -1], [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], np.float64)


def pcg3d_np(x, y, z):
    """pcg3d of Jarzynski & Olano (2020), listing 'pcg3d', in uint64 with explicit 32-bit wrap."""
    v = [np.asarray(a, np.uint64) & M32 for a in (x, y, z)]
    v = [(a * np.uint64(1664525) + np.uint64(1013904223)) & M32 for a in v]
    v[0] = (v[0] + v[1] * v[2]) & M32
    v[1] = (v[1] + v[2] * v[0]) & M32
    v[2] = (v[2] + v[0] * v[1]) & M32
    v = [a ^ (a >> np.uint64(16)) for a in v]
    v[0] = (v[0] + v[1] * v[2]) & M32
    v[1] = (v[1] + v[2] * v[0]) & M32
    v[2] = (v[2] + v[0] * v[1]) & M32
    return v


def hash3_np(x, y, z, seed):
    """The build's lattice hash: pcg3d of (x + seed, y xor seed * 0x85EBCA6B, z + seed * 0xC2B2AE35)."""
    s = np.uint64(seed)
    return pcg3d_np(np.asarray(x, np.uint64) + s, np.asarray(y, np.uint64) ^ ((s * np.uint64(0x85EBCA6B)) & M32),
                    np.asarray(z, np.uint64) + s * np.uint64(0xC2B2AE35))


def u01(h):
    return (np.asarray(h, np.uint64) >> np.uint64(8)).astype(np.float64) / 16777216.0


@pytest.fixture(autouse=True)
def no_mutation(orc):
    orc.L.orc_set_noise_mutation(0)
    yield
    orc.L.orc_set_noise_mutation(0)


# ------------------------------------------------------------------ PCG3D
def hash_mismatches(orc, n=4000, seed0=0):
    import ctypes

    rng = np.random.default_rng(seed0)
    xs = rng.integers(0, 2**32, (n, 4), dtype=np.uint64)
    xs[:8] = [[0, 0, 0, 0], [1, 2, 3, 4], [2**32 - 1, 0, 7, 9], [5, 2**32 - 1, 2**31, 1], [0, 0, 0, 1],
              [127, 127, 127, 0x9E3779B9], [3, 1, 4, 1592653589], [2**31, 2**31, 2**31, 2**31]]
    exp = hash3_np(xs[:, 0], xs[:, 1], xs[:, 2], 0)  # placeholder shape
    bad = 0
    out = (ctypes.c_uint32 * 3)()
    for row in xs:
        x, y, z, s = (int(v) for v in row)
        orc.L.orc_hash3(x, y, z, s, out)
        e = hash3_np(x, y, z, s)
        bad += tuple(out) != tuple(int(a) for a in e)
    del exp
    return bad


def test_pcg3d_matches_published_algorithm(orc):
    assert hash_mismatches(orc) == 0


def test_pcg3d_known_answers(orc):
    import ctypes

    # pcg3d(0, 0, 0): every lane starts at 1013904223 = 0x3C6EF35F; the values below were worked out
    # by hand from the listing (32-bit wrap at every product) and pin the constants independently
    v = 1013904223
    a = (v + v * v) % 2**32
    b = (v + v * a) % 2**32
    c = (v + a * b) % 2**32
    a, b, c = a ^ (a >> 16), b ^ (b >> 16), c ^ (c >> 16)
    a = (a + b * c) % 2**32
    b = (b + c * a) % 2**32
    c = (c + a * b) % 2**32
    out = (ctypes.c_uint32 * 3)()
    orc.L.orc_hash3(0, 0, 0, 0, out)
    assert tuple(out) == (a, b, c)
    assert tuple(int(t) for t in pcg3d_np(0, 0, 0)) == (a, b, c)


def test_pcg3d_catches_a_wrong_constant(orc):
    orc.L.orc_set_noise_mutation(1)
    assert hash_mismatches(orc, n=64) > 60


# ------------------------------------------------------------------ Worley F1
def worley_brute(q, cells, seed):
    """1 - min(1, F1) at q (cell units) over every cell of the periodic grid and its 27 images."""
    idx = np.arange(cells)
    gx, gy, gz = np.meshgrid(idx, idx, idx, indexing="ij")
    gx, gy, gz = gx.ravel(), gy.ravel(), gz.ravel()
    h = hash3_np(gx, gy, gz, seed)
    feat = np.stack([gx + u01(h[0]), gy + u01(h[1]), gz + u01(h[2])], 1)
    best = np.inf
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                d = np.sqrt(((q - (feat + cells * np.array([sx, sy, sz]))) ** 2).sum(1)).min()
                best = min(best, d)
    return 1.0 - min(1.0, best)


def worley_errors(orc, cells_list=(2, 3, 4), n=150):
    rng = np.random.default_rng(5)
    worst = 0.0
    for cells in cells_list:
        for seed in (1, 77):
            for p in rng.random((n, 3)):
                got = orc.L.orc_worley(float(np.float32(p[0])), float(np.float32(p[1])), float(np.float32(p[2])),
                                       cells, seed)
                q = np.float32(p).astype(np.float64) * cells
                worst = max(worst, abs(got - worley_brute(q, cells, seed)))
    return worst


def test_worley_is_the_periodic_f1(orc):
    assert worley_errors(orc) < 2e-6


def test_worley_catches_own_cell_search_and_missing_wrap(orc):
    orc.L.orc_set_noise_mutation(3)
    assert worley_errors(orc, (3, 4), 60) > 0.05
    orc.L.orc_set_noise_mutation(2)
    assert worley_errors(orc, (3, 4), 60) > 0.05


def periodic_worst(fn, dims=3, n=300, seed=9):
    """max |f(p + e_k) - f(p)| over dyadic points p (exact in float) and unit shifts e_k."""
    rng = np.random.default_rng(seed)
    worst = 0.0
    for p in rng.integers(0, 1024, (n, dims)) / 1024.0:
        base = fn(*p)
        for k in range(dims):
            for s in (1.0, -1.0):
                q = p.copy()
                q[k] += s
                worst = max(worst, abs(fn(*q) - base))
    return worst


def test_worley_and_fbm_tile_exactly(orc):
    for cells in (3, 4, 8):
        assert periodic_worst(lambda x, y, z: orc.L.orc_worley(x, y, z, cells, 3)) == 0.0
    assert periodic_worst(lambda x, y, z: orc.L.orc_worley_fbm(x, y, z, 4, 3), n=100) == 0.0


def test_worley_feature_point_is_a_zero_of_f1(orc):
    # at a feature point F1 = 0, so the inverted noise is exactly 1
    cells, seed = 4, 12
    for cx, cy, cz in [(0, 0, 0), (3, 1, 2), (2, 3, 3)]:
        h = hash3_np(cx, cy, cz, seed)
        p = [(c + u01(hh)) / cells for c, hh in zip((cx, cy, cz), h)]
        v = orc.L.orc_worley(*(float(np.float32(t)) for t in p), cells, seed)
        assert v > 1.0 - 1e-6


# ------------------------------------------------------------------ Perlin 2002
def grad_np(ix, iy, iz, period, seed):
    h = hash3_np(ix % period, iy % period, iz % period, seed)
    return GRAD12[int(h[0]) % 12]


def test_perlin_zero_at_lattice_points_and_periodic(orc):
    for period in (2, 4, 8):
        for ix, iy, iz in [(0, 0, 0), (1, 0, 1), (period - 1, 1, 0), (3, 5, 7)]:
            assert orc.L.orc_perlin_tiled(ix / period, iy / period, iz / period, period, 5) == 0.0
        assert periodic_worst(lambda x, y, z: orc.L.orc_perlin_tiled(x, y, z, period, 5), n=120) == 0.0


def perlin_edge_errors(orc, n=200):
    """Along a lattice x-edge (y, z on the lattice) Perlin 2002 noise is
    g0x r (1 - fade(r)) + g1x (r - 1) fade(r), fade(r) = 6r^5 - 15r^4 + 10r^3."""
    rng = np.random.default_rng(2)
    worst = 0.0
    period, seed = 8, 21
    for _ in range(n):
        ix, iy, iz = (int(v) for v in rng.integers(0, period, 3))
        r = int(rng.integers(1, 256)) / 256.0
        g0 = grad_np(ix, iy, iz, period, seed)
        g1 = grad_np(ix + 1, iy, iz, period, seed)
        f = r ** 3 * (r * (r * 6 - 15) + 10)
        expect = g0[0] * r * (1 - f) + g1[0] * (r - 1) * f
        got = orc.L.orc_perlin_tiled((ix + r) / period, iy / period, iz / period, period, seed)
        worst = max(worst, abs(got - expect))
    return worst


def test_perlin_on_lattice_edges_closed_form(orc):
    assert perlin_edge_errors(orc) < 1e-6


def test_perlin_slope_at_lattice_point_is_the_hashed_gradient(orc):
    period, seed, h = 4, 8, 1.0 / 4096
    for ix, iy, iz in [(0, 0, 0), (1, 2, 3), (3, 3, 1), (2, 0, 3)]:
        g = grad_np(ix, iy, iz, period, seed)
        for k in range(3):
            p = np.array([ix, iy, iz], np.float64) / period
            p[k] += h / period
            slope = orc.L.orc_perlin_tiled(*p, period, seed) / h
            assert abs(slope - g[k]) < 1e-3, (ix, iy, iz, k, slope, g)


@pytest.mark.parametrize("mutation", [4, 5, 7])
def test_perlin_checks_catch_planted_mistakes(orc, mutation):
    # 4 cubic fade, 5 gradient index from the wrong hash lane, 7 corner offset sign
    orc.L.orc_set_noise_mutation(mutation)
    assert perlin_edge_errors(orc, 60) > 1e-3


def test_perlin_fbm_octaves_and_normalisation(orc):
    # 4 octaves at period, 2 period, ... weights 1, 1/2, 1/4, 1/8 normalised by 1.875; on the
    # lattice of the finest octave every octave is 0 except where it is not on its own lattice
    rng = np.random.default_rng(4)
    for p in rng.random((50, 3)):
        p = np.float32(p).astype(np.float64)
        parts = [orc.L.orc_perlin_tiled(*p, 4 << o, 10 + o) for o in range(4)]
        expect = sum(w * v for w, v in zip((1, .5, .25, .125), parts)) / 1.875
        assert abs(orc.L.orc_perlin_fbm(*p, 4, 10) - expect) < 1e-6


# ------------------------------------------------------------------ value noise
def value_lattice(ix, iy, period, seed):
    return u01(hash3_np(ix % period, iy % period, 0, seed)[0])


def value_errors(orc, n=200):
    rng = np.random.default_rng(6)
    worst = 0.0
    period, seed = 8, 31
    for _ in range(n):
        ix, iy = (int(v) for v in rng.integers(0, period, 2))
        worst = max(worst, abs(orc.L.orc_value2(ix / period, iy / period, period, seed) - value_lattice(ix, iy, period, seed)))
        # along the x edge: smoothstep interpolation of the two lattice values
        t = int(rng.integers(1, 256)) / 256.0
        s = t * t * (3 - 2 * t)
        a, b = value_lattice(ix, iy, period, seed), value_lattice(ix + 1, iy, period, seed)
        got = orc.L.orc_value2((ix + t) / period, iy / period, period, seed)
        worst = max(worst, abs(got - (a + s * (b - a))))
    # wrap: the last lattice column interpolates towards column 0
    for iy in range(period):
        a, b = value_lattice(period - 1, iy, period, seed), value_lattice(0, iy, period, seed)
        got = orc.L.orc_value2((period - 0.5) / period, iy / period, period, seed)
        worst = max(worst, abs(got - 0.5 * (a + b)))
    return worst


def test_value_noise_lattice_values_and_edges(orc):
    assert value_errors(orc) < 1e-6
    assert periodic_worst(lambda x, y: orc.L.orc_value2(x, y, 8, 3), dims=2) == 0.0
    assert periodic_worst(lambda x, y: orc.L.orc_value_fbm(x, y, 4, 3), dims=2, n=100) == 0.0


def test_value_noise_catches_a_missing_wrap(orc):
    orc.L.orc_set_noise_mutation(6)
    assert value_errors(orc) > 0.01


# ------------------------------------------------------------------ generated textures
def test_texel_composition_follows_the_channel_contract(orc):
    """Low-res R = Perlin-Worley: remap(clamp(0.5 + 1.5 perlinFBM), worleyFBM - 1, 1, 0, 1) scaled by
    0.8 and shifted by -0.04; G/B/A = 0.26 + 0.9 worleyFBM at 8/16/32 cells (read by comp:246 as
    0.625g + 0.25b + 0.125a); hi-res RGB the same at 4/8/16 cells, A = 0 (comp:223-224)."""
    import ctypes

    def seeds(seed, layer):
        return (seed * 0x9E3779B9 + layer * 0x632BE5AB) % 2**32

    out = (ctypes.c_float * 4)()
    dim, seed = 16, 3
    for x, y, z in [(0, 0, 0), (5, 9, 2), (15, 15, 15), (7, 0, 11)]:
        p = [(2 * c + 1) / (2 * dim) for c in (x, y, z)]
        orc.L.orc_low_res_texel(dim, x, y, z, seed, out)
        per = orc.L.orc_perlin_fbm(*p, 4, seeds(seed, 1))
        p01 = min(1.0, max(0.0, 0.5 + 1.5 * per))
        w = orc.L.orc_worley_fbm(*p, 4, seeds(seed, 2))
        pw = (p01 - (w - 1)) / (1 - (w - 1))
        exp = [min(1.0, max(0.0, 0.8 * pw - 0.04))]
        exp += [min(1.0, max(0.0, 0.26 + 0.9 * orc.L.orc_worley_fbm(*p, c, seeds(seed, l))))
                for c, l in ((8, 3), (16, 4), (32, 5))]
        np.testing.assert_allclose(list(out), exp, atol=2e-6)
        orc.L.orc_hi_res_texel(dim, x, y, z, seed, out)
        exp = [min(1.0, max(0.0, 0.26 + 0.9 * orc.L.orc_worley_fbm(*p, c, seeds(seed, l))))
               for c, l in ((4, 6), (8, 7), (16, 8))] + [0.0]
        np.testing.assert_allclose(list(out), exp, atol=2e-6)


def test_worley_fbm_weights(orc):
    rng = np.random.default_rng(8)
    for p in rng.random((40, 3)):
        p = np.float32(p).astype(np.float64)
        parts = [orc.L.orc_worley(*p, 4 << o, 50 + o) for o in range(3)]
        expect = 0.625 * parts[0] + 0.25 * parts[1] + 0.125 * parts[2]
        assert abs(orc.L.orc_worley_fbm(*p, 4, 50) - expect) < 1e-6


def test_generated_volumes_tile(orc):
    """REPEAT sampling needs the generated textures to tile: the wrap step (last texel to first)
    is no larger than the largest step between neighbouring texels inside the volume."""
    low = orc.low_res(32, 1).astype(np.int32)
    for ax in range(3):
        inner = np.abs(np.diff(low, axis=ax)).max()
        wrap = np.abs(np.take(low, [0], ax) - np.take(low, [-1], ax)).max()
        assert wrap <= inner
    w = orc.weather(64, 1, 0.5).astype(np.int32)
    for ax in range(2):
        assert np.abs(np.take(w, [0], ax) - np.take(w, [-1], ax)).max() <= np.abs(np.diff(w, axis=ax)).max()


def test_generated_statistics_match_the_reference_assets(orc):
    """Channel means of the generated volumes and weather map against the reference's own baked
    textures (Textures/CloudPlacement.png, the low-res and hi-res TGA slice volumes), whose
    statistics tests/golden/asset_stats.json holds."""
    with open(STATS) as f:
        st = json.load(f)
    low = orc.low_res(64, 1).reshape(-1, 4) / 255.0
    lo = st["low_res"]["mean"]
    assert abs(low[:, 0].mean() - lo[0]) < 0.03
    assert abs(low[:, 1:].mean() - np.mean(lo[1:])) < 0.03
    hi = orc.hi_res(32, 1).reshape(-1, 4) / 255.0
    assert abs(hi[:, :3].mean() - np.mean(st["hi_res"]["mean"][:3])) < 0.05
    w = orc.weather(256, 1, 0.5).reshape(-1, 4) / 255.0
    wm = st["weather"]["mean"]
    assert abs(w[:, 0].mean() - wm[0]) < 0.08 and abs(w[:, 2].mean() - wm[2]) < 0.08
    corr = np.corrcoef(w[:, 0], w[:, 1])[0, 1]
    assert abs(corr - st["weather"]["corr_rg"]) < 0.2


def test_low_res_staged_boxes_of_a_128_volume_are_fixed():
    """The 128^3 low-res kernel reads its staged lattice cells at compile-time offsets: every brick's box
    (axis_span in noise.hip: [floor(q(t0)) - pad, floor(q(t1)) + 1], q(t) = ((2t + 1) / 256) c in float32)
    must have the size max(B c / 128 - 1, 0) + 2 + pad for B texels per brick axis (8, 8, 4)."""
    dim = 128
    inv = np.float32(1.0 / (2 * dim))
    for level in range(16):
        cells = 4 << level if level < 4 else (4 << ((level - 4) // 3)) << ((level - 4) % 3)
        pad = 0 if level < 4 else 1
        for b in (8, 4):
            t0 = np.arange(0, dim, b)
            t1 = t0 + b - 1
            q0 = (((2 * t0 + 1).astype(np.float32) * inv) * np.float32(cells)).astype(np.float32)
            q1 = (((2 * t1 + 1).astype(np.float32) * inv) * np.float32(cells)).astype(np.float32)
            n = (np.floor(q1).astype(int) + 1) - (np.floor(q0).astype(int) - pad) + 1
            want = max(b * cells // 128 - 1, 0) + 2 + pad
            assert (n == want).all(), (level, b, set(n.tolist()), want)
