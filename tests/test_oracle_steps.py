"""The march loop of compute-clouds.comp:386-482, pinned by a step sequence worked out by hand.

Scene: one vertical ray (the centre pixel of a 2 x 2 image whose camera looks straight up) from
194 m above the ground, in the scenario shell (Earth radius 6.371e6 m, shell 1500-4000 m, coarse
step 0.05 x 2500 m, fine step x 0.3, 10-miss revert). Uniform textures make the density a function
of the relative height h alone:

* weather B = 0 (cloud type 0): the stratus profile, 10h (3 - 10h) on (0, 0.3) (comp:193-204);
* low-res R = 79, GBA = 0: density = stratus(h) x (79/255 - 0.3) / 0.7, erosion 0 (comp:238-250);
  the 1e-4 early out (comp:243) ends the layer at 749.4 m above the shell base;
* hi-res RGB = 255: the detail erosion is 1 - clamp(10h, 0, 1) (comp:223-226), so a fine sample
  below h* = 0.09724 (243.1 m) has detail density 0: an uncounted `continue` (comp:437).

The ray height above the shell base is t - 1306 m: the reference's `t = length(p - ro')` (comp:173)
measures the entry point's distance from the world origin, not from the camera, so the march
starts 193.5 m into the shell. The expected trips, by hand:

1. coarse trip at 193.5 m: density > 0, the first hit: back-step 125 m, step 37.5 m (comp:428-433);
2. fine trips from 106 m every 37.5 m: 106, 143.5, 181, 218.5 m lie below h*: uncounted `continue`;
3. 256 ... 743.5 m: 14 fine hits, each with a six-sample light cone (comp:440-466);
4. 781 ... 1118.5 m: 10 fine misses; the tenth reverts to coarse steps of 37.5 / 0.3 (comp:468-474);
5. 1243.5 ... 2618.5 m: 12 coarse misses, until t passes the outer shell (t = 3999.5, 2693.5 m).

Sample counts: density = one per trip + six per fine hit (comp:431-453); detail = one per fine trip
whose low-res density is positive, + one per cone sample whose density is positive.
"""
import ctypes
import math

import numpy as np
import pytest

R = 6.371e6
BASE, TOP = 1500.0, 4000.0
THICK = TOP - BASE
R79 = (np.float32(79) / np.float32(255) - 0.3) / 0.7


def stratus(h):
    return max(0.0, 10 * h * (3 - 10 * h))


def density(h):
    d = stratus(h) * R79
    return d if d >= 1e-4 else 0.0


def detail(h):
    d = density(h)
    e = 1.0 - min(1.0, max(0.0, 10 * h))
    if d <= 0:
        return 0.0
    return min(1.0, max(0.0, (d - e) / (1 - e))) if e < 1 else 0.0


def uniform_textures(orc):
    weather = np.zeros((8, 8, 4), np.uint8)
    weather[..., 0] = 128
    weather[..., 3] = 255
    curl = np.full((8, 8, 4), 128, np.uint8)
    low = np.zeros((4, 4, 4, 4), np.uint8)
    low[..., 0] = 79
    hi = np.full((4, 4, 4, 4), 255, np.uint8)
    hi[..., 3] = 0
    return orc.textures(weather, curl, low, hi)


def vertical_camera(mm, altitude):
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(altitude=altitude, aspect=1.0, sun_elevation_deg=25.0))
    v = [0.0] * 16
    # column-major view matrix: rows 0/1/2 = right / up / backward (camera.cpp:33-35)
    right, up, back = (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, -1.0, 0.0)
    for c in range(3):
        v[4 * c + 0], v[4 * c + 1], v[4 * c + 2] = right[c], up[c], back[c]
    v[15] = 1.0
    cam.view[:] = v
    return cam, sun, sky


def trace(orc, cam, sun, sky, tex, prm):
    trips = np.zeros((4096, 4), np.float32)
    cnt = np.zeros(2, np.int32)
    out = np.zeros(4, np.float32)
    n = orc.L.orc_march_trace_f32(ctypes.addressof(cam), ctypes.addressof(sun), ctypes.addressof(sky),
                                  ctypes.addressof(tex), ctypes.addressof(prm), 2, 2, 1, 1,
                                  trips.ctypes.data_as(ctypes.c_void_p), 4096,
                                  cnt.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    assert 0 < n <= 4096
    return trips[:n], cnt, out


def cone_heights(sun, y_rel, step):
    """Relative heights of the six cone samples of a fine hit at height y_rel (vertical ray on the
    shell axis): the sample sits at 3 step x (b0 c0 + b1 c1 + b2 c2) from the hit (comp:392-401,
    441-446), and its height is its distance from the Earth centre minus the shell radius."""
    db = np.array(sun.directionBasis, np.float64).reshape(4, 4)[:3, :3]  # columns N, sun, B
    tab = [(0, .6, 0), (0, .5, .05), (.1, .75, 0), (.2, 2.5, .3), (0, 6, 0), (-.1, 1, -.2)]
    hs = []
    for c in tab:
        off = 3 * step * (db[0] * c[0] + db[1] * c[1] + db[2] * c[2])
        p = np.array([0.0, R + BASE + y_rel, 0.0]) + off
        hs.append((np.linalg.norm(p) - (R + BASE)) / THICK)
    return hs


def expected(sun, y0, t_end_rel):
    """Event list (y_rel, event) and counts for a ray entering at y0 metres above the shell base."""
    ev = []
    y, step, fine, misses = y0, 0.05 * THICK, False, 0
    nd = nh = 0
    accum = 0.0
    while y < t_end_rel:
        h = min(1.0, max(0.0, y / THICK))
        nd += 1
        d = density(h)
        if d > 0:
            misses = 0
            if not fine:
                ev.append((y, 1))
                y -= step
                step *= 0.3
                fine = True
            else:
                nh += 1
                dt = detail(h)
                if dt < 1e-4:
                    ev.append((y, 2))
                else:
                    ev.append((y, 3))
                    for hc in cone_heights(sun, y, step):
                        nd += 1
                        if density(hc) > 0:
                            nh += 1
                    accum += dt
        elif fine:
            misses += 1
            ev.append((y, 5 if misses >= 10 else 4))
            if misses >= 10:
                fine = False
                step /= 0.3
        else:
            ev.append((y, 0))
        y += step
    return ev, nd, nh, accum


def test_hand_derived_step_sequence(mm, orc):
    cam, sun, sky = vertical_camera(mm, 194.0)
    prm = mm.march_params("scenario")
    trips, cnt, out = trace(orc, cam, sun, sky, uniform_textures(orc), prm)
    y_rel = trips[:, 0].astype(np.float64) - 1306.0
    events = trips[:, 3].astype(int).tolist()

    # the sequence written out above
    assert abs(y_rel[0] - 193.5) < 0.01 and events[0] == 1
    assert events[1:5] == [2, 2, 2, 2]
    np.testing.assert_allclose(y_rel[1:5], [106.0, 143.5, 181.0, 218.5], atol=0.01)
    assert events[5:19] == [3] * 14
    np.testing.assert_allclose(y_rel[5:19], 256.0 + 37.5 * np.arange(14), atol=0.01)
    assert events[19:29] == [4] * 9 + [5]
    np.testing.assert_allclose(y_rel[19:29], 781.0 + 37.5 * np.arange(10), atol=0.01)
    assert events[29:] == [0] * 12 and len(events) == 41
    np.testing.assert_allclose(y_rel[29:], 1243.5 + 125.0 * np.arange(12), atol=0.02)
    t_end = 4000.0 - (194.0 + R) / (2 * (R + 4000.0))
    assert y_rel[-1] < t_end - 1306.0 <= y_rel[-1] + 125.0

    # the same list by the rules of comp:408-482, and the sample counts
    ev, nd, nh, accum = expected(sun, 193.5, t_end - 1306.0)
    assert [e for _, e in ev] == events
    np.testing.assert_allclose([y for y, _ in ev], y_rel, atol=0.02)
    # margins: no trip or cone sample within 2 m of a decision height (float32 heights are 0.5 m
    # apart at Earth radius), so the hand derivation holds whatever the rounding
    edges = [749.4, 243.1, 250.0]
    for y, _ in ev:
        assert min(abs(y - e) for e in edges) > 2.0
        if 262 <= y <= 712:
            for hc in cone_heights(sun, y, 37.5):
                assert abs(hc * THICK - 749.4) > 1.0 and hc * THICK > 1.0
    assert (int(cnt[0]), int(cnt[1])) == (nd, nh)
    assert len(events) + 6 * 14 == nd
    # opacity: accum = sum of the hits' detail densities; the ray looks straight up, so the horizon
    # fade is 1 and alpha = glow x (1 - accum) (comp:484-496). The float32 heights are 0.5 m apart,
    # which moves each hit's density by at most 0.42 x 0.25 / 2500 = 4.2e-5 (14 hits: 6e-4), while
    # a hit too many or too few moves accum by at least 0.02
    d = float(np.dot(np.array(sun.directionBasis[4:7], np.float64), [0, 1, 0]))
    glow = (0.18 * d) ** 3
    assert 0.3 < accum < 0.35
    assert abs((1 - out[3] / glow) - accum) < 1e-3


def test_step_sequence_reference_preset_counts_only_counted_steps(mm, orc):
    """With MAX_STEPS = 100 counted iterations (comp:286, 481) instead of the per-mode bounds: the
    first-hit trip and the thin fine trips `continue` past the counter, so the same ray ends after
    100 counted trips, not 100 trips."""
    cam, sun, sky = vertical_camera(mm, 194.0)
    prm = mm.march_params("scenario")
    prm.flags &= ~mm.MARCH_BOUNDED
    prm.max_steps = 8
    prm.max_coarse_steps = prm.max_fine_steps = -1
    trips, cnt, _ = trace(orc, cam, sun, sky, uniform_textures(orc), prm)
    events = trips[:, 3].astype(int).tolist()
    # counted trips: hits, misses; uncounted: the first hit and the four thin trips. The loop breaks
    # when ++steps exceeds 8, i.e. on the ninth counted trip
    counted = [e for e in events if e not in (1, 2)]
    assert events[:5] == [1, 2, 2, 2, 2] and len(counted) == 9
    assert len(events) == 14


@pytest.mark.parametrize("altitude", [119.0, 156.5, 230.0])
def test_step_sequence_other_entry_heights(mm, orc, altitude):
    """Other entry heights (the march starts altitude - 0.5 m into the shell): the trace equals the
    rule list, and the counts follow."""
    cam, sun, sky = vertical_camera(mm, altitude)
    prm = mm.march_params("scenario")
    trips, cnt, _ = trace(orc, cam, sun, sky, uniform_textures(orc), prm)
    shift = 1500.0 - altitude
    y_rel = trips[:, 0].astype(np.float64) - shift
    r_in = (altitude + R) / (2 * (R + 1500.0))
    r_out = (altitude + R) / (2 * (R + 4000.0))
    y0 = (1500.0 - r_in) - shift
    ev, nd, nh, _ = expected(sun, y0, (4000.0 - r_out) - shift)
    assert [e for _, e in ev] == trips[:, 3].astype(int).tolist()
    np.testing.assert_allclose([y for y, _ in ev], y_rel, atol=0.02)
    for y, _ in ev:
        assert min(abs(y - e) for e in (749.4, 243.1, 250.0)) > 2.0
    assert (int(cnt[0]), int(cnt[1])) == (nd, nh)
