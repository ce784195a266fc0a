"""Many views in one launch (no counterpart in the reference): seeded view sets, HIP-graph
capture of the launch, and the autotuner that picks the tile shape and the coarse step length.

View set (seeded per view): resolution log-uniform from 256x256 to 1920x1080, camera altitude
uniform in [0, 3000] m, sun elevation uniform in [2, 70] deg, coverage uniform in [0.2, 0.9]; each
view has its own 512^2 weather map, all views share the 128^3 / 32^3 volumes and the curl map.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np


@dataclass
class ViewSpec:
    width: int
    height: int
    altitude: float
    sun_deg: float
    coverage: float
    weather_seed: int


def view_specs(n: int = 64, seed: int = 1, lo=(256, 256), hi=(1920, 1080)) -> list[ViewSpec]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = rng.uniform()
        w = int(round(lo[0] * (hi[0] / lo[0]) ** t))
        h = int(round(lo[1] * (hi[1] / lo[1]) ** t))
        out.append(ViewSpec(w, h, float(rng.uniform(0.0, 3000.0)), float(rng.uniform(2.0, 70.0)),
                            float(rng.uniform(0.2, 0.9)), 1000 + seed * 131 + i))
    return out


class _Bound:
    """A texture set given by its C descriptor only (what march_clouds reads)."""

    def __init__(self, c):
        self.c = c


class ViewSet:
    """Device resources of a view set: shared volumes, per-view weather maps and output images."""

    def __init__(self, mm, specs, fset: str = "fast", noise_seed: int = 1, stream=None):
        import torch

        self.mm = mm
        self.specs = specs
        stream = mm.current_stream() if stream is None else stream
        self.shared = mm.NoiseSet(fset)
        self.shared.generate(mm.noise_params(seed=noise_seed, what=mm.GEN_CURL | mm.GEN_LOW_RES | mm.GEN_HI_RES),
                             stream)
        fmts, layout = mm.FORMAT_SETS[fset]
        self.weathers, self.outs, self.ubos = [], [], []
        for s in specs:
            wt = mm.device_texture(512, 512, 1, fmts[0], layout)
            tmp = mm.NoiseSetC(wt.desc, self.shared.curl.desc, self.shared.low_res.desc, self.shared.hi_res.desc)
            p = mm.noise_params(weather_seed=s.weather_seed, coverage=s.coverage, what=mm.GEN_WEATHER)
            mm.check(mm.lib().mm_generate_noise(ctypes.byref(p), ctypes.byref(tmp), None, stream), "mm_generate_noise")
            self.weathers.append(wt)
            self.outs.append(torch.zeros(s.width * s.height * 4, dtype=torch.float32, device="cuda"))
            self.ubos.append(mm.build_ubos(mm.scenario_scene(altitude=s.altitude, aspect=s.width / s.height,
                                                             sun_elevation_deg=s.sun_deg)))
        torch.cuda.synchronize()

    def views(self, phase: int = -1):
        mm = self.mm
        return [mm.make_view(*self.ubos[i], self.weathers[i], self.outs[i], s.width, s.height, phase=phase)
                for i, s in enumerate(self.specs)]

    def pixels(self) -> int:
        return sum(s.width * s.height for s in self.specs)

    def march_one_by_one(self, params, outs=None):
        """Reference result: each view by its own mm_march_clouds."""
        mm = self.mm
        outs = self.outs if outs is None else outs
        for i, s in enumerate(self.specs):
            cam, sun, sky = self.ubos[i]
            ns = _Bound(mm.NoiseSetC(self.weathers[i].desc, self.shared.curl.desc, self.shared.low_res.desc,
                                     self.shared.hi_res.desc))
            mm.march_clouds(cam, sun, sky, ns, outs[i], s.width, s.height, params=params)


def capture(plan, counts=None):
    """HIP graph of one mm_march_views launch."""
    import torch

    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.launch(counts)  # warm-up outside the graph on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.launch(counts)
    return g


def time_launch(fn, reps: int = 10, warmup: int = 2) -> float:
    import torch

    for _ in range(warmup):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(mm, vs: ViewSet, shapes=((1, 8), (1, 16), (1, 32), (2, 8), (4, 8)), steps=(0.05, 0.0625, 0.075),
             max_mean_err_255: float = 1.0, reps: int = 5, preset: str = "scenario", base_params=None):
    """Time every (kernel shape, coarse step) candidate on the view set. A shape is (lanes per ray,
    tile width): one lane per ray in 64-ray tiles 8, 16 or 32 wide, or the ray-parallel kernel's
    8 x (8 / lanes) tiles; it does not change the image. The coarse step does, so a candidate is
    admissible only when its mean per-channel error against the reference step (steps[0]) stays
    within max_mean_err_255 (1/255 units, after the display tonemap). Returns (best candidate, table)."""
    import torch

    from . import training

    ref_outs = None
    table = []
    for st in steps:
        outs_err = None
        for lanes, tw in shapes:
            prm = mm.march_params(preset, coarse_step_frac=st, tile_width=tw, lanes_per_ray=lanes) \
                if base_params is None else base_params(st, tw, lanes)
            plan = mm.ViewPlan(vs.views(), vs.shared, prm)
            ms = time_launch(lambda: plan.launch(), reps=reps)
            torch.cuda.synchronize()
            if outs_err is None:
                snap = [o.clone() for o in vs.outs]
                if ref_outs is None:
                    ref_outs = snap
                    err = 0.0
                else:
                    errs = [training.image_error_255(a, b)[0] for a, b in zip(snap, ref_outs)]
                    err = float(np.mean(errs))
                outs_err = err
            plan.close()
            table.append({"lanes_per_ray": lanes, "tile_width": tw, "coarse_step_frac": st, "ms": ms,
                          "mean_err_255": outs_err})
    ok = [r for r in table if r["mean_err_255"] <= max_mean_err_255]
    best = min(ok, key=lambda r: r["ms"])
    return best, table
