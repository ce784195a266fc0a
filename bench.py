"""Benchmark of the volumetric-cloud path on one MI355X.

One step = one frame of the 1920x1080 scenario workload: every noise texture regenerated on the GPU
(weather 512^2, curl 128^2, low-res 128^3, hi-res 32^3), then the full-frame cloud march (all
pixels; its tiles longest first by the previous step's tile costs unless --no-schedule). Inputs live
in HBM before timing starts. Prints one JSON line with ms/frame, density
samples/s (the value), the march kernel's roofline position and a CPU baseline (the FP32 oracle
on this host's cores, on a bounded sample of the same frame).
"""
import argparse
import glob
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "project-marshmallow_amd")
METRIC = "ms/frame @1920x1080 + density samples/s; surrogate vs plain at ≤1/255 err"
HBM_PEAK_GBS = 8000.0
# Bytes per texel of the weather map, low-res volume, curl map and hi-res volume in each format set.
TEXEL_BYTES = {"fast": (2, 4, 4, 4), "ref": (4, 4, 4, 4), "prec": (2, 4, 8, 8)}
# Algorithmic texel-footprint bytes (4 weather texels + 8 low-res texels per density sample,
# 4 curl texels + 8 hi-res texels per detail sample), per format set, plus 16 B of output per pixel.
BYTES_PER_SAMPLE = {k: {"density": 4 * w + 8 * l, "detail": 4 * c + 8 * h} for k, (w, l, c, h) in TEXEL_BYTES.items()}
# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies a SIMD for one quad-cycle (4 clocks)
SIMDS = 1024


def load_package():
    if "marshmallow_amd" in sys.modules:
        return sys.modules["marshmallow_amd"]
    spec = importlib.util.spec_from_file_location(
        "marshmallow_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["marshmallow_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _same_kernel(profiled, kernel):
    """True when no kernel is asked for, or when the profiled instantiation (`name<...>`) is of it."""
    return kernel is None or (profiled or "").split("<")[0] == kernel


def traffic_from_profiles(fset, width, height, kernel=None):
    """HBM bytes per march launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated), if a summary for this workload (and, when named, this march kernel) is committed
    under profiles/."""
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*pmc*.json"))):
        try:
            with open(path) as f:
                d = json.load(f)
        except (OSError, ValueError):
            continue
        if (d.get("format_set") == fset and d.get("width") == width and d.get("height") == height
                and _same_kernel(d.get("kernel"), kernel)):
            return d.get("hbm_bytes_per_launch")
    return None


def valu_from_profiles(fset, width, height, kernel=None):
    """VALU position of the march from a committed rocprofv3 --pmc summary of this workload, if present:
    VALU instructions per launch, the issue fraction SQ_INSTS_VALU / (SIMDs x GRBM_GUI_ACTIVE / 4)
    (above 1 where the SIMDs dual-issue), the share of VALU quad-cycles that dual-issued, and the lane
    utilisation SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)."""
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*valu*.json"))):
        try:
            with open(path) as f:
                d = json.load(f)
        except (OSError, ValueError):
            continue
        if (d.get("format_set") == fset and d.get("width") == width and d.get("height") == height
                and _same_kernel((d.get("valu") or {}).get("kernel"), kernel)):
            return d.get("valu")
    return None


def valu_live(p, kernel_ms):
    """The VALU-issue axis of the march at this run's kernel time: the profiled VALU instructions per
    launch over the time, against the issue rate of 1024 SIMDs at the profiled clock (one wave64 VALU
    instruction per SIMD every 4 cycles; dual issue can exceed it)."""
    if not p or "insts_per_launch" not in p or "clock_mhz" not in p:
        return p
    achieved = p["insts_per_launch"] / (kernel_ms * 1e-3) / 1e9
    peak = SIMDS * p["clock_mhz"] * 1e6 / 4 / 1e9
    return dict(p, achieved=round(achieved, 1), peak=round(peak, 1), unit="G wave-instructions/s",
                frac=round(achieved / peak, 4))


def cpu_baseline(mm, ns, cam, sun, sky, prm, W, H, budget_s):
    """FP32 oracle (all host threads) on whole 1/16 pixel phases of the same frame until the budget
    is spent; returns density samples/s."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    orc = oracle.Oracle()
    ref = mm.NoiseSet("ref")
    ref.generate(mm.noise_params(seed=1, weather_seed=1, coverage=0.5), mm.current_stream())
    import torch

    torch.cuda.synchronize()
    t = {k: getattr(ref, k).tensor.cpu().numpy() for k in ("weather", "curl", "low_res", "hi_res")}
    tex = orc.textures(t["weather"].reshape(512, 512, 4), t["curl"].reshape(128, 128, 4),
                       t["low_res"].reshape(128, 128, 128, 4), t["hi_res"].reshape(32, 32, 32, 4))
    order = [0, 5, 10, 15, 2, 7, 8, 13, 1, 4, 11, 14, 3, 6, 9, 12]
    samples, pixels, elapsed, done = 0, 0, 0.0, 0
    for p in order:
        t0 = time.perf_counter()
        _, cnt = orc.march(cam, sun, sky, tex, prm, W, H, phase=p, counts=True)
        elapsed += time.perf_counter() - t0
        samples += int(cnt[..., 0].sum())
        pixels += len(range(p % 4, W, 4)) * len(range(p // 4, H, 4))
        done += 1
        if elapsed >= budget_s:
            break
    threads = int(orc.L.orc_max_threads())
    return {"value": samples / elapsed, "unit": "density samples/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle march of {done}/16 pixel phases of the same 1920x1080 frame "
                      f"({pixels} px, {samples} density samples, {elapsed:.1f} s)",
            "ms_per_frame_projected": None if not samples else None}


SURROGATE_FLOP_PER_JOB = 2 * (16 * 64 + 64 * 64 + 64 * 64 + 64 * 2)  # one forward pass (bias adds excluded)
BF16_PEAK_TFLOPS = 2500.0


def surrogate_leg(mm, ns, cam, sun, sky, prm, W, H, steps, warmup, stream, train_steps, scheduled=True):
    """Learned light cone at this frame. Training to the config's spec outside the timed region: 2e7
    recorded cone jobs over 64 seeded 1920x1080 views (weather seeds 1-64, sun U[2, 70] deg, coverage
    U[0.2, 0.9]), Adam, batch 8192. The threshold is the largest of a fixed ladder (0 = plain march,
    1e30 = every cone replaced) whose image stays within 1/255 mean per-channel error of the plain
    march; the surrogate-lit march is then timed like the plain one. The light cone's share of the
    frame's samples comes from the work counters (plain march against every cone replaced, whose
    network still takes three density probes per job) and bounds the speed-up a free network could
    reach: 1 / (1 - share)."""
    import torch
    from marshmallow_amd import training

    cfg = training.TrainConfig(steps=train_steps)
    model, info = training.train_surrogate(mm, cfg)
    plain = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    surr = torch.empty_like(plain)
    cp = torch.zeros(4, dtype=torch.int64, device="cuda")
    mm.march_clouds(cam, sun, sky, ns, plain, W, H, params=prm, stream=stream, counts=cp)
    cp = [int(v) for v in cp.cpu()]
    best, every = None, None
    for thr in (0.0, 0.0001, 0.0002, 0.0005, 0.001, 0.002, 0.005, 0.01, 0.02, 1e30):
        ps = mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=thr, surrogate=model.ptr)
        c = torch.zeros(4, dtype=torch.int64, device="cuda")
        mm.march_clouds(cam, sun, sky, ns, surr, W, H, params=ps, counts=c, stream=stream)
        torch.cuda.synchronize()
        mean, mx, frac = training.image_error_255(surr, plain)
        c = [int(v) for v in c.cpu()]
        row = {"threshold": thr, "mean_err_255": round(mean, 4), "max_err_255": round(mx, 3),
               "frac_within_1": round(frac, 5), "replaced_frac": round(1.0 - c[3] / max(c[2], 1), 4),
               "jobs_on_network": c[2] - c[3], "params": ps}
        if mean <= 1.0:
            best = row
        if thr == 1e30:
            every = dict(row, counts=c)

    def timed(p):
        # plain and surrogate-lit marches alike: with a tile schedule of their own unless the bench runs
        # unscheduled (the warm-up launches record the first tile costs)
        sch = None if scheduled is False else mm.MarchSchedule()
        for _ in range(max(warmup, 1)):
            mm.march_clouds(cam, sun, sky, ns, surr, W, H, params=p, stream=stream, schedule=sch)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            mm.march_clouds(cam, sun, sky, ns, surr, W, H, params=p, stream=stream, schedule=sch)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    t_plain = timed(prm)
    t_surr = timed(best["params"])
    t_every = timed(every["params"])
    ce = every["counts"]
    cone_density = cp[0] - (ce[0] - 3 * ce[2])  # the network's probes are three density samples per job
    cone_detail = cp[1] - ce[1]
    share = (cone_density + cone_detail) / max(cp[0] + cp[1], 1)
    flop = best["jobs_on_network"] * SURROGATE_FLOP_PER_JOB
    for r in (best, every):
        r.pop("params")
    every.pop("counts")
    surrogate_leg.model = model  # kept for the views leg
    return {"threshold": best["threshold"], "mean_err_255": best["mean_err_255"], "max_err_255": best["max_err_255"],
            "frac_within_1": best["frac_within_1"], "replaced_frac": best["replaced_frac"],
            "within_1_255": best["mean_err_255"] <= 1.0, "march_ms_plain": round(t_plain, 4),
            "march_ms_surrogate": round(t_surr, 4), "speedup": round(t_plain / t_surr, 3),
            "every_cone_replaced": dict(every, march_ms=round(t_every, 4), speedup=round(t_plain / t_every, 3)),
            "cone_sample_share": round(share, 4), "speedup_ceiling_free_network": round(1.0 / (1.0 - share), 3),
            "gap_to_2x": round(2.0 / (t_plain / t_surr), 3),
            "mfma": {"flop_per_launch": flop, "achieved_tflops": round(flop / (t_surr * 1e-3) / 1e12, 2),
                     "frac_of_bf16_peak": round(flop / (t_surr * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 5)},
            "training": info}


def views_leg(mm, model, reps):
    """The many-view workload: 64 seeded views of log-uniform size from 256x256 to 1920x1080 (camera
    altitude, sun and coverage per view, one weather map each, shared volumes), marched one dispatch
    per view like the reference (VulkanApplication.cpp:1061-1071) and as one launch over every view's
    tiles replayed from a HIP graph, with the plain cone and (given a trained model) with every cone
    on the surrogate; density samples/s from an instrumented launch; then one autotuner pass over
    tile shape and coarse step (plain cones)."""
    import torch
    from marshmallow_amd import multiview

    specs = multiview.view_specs(n=64, seed=1)
    vs = multiview.ViewSet(mm, specs)
    res = {"views": len(specs), "pixels": vs.pixels()}
    modes = [("plain", mm.march_params("scenario"))]
    if model is not None:
        modes.append(("surrogate", mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=1e30,
                                                   surrogate=model.ptr)))
    for name, prm in modes:
        one = multiview.time_launch(lambda: vs.march_one_by_one(prm), reps=reps)
        plan = mm.ViewPlan(vs.views(), vs.shared, prm)
        g = multiview.capture(plan)
        ms = multiview.time_launch(lambda: g.replay(), reps=reps)
        cnt = torch.zeros(4, dtype=torch.int64, device="cuda")
        plan.launch(cnt)
        torch.cuda.synchronize()
        plan.close()
        res[name] = {"ms_graph_launch": round(ms, 3), "ms_one_dispatch_per_view": round(one, 3),
                     "density_samples": int(cnt[0]), "density_samples_per_s": round(int(cnt[0]) / (ms * 1e-3), 1),
                     "cone_jobs": int(cnt[2]), "jobs_on_plain_cone": int(cnt[3])}
    if "surrogate" in res:
        res["surrogate"]["speedup_vs_plain_graph"] = round(res["plain"]["ms_graph_launch"] /
                                                           res["surrogate"]["ms_graph_launch"], 3)
    best, table = multiview.autotune(mm, vs, reps=3)
    res["autotune_best"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in best.items()}
    res["autotune_candidates"] = len(table)
    return res


def post_leg(mm, cam, sun, out, W, H, steps, warmup):
    """Post chain (god rays -> radial blur -> tonemap, mm_post_process) on the marched frame, timed on
    its own. Algorithmic HBM bytes per pixel: the RGBA32F image read once (16) and the RGBA8 frame
    written (4); the passes move 52 (alpha plane 16 + 4, god rays 4 + 4, radial blur + tonemap
    16 + 4 + 4)."""
    import torch

    alpha = torch.empty(2 * W * H, dtype=torch.float32, device="cuda")
    rgba8 = torch.empty(W * H * 4, dtype=torch.uint8, device="cuda")
    for _ in range(warmup):
        mm.post_process(cam, sun, out, W, H, alpha, None, rgba8)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        mm.post_process(cam, sun, out, W, H, alpha, None, rgba8)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    alg, moved = W * H * (16 + 4), W * H * 52
    return {"post_ms": round(ms, 4), "algorithmic_bytes": alg, "achieved_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "pass_bytes": moved,
            "pass_gbs": round(moved / (ms * 1e-3) / 1e9, 1)}


def noise_block(ns, noise_ms):
    """The four generators' position against HBM: the bytes of the four textures they write per
    step over the step's generation time (their bound is VALU issue and LDS reads; their writes are
    a few MiB)."""
    written = sum(int(t.desc.bytes) for t in (ns.weather, ns.curl, ns.low_res, ns.hi_res))
    gbs = written / (noise_ms * 1e-3) / 1e9
    return {"ms": round(noise_ms, 4), "bytes_written": written, "write_gbs": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}


def march_kernel_name(args):
    """The kernel a full-frame plain march launches: lock-step, cone-batched (1 lane per ray: the
    default of a scheduled march) or ray-parallel (2 lanes per ray: the unscheduled default)."""
    if args.no_batch:
        return "march_kernel"
    lanes = args.lanes_per_ray or (2 if args.no_schedule else 1)
    return "march_batched_kernel" if lanes == 1 else "march_rp_kernel"


def lanes_and_tiles(args):
    """(lanes per ray, tile width) the launch runs with (the library's defaults when 0)."""
    if args.no_batch:
        return 1, args.tile_width or 8
    lanes = args.lanes_per_ray or (2 if args.no_schedule else 1)
    if lanes > 1:
        return lanes, 8
    return 1, args.tile_width or (8 if args.no_schedule else 16)


def amortized_leg(mm, ns, cam, sun, sky, prm, W, H, warmup):
    """The reference's per-frame compute cost model (VulkanApplication.cpp:1046-1071): the previous
    frame reprojected at full resolution, then one 1/16 pixel phase marched (mm_render_frame), over
    16 consecutive frames (every phase once), textures resident. Reports ms per frame."""
    import torch

    frames = mm.CloudFrames(W, H, ns, params=prm)
    sun = type(sun).from_buffer_copy(sun)  # the frames advance the phase in sun.color.a
    for _ in range(max(warmup, 16)):
        frames.frame(cam, sun, sky)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(32):
        frames.frame(cam, sun, sky)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 32
    # the single-phase march alone (one reference dispatch: every 16th pixel)
    img = frames.images[0]
    e0.record()
    for p in range(32):
        mm.march_clouds(cam, sun, sky, ns, img, W, H, phase=p % 16, params=prm)
    e1.record()
    torch.cuda.synchronize()
    return {"ms_per_frame": round(ms, 4), "phase_march_ms": round(e0.elapsed_time(e1) / 32, 4), "frames": 32,
            "what": "reprojection of the previous frame + 1/16 of the pixels marched (phase cycles 0..15)"}


SWEEP = {"coverage": (0.3, 0.5, 0.8), "sun_deg": (5.0, 25.0, 60.0), "time_s": (0.0, 150.0, 300.0, 450.0, 600.0),
         "weather_seed": (1, 2, 3, 4, 5)}


def sweep_leg(mm, args, prm, W, H):
    """The 1920x1080 workload over its scene sweep (coverage x sun elevation x wind time x weather
    seed, 225 scenes): per scene one step = every noise texture regenerated + the full-frame march,
    timed like the headline step (HIP events, args.steps steps after args.warmup), the sample counts
    from one instrumented launch. Both roofline axes per scene: the HBM one from the scene's own
    sample counts, the VALU one estimated from the profiled headline scene's instructions per sample
    (density + detail) times the scene's samples. Prints one JSON line per scene and returns the summary."""
    import itertools

    import torch

    ns = mm.NoiseSet(args.fset)
    out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    stream = mm.current_stream()
    b = BYTES_PER_SAMPLE[args.fset]
    vp = valu_from_profiles(args.fset, W, H, march_kernel_name(args))
    rows = []
    for cov, sun_deg, t, seed in itertools.product(*SWEEP.values()):
        cam, sun, sky = mm.build_ubos(mm.scenario_scene(altitude=200.0, pitch_deg=20.0, fov_deg=60.0, aspect=W / H,
                                                        sun_elevation_deg=sun_deg, time=t))
        nparams = mm.noise_params(seed=1, weather_seed=seed, coverage=cov)
        ns.generate(nparams, stream)
        counts = torch.zeros(4, dtype=torch.int64, device="cuda")
        mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm, stream=stream, counts=counts)
        # one tile schedule per scene, like a renderer's per-view state (the warm-up records its costs)
        sched = None if args.no_schedule else mm.MarchSchedule()
        for _ in range(max(args.warmup, 1 if sched else 0)):
            ns.generate(nparams, stream)
            mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm, stream=stream, schedule=sched)
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
        torch.cuda.synchronize()
        for e in ev:
            e[0].record()
            ns.generate(nparams, stream)
            e[1].record()
            mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm, stream=stream, schedule=sched)
            e[2].record()
        torch.cuda.synchronize()
        frame = sum(e[0].elapsed_time(e[2]) for e in ev) / args.steps
        march = sum(e[1].elapsed_time(e[2]) for e in ev) / args.steps
        nd, nh = int(counts[0]), int(counts[1])
        alg = nd * b["density"] + nh * b["detail"] + W * H * 16
        row = {"coverage": cov, "sun_deg": sun_deg, "time_s": t, "weather_seed": seed, "ms_per_frame": round(frame, 4),
               "march_ms": round(march, 4), "density_samples": nd, "detail_samples": nh,
               "density_samples_per_s": round(nd / (frame * 1e-3), 1),
               "roofline_hbm": {"achieved_gbs": round(alg / (march * 1e-3) / 1e9, 1),
                                "frac": round(alg / (march * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
        if vp and "samples_per_launch" in vp:
            est = valu_live(dict(vp, insts_per_launch=vp["insts_per_launch"] * (nd + nh) / vp["samples_per_launch"]), march)
            row["roofline_valu_est"] = {"insts_per_launch": round(est["insts_per_launch"]), "frac": est["frac"]}
        rows.append(row)
        # per-scene progress on stderr: stdout carries the one result line
        print(json.dumps({"sweep_scene": row}), file=sys.stderr, flush=True)
    ms = sorted(r["ms_per_frame"] for r in rows)
    sps = sorted(r["density_samples_per_s"] for r in rows)
    fr = sorted(r["roofline_hbm"]["frac"] for r in rows)
    fv = sorted(r["roofline_valu_est"]["frac"] for r in rows if "roofline_valu_est" in r)
    pick = lambda v, q: v[min(len(v) - 1, int(q * (len(v) - 1) + 0.5))]
    return {"scenes": len(rows), "axes": {k: list(v) for k, v in SWEEP.items()},
            "ms_per_frame": {"min": ms[0], "median": pick(ms, 0.5), "max": ms[-1]},
            "density_samples_per_s": {"min": sps[0], "median": pick(sps, 0.5), "max": sps[-1]},
            "roofline_hbm_frac": {"min": fr[0], "median": pick(fr, 0.5), "max": fr[-1]},
            "roofline_valu_frac_est": {"min": fv[0], "median": pick(fv, 0.5), "max": fv[-1]} if fv else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--fset", default="fast", choices=["fast", "ref", "prec"])
    ap.add_argument("--no-batch", action="store_true", help="evaluate light cones in place (lock-step kernel)")
    ap.add_argument("--tile-width", type=int, default=0, help="8, 16 or 32 (0 = library default)")
    ap.add_argument("--lanes-per-ray", type=int, default=0, help="1, 2, 4 or 8 (0 = library default)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work (0 = skip)")
    ap.add_argument("--no-post", action="store_true", help="skip the post-chain leg")
    ap.add_argument("--no-amortized", action="store_true", help="skip the 1/16-phase frame leg")
    ap.add_argument("--surrogate-steps", type=int, default=20000,
                    help="Adam steps of the learned light-cone surrogate leg (0 = skip the leg)")
    ap.add_argument("--no-views", action="store_true", help="skip the 64-view graph-launch leg")
    ap.add_argument("--no-schedule", action="store_true",
                    help="march without the tile schedule (launch-order tiles, two lanes per ray by default)")
    ap.add_argument("--sweep", action="store_true",
                    help="also time the workload's 225-scene sweep (coverage x sun x wind time x weather seed)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device-to-device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 5}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the frame shards by row bands "
              "or views with no exchange; see the sharding note in the design doc)", file=sys.stderr)
        sys.exit(2)

    import torch

    mm = load_package()
    torch.cuda.set_device(0)
    W, H = args.width, args.height
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(altitude=200.0, pitch_deg=20.0, fov_deg=60.0, aspect=W / H,
                                                    sun_elevation_deg=25.0, time=0.0))
    prm = mm.march_params("scenario")
    if args.no_batch:
        prm.flags |= mm.MARCH_NO_BATCH
    prm.tile_width = args.tile_width
    prm.lanes_per_ray = args.lanes_per_ray
    ns = mm.NoiseSet(args.fset)
    nparams = mm.noise_params(seed=1, weather_seed=1, coverage=0.5)
    out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    stream = mm.current_stream()

    # instrumented run (outside the timed region): sample counts of this frame
    ns.generate(nparams, stream)
    counts = torch.zeros(6, dtype=torch.int64, device="cuda")
    mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm, stream=stream, counts=counts)
    torch.cuda.synchronize()
    n_density, n_detail, n_jobs, _, n_weather, n_low = (int(v) for v in counts.cpu())

    # the tile schedule: each launch orders its tiles longest first from the previous launch's tile
    # costs (the warm-up steps record the first ones); the image does not depend on it
    sched = None if args.no_schedule else mm.MarchSchedule()
    for _ in range(max(args.warmup, 1 if sched else 0)):
        ns.generate(nparams, stream)
        mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm, stream=stream, schedule=sched)
    torch.cuda.synchronize()

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for i in range(args.steps):
        ev[i][0].record()
        ns.generate(nparams, stream)
        ev[i][1].record()
        mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm, stream=stream, schedule=sched)
        ev[i][2].record()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = start.elapsed_time(end) / args.steps
    noise_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps
    march_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / args.steps
    step_ms_sorted = sorted(e[0].elapsed_time(e[2]) for e in ev)
    march_ms_sorted = sorted(e[1].elapsed_time(e[2]) for e in ev)
    median = lambda v: (v[(len(v) - 1) // 2] + v[len(v) // 2]) / 2
    assert torch.isfinite(out).all().item()
    unscheduled = None
    if sched is not None:
        # the same frame's march without the schedule (launch-order tiles, the unscheduled default
        # shape), for comparison; outside the timed steps
        prm_u = mm.march_params("scenario", tile_width=args.tile_width, lanes_per_ray=args.lanes_per_ray)
        prm_u.flags = prm.flags
        for _ in range(2):
            mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm_u, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            mm.march_clouds(cam, sun, sky, ns, out, W, H, params=prm_u, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        unscheduled = round(e0.elapsed_time(e1) / args.steps, 4)

    b = BYTES_PER_SAMPLE[args.fset]
    alg_bytes = n_density * b["density"] + n_detail * b["detail"] + W * H * 16
    achieved = alg_bytes / (march_ms * 1e-3) / 1e9
    # the same accounting over the samples that read texels: a density sample outside the layer
    # heights reads nothing, one on a zero layer profile reads only the weather map
    tw, tl, _, _ = TEXEL_BYTES[args.fset]
    gathered = n_weather * 4 * tw + n_low * 8 * tl + n_detail * b["detail"] + W * H * 16
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(args.fset, W, H, march_kernel_name(args)),
                "kernel": march_kernel_name(args), "kernel_ms": round(march_ms, 4),
                "algorithmic_bytes_per_launch": alg_bytes,
                "bytes_per_density_sample": b["density"], "bytes_per_detail_sample": b["detail"],
                "gathered": {"bytes_per_launch": gathered, "achieved": round(gathered / (march_ms * 1e-3) / 1e9, 1),
                             "frac": round(gathered / (march_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "weather_reads": n_weather, "low_res_reads": n_low},
                "valu": valu_live(valu_from_profiles(args.fset, W, H, march_kernel_name(args)), march_ms)}
    rec = {
        "metric": METRIC,
        "value": round(n_density / (ms_step * 1e-3), 1),
        "unit": "density samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded procedural noise volumes and weather map, regenerated on the GPU every step",
        "config": {"workload": f"{W}x{H} scenario frame: all noise textures regenerated, then full-frame FP32 march "
                               f"(Earth radius 6.371e6 m, shell 1500-4000 m, camera 200 m pitched +20 deg, 60 deg "
                               f"FOV, sun 25 deg, coverage 0.5, t = 0, weather seed 1)",
                   "width": W, "height": H, "format_set": args.fset,
                   "kernel": march_kernel_name(args), "coarse_steps_max": 128, "fine_steps_max": 256,
                   "tile_width": lanes_and_tiles(args)[1], "lanes_per_ray": lanes_and_tiles(args)[0],
                   "tile_schedule": None if sched is None else
                   "longest first from the previous step's tile costs (mm_march_clouds_scheduled)"},
        "ms_per_frame": round(ms_step, 4),
        "frames_per_s": round(1000.0 / ms_step, 2),
        "march_ms": round(march_ms, 4),
        "march_ms_unscheduled": unscheduled,
        "median_ms": {"frame": round(median(step_ms_sorted), 4), "march": round(median(march_ms_sorted), 4)},
        "noise_ms": round(noise_ms, 4),
        "noise": noise_block(ns, noise_ms),
        "density_samples_per_frame": n_density,
        "detail_samples_per_frame": n_detail,
        "cone_jobs_per_frame": n_jobs,
        "wall_s": round(wall, 3),
        "roofline": roofline,
        "cpu_baseline": None,
        "surrogate": None,
        "post": None,
        "amortized_frame": None,
        "views": None,
        "sweep": None,
    }
    if args.sweep:
        rec["sweep"] = sweep_leg(mm, args, prm, W, H)
    if not args.no_amortized:
        rec["amortized_frame"] = amortized_leg(mm, ns, cam, sun, sky, prm, W, H, args.warmup)
    if not args.no_post:
        rec["post"] = post_leg(mm, cam, sun, out, W, H, args.steps, args.warmup)
    if args.surrogate_steps > 0 and not args.no_batch:
        rec["surrogate"] = surrogate_leg(mm, ns, cam, sun, sky, prm, W, H, args.steps, args.warmup, stream,
                                         args.surrogate_steps, not args.no_schedule)
    if not args.no_views and args.fset == "fast" and not args.no_batch:
        rec["views"] = views_leg(mm, getattr(surrogate_leg, "model", None), 5)
    if args.cpu_budget > 0:
        cb = cpu_baseline(mm, ns, cam, sun, sky, prm, W, H, args.cpu_budget)
        cb["ms_per_frame_projected"] = round(n_density / cb["value"] * 1e3, 1)
        rec["cpu_baseline"] = cb
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
