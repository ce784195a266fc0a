"""Training pipeline of the learned light-cone surrogate (no counterpart in the reference).

Data: the plain march records (features, density along the light) for its light-cone jobs
(mm_march_clouds_collect) over seeded views -- weather seed, sun elevation uniform in [2, 70] deg
and coverage uniform in [0.2, 0.9] per view -- thinned per view to a fixed row budget.
Training: the library's MFMA kernels (mm_surrogate_fit_stats, mm_surrogate_train: Adam).
Evaluation: the surrogate-lit march against the plain march of the same view, error per channel in
1/255 after the display tonemap (RGB) and raw (alpha), plus the fraction of cone jobs the network
replaced (from the march's work counters).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class TrainConfig:
    width: int = 1920
    height: int = 1080
    train_seeds: tuple = tuple(range(1, 65))       # weather-map seeds, one view each
    noise_seed: int = 1                             # the shared 128^3 / 32^3 volumes
    sun_deg: tuple = (2.0, 70.0)
    coverage: tuple = (0.2, 0.9)
    rows_per_view: int = 312_500                    # 64 views x 312500 = 2e7 rows
    steps: int = 20_000
    batch: int = 8192
    lr: float = 1e-3
    lr_final: float = 1e-4                          # geometric decay over the steps
    seed: int = 1
    chunk: int = 500                                # Adam steps per library call
    view_seed: int = 1234                           # draws the per-view sun / coverage
    extra: dict = field(default_factory=dict)


def view_params(cfg: TrainConfig, i: int):
    rng = np.random.default_rng(cfg.view_seed + 7919 * i)
    return float(rng.uniform(*cfg.sun_deg)), float(rng.uniform(*cfg.coverage))


def train_surrogate(mm, cfg: TrainConfig, rows=None, model=None, stream=None):
    """Collect (unless rows are given), fit feature statistics, run Adam. Returns (model, info)."""
    import torch

    t0 = time.time()
    jobs = None
    if rows is None:
        rows, jobs = collect_dataset(mm, cfg, stream)
    t1 = time.time()
    model = mm.Surrogate(seed=cfg.seed) if model is None else model
    model.fit_stats(rows)
    losses = []
    done = 0
    while done < cfg.steps:
        k = min(cfg.chunk, cfg.steps - done)
        frac = done / max(1, cfg.steps)
        lr = cfg.lr * (cfg.lr_final / cfg.lr) ** frac
        lo = torch.zeros(k, dtype=torch.float32, device="cuda")
        model.train(rows, steps=k, batch=cfg.batch, lr=lr, seed=cfg.seed + done, loss_out=lo, stream=stream)
        losses.append(lo.cpu().numpy() / cfg.batch)
        done += k
    torch.cuda.synchronize()
    l = np.concatenate(losses) if losses else np.zeros(0)
    info = {"rows": int(rows.shape[0]), "cone_jobs_seen": jobs, "collect_s": round(t1 - t0, 2),
            "train_s": round(time.time() - t1, 2), "steps": cfg.steps, "batch": cfg.batch,
            "loss_first": float(l[:10].mean()) if l.size else None,
            "loss_last": float(l[-50:].mean()) if l.size else None}
    return model, info


def collect_dataset(mm, cfg: TrainConfig, stream=None):
    """Rows [N, 17] float32 on the device, at most rows_per_view per view. Each view first counts
    its cone jobs, then records about rows_per_view of them (hash thinning) into its own window;
    the windows' filled prefixes are gathered at the end. Returns (rows, cone jobs seen)."""
    import torch

    W, H = cfg.width, cfg.height
    stream = mm.current_stream() if stream is None else stream
    prm = mm.march_params("scenario")
    out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    per = cfg.rows_per_view
    rows = torch.empty((len(cfg.train_seeds) * per, mm.SURROGATE_ROW), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    ns = mm.NoiseSet("fast")
    filled = []
    jobs = 0
    for i, ws in enumerate(cfg.train_seeds):
        sun, cov = view_params(cfg, i)
        ns.generate(mm.noise_params(seed=cfg.noise_seed, weather_seed=ws, coverage=cov), stream)
        cam, sunu, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H, sun_elevation_deg=sun))
        win = rows[i * per: (i + 1) * per]
        cnt.zero_()
        mm.march_collect(cam, sunu, sky, ns, out, W, H, win[:1], cnt, keep_one_in=1, seed=ws, params=prm,
                         stream=stream)
        total = int(cnt.item())
        jobs += total
        if total == 0:
            filled.append(0)
            continue
        keep = max(1, math.ceil(total / (0.9 * per)))
        cnt.zero_()
        mm.march_collect(cam, sunu, sky, ns, out, W, H, win, cnt, keep_one_in=keep, seed=ws, params=prm,
                         stream=stream)
        filled.append(min(int(cnt.item()), per))
    torch.cuda.synchronize()
    parts = [rows[i * per: i * per + f] for i, f in enumerate(filled) if f > 0]
    return (torch.cat(parts) if parts else rows[:0]), jobs


def tonemap(x):
    """Uncharted-2 display tonemap (tonemap.frag:11-28: exposure 0.7, white 50.2, gamma 1/2.2) of a
    torch tensor; negative inputs clamp to 0."""
    import torch

    x = torch.clamp(x.double(), min=0.0)

    def uc2(v):
        return ((v * (0.15 * v + 0.1 * 0.5) + 0.2 * 0.02) / (v * (0.15 * v + 0.5) + 0.2 * 0.3)) - 0.02 / 0.3

    c = uc2(0.7 * x) / uc2(torch.tensor(50.2, dtype=torch.float64))
    return torch.clamp(c, min=0.0) ** (1.0 / 2.2)


def image_error_255(a, b):
    """Per-channel |a - b| in 1/255 after the tonemap (RGB) and raw (alpha): (mean, max, frac<=1)."""
    import torch

    A = a.view(-1, 4)
    B = b.view(-1, 4)
    e = torch.cat([(tonemap(A[:, :3]) - tonemap(B[:, :3])).abs(), (A[:, 3:].double() - B[:, 3:].double()).abs()],
                  1) * 255.0
    return float(e.mean()), float(e.max()), float((e.max(1).values <= 1.0).double().mean())


def evaluate(mm, model, width=1920, height=1080, weather_seed=1, sun_deg=25.0, coverage=0.5, threshold=0.02,
             noise_seed=1, stream=None):
    """Surrogate-lit vs plain march of one view."""
    import torch

    stream = mm.current_stream() if stream is None else stream
    W, H = width, height
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=noise_seed, weather_seed=weather_seed, coverage=coverage), stream)
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H, sun_elevation_deg=sun_deg))
    plain = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    surr = torch.empty_like(plain)
    prm = mm.march_params("scenario")
    ps = mm.march_params("scenario", cone_mode=mm.CONE_SURROGATE, surrogate_threshold=threshold, surrogate=model.ptr)
    c0 = torch.zeros(4, dtype=torch.int64, device="cuda")
    c1 = torch.zeros(4, dtype=torch.int64, device="cuda")
    mm.march_clouds(cam, sun, sky, ns, plain, W, H, params=prm, counts=c0, stream=stream)
    mm.march_clouds(cam, sun, sky, ns, surr, W, H, params=ps, counts=c1, stream=stream)
    torch.cuda.synchronize()
    mean, mx, frac = image_error_255(surr, plain)
    c0 = [int(v) for v in c0.cpu()]
    c1 = [int(v) for v in c1.cpu()]
    jobs = max(c1[2], 1)
    return {"mean_err_255": mean, "max_err_255": mx, "frac_within_1": frac,
            "replaced_frac": 1.0 - c1[3] / jobs, "cone_jobs": c1[2],
            "density_samples": (c0[0], c1[0]), "detail_samples": (c0[1], c1[1])}
