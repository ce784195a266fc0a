"""bench.py's command-line contract that needs no GPU: a multi-GPU request exits with a message
(every box here has one GPU and this build has no multi-GPU path), and the helpers that name the
measured kernel and read the committed HBM-traffic summary behave as documented."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_kernel_name_follows_the_options():
    b = _bench()
    args = lambda **kw: type("A", (), dict({"no_batch": False, "lanes_per_ray": 0, "tile_width": 0,
                                            "no_schedule": False}, **kw))
    # scheduled (the default): one lane per ray in 16 x 4 tiles; unscheduled: two lanes per ray
    assert b.march_kernel_name(args()) == "march_batched_kernel"
    assert b.lanes_and_tiles(args()) == (1, 16)
    assert b.march_kernel_name(args(no_schedule=True)) == "march_rp_kernel"
    assert b.lanes_and_tiles(args(no_schedule=True)) == (2, 8)
    assert b.march_kernel_name(args(lanes_per_ray=2)) == "march_rp_kernel"
    assert b.march_kernel_name(args(lanes_per_ray=1, no_schedule=True)) == "march_batched_kernel"
    assert b.lanes_and_tiles(args(lanes_per_ray=1, no_schedule=True)) == (1, 8)
    assert b.march_kernel_name(args(no_batch=True)) == "march_kernel"


def test_traffic_comes_from_the_committed_summary():
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "march_fast_1080p_pmc.json")) as f:
        d = json.load(f)
    assert b.traffic_from_profiles("fast", 1920, 1080) == d["hbm_bytes_per_launch"] > 0
    assert b.traffic_from_profiles("fast", 640, 480) is None
    # a summary is used only for the kernel it profiled
    assert b.traffic_from_profiles("fast", 1920, 1080, d["kernel"].split("<")[0]) == d["hbm_bytes_per_launch"]
    assert b.traffic_from_profiles("fast", 1920, 1080, "march_kernel") is None
    assert b.valu_from_profiles("fast", 1920, 1080, "march_kernel") is None
    assert b.valu_from_profiles("fast", 1920, 1080, d["kernel"].split("<")[0])["insts_per_launch"] > 0


def test_noise_block_counts_the_written_textures():
    """The generators' HBM position: the four textures' bytes over the generation time."""
    b = _bench()
    tex = lambda n: type("T", (), {"desc": type("D", (), {"bytes": n})})
    ns = type("N", (), {"weather": tex(1 << 20), "curl": tex(2 << 20), "low_res": tex(4 << 20), "hi_res": tex(1 << 20)})
    blk = b.noise_block(ns, 0.25)
    assert blk["bytes_written"] == 8 << 20
    assert blk["write_gbs"] == round((8 << 20) / 0.25e-3 / 1e9, 1)  # 33.6 GB/s
    assert blk["frac_of_hbm_peak"] == round((8 << 20) / 0.25e-3 / 1e9 / 8000.0, 4)
