"""The C ABI from a plain C host (tests/c/mm_c_host.c, built by __graft_entry__.build): no Python or
PyTorch in that process. On the CPU it checks the calls that need no device; on the GPU it generates
the noise and marches a frame, and the image must equal, bit for bit, the same call made through the
Python binding."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "c", "mm_c_host")


@pytest.fixture(scope="module")
def c_host():
    if not os.path.exists(BIN):
        import __graft_entry__

        __graft_entry__.build_c_host()
    return BIN


def test_c_host_checks_without_a_device(c_host):
    r = subprocess.run([c_host, "check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_c_host_frame_equals_the_python_binding(c_host, mm, gpu, tmp_path):
    import torch

    W, H = 320, 180
    path = str(tmp_path / "frame.f32")
    r = subprocess.run([c_host, "run", str(W), str(H), path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(path, np.float32).reshape(H, W, 4)
    ns = mm.NoiseSet("fast")
    ns.generate(mm.noise_params(seed=1, weather_seed=1, coverage=0.5), mm.current_stream())
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(altitude=200.0, pitch_deg=20.0, fov_deg=60.0, aspect=W / H,
                                                    sun_elevation_deg=25.0, time=0.0))
    out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    mm.march_clouds(cam, sun, sky, ns, out, W, H, params=mm.march_params("scenario"))
    torch.cuda.synchronize()
    assert np.array_equal(got, out.cpu().numpy().reshape(H, W, 4))
