"""Cloud shadows on surface points (model.frag's shadow march) on the GPU against the oracle."""
import numpy as np
import pytest

from test_gpu_march import tex, upload  # noqa: F401  (tex is a fixture)

pytestmark = pytest.mark.gpu


def _points(n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 3), np.float32)
    p[:, 0] = rng.uniform(-20000, 20000, n)
    p[:, 2] = rng.uniform(-20000, 20000, n)
    p[:, 1] = rng.uniform(0, 50, n)
    return p


@pytest.mark.parametrize("fset", ["ref", "fast"])
@pytest.mark.parametrize("sun_deg,time", [(40.0, 0.0), (12.0, 300.0)])
def test_shadow_matches_oracle(mm, gpu, orc, tex, fset, sun_deg, time):
    import torch

    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=1.0, sun_elevation_deg=sun_deg, time=time))
    ns = upload(mm, tex, fset)
    prm = mm.shadow_params()
    pts = _points(20011, 5)  # ragged last block
    out = torch.full((len(pts),), -1.0, dtype=torch.float32, device="cuda")
    mm.cloud_shadow(cam, sun, sky, ns, torch.from_numpy(pts).cuda(), out, params=prm)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = orc.cloud_shadow(cam, sun, sky, tex["weather"], tex["low"], pts, prm)
    d = np.abs(got - ref)
    print(fset, sun_deg, "nonzero", float((ref > 0).mean()), "max diff", float(d.max()), flush=True)
    assert (ref > 0).mean() > 0.01            # the case exercises clouds
    print("exact:", float(np.mean(d == 0)), "frac <= 1e-6:", float(np.mean(d <= 1e-6)), flush=True)
    # same UNORM8 channels in both sets, same operations in the same order, correctly rounded
    # sqrt/div, one shared pow formula: equal on every point
    assert np.array_equal(got, ref)


def test_shadow_empty_and_errors(mm, gpu, tex):
    import torch

    cam, sun, sky = mm.build_ubos(mm.scenario_scene())
    ns = upload(mm, tex, "fast")
    out = torch.zeros(4, dtype=torch.float32, device="cuda")
    mm.cloud_shadow(cam, sun, sky, ns, torch.zeros((0, 3), dtype=torch.float32, device="cuda"), out)
    prec = mm.NoiseSet("prec")
    with pytest.raises(mm.MMError) as e:
        mm.cloud_shadow(cam, sun, sky, prec, torch.zeros((4, 3), dtype=torch.float32, device="cuda"), out)
    assert e.value.code == mm.MM_E_FORMAT
