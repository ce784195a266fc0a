"""Asset ingestion feeding the march: textures written as the reference stores its assets (TGA
slices per z for the volumes, RLE with a bottom-left origin; PNG for the 2D maps), read back by the
library's loaders, uploaded into the reference and packed format sets, and marched on the GPU
against the oracle's march of the same pixels."""
import numpy as np
import pytest

from test_assets import png_bytes, tga_bytes
from test_gpu_march import check_parity, gpu_march, tex  # noqa: F401  (tex is a fixture)

pytestmark = pytest.mark.gpu


def _write_assets(tmp_path, tex):
    (tmp_path / "weather.png").write_bytes(png_bytes(tex["weather"].astype(np.int64), 6, 8))
    (tmp_path / "curl.tga").write_bytes(tga_bytes(np.ascontiguousarray(tex["curl"][..., [2, 1, 0, 3]]), 10, 32, False))
    for key, name in (("low", "low"), ("hi", "hi")):
        for z in range(tex[key].shape[0]):
            stored = np.ascontiguousarray(tex[key][z][..., [2, 1, 0, 3]])
            (tmp_path / f"{name}({z}).tga").write_bytes(tga_bytes(stored, 10, 32, False))


@pytest.mark.parametrize("fset", ["ref", "fast"])
def test_ingested_assets_march_like_the_oracle(mm, gpu, orc, tex, tmp_path, fset):
    _write_assets(tmp_path, tex)
    loaded = {"weather": mm.load_image(str(tmp_path / "weather.png")),
              "curl": mm.load_image(str(tmp_path / "curl.tga")),
              "low": mm.load_volume(str(tmp_path / "low"), tex["low"].shape[0]),
              "hi": mm.load_volume(str(tmp_path / "hi"), tex["hi"].shape[0])}
    for k in loaded:
        assert np.array_equal(loaded[k], tex[k].reshape(loaded[k].shape)), k
    ns = mm.NoiseSet(fset)
    for kind, (slot, key) in enumerate((("weather", "weather"), ("curl", "curl"), ("low_res", "low"),
                                        ("hi_res", "hi"))):
        mm.upload_rgba8(getattr(ns, slot), loaded[key], kind)
    W, H = 192, 108
    cam, sun, sky = mm.build_ubos(mm.scenario_scene(aspect=W / H))
    prm = mm.march_params("scenario")
    g = gpu_march(mm, ns, cam, sun, sky, W, H, params=prm)
    r = orc.march(cam, sun, sky, orc.textures(tex["weather"], tex["curl"], tex["low"], tex["hi"]), prm, W, H)
    check_parity(g, r, f"assets/{fset}")
