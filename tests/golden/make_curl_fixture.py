"""Writes tests/golden/curl_fbm_fixture.json from the reference's own curl-noise output.

The reference's GenerateCurlNoise (ImageUtils.cpp:176-223) wrote Textures/CurlNoiseFBM.tga (RLE
TGA, 128x128 RGBA8; its RGB is byte-identical to CurlNoiseFBM.png). The reference has no licence
file, so the texture itself is not copied: the fixture keeps the SHA-256 of the full 128x128x3 RGB
plane (enough to prove byte equality of a whole regenerated map), per-channel statistics, and 512
seeded sample texels with their coordinates. Run in a container where /root/reference exists.
"""
import hashlib
import json
import os

import numpy as np
from PIL import Image

SRC = "/root/reference/SkyEngine/SkyEngine/Textures/CurlNoiseFBM.tga"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "curl_fbm_fixture.json")


def main():
    img = np.array(Image.open(SRC))
    assert img.shape == (128, 128, 4) and img.dtype == np.uint8
    rgb = np.ascontiguousarray(img[..., :3])
    rng = np.random.default_rng(1)
    idx = rng.choice(128 * 128, size=512, replace=False)
    samples = [[int(i // 128), int(i % 128)] + [int(v) for v in rgb.reshape(-1, 3)[i]] for i in sorted(idx)]
    fx = {
        "source": "SkyEngine/SkyEngine/Textures/CurlNoiseFBM.tga (output of ImageUtils.cpp:176-223)",
        "shape": [128, 128, 3],
        "rgb_sha256": hashlib.sha256(rgb.tobytes()).hexdigest(),
        "alpha_all_255": bool((img[..., 3] == 255).all()),
        "channel_mean": [float(m) for m in rgb.reshape(-1, 3).mean(0)],
        "channel_min": [int(m) for m in rgb.reshape(-1, 3).min(0)],
        "channel_max": [int(m) for m in rgb.reshape(-1, 3).max(0)],
        "samples_row_col_r_g_b": samples,
    }
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=0)
    print("wrote", OUT, fx["rgb_sha256"])


if __name__ == "__main__":
    main()
