"""Writes tests/golden/asset_stats.json: channel statistics of the reference's baked cloud textures.

The reference ships its noise volumes and weather map as assets (README.md:68; loaded by
Texture.cpp:212-246 and 502-538): Textures/CloudPlacement.png (512^2 RGBA16),
Textures/3DTextures/lowResCloudShape/lowResCloud*.tga (128 slices of 128^2) and
Textures/3DTextures/hiResCloudShape/hiResClouds *.tga (32 slices of 32^2). The build defines its
own generators for these textures and tunes them to these statistics; the fixture keeps only the
numbers (means, standard deviations, the R-G correlation of the placement map), not the data.

Run in a checkout where the reference is mounted at /root/reference:
    python tests/golden/make_asset_stats.py
The files are decoded with this repository's own loaders (mm_load_image_rgba8 / mm_load_volume_rgba8).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import load_package  # noqa: E402

TEX = "/root/reference/SkyEngine/SkyEngine/Textures"


def stats(a):
    f = a.reshape(-1, 4).astype(np.float64) / 255.0
    return {"mean": [round(float(v), 5) for v in f.mean(0)], "std": [round(float(v), 5) for v in f.std(0)]}


def main():
    mm = load_package()
    low = mm.load_volume(os.path.join(TEX, "3DTextures", "lowResCloudShape", "lowResCloud"), 128)
    hi = mm.load_volume(os.path.join(TEX, "3DTextures", "hiResCloudShape", "hiResClouds "), 32)
    wm = mm.load_image(os.path.join(TEX, "CloudPlacement.png"))
    f = wm.reshape(-1, 4).astype(np.float64)
    out = {"source": "reference Textures/ (CloudPlacement.png, lowResCloud*.tga x128, hiResClouds *.tga x32)",
           "low_res": stats(low), "hi_res": stats(hi),
           "weather": dict(stats(wm), corr_rg=round(float(np.corrcoef(f[:, 0], f[:, 1])[0, 1]), 4))}
    with open(os.path.join(HERE, "asset_stats.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
