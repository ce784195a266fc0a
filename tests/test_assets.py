"""Asset ingestion (mm_image_info / mm_load_image_rgba8 / mm_load_volume_rgba8), host side.

Pinned by round trips: random pixels are written with small TGA and PNG encoders here (the file
formats' published layouts), decoded by the library, and must come back as stb_image's
STBI_rgb_alpha would return them (RGBA8, rows top first; 16-bit samples -> high byte; 5-5-5 ->
v * 255 / 31; grey replicated; missing alpha 255). The reference's own files, where present,
pin the loader on real data: its TGA and PNG encodings of the curl map decode to the same pixels,
whose hash is the committed golden fixture.
"""
import hashlib
import json
import os
import struct
import zlib

import numpy as np
import pytest

REF_TEX = "/root/reference/SkyEngine/SkyEngine/Textures"
FIX = os.path.join(os.path.dirname(__file__), "golden", "curl_fbm_fixture.json")


# ---------------------------------------------------------------- encoders
def tga_bytes(stored: np.ndarray, image_type: int, bits: int, top_left: bool, cmap=None, rle_runs=True) -> bytes:
    """stored: (h, w, bytes-per-pixel) in file byte order, rows top first."""
    h, w, pb = stored.shape
    rows = stored if top_left else stored[::-1]
    px = rows.reshape(-1, pb)
    hdr = bytearray(18)
    hdr[0] = 3  # an image ID field, skipped by the decoder
    if cmap is not None:
        hdr[1] = 1
        hdr[3:5] = struct.pack("<H", 0)
        hdr[5:7] = struct.pack("<H", len(cmap))
        hdr[7] = cmap.shape[1] * 8
    hdr[2] = image_type
    hdr[12:14] = struct.pack("<H", w)
    hdr[14:16] = struct.pack("<H", h)
    hdr[16] = bits
    hdr[17] = 0x20 if top_left else 0x00
    body = bytearray(b"abc")
    if cmap is not None:
        body += cmap.tobytes()
    if image_type < 9:
        body += px.tobytes()
    else:
        i = 0
        n = len(px)
        while i < n:
            j = i
            while j + 1 < n and j - i < 127 and (px[j + 1] == px[i]).all():
                j += 1
            if rle_runs and j > i:
                body += bytes([0x80 | (j - i)]) + px[i].tobytes()
                i = j + 1
            else:
                k = min(n, i + 128)
                body += bytes([k - i - 1]) + px[i:k].tobytes()
                i = k
    return bytes(hdr) + bytes(body)


def _chunk(t: bytes, d: bytes) -> bytes:
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def png_bytes(samples: np.ndarray, ctype: int, depth: int, plte=None, trns=None, filters=(0, 1, 2, 3, 4)) -> bytes:
    """samples: (h, w, channels) integer samples; rows filtered with types cycling over `filters`."""
    h, w, ch = samples.shape
    if depth < 8:
        per = 8 // depth
        rows = []
        for y in range(h):
            v = samples[y, :, 0].astype(np.int64)
            pad = (-w) % per
            v = np.concatenate([v, np.zeros(pad, np.int64)]).reshape(-1, per)
            shifts = np.array([8 - depth * (k + 1) for k in range(per)])
            rows.append((v << shifts).sum(1).astype(np.uint8))
        raw = np.stack(rows)
        bpp = 1
    elif depth == 8:
        raw = samples.astype(np.uint8).reshape(h, w * ch)
        bpp = ch
    else:
        raw = samples.astype(">u2").view(np.uint8).reshape(h, w * ch * 2)
        bpp = 2 * ch
    out = bytearray()
    prev = np.zeros(raw.shape[1], np.int64)
    for y in range(h):
        f = filters[y % len(filters)]
        cur = raw[y].astype(np.int64)
        a = np.concatenate([np.zeros(bpp, np.int64), cur[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]])
        if f == 0:
            enc = cur
        elif f == 1:
            enc = cur - a
        elif f == 2:
            enc = cur - prev
        elif f == 3:
            enc = cur - ((a + prev) >> 1)
        else:
            p = a + prev - c
            pa, pb, pc = np.abs(p - a), np.abs(p - prev), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
            enc = cur - pred
        out += bytes([f]) + (enc & 0xFF).astype(np.uint8).tobytes()
        prev = cur
    b = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if plte is not None:
        b += _chunk(b"PLTE", plte.astype(np.uint8).tobytes())
    if trns is not None:
        b += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(bytes(out), 6)
    b += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])  # split IDAT
    return b + _chunk(b"IEND", b"")


def _load(mm, tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data)
    return mm.load_image(str(p))


# ---------------------------------------------------------------- TGA
@pytest.mark.parametrize("image_type", [2, 10])
@pytest.mark.parametrize("bits", [24, 32])
@pytest.mark.parametrize("top_left", [False, True])
def test_tga_truecolor_round_trip(mm, tmp_path, image_type, bits, top_left):
    rng = np.random.default_rng(bits + image_type)
    h, w = 13, 21
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgba[2:5, 3:15] = rgba[2, 3]  # runs for the RLE packets
    if bits == 24:
        rgba[..., 3] = 255
    stored = rgba[..., [2, 1, 0, 3]][..., : bits // 8]  # BGR(A) in the file
    got = _load(mm, tmp_path, "t.tga", tga_bytes(np.ascontiguousarray(stored), image_type, bits, top_left))
    assert np.array_equal(got, rgba)


@pytest.mark.parametrize("image_type", [3, 11])
def test_tga_grey_round_trip(mm, tmp_path, image_type):
    g = np.random.default_rng(5).integers(0, 256, (9, 17, 1), dtype=np.uint8)
    got = _load(mm, tmp_path, "g.tga", tga_bytes(g, image_type, 8, False))
    assert np.array_equal(got[..., :3], np.repeat(g, 3, 2)) and (got[..., 3] == 255).all()


def test_tga_rgb555_and_colour_map(mm, tmp_path):
    rng = np.random.default_rng(7)
    v = rng.integers(0, 1 << 15, (6, 10), dtype=np.uint16)
    got = _load(mm, tmp_path, "p.tga", tga_bytes(v.view(np.uint8).reshape(6, 10, 2), 2, 16, True))
    for k, sh in enumerate((10, 5, 0)):
        assert np.array_equal(got[..., k], (((v >> sh) & 31).astype(int) * 255 // 31))
    cmap = rng.integers(0, 256, (40, 3), dtype=np.uint8)  # BGR entries
    idx = rng.integers(0, 40, (5, 8, 1), dtype=np.uint8)
    got = _load(mm, tmp_path, "c.tga", tga_bytes(idx, 9, 8, True, cmap=cmap))
    assert np.array_equal(got[..., :3], cmap[idx[..., 0]][..., ::-1])


def test_tga_truncated_and_missing(mm, tmp_path):
    data = tga_bytes(np.zeros((4, 4, 4), np.uint8), 10, 32, True)
    p = tmp_path / "bad.tga"
    p.write_bytes(data[:-3])
    with pytest.raises(mm.MMError) as e:
        mm.load_image(str(p))
    assert e.value.code == mm.MM_E_FORMAT
    with pytest.raises(mm.MMError) as e:
        mm.load_image(str(tmp_path / "absent.tga"))
    assert e.value.code == mm.MM_E_IO


def test_oversized_header_is_rejected_before_allocating(mm, tmp_path):
    hdr = bytearray(18)
    hdr[2] = 2
    hdr[12:16] = struct.pack("<HH", 65535, 65535)  # 4.3 G pixels claimed, no data
    hdr[16] = 32
    p = tmp_path / "huge.tga"
    p.write_bytes(bytes(hdr))
    with pytest.raises(mm.MMError) as e:
        mm.load_image(str(p))
    assert e.value.code == mm.MM_E_DIM


def test_volume_slices(mm, tmp_path):
    rng = np.random.default_rng(9)
    vol = rng.integers(0, 256, (5, 6, 7, 4), dtype=np.uint8)
    for i in range(5):
        stored = np.ascontiguousarray(vol[i][..., [2, 1, 0, 3]])
        (tmp_path / f"slice ({i}).tga").write_bytes(tga_bytes(stored, 10, 32, False))
    got = mm.load_volume(str(tmp_path / "slice "), 5)
    assert np.array_equal(got, vol)  # slice i at offset i * w * h * 4 (Texture.cpp:510-523)
    (tmp_path / "slice (2).tga").write_bytes(tga_bytes(np.zeros((3, 3, 4), np.uint8), 10, 32, False))
    with pytest.raises(mm.MMError) as e:
        mm.load_volume(str(tmp_path / "slice "), 5)
    assert e.value.code == mm.MM_E_DIM


# ---------------------------------------------------------------- PNG
@pytest.mark.parametrize("ctype,depth", [(0, 8), (0, 16), (2, 8), (2, 16), (4, 8), (4, 16), (6, 8), (6, 16),
                                         (0, 1), (0, 2), (0, 4)])
def test_png_round_trip(mm, tmp_path, ctype, depth):
    ch = {0: 1, 2: 3, 4: 2, 6: 4}[ctype]
    rng = np.random.default_rng(ctype * 100 + depth)
    h, w = 11, 19
    s = rng.integers(0, 1 << depth, (h, w, ch))
    got = _load(mm, tmp_path, "i.png", png_bytes(s, ctype, depth))
    to8 = (s >> 8) if depth == 16 else (s * (255 // ((1 << depth) - 1)) if depth < 8 else s)
    if ctype in (0, 4):
        want = np.concatenate([np.repeat(to8[..., :1], 3, 2), to8[..., 1:] if ctype == 4 else np.full((h, w, 1), 255)], 2)
    elif ctype == 2:
        want = np.concatenate([to8, np.full((h, w, 1), 255)], 2)
    else:
        want = to8
    assert np.array_equal(got, want.astype(np.uint8))


@pytest.mark.parametrize("depth", [8, 4, 2, 1])
def test_png_palette_with_transparency(mm, tmp_path, depth):
    rng = np.random.default_rng(depth)
    n = 1 << depth
    plte = rng.integers(0, 256, (n, 3))
    trns = list(rng.integers(0, 256, n // 2 + 1))  # entries past the tRNS list are opaque
    idx = rng.integers(0, n, (7, 13, 1))
    got = _load(mm, tmp_path, "p.png", png_bytes(idx, 3, depth, plte=plte, trns=trns))
    assert np.array_equal(got[..., :3], plte[idx[..., 0]])
    alpha = np.array([trns[i] if i < len(trns) else 255 for i in idx.reshape(-1)]).reshape(7, 13)
    assert np.array_equal(got[..., 3], alpha)


def test_png_colour_key_transparency(mm, tmp_path):
    s = np.random.default_rng(2).integers(0, 4, (6, 9, 3)) * 60
    s[1, 2] = (60, 120, 180)
    got = _load(mm, tmp_path, "k.png", png_bytes(s, 2, 8, trns=[0, 60, 0, 120, 0, 180]))
    key = (s == [60, 120, 180]).all(2)
    assert (got[..., 3][key] == 0).all() and (got[..., 3][~key] == 255).all()


def test_png_interlaced_is_rejected(mm, tmp_path):
    data = bytearray(png_bytes(np.zeros((4, 4, 3), np.int64), 2, 8))
    data[28] = 1  # IHDR interlace byte (the CRC is not checked)
    p = tmp_path / "a7.png"
    p.write_bytes(bytes(data))
    with pytest.raises(mm.MMError) as e:
        mm.load_image(str(p))
    assert e.value.code == mm.MM_E_FORMAT


# ---------------------------------------------------------------- the reference's own files
needs_ref = pytest.mark.skipif(not os.path.isdir(REF_TEX), reason="reference assets not present")


@needs_ref
def test_reference_curl_map_tga_and_png_agree_with_the_golden_hash(mm):
    a = mm.load_image(os.path.join(REF_TEX, "CurlNoiseFBM.tga"))  # RLE, bottom-left origin
    b = mm.load_image(os.path.join(REF_TEX, "CurlNoiseFBM.png"))  # 8-bit RGB PNG
    assert a.shape == (128, 128, 4) and np.array_equal(a[..., :3], b[..., :3])
    with open(FIX) as f:
        fx = json.load(f)
    assert hashlib.sha256(np.ascontiguousarray(a[..., :3]).tobytes()).hexdigest() == fx["rgb_sha256"]


@needs_ref
def test_reference_volumes_and_placement_map(mm):
    low = mm.load_volume(os.path.join(REF_TEX, "3DTextures", "lowResCloudShape", "lowResCloud"), 128)
    hi = mm.load_volume(os.path.join(REF_TEX, "3DTextures", "hiResCloudShape", "hiResClouds "), 32)
    assert low.shape == (128, 128, 128, 4) and hi.shape == (32, 32, 32, 4)
    m = low.reshape(-1, 4).mean(0) / 255
    assert abs(m[0] - 0.50) < 0.01 and (abs(m[1:] - 0.69) < 0.01).all()
    placement = mm.load_image(os.path.join(REF_TEX, "CloudPlacement.png"))  # 16-bit RGBA
    assert placement.shape == (512, 512, 4)
    # 16-bit samples keep their high byte: an independent decode of the PNG's first row
    data = open(os.path.join(REF_TEX, "CloudPlacement.png"), "rb").read()
    pos, idat = 8, b""
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        if data[pos + 4:pos + 8] == b"IDAT":
            idat += data[pos + 8:pos + 8 + n]
        pos += 12 + n
    raw = zlib.decompress(idat)
    enc = np.frombuffer(raw[1:1 + 512 * 8], np.uint8).astype(np.int64)
    assert raw[0] in (0, 1, 2)  # the first row has no previous row: None, Sub or Up
    if raw[0] == 1:  # Sub: each byte adds the byte one pixel (8 bytes) to its left
        enc = (np.cumsum(enc.reshape(512, 8), axis=0) & 0xFF).reshape(-1)
    assert np.array_equal(placement[0].reshape(-1), enc[0::2].astype(np.uint8))
