// Asset ingestion: the reference's texture files decoded to RGBA8 on the host, with the semantics
// of its loader -- stb_image 2.16 with STBI_rgb_alpha (Texture.cpp:212-246 for 2D images,
// Texture.cpp:502-538 for volumes stored as one TGA per z slice, "<prefix>(<i>).tga", slice i at
// byte offset i * width * height * 4, rows top first, x fastest).
//
// TGA: image types 1/2/3 and their run-length forms 9/10/11; 8-, 15/16-, 24- and 32-bit pixels;
// BGR(A) stored, RGB(A) returned; a bottom-left origin (descriptor bit 5 clear) is flipped so the
// first row returned is the top one; the right-to-left bit is ignored, as stb does.
// PNG: every colour type, bit depths 1/2/4/8/16 (16-bit samples keep their high byte, as stb's
// 8-bit interface does), tRNS transparency, filters 0-4, zlib via the system library; interlaced
// (Adam7) files are rejected. Colour-management chunks are ignored, as stb does.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mm.h"

namespace {

// Images above this many pixels are rejected before any allocation (16384 x 16384).
constexpr size_t kMaxPixels = (size_t)1 << 28;

struct Image {
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;
};

bool read_file(const char* path, std::vector<uint8_t>& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return false; }
    const long n = ftell(f);
    if (n < 0 || fseek(f, 0, SEEK_SET) != 0) { fclose(f); return false; }
    out.resize((size_t)n);
    const bool ok = n == 0 || fread(out.data(), 1, (size_t)n, f) == (size_t)n;
    fclose(f);
    return ok;
}

// ---------------------------------------------------------------- TGA
int decode_tga(const std::vector<uint8_t>& b, Image& im) {
    if (b.size() < 18) return MM_E_FORMAT;
    const int idLen = b[0], cmapType = b[1], type = b[2];
    const int cmapFirst = b[3] | (b[4] << 8), cmapLen = b[5] | (b[6] << 8), cmapBits = b[7];
    const int w = b[12] | (b[13] << 8), h = b[14] | (b[15] << 8), bits = b[16], desc = b[17];
    const bool rle = type >= 9;
    const int base = rle ? type - 8 : type;
    if (base < 1 || base > 3 || w <= 0 || h <= 0) return MM_E_FORMAT;
    if ((size_t)w * h > kMaxPixels) return MM_E_DIM;
    if (base == 1 && (cmapType != 1 || bits != 8)) return MM_E_FORMAT;
    if (base == 2 && bits != 15 && bits != 16 && bits != 24 && bits != 32) return MM_E_FORMAT;
    if (base == 3 && bits != 8 && bits != 16) return MM_E_FORMAT;
    size_t pos = 18 + (size_t)idLen;
    // colour map entries (type 1 only; skipped for the others)
    std::vector<uint8_t> cmap;
    int cmapBytes = 0;
    if (cmapType == 1) {
        cmapBytes = (cmapBits + 7) / 8;
        const size_t n = (size_t)cmapLen * cmapBytes;
        if (pos + n > b.size()) return MM_E_FORMAT;
        if (base == 1) {
            if (cmapBits != 15 && cmapBits != 16 && cmapBits != 24 && cmapBits != 32) return MM_E_FORMAT;
            cmap.assign(b.begin() + pos, b.begin() + pos + n);
        }
        pos += n;
    }
    const int pb = (bits + 7) / 8;  // bytes per stored pixel
    // one stored pixel (or colour-map entry) -> RGBA
    auto to_rgba = [&](const uint8_t* p, int nbits, bool gray, uint8_t* o) {
        if (gray) {
            o[0] = o[1] = o[2] = p[0];
            o[3] = nbits == 16 ? p[1] : 255;
        } else if (nbits == 15 || nbits == 16) {
            const int v = p[0] | (p[1] << 8);
            const int r = (v >> 10) & 31, g = (v >> 5) & 31, bl = v & 31;
            o[0] = (uint8_t)((r * 255) / 31);
            o[1] = (uint8_t)((g * 255) / 31);
            o[2] = (uint8_t)((bl * 255) / 31);
            o[3] = 255;
        } else {
            o[0] = p[2];
            o[1] = p[1];
            o[2] = p[0];
            o[3] = nbits == 32 ? p[3] : 255;
        }
    };
    const size_t npx = (size_t)w * h;
    std::vector<uint8_t> stored(npx * pb);
    if (!rle) {
        if (pos + stored.size() > b.size()) return MM_E_FORMAT;
        memcpy(stored.data(), b.data() + pos, stored.size());
    } else {
        size_t k = 0;
        while (k < npx) {
            if (pos >= b.size()) return MM_E_FORMAT;
            const int hdr = b[pos++];
            const size_t cnt = (size_t)(hdr & 0x7f) + 1;
            if (k + cnt > npx) return MM_E_FORMAT;
            if (hdr & 0x80) {
                if (pos + pb > b.size()) return MM_E_FORMAT;
                for (size_t j = 0; j < cnt; j++) memcpy(&stored[(k + j) * pb], b.data() + pos, pb);
                pos += pb;
            } else {
                if (pos + cnt * pb > b.size()) return MM_E_FORMAT;
                memcpy(&stored[k * pb], b.data() + pos, cnt * pb);
                pos += cnt * pb;
            }
            k += cnt;
        }
    }
    im.w = w;
    im.h = h;
    im.rgba.assign(npx * 4, 0);
    const bool flip = ((desc >> 5) & 1) == 0;  // bottom-left origin: the last stored row is the top
    for (int y = 0; y < h; y++) {
        const int sy = flip ? h - 1 - y : y;
        for (int x = 0; x < w; x++) {
            const uint8_t* p = &stored[((size_t)sy * w + x) * pb];
            uint8_t* o = &im.rgba[((size_t)y * w + x) * 4];
            if (base == 1) {
                const int idx = p[0] - cmapFirst;
                if (idx < 0 || idx >= cmapLen) return MM_E_FORMAT;
                to_rgba(&cmap[(size_t)idx * cmapBytes], cmapBits, false, o);
            } else {
                to_rgba(p, bits, base == 3, o);
            }
        }
    }
    return MM_OK;
}

// ---------------------------------------------------------------- PNG
uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

int decode_png(const std::vector<uint8_t>& b, Image& im) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (b.size() < 8 || memcmp(b.data(), sig, 8) != 0) return MM_E_FORMAT;
    size_t pos = 8;
    int w = 0, h = 0, depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool seenIhdr = false;
    while (pos + 12 <= b.size()) {
        const uint32_t n = be32(&b[pos]);
        const uint8_t* t = &b[pos + 4];
        if (pos + 12 + (size_t)n > b.size()) return MM_E_FORMAT;
        const uint8_t* d = &b[pos + 8];
        if (!memcmp(t, "IHDR", 4)) {
            if (n < 13) return MM_E_FORMAT;
            w = (int)be32(d);
            h = (int)be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            seenIhdr = true;
        } else if (!memcmp(t, "PLTE", 4)) {
            plte.assign(d, d + n);
        } else if (!memcmp(t, "tRNS", 4)) {
            trns.assign(d, d + n);
        } else if (!memcmp(t, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + n);
        } else if (!memcmp(t, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)n;
    }
    if (!seenIhdr || w <= 0 || h <= 0 || w > (1 << 16) || h > (1 << 16)) return MM_E_FORMAT;
    if ((size_t)w * h > kMaxPixels) return MM_E_DIM;
    if (interlace != 0) return MM_E_FORMAT;
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: return MM_E_FORMAT;
    }
    const bool lowOk = ctype == 0 || ctype == 3;
    if (!(depth == 8 || depth == 16 || (lowOk && (depth == 1 || depth == 2 || depth == 4)))) return MM_E_FORMAT;
    if (ctype == 3 && (depth == 16 || plte.size() < 3)) return MM_E_FORMAT;
    const size_t rowBytes = ((size_t)w * channels * depth + 7) / 8;
    const size_t bpp = std::max<size_t>(1, (size_t)channels * depth / 8);  // filter byte distance
    std::vector<uint8_t> raw((rowBytes + 1) * h);
    {
        z_stream zs;
        memset(&zs, 0, sizeof(zs));
        if (inflateInit(&zs) != Z_OK) return MM_E_FORMAT;
        zs.next_in = idat.data();
        zs.avail_in = (uInt)idat.size();
        zs.next_out = raw.data();
        zs.avail_out = (uInt)raw.size();
        const int r = inflate(&zs, Z_FINISH);
        const size_t got = raw.size() - zs.avail_out;
        inflateEnd(&zs);
        if ((r != Z_STREAM_END && r != Z_OK && r != Z_BUF_ERROR) || got != raw.size()) return MM_E_FORMAT;
    }
    // unfilter in place, row by row
    std::vector<uint8_t> img(rowBytes * h);
    for (int y = 0; y < h; y++) {
        const uint8_t f = raw[(rowBytes + 1) * y];
        const uint8_t* src = &raw[(rowBytes + 1) * y + 1];
        uint8_t* cur = &img[rowBytes * y];
        const uint8_t* prev = y ? &img[rowBytes * (y - 1)] : nullptr;
        for (size_t i = 0; i < rowBytes; i++) {
            const int a = i >= bpp ? cur[i - bpp] : 0, up = prev ? prev[i] : 0, c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int v = src[i];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) >> 1; break;
                case 4: v += paeth(a, up, c); break;
                default: return MM_E_FORMAT;
            }
            cur[i] = (uint8_t)v;
        }
    }
    // one sample (channel k of pixel x in row) as stored, 16-bit samples whole
    auto sample = [&](const uint8_t* row, int x, int k) -> int {
        if (depth == 16) {
            const uint8_t* p = row + ((size_t)x * channels + k) * 2;
            return (p[0] << 8) | p[1];
        }
        if (depth == 8) return row[(size_t)x * channels + k];
        const size_t bit = (size_t)x * depth;  // 1/2/4-bit: one channel
        return (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
    };
    auto to8 = [&](int v) -> uint8_t {
        if (depth == 16) return (uint8_t)(v >> 8);
        if (depth == 8) return (uint8_t)v;
        return (uint8_t)(v * (255 / ((1 << depth) - 1)));  // gray 1/2/4 bits scaled to 8
    };
    // tRNS key of grey / RGB images, compared on the stored samples
    int key[3] = {-1, -1, -1};
    if (!trns.empty() && ctype == 0 && trns.size() >= 2) key[0] = (trns[0] << 8) | trns[1];
    if (!trns.empty() && ctype == 2 && trns.size() >= 6)
        for (int k = 0; k < 3; k++) key[k] = (trns[2 * k] << 8) | trns[2 * k + 1];
    im.w = w;
    im.h = h;
    im.rgba.assign((size_t)w * h * 4, 0);
    for (int y = 0; y < h; y++) {
        const uint8_t* row = &img[rowBytes * y];
        for (int x = 0; x < w; x++) {
            uint8_t* o = &im.rgba[((size_t)y * w + x) * 4];
            switch (ctype) {
                case 0: {
                    const int g = sample(row, x, 0);
                    o[0] = o[1] = o[2] = to8(g);
                    o[3] = (key[0] >= 0 && g == key[0]) ? 0 : 255;
                    break;
                }
                case 2: {
                    int s[3];
                    for (int k = 0; k < 3; k++) s[k] = sample(row, x, k);
                    for (int k = 0; k < 3; k++) o[k] = to8(s[k]);
                    o[3] = (key[0] >= 0 && s[0] == key[0] && s[1] == key[1] && s[2] == key[2]) ? 0 : 255;
                    break;
                }
                case 3: {
                    const int i = sample(row, x, 0);
                    if ((size_t)i * 3 + 2 >= plte.size()) return MM_E_FORMAT;
                    o[0] = plte[i * 3];
                    o[1] = plte[i * 3 + 1];
                    o[2] = plte[i * 3 + 2];
                    o[3] = (size_t)i < trns.size() ? trns[i] : 255;
                    break;
                }
                case 4: {
                    o[0] = o[1] = o[2] = to8(sample(row, x, 0));
                    o[3] = to8(sample(row, x, 1));
                    break;
                }
                case 6: {
                    for (int k = 0; k < 4; k++) o[k] = to8(sample(row, x, k));
                    break;
                }
            }
        }
    }
    return MM_OK;
}

int decode(const char* path, Image& im) {
    // no exception crosses the C boundary: allocation failures become an error code
    try {
        std::vector<uint8_t> b;
        if (!read_file(path, b)) return MM_E_IO;
        if (b.size() >= 8 && b[0] == 137 && b[1] == 'P' && b[2] == 'N' && b[3] == 'G') return decode_png(b, im);
        return decode_tga(b, im);
    } catch (...) {
        return MM_E_SIZE;
    }
}

}  // namespace

extern "C" {

int mm_image_info(const char* path, int32_t* width, int32_t* height) {
    if (!path || !width || !height) return MM_E_NULL;
    Image im;
    const int e = decode(path, im);
    if (e) return e;
    *width = im.w;
    *height = im.h;
    return MM_OK;
}

int mm_load_image_rgba8(const char* path, uint8_t* rgba, size_t bytes, int32_t* width, int32_t* height) {
    if (!path || !rgba || !width || !height) return MM_E_NULL;
    Image im;
    const int e = decode(path, im);
    if (e) return e;
    if (bytes < im.rgba.size()) return MM_E_SIZE;
    memcpy(rgba, im.rgba.data(), im.rgba.size());
    *width = im.w;
    *height = im.h;
    return MM_OK;
}

int mm_load_volume_rgba8(const char* prefix, int32_t depth, uint8_t* rgba, size_t bytes, int32_t* width,
                         int32_t* height) {
    if (!prefix || !rgba || !width || !height) return MM_E_NULL;
    if (depth <= 0) return MM_E_DIM;
    int w = 0, h = 0;
    try {
        for (int i = 0; i < depth; i++) {
            char suffix[32];
            snprintf(suffix, sizeof(suffix), "(%d).tga", i);
            const std::string path = std::string(prefix) + suffix;
            Image im;
            const int e = decode(path.c_str(), im);
            if (e) return e;
            if (i == 0) {
                w = im.w;
                h = im.h;
                if (bytes < (size_t)w * h * 4 * depth) return MM_E_SIZE;
            } else if (im.w != w || im.h != h) {
                return MM_E_DIM;
            }
            memcpy(rgba + (size_t)i * w * h * 4, im.rgba.data(), im.rgba.size());
        }
    } catch (...) {
        return MM_E_SIZE;
    }
    *width = w;
    *height = h;
    return MM_OK;
}

}  // extern "C"
