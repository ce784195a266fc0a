// CPU restatement of the temporal reprojection pass, SkyEngine/SkyEngine/Shaders/reproject.comp:91-153
// (test infrastructure: the checker of the HIP kernel, never linked into the product).
//
// Per target pixel: the ray through the pixel (same generation as compute-clouds.comp:305-322) is
// intersected with the inner shell sphere (reproject.comp:56-87: `.point` of the intersection,
// vec3(0) on a miss); the point goes to the previous camera's view space
// (`cameraPrev.view * vec4(p, 1)`), is de-normalised to <u, v, 1>, mapped to the previous frame's uv
// with the *current* camera's aspect/tan (reproject.comp:130-135), and the source image is sampled
// nearest at 10 points along the motion vector (average of RGBA), alpha replaced by the centre tap.
//
// Choices where GLSL leaves latitude, shared with the HIP kernel:
//   * round() rounds half away from zero;
//   * the float -> int conversion before the clamp saturates, NaN -> 0 (clamp then applies);
//   * mat4 * vec4 sums the columns left to right: ((c0 x + c1 y) + c2 z) + c3.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

namespace {

template <class T>
struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return v3(s * a.x, s * a.y, s * a.z); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / sqrt(dot(a, a))); }

// raySphereIntersection(...).point (reproject.comp:56-87); vec3(0) when there is no hit.
template <class I>
V3<I> sphere_point(V3<I> ro, V3<I> rd, V3<I> c, I w) {
    ro = v3((ro.x - c.x) / w, (ro.y - c.y) / w, (ro.z - c.z) / w);
    I A = dot(rd, rd);
    I B = I(2) * dot(rd, ro);
    I C = dot(ro, ro) - I(0.25);
    I disc = (B * B) - ((I(4) * A) * C);
    if (disc < I(0)) return v3(I(0), I(0), I(0));
    I t = ((-sqrt(disc) - B) / A) * I(0.5);
    if (t < I(0)) t = ((sqrt(disc) - B) / A) * I(0.5);
    if (!(t >= I(0))) return v3(I(0), I(0), I(0));
    V3<I> p = ro + rd * t;
    p = p * w;
    return p + c;
}

inline int clamp_index(float v, int n) {
    // saturating float -> int (NaN -> 0), then clamp to [0, n - 1]
    float f = v != v ? 0.0f : fminf(fmaxf(v, 0.0f), (float)(n - 1));
    return (int)f;
}

}  // namespace

extern "C" int orc_reproject(const mm_camera_ubo* cam, const mm_camera_ubo* prev, const float* src,
                             float* dst, int width, int height, int pitch, const mm_march_params* prm, int y0,
                             int y1) {
    if (!cam || !prev || !src || !dst || !prm) return -1;
    const float* V = cam->view;
    const V3<float> camLook = v3(V[2], V[6], V[10]);
    const V3<float> camRight = v3(V[0], V[4], V[8]);
    const V3<float> camUp = v3(V[1], V[5], V[9]);
    const V3<float> cameraPos = v3(cam->cameraPosition[0], cam->cameraPosition[1], cam->cameraPosition[2]);
    const V3<float> refPoint = cameraPos - camLook;
    const float aspect = cam->cameraParams[0], tanH = cam->cameraParams[1];
    const float* P = prev->view;
    const float dimx = (float)width, dimy = (float)height;
    for (int py = y0; py < y1; py++) {
        for (int px = 0; px < width; px++) {
            const float uvx = (float)px / dimx, uvy = (float)py / dimy;
            const float sx = uvx * 2.0f - 1.0f, sy = uvy * 2.0f - 1.0f;
            V3<float> p = (refPoint + (((aspect * sx) * tanH) * camRight)) - ((sy * tanH) * camUp);
            V3<float> rd = normalize(p - cameraPos);
            V3<float> center = cameraPos;
            center.y = prm->center_y;
            V3<float> hit;
            if (prm->flags & MM_MARCH_ISECT_FP64) {
                V3<double> h = sphere_point<double>(v3<double>(cameraPos.x, cameraPos.y, cameraPos.z),
                                                    v3<double>(rd.x, rd.y, rd.z), v3<double>(center.x, center.y, center.z),
                                                    (double)prm->inner_diameter);
                hit = v3((float)h.x, (float)h.y, (float)h.z);
            } else {
                hit = sphere_point<float>(cameraPos, rd, center, prm->inner_diameter);
            }
            // cameraPrev.view * vec4(hit, 1)
            V3<float> q;
            q.x = ((P[0] * hit.x + P[4] * hit.y) + P[8] * hit.z) + P[12];
            q.y = ((P[1] * hit.x + P[5] * hit.y) + P[9] * hit.z) + P[13];
            q.z = ((P[2] * hit.x + P[6] * hit.y) + P[10] * hit.z) + P[14];
            V3<float> d = normalize(q);
            const float nz = -d.z;
            d = v3(d.x / nz, d.y / nz, d.z / nz);
            float oldU = (d.x / tanH) / aspect;
            oldU = oldU * 0.5f + 0.5f;
            float oldV = (-d.y) / tanH;
            oldV = oldV * 0.5f + 0.5f;
            const float bx = oldU - uvx, by = oldV - uvy;
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
            for (int s = 0; s < 10; s++) {
                const float k = (float)s / 9.0f - 0.5f;
                const float ox = bx * k, oy = by * k;
                const int ix = clamp_index(roundf((oldU - ox) * dimx), width);
                const int iy = clamp_index(roundf((oldV - oy) * dimy), height);
                const float* t = src + ((size_t)iy * pitch + ix) * 4;
                for (int c = 0; c < 4; c++) acc[c] += t[c];
            }
            float* o = dst + ((size_t)py * pitch + px) * 4;
            for (int c = 0; c < 4; c++) o[c] = acc[c] / 10.0f;
            const int cx = clamp_index(roundf(oldU * dimx), width), cy = clamp_index(roundf(oldV * dimy), height);
            o[3] = src[((size_t)cy * pitch + cx) * 4 + 3];
        }
    }
    return 0;
}
