// Oracle: scalar restatement of the cloud march, SkyEngine/SkyEngine/Shaders/compute-clouds.comp
// (test infrastructure only; see oracle.h). Templated on the arithmetic type: float is the parity
// oracle, double the FP64 error reference. Expressions keep the shader's evaluation order, and
// GLSL built-ins follow the GLSL 4.50 definitions:
//   mix(x, y, a) = x * (1 - a) + y * a        smoothstep = t * t * (3 - 2 t), t clamped
//   clamp(x, lo, hi) = min(max(x, lo), hi)    (IEEE minNum/maxNum, so a NaN operand yields lo/hi)
//   normalize(v) = v * (1 / sqrt(dot(v, v)))  dot(a, b) = (a.x b.x + a.y b.y) + a.z b.z
// The hardware sampler (Texture.cpp:29-52, 315-338: LINEAR, REPEAT, RGBA8 UNORM, one mip level) is
// restated in sample_2d / sample_3d.
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <omp.h>
#include "oracle.h"

namespace {

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return v3(s * a.x, s * a.y, s * a.z); }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return v3(a.x / s, a.y / s, a.z / s); }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
template <class T> inline T length(V3<T> a) { return sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / sqrt(dot(a, a))); }

template <class T> inline T fmaxT(T a, T b) { return fmax(a, b); }
template <class T> inline T fminT(T a, T b) { return fmin(a, b); }
template <class T> inline T clampT(T x, T lo, T hi) { return fminT(fmaxT(x, lo), hi); }
template <class T> inline T mixT(T x, T y, T a) { return x * (T(1) - a) + y * a; }
template <class T> inline V3<T> mix3(V3<T> x, V3<T> y, T a) {
    return v3(mixT(x.x, y.x, a), mixT(x.y, y.y, a), mixT(x.z, y.z, a));
}
template <class T> inline T smoothstepT(T e0, T e1, T x) {
    T t = clampT((x - e0) / (e1 - e0), T(0), T(1));
    return t * t * (T(3) - T(2) * t);
}
template <class T> inline T powT(T x, T y) { return pow(x, y); }
template <> inline float powT<float>(float x, float y) { return powf(x, y); }

// compute-clouds.comp:65-71
template <class T> inline T remap(T v, T a, T b, T c, T d) { return c + (((v - a) / (b - a)) * (d - c)); }
template <class T> inline T remapClamped(T v, T a, T b, T c, T d) {
    return clampT(c + (((v - a) / (b - a)) * (d - c)), c, d);
}
// remap / remapClamped whose range arguments are literals of the shader. A Vulkan driver folds the
// constant divisor into its reciprocal and multiplies (SPIR-V OpFDiv carries 2.5 ulp precision, and
// the compilers lower x / c to x * (1 / c) with 1 / c folded once): the float restatement does the
// same, rounding 1 / (b - a) once. The FP64 error reference keeps the quotient.
template <class T> inline T remapK(T v, T a, T b, T c, T d) { return c + (((v - a) * (T(1) / (b - a))) * (d - c)); }
template <> inline double remapK<double>(double v, double a, double b, double c, double d) {
    return remap<double>(v, a, b, c, d);
}
template <class T> inline T remapClampedK(T v, T a, T b, T c, T d) { return clampT(remapK<T>(v, a, b, c, d), c, d); }

// pow of the coverage exponent (comp:206-208). GLSL defines pow's precision as that of
// exp2(y * log2(x)) and the drivers evaluate it with hardware log2 / exp2 whose rounding no
// specification fixes. The restatement and the GPU kernel use one written-out evaluation, so both
// round alike: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log2(m) = f P(f) with f = m - 1 (degree-8
// polynomial, |error| < 3e-8, inside GLSL's 2^-21 bound for log2 on [0.5, 2]), z = y log2(x),
// exp2(z) = 2^n Q(z - n) with n = rint(z) (degree-6 polynomial, relative error < 2e-8). pow(x, 1) = x
// and x <= 0 gives 0. The FP64 error reference keeps the library pow.
inline float pow_rt(float x, float y) {
    if (y == 1.0f) return x;
    if (!(x > 0.0f)) return 0.0f;
    int e = 0;
    float m = frexpf(x, &e);
    if (m < 0.70710678f) {
        m = m + m;
        e = e - 1;
    }
    const float f = m - 1.0f;
    float p = 1.231096834e-01f;
    p = fmaf(p, f, -2.058618814e-01f);
    p = fmaf(p, f, 2.160782218e-01f);
    p = fmaf(p, f, -2.391698807e-01f);
    p = fmaf(p, f, 2.879032493e-01f);
    p = fmaf(p, f, -3.606932759e-01f);
    p = fmaf(p, f, 4.809107482e-01f);
    p = fmaf(p, f, -7.213474512e-01f);
    p = fmaf(p, f, 1.442695022e+00f);
    const float l = fmaf(f, p, (float)e);
    const float z = y * l;
    const float n = rintf(z);
    const float r = z - n;
    float q = 1.546144485e-04f;
    q = fmaf(q, r, 1.340042800e-03f);
    q = fmaf(q, r, 9.618056938e-03f);
    q = fmaf(q, r, 5.550327152e-02f);
    q = fmaf(q, r, 2.402265072e-01f);
    q = fmaf(q, r, 6.931471825e-01f);
    q = fmaf(q, r, 1.000000000e+00f);
    return ldexpf(q, (int)n);
}
template <class T> inline T coveragePow(T x, T y) { return pow(x, y); }
template <> inline float coveragePow<float>(float x, float y) { return pow_rt(x, y); }

// ---- sampler (Texture.cpp:29-52, 315-338) ----
// Unnormalised coordinate x = u * n - 0.5; with `bits` sub-texel bits it is rounded to the nearest
// multiple of 2^-bits (the fixed-point conversion of the texture unit), then split into the integer
// texel and the fraction. REPEAT wraps the texel index modulo n.
template <class T> inline void axis(T u, int n, int bits, int& i0, int& i1, T& f) {
    T x = u * T(n) - T(0.5);
    if (bits > 0) {
        T s = T(1 << bits);
        x = floor(x * s + T(0.5)) / s;
    }
    T fl = floor(x);
    f = x - fl;
    long long ii = (fl == fl) ? (long long)fl : 0;  // NaN coordinate -> texel 0
    long long m = ii % n;
    if (m < 0) m += n;
    i0 = (int)m;
    i1 = (i0 + 1 == n) ? 0 : i0 + 1;
}
// Filter lerps as one fused multiply-add, a + t (b - a) rounded once. With 8-bit sub-texel weights
// every bilinear step of UNORM8 texels is exact and the trilinear result is the exact weighted sum
// rounded once (the last lerp's single rounding).
template <class T> inline T lerpT(T a, T b, T t) { return fma(t, b - a, a); }

// Bilinear fetch of channels [c0, c0 + nc) of an RGBA8 texture; values are filtered as the integer
// byte values and scaled by 1/255 once (UNORM8).
template <class T>
inline void sample2d(const uint8_t* tex, int w, int h, T u, T v, int bits, T* out) {
    int x0, x1, y0, y1;
    T fx, fy;
    axis<T>(u, w, bits, x0, x1, fx);
    axis<T>(v, h, bits, y0, y1, fy);
    const uint8_t* r0 = tex + (size_t)y0 * w * 4;
    const uint8_t* r1 = tex + (size_t)y1 * w * 4;
    for (int c = 0; c < 4; c++) {
        T a = lerpT(T(r0[x0 * 4 + c]), T(r0[x1 * 4 + c]), fx);
        T b = lerpT(T(r1[x0 * 4 + c]), T(r1[x1 * 4 + c]), fx);
        out[c] = lerpT(a, b, fy) * (T(1) / T(255));
    }
}

template <class T>
inline void sample3d(const uint8_t* tex, int n, T u, T v, T w, int bits, T* out) {
    int x0, x1, y0, y1, z0, z1;
    T fx, fy, fz;
    axis<T>(u, n, bits, x0, x1, fx);
    axis<T>(v, n, bits, y0, y1, fy);
    axis<T>(w, n, bits, z0, z1, fz);
    const size_t n2 = (size_t)n * n;
    const uint8_t* p00 = tex + (z0 * n2 + (size_t)y0 * n) * 4;
    const uint8_t* p10 = tex + (z0 * n2 + (size_t)y1 * n) * 4;
    const uint8_t* p01 = tex + (z1 * n2 + (size_t)y0 * n) * 4;
    const uint8_t* p11 = tex + (z1 * n2 + (size_t)y1 * n) * 4;
    for (int c = 0; c < 4; c++) {
        T a0 = lerpT(T(p00[x0 * 4 + c]), T(p00[x1 * 4 + c]), fx);
        T b0 = lerpT(T(p10[x0 * 4 + c]), T(p10[x1 * 4 + c]), fx);
        T a1 = lerpT(T(p01[x0 * 4 + c]), T(p01[x1 * 4 + c]), fx);
        T b1 = lerpT(T(p11[x0 * 4 + c]), T(p11[x1 * 4 + c]), fx);
        T s0 = lerpT(a0, b0, fy);
        T s1 = lerpT(a1, b1, fy);
        out[c] = lerpT(s0, s1, fz) * (T(1) / T(255));
    }
}

// ---- scene state of one dispatch ----
template <class T> struct Scene {
    const mm_camera_ubo* cam;
    const mm_sun_ubo* sun;
    const mm_sky_ubo* sky;
    const orc_textures* tex;
    const mm_march_params* p;
    int width, height;
};

template <class T> inline V3<T> ubo3(const float* f) { return v3(T(f[0]), T(f[1]), T(f[2])); }

// compute-clouds.comp:73-77
template <class T> inline T hgPhase(T cosTheta, T g) {
    T g2 = g * g;
    T inv = T(1) / powT<T>(((T(1) - (T(2) * g) * cosTheta) + g2), T(1.5));
    return T(0.07957747154594767) * ((T(1) - g2) * inv);
}
// compute-clouds.comp:84-86
template <class T> inline T rayleighPhase(T c) { return T(0.05968310365946075) * (T(1) + c * c); }

// compute-clouds.comp:88-127 (sun disk term forced to 0 at :120; sun.location unused)
template <class T> V3<T> atmosphere(const Scene<T>& s, V3<T> dir, V3<T> sunDir) {
    const T PI = T(3.14159265f);
    T sunE = T(s.sun->intensity);
    V3<T> BetaR = ubo3<T>(s.sky->betaR), BetaM = ubo3<T>(s.sky->betaV);
    T zenith = acos(fmaxT(T(0), dir.y));
    T inverse = T(1) / (cos(zenith) + T(0.15) * powT<T>(T(93.885) - ((zenith * T(180)) / PI), T(-1.253)));
    T sR = T(8.4e3) * inverse;
    T sM = T(1.25e3) * inverse;
    V3<T> e = (BetaR * (-sR)) + (BetaM * sM);  // -BetaR * sR + BetaM * sM
    V3<T> fex = v3<T>(exp(e.x), exp(e.y), exp(e.z));
    T cosTheta = dot(sunDir, dir);
    T rPhase = rayleighPhase<T>(cosTheta * T(0.5) + T(0.5));
    V3<T> betaRTheta = BetaR * rPhase;
    T mPhase = hgPhase<T>(cosTheta, T(s.sky->mie_directional));
    V3<T> betaMTheta = BetaM * mPhase;
    T yDot = T(1) - sunDir.y;
    yDot = yDot * (((yDot * yDot) * yDot) * yDot);
    V3<T> betas = (betaRTheta + betaMTheta) / (BetaR + BetaM);
    V3<T> one = v3<T>(1, 1, 1);
    V3<T> a = (betas * sunE) * (one - fex);
    V3<T> Lin = v3<T>(powT<T>(a.x, T(1.5)), powT<T>(a.y, T(1.5)), powT<T>(a.z, T(1.5)));
    V3<T> b = (betas * sunE) * fex;
    V3<T> pb = v3<T>(powT<T>(b.x, T(0.5)), powT<T>(b.y, T(0.5)), powT<T>(b.z, T(0.5)));
    Lin = Lin * mix3(one, pb, clampT(yDot, T(0), T(1)));
    V3<T> L0 = fex * T(0.1);
    // sunDisk = 0 (comp:120): L0 += (sunE * 15000 * fex) * 0 leaves L0 unchanged for finite inputs
    V3<T> color = ((Lin + L0) * T(0.04)) + v3<T>(T(0), T(0.0003), T(0.00075));
    return color;
}

// compute-clouds.comp:147-177, evaluated in arithmetic type I. `t` of a miss is undefined in the
// reference; here it is 0 (valid = false).
template <class I> struct Isect {
    bool valid;
    I t;
};
template <class I>
Isect<I> raySphere(V3<I> ro, V3<I> rd, V3<I> c, I w) {
    Isect<I> r{false, I(0)};
    ro = (ro - c) / w;
    I A = dot(rd, rd);
    I B = I(2) * dot(rd, ro);
    I C = dot(ro, ro) - I(0.25);
    I disc = (B * B) - ((I(4) * A) * C);
    if (disc < I(0)) return r;
    I t = ((-sqrt(disc) - B) / A) * I(0.5);
    if (t < I(0)) t = ((sqrt(disc) - B) / A) * I(0.5);
    if (t >= I(0)) {
        r.valid = true;
        V3<I> p = ro + rd * t;
        p = p * w;
        p = p + c;
        r.t = length(p - ro);  // comp:173: world-space point minus the normalised origin
    }
    return r;
}

// compute-clouds.comp:180-188
template <class T> inline V3<T> projectedShellPoint(const Scene<T>& s, V3<T> pt, V3<T> center) {
    return (T(s.p->proj_radius) * normalize(pt - center)) + center;
}
// The shell thickness is a constant expression of the shader (comp:359): its division folds into a
// multiplication by the reciprocal, like remapK's (the FP64 reference keeps the quotient).
template <class T> inline T relativeHeight(V3<T> pt, V3<T> proj, T thickness) {
    return clampT(length(pt - proj) * (T(1) / thickness), T(0), T(1));
}
template <> inline double relativeHeight<double>(V3<double> pt, V3<double> proj, double thickness) {
    return clampT(length(pt - proj) / thickness, 0.0, 1.0);
}

// compute-clouds.comp:193-204
template <class T> T cloudLayerDensity(T h, T type) {
    h = clampT(h, T(0), T(1));
    T cumulus = fmaxT(T(0), remapK<T>(h, T(0), T(0.2), T(0), T(1)) * remapK<T>(h, T(0.7), T(0.9), T(1), T(0)));
    T strato = fmaxT(T(0), remapK<T>(h, T(0), T(0.2), T(0), T(1)) * remapK<T>(h, T(0.2), T(0.7), T(1), T(0)));
    T stratus = fmaxT(T(0), remapK<T>(h, T(0), T(0.1), T(0), T(1)) * remapK<T>(h, T(0.2), T(0.3), T(1), T(0)));
    T d1 = mixT(stratus, strato, clampT(type * T(2), T(0), T(1)));
    T d2 = mixT(strato, cumulus, clampT((type - T(0.5)) * T(2), T(0), T(1)));
    return mixT(d1, d2, type);
}

// compute-clouds.comp:206-208
template <class T> inline T heightBiasCoverage(T coverage, T height) {
    return coveragePow<T>(coverage, clampT(remapK<T>(height, T(0.7), T(0.8), T(1), T(0.8)), T(0.8), T(1)));
}

template <class T> struct Counters {
    int density, detail;
};

// compute-clouds.comp:214-228
template <class T>
T cloudHiRes(const Scene<T>& s, V3<T> pos, T curlStrength, T origDensity, T h, Counters<T>& n) {
    n.detail++;
    T c[4], d[4];
    T cs = T(s.p->curl_scale);
    sample2d<T>(s.tex->curl, s.tex->curl_w, s.tex->curl_h, cs * pos.x, cs * pos.z, s.p->subtexel_bits, c);
    V3<T> curl = v3<T>(T(2) * c[0] - T(1), T(2) * c[1] - T(1), T(2) * c[2] - T(1));
    pos = pos + ((T(1.9) * curlStrength) * curl);
    T hs = T(s.p->hi_scale);
    sample3d<T>(s.tex->hi, s.tex->hi_dim, hs * pos.x, hs * pos.y, hs * pos.z, s.p->subtexel_bits, d);
    T erosion = ((T(0.625) * d[0]) + (T(0.25) * d[1])) + (T(0.125) * d[2]);
    erosion = mixT(erosion, T(1) - erosion, clampT(h * T(10), T(0), T(1)));
    return remapClamped<T>(origDensity, T(1) * erosion, T(1), T(0), T(1));
}

// compute-clouds.comp:231-253
template <class T>
T cloudTest(const Scene<T>& s, V3<T> pos, T h, V3<T> center, Counters<T>& n) {
    n.density++;
    V3<T> proj = projectedShellPoint<T>(s, pos, center);
    T ws = T(s.p->weather_scale);
    T info[4], dn[4];
    sample2d<T>(s.tex->weather, s.tex->weather_w, s.tex->weather_h,
                ws * (proj.x - T(s.cam->cameraPosition[0])), ws * (proj.z - T(s.cam->cameraPosition[2])),
                s.p->subtexel_bits, info);
    T layer = cloudLayerDensity<T>(h, info[2]);
    T ls = T(s.p->low_scale);
    sample3d<T>(s.tex->low, s.tex->low_dim, ls * pos.x, ls * pos.y, ls * pos.z, s.p->subtexel_bits, dn);
    T density = layer * remapClampedK<T>(dn[0], T(0.3), T(1), T(0), T(1));
    if (density < T(0.0001)) return T(0);
    T coverage = heightBiasCoverage<T>(h, fminT(T(0.85), info[0]));  // arguments as in comp:245
    T erosion = ((T(0.625) * dn[1]) + (T(0.25) * dn[2])) + (T(0.125) * dn[3]);
    erosion = remapClamped<T>(erosion, coverage, T(1), T(0), T(1));
    density = remapClamped<T>(density, erosion, T(1), T(0), T(1));
    return density;
}

// compute-clouds.comp:256-277 (columns of the returned matrix)
template <class T> void fromAngleAxis(V3<T> a, T angle, V3<T> col[3]) {
    T cost = cos(angle), sint = sin(angle);
    T omc = T(1) - cost;
    col[0] = v3<T>(cost + a.x * a.x * omc, a.y * a.x * omc + a.z * sint, a.z * a.x * omc - a.y * sint);
    col[1] = v3<T>(a.x * a.y * omc - a.z * sint, cost + a.y * a.y * omc, a.z * a.y * omc + a.x * sint);
    col[2] = v3<T>(a.x * a.z * omc + a.y * sint, a.y * a.z * omc - a.x * sint, cost + a.z * a.z * omc);
}
template <class T> inline V3<T> matmul(const V3<T> col[3], V3<T> v) {
    return ((col[0] * v.x) + (col[1] * v.y)) + (col[2] * v.z);
}

template <class T> inline T pow_chain4(T d) { return d * ((d * d) * d); }

// Bilinear fetch of the night-sky map (binding 5); black when absent.
template <class T> V3<T> nightSky(const Scene<T>& s, T u, T v) {
    if (!s.tex->night) return v3<T>(0, 0, 0);
    T c[4];
    sample2d<T>(s.tex->night, s.tex->night_w, s.tex->night_h, u, v, s.p->subtexel_bits, c);
    return v3<T>(c[0], c[1], c[2]);
}

// Optional record of one pixel's march loop, one entry per loop trip (orc_march_trace_f32):
// t and the step size at the top of the trip, the low-res density, and what the trip did.
enum TripEvent {
    TRIP_COARSE_MISS = 0,   // coarse mode, density 0
    TRIP_FIRST_HIT = 1,     // coarse mode, density > 0: back-step, fine step, uncounted continue (comp:428-433)
    TRIP_FINE_THIN = 2,     // fine mode, detail density < 1e-4: uncounted continue (comp:437)
    TRIP_FINE_HIT = 3,      // fine mode, light cone and accumulation (comp:440-466)
    TRIP_FINE_MISS = 4,     // fine mode, density 0, misses below the limit
    TRIP_REVERT = 5,        // fine mode, the miss that reaches the limit: back to coarse (comp:468-474)
};
struct Trace {
    float* trips;  // [max][4]
    int max, n;
};
template <class T> inline void trace_trip(Trace* tr, T t, T step, T dens, int ev) {
    if (!tr) return;
    if (tr->n < tr->max) {
        float* e = tr->trips + 4 * (size_t)tr->n;
        e[0] = (float)t; e[1] = (float)step; e[2] = (float)dens; e[3] = (float)ev;
    }
    tr->n++;
}

// compute-clouds.comp:288-500 for one pixel.
// state (optional): the opacity accumulator after the horizon fade and the 0.999 cap (comp:484-486) and
// the light-energy accumulator `transmittance` (comp:464), the two transmittance forms of a pixel.
template <class T>
void marchPixel(const Scene<T>& s, int px, int py, T out[4], int* cnt, Trace* tr = nullptr, T* state = nullptr) {
    const mm_march_params* P = s.p;
    Counters<T> n{0, 0};
    T timeOffset = T(s.sky->wind[3]);
    // comp:305 uv = pixel / (WIDTH, HEIGHT); the build uses the image size it is given
    T uvx = T(px) / T(s.width), uvy = T(py) / T(s.height);
    T sx = uvx * T(2) - T(1), sy = uvy * T(2) - T(1);
    const float* V = s.cam->view;
    V3<T> camLook = v3<T>(T(V[2]), T(V[6]), T(V[10]));
    V3<T> camRight = v3<T>(T(V[0]), T(V[4]), T(V[8]));
    V3<T> camUp = v3<T>(T(V[1]), T(V[5]), T(V[9]));
    const V3<T> cameraPos = ubo3<T>(s.cam->cameraPosition);
    V3<T> refPoint = cameraPos - camLook;
    T aspect = T(s.cam->cameraParams[0]), tanH = T(s.cam->cameraParams[1]);
    V3<T> p = (refPoint + (((aspect * sx) * tanH) * camRight)) - ((sy * tanH) * camUp);
    V3<T> rd = normalize(p - cameraPos);
    V3<T> sunDir = normalize(ubo3<T>(s.sun->directionBasis + 4));

    const T SUN_COS = T(0.999956676946448443553574619906976478926848692873900859324);
    T dotToSun = fmaxT(T(0), dot(sunDir, rd));
    T skyAmbient = dotToSun * T(0.18);
    skyAmbient = skyAmbient * (skyAmbient * skyAmbient);
    T sunDisk = smoothstepT<T>(SUN_COS, SUN_COS + T(0.00003), dotToSun);
    dotToSun = pow_chain4(dotToSun);
    dotToSun = pow_chain4(dotToSun);
    dotToSun = pow_chain4(dotToSun);
    dotToSun = pow_chain4(dotToSun);
    dotToSun = dotToSun * (dotToSun * dotToSun);
    const bool night = T(s.sun->direction[1]) < T(0);
    if (night) {
        T d = dotToSun;
        dotToSun = d * ((((((d * d) * d) * d) * d) * d) * d);
    }
    sunDisk = fmaxT(sunDisk, dotToSun);
    sunDisk = fmaxT(T(0), sunDisk);

    T fc[4] = {0, 0, 0, 0};
    V3<T> bg = v3<T>(0, 0, 0);
    if (!night) {
        bg = atmosphere<T>(s, rd, sunDir);
        fc[3] = fmaxT(skyAmbient, sunDisk);
        fc[0] = bg.x; fc[1] = bg.y; fc[2] = bg.z;
    }
    if (rd.y < T(0)) {  // dot(rayDirection, vec3(0, 1, 0)) < 0
        for (int k = 0; k < 4; k++) out[k] = fc[k];
        if (cnt) { cnt[0] = 0; cnt[1] = 0; }
        if (state) { state[0] = T(0); state[1] = T(1); }
        return;
    }

    V3<T> earthCenter = cameraPos;
    earthCenter.y = T(P->center_y);
    const T thickness = T(P->thickness);
    T tStart, tEnd;
    if (P->flags & MM_MARCH_ISECT_FP64) {
        V3<double> ro = v3<double>(cameraPos.x, cameraPos.y, cameraPos.z);
        V3<double> rdd = v3<double>(rd.x, rd.y, rd.z);
        V3<double> c = v3<double>(earthCenter.x, earthCenter.y, earthCenter.z);
        Isect<double> in = raySphere<double>(ro, rdd, c, (double)P->inner_diameter);
        Isect<double> ou = raySphere<double>(ro, rdd, c, (double)P->outer_diameter);
        tStart = in.valid ? T(in.t) : T(0);
        tEnd = ou.valid ? T(ou.t) : T(0);
    } else {
        Isect<T> in = raySphere<T>(cameraPos, rd, earthCenter, T(P->inner_diameter));
        Isect<T> ou = raySphere<T>(cameraPos, rd, earthCenter, T(P->outer_diameter));
        tStart = in.valid ? in.t : T(0);
        tEnd = ou.valid ? ou.t : T(0);
    }

    if (night) {  // comp:365-384
        V3<T> axisv = normalize(v3<T>(1, 0, 1));
        V3<T> rot[3];
        fromAngleAxis<T>(axisv, T(s.sun->direction[1]) * T(0.5), rot);
        V3<T> rrd = matmul<T>(rot, rd);
        V3<T> rro = matmul<T>(rot, cameraPos);
        V3<T> point = (rrd * tEnd) + rro;
        V3<T> proj = projectedShellPoint<T>(s, point, earthCenter);
        T u = T(0.00002) * (proj.x - cameraPos.x) + T(0.35);
        T v = T(0.00002) * (proj.z - cameraPos.z) + T(0.35);
        bg = nightSky<T>(s, u, v);
        bg = v3<T>(bg.x * (sqrt(bg.x) * T(0.75)), bg.y * (sqrt(bg.y) * T(0.75)), bg.z * (sqrt(bg.z) * T(0.75)));
        bg = v3<T>(powT<T>(bg.x, T(2.2)), powT<T>(bg.y, T(2.2)), powT<T>(bg.z, T(2.2)));
        bg = bg * T(10);
        T falloff = powT<T>(rd.y, T(6));
        bg = bg * falloff;
        bg = mix3(v3<T>(T(0.3), T(0.6), T(4.0)) * T(0.05), bg, powT<T>(rd.y, T(0.03125)));
        bg = bg + v3<T>(sunDisk, sunDisk, sunDisk);
        fc[3] = sunDisk;
    }

    T cosTheta = dot(rd, sunDir);
    T accum = T(0), transmittance = T(1);
    T stepSize = T(P->coarse_step_frac) * thickness;
    const float* DB = s.sun->directionBasis;
    V3<T> b0 = ubo3<T>(DB), b1 = ubo3<T>(DB + 4), b2 = ubo3<T>(DB + 8);
    const T coneTab[6][3] = {{0, T(0.6), 0}, {0, T(0.5), T(0.05)}, {T(0.1), T(0.75), 0},
                             {T(0.2), T(2.5), T(0.3)}, {0, 6, 0}, {T(-0.1), 1, T(-0.2)}};
    V3<T> samples[6];
    for (int i = 0; i < 6; i++) samples[i] = ((b0 * coneTab[i][0]) + (b1 * coneTab[i][1])) + (b2 * coneTab[i][2]);

    bool noHits = true;
    int misses = 0, steps = 0, modeCount[2] = {0, 0};
    const bool bounded = (P->flags & MM_MARCH_BOUNDED) != 0;
    const int modeLimit[2] = {P->max_coarse_steps, P->max_fine_steps};
    const V3<T> wind = ubo3<T>(s.sky->wind);
    const T ws = T(P->wind_strength);
    const V3<T> windH = v3<T>(T(0.1), T(0.05), T(0));
    const T HG = fmaxT(hgPhase<T>(cosTheta, T(0.6)), T(0.7) * hgPhase<T>(cosTheta, T(0.99) - T(0.1)));

    int iterations = 0;
    for (T t = tStart; t < tEnd; t += stepSize) {
        if (++iterations > MM_MAX_MARCH_ITERATIONS) break;  // same hard bound as the GPU kernel
        const int mode = noHits ? 0 : 1;
        if (bounded && modeLimit[mode] >= 0 && ++modeCount[mode] > modeLimit[mode]) break;
        V3<T> currentPos = cameraPos + (t * rd);
        V3<T> currentProj = projectedShellPoint<T>(s, currentPos, earthCenter);
        T rHeight = relativeHeight<T>(currentPos, currentProj, thickness);
        V3<T> windOffset = (ws * (wind + (rHeight * windH))) * (timeOffset + rHeight * T(200));
        T density = cloudTest<T>(s, currentPos + windOffset, rHeight, earthCenter, n);
        T loDensity = density;
        if (density > T(0)) {
            misses = 0;
            if (noHits) {
                trace_trip(tr, t, stepSize, density, TRIP_FIRST_HIT);
                t -= stepSize;
                stepSize *= T(P->fine_step_mul);
                noHits = false;
                continue;
            }
            const T tripStep = stepSize;
            density = cloudHiRes<T>(s, currentPos + windOffset, stepSize, density, rHeight, n);
            if (density < T(0.0001)) {
                trace_trip(tr, t, tripStep, loDensity, TRIP_FINE_THIN);
                continue;
            }
            trace_trip(tr, t, tripStep, loDensity, TRIP_FINE_HIT);
            T dal = T(0);
            for (int i = 0; i < 6; i++) {
                V3<T> lsPos = currentPos + ((T(3) * stepSize) * samples[i]);
                V3<T> lsProj = projectedShellPoint<T>(s, lsPos, earthCenter);
                T lsHeight = relativeHeight<T>(lsPos, lsProj, thickness);
                windOffset = (ws * (wind + (lsHeight * windH))) * (timeOffset + lsHeight * T(200));
                T lsDensity = cloudTest<T>(s, lsPos + windOffset, lsHeight, earthCenter, n);
                if (lsDensity > T(0)) {
                    lsDensity = cloudHiRes<T>(s, lsPos + windOffset, stepSize, lsDensity, lsHeight, n);
                    dal += lsDensity;
                }
            }
            T beers = exp(-dal);
            T beersMod = fmaxT(beers, T(0.7) * exp(T(-0.25) * dal));
            beers = mixT(beers, beersMod, (-cosTheta) * T(0.5) + T(0.5));
            T inScatter = T(0.09) + powT<T>(loDensity, remapClampedK<T>(rHeight, T(0.3), T(0.85), T(0.5), T(2)));
            inScatter *= powT<T>(remapClampedK<T>(rHeight, T(0.07), T(0.34), T(0.1), T(1)), T(0.8));
            transmittance = mixT(transmittance, (inScatter * HG) * beers, T(1) - accum);
            accum += density;
        } else if (!noHits) {
            misses++;
            trace_trip(tr, t, stepSize, density, misses >= P->miss_limit ? TRIP_REVERT : TRIP_FINE_MISS);
            if (misses >= P->miss_limit) {
                noHits = true;
                stepSize /= T(P->fine_step_mul);
            }
        } else {
            trace_trip(tr, t, stepSize, density, TRIP_COARSE_MISS);
        }
        if (accum > T(P->opacity_cutoff)) {
            accum = T(1);
            break;
        }
        if (P->max_steps >= 0 && ++steps > P->max_steps) break;
        if (!bounded && modeLimit[mode] >= 0 && ++modeCount[mode] > modeLimit[mode]) break;
    }

    accum *= smoothstepT<T>(T(0), T(1), fminT(T(1), remapK<T>(rd.y, T(0), T(0.1), T(0), T(1))));
    accum = fminT(accum, T(0.999));
    if (state) { state[0] = accum; state[1] = transmittance; }
    const V3<T> sunCol = ubo3<T>(s.sun->color);
    const T I = T(s.sun->intensity);
    V3<T> cloud;
    if (!night) {
        T e = exp(-transmittance);
        V3<T> amb = (bg * T(0.08)) * e;
        V3<T> direct = v3<T>(I, I, I) * fmaxT(T(0), transmittance);
        cloud = sunCol * (direct + amb);
    } else {
        T e = exp(-transmittance);
        V3<T> nightAmb = (v3<T>(T(0.3), T(0.6), T(4.0)) * T(0.05)) * powT<T>(rd.y, T(0.03125));
        V3<T> amb = (nightAmb * T(0.08)) * e;
        V3<T> direct = v3<T>(I, I, I) * fmaxT(T(0), transmittance);
        cloud = sunCol * (direct + amb);
    }
    V3<T> rgb = mix3(bg, cloud, accum);
    out[0] = rgb.x; out[1] = rgb.y; out[2] = rgb.z;
    out[3] = fc[3] * fmaxT(T(1) - accum, T(0));
    if (cnt) { cnt[0] = n.density; cnt[1] = n.detail; }
}

template <class T>
int validate(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky, const orc_textures* tex,
             const mm_march_params* prm, int width, int height) {
    if (!cam || !sun || !sky || !tex || !prm) return MM_E_NULL;
    if (!tex->weather || !tex->curl || !tex->low || !tex->hi) return MM_E_NULL;
    if (width <= 0 || height <= 0) return MM_E_DIM;
    return MM_OK;
}

template <class T>
int march_rows(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky, const orc_textures* tex,
               const mm_march_params* prm, int width, int height, int phase, int y0, int y1, T* out,
               int32_t* counts, int threads, T* state = nullptr) {
    int e = validate<T>(cam, sun, sky, tex, prm, width, height);
    if (e) return e;
    if (phase < -1 || phase > 15) return MM_E_PHASE;
    Scene<T> s{cam, sun, sky, tex, prm, width, height};
    if (threads > 0) omp_set_num_threads(threads);
    if (y0 < 0) y0 = 0;
    if (y1 > height) y1 = height;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = y0; y < y1; y++) {
        if (phase >= 0 && (y & 3) != phase / 4) continue;
        for (int x = 0; x < width; x++) {
            if (phase >= 0 && (x & 3) != phase % 4) continue;
            size_t i = (size_t)y * width + x;
            int c[2];
            marchPixel<T>(s, x, y, out + 4 * i, c, nullptr, state ? state + 2 * i : nullptr);
            if (counts) { counts[2 * i] = c[0]; counts[2 * i + 1] = c[1]; }
        }
    }
    return MM_OK;
}

template <class T>
int march_list(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky, const orc_textures* tex,
               const mm_march_params* prm, int width, int height, const int32_t* xy, int n, T* out, int threads,
               T* state = nullptr) {
    int e = validate<T>(cam, sun, sky, tex, prm, width, height);
    if (e) return e;
    Scene<T> s{cam, sun, sky, tex, prm, width, height};
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++)
        marchPixel<T>(s, xy[2 * i], xy[2 * i + 1], out + 4 * (size_t)i, nullptr, nullptr,
                      state ? state + 2 * (size_t)i : nullptr);
    return MM_OK;
}

}  // namespace

extern "C" {

void orc_sample2d(const uint8_t* tex, int w, int h, float u, float v, int bits, float out[4]) {
    sample2d<float>(tex, w, h, u, v, bits, out);
}
void orc_sample3d(const uint8_t* tex, int n, float u, float v, float w, int bits, float out[4]) {
    sample3d<float>(tex, n, u, v, w, bits, out);
}

int orc_march_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky, const orc_textures* tex,
                  const mm_march_params* prm, int width, int height, int phase, int y0, int y1, float* out,
                  int32_t* counts, int threads) {
    return march_rows<float>(cam, sun, sky, tex, prm, width, height, phase, y0, y1, out, counts, threads);
}
int orc_march_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky, const orc_textures* tex,
                  const mm_march_params* prm, int width, int height, int phase, int y0, int y1, double* out,
                  int32_t* counts, int threads) {
    return march_rows<double>(cam, sun, sky, tex, prm, width, height, phase, y0, y1, out, counts, threads);
}
int orc_march_pixels_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                         const orc_textures* tex, const mm_march_params* prm, int width, int height,
                         const int32_t* xy, int n, float* out, int threads) {
    return march_list<float>(cam, sun, sky, tex, prm, width, height, xy, n, out, threads);
}
int orc_march_pixels_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                         const orc_textures* tex, const mm_march_params* prm, int width, int height,
                         const int32_t* xy, int n, double* out, int threads) {
    return march_list<double>(cam, sun, sky, tex, prm, width, height, xy, n, out, threads);
}
int orc_march_state_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                        const orc_textures* tex, const mm_march_params* prm, int width, int height, float* out,
                        float* state, int threads) {
    if (!state) return MM_E_NULL;
    return march_rows<float>(cam, sun, sky, tex, prm, width, height, -1, 0, height, out, nullptr, threads, state);
}
int orc_march_state_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                        const orc_textures* tex, const mm_march_params* prm, int width, int height, double* out,
                        double* state, int threads) {
    if (!state) return MM_E_NULL;
    return march_rows<double>(cam, sun, sky, tex, prm, width, height, -1, 0, height, out, nullptr, threads, state);
}
int orc_march_pixels_state_f64(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                               const orc_textures* tex, const mm_march_params* prm, int width, int height,
                               const int32_t* xy, int n, double* out, double* state, int threads) {
    if (!state) return MM_E_NULL;
    return march_list<double>(cam, sun, sky, tex, prm, width, height, xy, n, out, threads, state);
}
int orc_max_threads(void) { return omp_get_max_threads(); }

int orc_march_trace_f32(const mm_camera_ubo* cam, const mm_sun_ubo* sun, const mm_sky_ubo* sky,
                        const orc_textures* tex, const mm_march_params* prm, int width, int height, int px, int py,
                        float* trips, int max_trips, int32_t* counts, float* out) {
    int e = validate<float>(cam, sun, sky, tex, prm, width, height);
    if (e) return e;
    if (!trips || !counts || !out || max_trips < 0) return MM_E_NULL;
    Scene<float> s{cam, sun, sky, tex, prm, width, height};
    Trace tr{trips, max_trips, 0};
    int c[2];
    marchPixel<float>(s, px, py, out, c, &tr);
    counts[0] = c[0];
    counts[1] = c[1];
    return tr.n;
}

}  // extern "C"

// Single-function entry points, so tests can pin the pieces of the restatement against closed forms.
extern "C" {
float orc_hg_phase(float cosTheta, float g) { return hgPhase<float>(cosTheta, g); }
double orc_hg_phase_f64(double cosTheta, double g) { return hgPhase<double>(cosTheta, g); }
float orc_rayleigh_phase(float c) { return rayleighPhase<float>(c); }
float orc_layer_density(float h, float type) { return cloudLayerDensity<float>(h, type); }
float orc_height_bias_coverage(float coverage, float height) { return heightBiasCoverage<float>(coverage, height); }
float orc_remap_clamped(float v, float a, float b, float c, float d) { return remapClamped<float>(v, a, b, c, d); }
float orc_pow_rt(float x, float y) { return pow_rt(x, y); }
float orc_ray_sphere_f32(const float ro[3], const float rd[3], const float c[3], float w, int* valid) {
    Isect<float> r = raySphere<float>(v3(ro[0], ro[1], ro[2]), v3(rd[0], rd[1], rd[2]), v3(c[0], c[1], c[2]), w);
    *valid = r.valid ? 1 : 0;
    return r.t;
}
double orc_ray_sphere_f64(const double ro[3], const double rd[3], const double c[3], double w, int* valid) {
    Isect<double> r = raySphere<double>(v3(ro[0], ro[1], ro[2]), v3(rd[0], rd[1], rd[2]), v3(c[0], c[1], c[2]), w);
    *valid = r.valid ? 1 : 0;
    return r.t;
}
}
