// Cloud raymarch kernel for gfx950 (CDNA4), the HIP counterpart of the compute dispatch of
// SkyEngine/SkyEngine/Shaders/compute-clouds.comp:288-500.
//
// One work-item per marched pixel, 64-lane wavefronts covering 64-pixel tiles (8x8, 16x4 or 32x2;
// rays of a wave are spatially coherent, so their step counts and texel footprints are close). Uniform-derived
// values are hoisted to the host (MarchConsts). Texture filtering is done in software: LINEAR +
// REPEAT with the texture unit's fixed-point sub-texel weights (8 bits) or exact float weights,
// over one of three texel format sets:
//   FSET_REF : every texture RGBA8 UNORM, linear layout (the reference's bindings, drop-in)
//   FSET_FAST: weather RG8 (the two channels the march reads), curl / low-res / hi-res RGBA8, apron
//              layout (N+1 texels per axis) so a filter footprint never wraps; the same arithmetic
//              as FSET_REF, so the same image
//   FSET_PREC: weather RG8, curl FP16x4, low-res FP8 e4m3 x4 with per-channel scale, hi-res FP16x4
// Built with -ffp-contract=off and IEEE division/sqrt. Every operation that decides a ray's step
// sequence (positions, heights, filters, layer profile, coverage pow, erosion and density remaps)
// rounds as the oracle's restatement does, so the density samples and sample counts are the
// restatement's bit for bit; the lighting's exp / pow use the hardware exp2 / log2 (GLSL precision).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include "mm_internal.hpp"
#include "surrogate_dev.hpp"
#include "cr_math.hpp"

namespace {

// The FSET template argument of the kernels: a format set of mm.h (FSET_REF, FSET_FAST, FSET_PREC), or'ed
// with kFloatWeights for the float sampler (subtexel_bits 0). Without the flag the sub-texel weights
// are rounded to 1/256 like the hardware sampler's, and 8-bit channels are filtered in integers.
constexpr int kFloatWeights = 8;
// Step-budget flag of the template argument: the launch's params are bounded (MM_MARCH_BOUNDED) with
// no counted-step limit (max_steps < 0) and both coarse and fine budgets set (>= 0), as in the
// scenario preset. The loop's step checks then compile to the two budget tests alone; the other
// cases run the general checks. Fewer live uniform values leave the kernels without SGPR spills.
constexpr int kBoundedSteps = 16;
template <int F> constexpr int kSet = F & 7;
template <int F> constexpr int kQ8 = (F & kFloatWeights) ? 0 : 1;
template <int F> constexpr bool kBS = (F & kBoundedSteps) != 0;

using mm::MarchConsts;
using mm::MarchTextures;
using mm::TexView;
using mm::ViewDesc;

// Work counters of an instrumented launch, per lane: density samples (cloudTest evaluations), those
// that read the weather map, and those that read the low-res volume, in three 21-bit fields of one
// 64-bit value (a lane takes far fewer than 2^21 samples: at most 16384 loop trips plus its light
// cones).
typedef unsigned long long Tally;
constexpr Tally kTallyWeather = 1ull << 21, kTallyLow = 1ull << 42, kTallyMask = (1ull << 21) - 1;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) {
    V3 r;
    r.x = x; r.y = y; r.z = z;
    return r;
}
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 mul3(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
__device__ __forceinline__ V3 ld3(const float* f) { return mk(f[0], f[1], f[2]); }
__device__ __forceinline__ float sqrt_(float x) { return sqrt_rn(x); }
// normalize with a correctly rounded 1 / sqrt: getProjectedShellPoint scales the unit vector by the
// shell radius (6.4e6 m in the Earth-radius scenario), so one ulp of the reciprocal moves the
// projected point by ~0.5 m = 2e-4 of the shell height and flips density thresholds downstream.
// Every normalised or measured vector here is a ray direction or a difference of positions at metre
// scale (squared lengths 0 or far above 2^-96, lengths within [2^-125, 2^125]), so the unscaled forms
// of the correctly rounded square root and reciprocal apply.
__device__ __forceinline__ V3 normalize3(V3 a) { return a * rcp_rn(sqrt_rn_normal(dot3(a, a))); }
__device__ __forceinline__ float length3(V3 a) { return sqrt_rn_normal(dot3(a, a)); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
// Filter lerps: one fused multiply-add, a + t (b - a) rounded once. With 8-bit sub-texel weights every
// bilinear step of UNORM8 texels is exact, so a trilinear filter returns the exact weighted sum rounded
// once.
__device__ __forceinline__ float lerpf(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }
__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float exp2_(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float log2_(float x) { return __builtin_amdgcn_logf(x); }
// GLSL-grade pow/exp: exp2(y * log2(x)), exp2(x * log2(e)), with the hardware exp2 / log2.
__device__ __forceinline__ float powg(float x, float y) { return exp2_(y * log2_(x)); }
__device__ __forceinline__ float expg(float x) { return exp2_(x * 1.4426950408889634f); }

// pow of the coverage exponent (comp:206-208), written out so the scalar restatement rounds alike
// (the hardware log2 / exp2 have no published rounding): x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log2(m) = f P(f) with f = m - 1 (degree 8, |error| < 3e-8), z = y log2(x), exp2(z) = 2^n Q(z - n)
// with n = rint(z) (degree 6, relative error < 2e-8); GLSL bounds pow by exp2(y log2 x), whose log2
// may be off by 2^-21 on [0.5, 2]. pow(x, 1) = x; x <= 0 gives 0.
__device__ __forceinline__ float pow_rt(float x, float y) {
    if (y == 1.0f) return x;
    if (!(x > 0.0f)) return 0.0f;
    int e = 0;
    float m = __builtin_frexpf(x, &e);
    if (m < 0.70710678f) {
        m = m + m;
        e = e - 1;
    }
    const float f = m - 1.0f;
    float p = 1.231096834e-01f;
    p = __builtin_fmaf(p, f, -2.058618814e-01f);
    p = __builtin_fmaf(p, f, 2.160782218e-01f);
    p = __builtin_fmaf(p, f, -2.391698807e-01f);
    p = __builtin_fmaf(p, f, 2.879032493e-01f);
    p = __builtin_fmaf(p, f, -3.606932759e-01f);
    p = __builtin_fmaf(p, f, 4.809107482e-01f);
    p = __builtin_fmaf(p, f, -7.213474512e-01f);
    p = __builtin_fmaf(p, f, 1.442695022e+00f);
    const float l = __builtin_fmaf(f, p, (float)e);
    const float z = y * l;
    const float n = __builtin_rintf(z);
    const float r = z - n;
    float q = 1.546144485e-04f;
    q = __builtin_fmaf(q, r, 1.340042800e-03f);
    q = __builtin_fmaf(q, r, 9.618056938e-03f);
    q = __builtin_fmaf(q, r, 5.550327152e-02f);
    q = __builtin_fmaf(q, r, 2.402265072e-01f);
    q = __builtin_fmaf(q, r, 6.931471825e-01f);
    q = __builtin_fmaf(q, r, 1.000000000e+00f);
    return __builtin_ldexpf(q, (int)n);
}

// remap with a constant range: (v - a) * (1 / (b - a)) * (d - c) + c
__device__ __forceinline__ float remapk(float v, float a, float inv_ba, float c, float dc) {
    return c + (((v - a) * inv_ba) * dc);
}
// remapClamped(v, a, 1, 0, 1) for a in [0, 1]. The IEEE quotient is only needed when v > a: otherwise
// (v - a) / (1 - a) <= 0 (or 0/0 when v = a = 1) and the clamp yields 0 exactly, so the division
// sits behind a branch that whole waves of empty samples skip.
__device__ __forceinline__ float remapClamped01(float v, float a) {
    if (!(v > a)) return 0.0f;
    // v > a: the quotient is positive, so the shader's 0 + (...) * 1 leaves it unchanged
    return clampf(__fdiv_rn(v - a, 1.0f - a), 0.0f, 1.0f);
}

// ------------------------------------------------------------------ sampler
// x = u * n - 0.5, optionally rounded to 1/256 (fixed-point sub-texel weights), i0 = floor(x) mod n
// (n a power of two: two's-complement AND is the floor modulo), f = x - floor(x).
// u * n is exact (n a power of two), so fma(u, n, -0.5) rounds like u * n - 0.5. With the 1/256 rounding,
// k = floor(x * 256 + 0.5) (x * 256 exact) gives floor(k / 256) = k >> 8 and the fraction
// (k & 255) / 256 exactly as the float steps do, for texel coordinates below 2^23 in magnitude
// (8.4 M texels; the presets stay below 1e5). With the rounding, w = (256 - fk) | fk << 16 holds the
// integer weights of the two texels (fk = k & 255) for the integer filter below. k comes from one
// v_cvt_flr_i32_f32 (floor and conversion), i0 from one v_bfe_u32 (bits 8 .. 8 + log2 n of k, the
// same as (k >> 8) & (n - 1) in two's complement); lg = log2 n.
__device__ __forceinline__ void axis_p2(float u, int n, int lg, int q8, int& i0, float& f, uint32_t& w) {
    const float x = __builtin_fmaf(u, (float)n, -0.5f);
    if (q8) {
        int k;
        asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(k) : "v"(__builtin_fmaf(x, 256.0f, 0.5f)));
        i0 = (int)__builtin_amdgcn_ubfe((uint32_t)k, 8u, (uint32_t)lg);
        const uint32_t fk = (uint32_t)k & 255u;
        f = (float)fk * (1.0f / 256.0f);
        w = __umul24(fk, 65535u) + 256u;  // (256 - fk) | fk << 16
    } else {
        const float fl = floorf(x);
        f = x - fl;
        i0 = ((int)fl) & (n - 1);
        w = 0u;
    }
}

__device__ __forceinline__ uint32_t ld_u32(const uint8_t* p) { return *reinterpret_cast<const uint32_t*>(p); }
__device__ __forceinline__ uint32_t ld_u16(const uint8_t* p) { return *reinterpret_cast<const uint16_t*>(p); }
__device__ __forceinline__ uint2 ld_u64(const uint8_t* p) { return *reinterpret_cast<const uint2*>(p); }

template <int BYTES> struct TexelT;
template <> struct TexelT<2> { typedef uint32_t type; };
template <> struct TexelT<4> { typedef uint32_t type; };
template <> struct TexelT<8> { typedef uint2 type; };

template <int BYTES>
__device__ __forceinline__ typename TexelT<BYTES>::type ld_texel(const uint8_t* p) {
    if constexpr (BYTES == 2) return ld_u16(p);
    else if constexpr (BYTES == 4) return ld_u32(p);
    else return ld_u64(p);
}

// Two x-adjacent texels in one load (apron layouts: texel x0 + 1 is always stored after x0). The
// pair is only aligned to the texel size; gfx950 serves such under-aligned global loads.
typedef uint32_t __attribute__((aligned(2))) u32_a2;
struct alignas(4) u64_a4 { uint32_t x, y; };
struct alignas(8) u128_a8 { uint32_t x, y, z, w; };
template <int BYTES>
__device__ __forceinline__ void ld_pair(const uint8_t* p, typename TexelT<BYTES>::type& a,
                                        typename TexelT<BYTES>::type& b) {
    if constexpr (BYTES == 2) {
        // a keeps the whole word (texel x0 in bytes 0-1, x0 + 1 in bytes 2-3): the byte decoders only
        // read bytes 0-1 of it, and the integer filter takes both texels' channels from it directly
        const uint32_t v = *reinterpret_cast<const u32_a2*>(p);
        a = v;
        b = v >> 16;
    } else if constexpr (BYTES == 4) {
        const u64_a4 v = *reinterpret_cast<const u64_a4*>(p);
        a = v.x;
        b = v.y;
    } else {
        const u128_a8 v = *reinterpret_cast<const u128_a8*>(p);
        a = make_uint2(v.x, v.y);
        b = make_uint2(v.z, v.w);
    }
}

// A gathered footprint: texels, float fractions and, with 1/256 weights (q8), the integer weight pairs.
template <int BYTES> struct G2 {
    typename TexelT<BYTES>::type t[4];
    float fx, fy;
    uint32_t wx, wy;
    int q8;
};
template <int BYTES> struct G3 {
    typename TexelT<BYTES>::type t[8];
    float fx, fy, fz;
    uint32_t wx, wy, wz;
    int q8;
};

// Gather of the 2x2 bilinear footprint. APRON: the texture stores n+1 texels per axis with texel n
// equal to texel 0, so the +1 neighbour never wraps.
template <int BYTES, bool APRON>
__device__ __forceinline__ G2<BYTES> gather2(const TexView& tv, float u, float v, int q8) {
    G2<BYTES> g;
    g.q8 = q8;
    int x0, y0;
    axis_p2(u, tv.w, tv.lw, q8, x0, g.fx, g.wx);
    axis_p2(v, tv.h, tv.lh, q8, y0, g.fy, g.wy);
    int x1, y1;
    if (APRON) { x1 = x0 + 1; y1 = y0 + 1; }
    else { x1 = (x0 + 1) & (tv.w - 1); y1 = (y0 + 1) & (tv.h - 1); }
    // 32-bit byte offsets from the uniform base (textures are below 4 GiB, checked at the boundary):
    // the loads take the scalar-base + vector-offset form, with no 64-bit address arithmetic
    // (row and slice products fit 24-bit operands: dims <= 8192, slices <= 1025^2 texels)
    const uint32_t r0 = __umul24((uint32_t)y0, (uint32_t)tv.stride), r1 = __umul24((uint32_t)y1, (uint32_t)tv.stride);
    if constexpr (APRON) {
        ld_pair<BYTES>(tv.data + (r0 + (uint32_t)x0) * BYTES, g.t[0], g.t[1]);
        ld_pair<BYTES>(tv.data + (r1 + (uint32_t)x0) * BYTES, g.t[2], g.t[3]);
    } else {
        g.t[0] = ld_texel<BYTES>(tv.data + (r0 + (uint32_t)x0) * BYTES);
        g.t[1] = ld_texel<BYTES>(tv.data + (r0 + (uint32_t)x1) * BYTES);
        g.t[2] = ld_texel<BYTES>(tv.data + (r1 + (uint32_t)x0) * BYTES);
        g.t[3] = ld_texel<BYTES>(tv.data + (r1 + (uint32_t)x1) * BYTES);
    }
    return g;
}

template <int BYTES, bool APRON>
__device__ __forceinline__ G3<BYTES> gather3(const TexView& tv, float u, float v, float w, int q8) {
    G3<BYTES> g;
    g.q8 = q8;
    int x0, y0, z0;
    const int n = tv.w;
    axis_p2(u, n, tv.lw, q8, x0, g.fx, g.wx);
    axis_p2(v, n, tv.lw, q8, y0, g.fy, g.wy);
    axis_p2(w, n, tv.lw, q8, z0, g.fz, g.wz);

    int x1, y1, z1;
    if (APRON) { x1 = x0 + 1; y1 = y0 + 1; z1 = z0 + 1; }
    else { x1 = (x0 + 1) & (n - 1); y1 = (y0 + 1) & (n - 1); z1 = (z0 + 1) & (n - 1); }
    const uint8_t* b = tv.data;
    const uint32_t s0 = __umul24((uint32_t)z0, (uint32_t)tv.sliceStride), s1 = __umul24((uint32_t)z1, (uint32_t)tv.sliceStride);
    const uint32_t q0 = __umul24((uint32_t)y0, (uint32_t)tv.stride), q1 = __umul24((uint32_t)y1, (uint32_t)tv.stride);
    const uint32_t r00 = s0 + q0, r10 = s0 + q1, r01 = s1 + q0, r11 = s1 + q1;
    const uint32_t ux0 = (uint32_t)x0, ux1 = (uint32_t)x1;
    if constexpr (APRON) {
        ld_pair<BYTES>(b + (r00 + ux0) * BYTES, g.t[0], g.t[1]);
        ld_pair<BYTES>(b + (r10 + ux0) * BYTES, g.t[2], g.t[3]);
        ld_pair<BYTES>(b + (r01 + ux0) * BYTES, g.t[4], g.t[5]);
        ld_pair<BYTES>(b + (r11 + ux0) * BYTES, g.t[6], g.t[7]);
    } else {
        g.t[0] = ld_texel<BYTES>(b + (r00 + ux0) * BYTES);
        g.t[1] = ld_texel<BYTES>(b + (r00 + ux1) * BYTES);
        g.t[2] = ld_texel<BYTES>(b + (r10 + ux0) * BYTES);
        g.t[3] = ld_texel<BYTES>(b + (r10 + ux1) * BYTES);
        g.t[4] = ld_texel<BYTES>(b + (r01 + ux0) * BYTES);
        g.t[5] = ld_texel<BYTES>(b + (r01 + ux1) * BYTES);
        g.t[6] = ld_texel<BYTES>(b + (r11 + ux0) * BYTES);
        g.t[7] = ld_texel<BYTES>(b + (r11 + ux1) * BYTES);
    }
    return g;
}

template <int BYTES, class F>
__device__ __forceinline__ float filt2(const G2<BYTES>& g, F ch) {
    float a = lerpf(ch(g.t[0]), ch(g.t[1]), g.fx);
    float b = lerpf(ch(g.t[2]), ch(g.t[3]), g.fx);
    return lerpf(a, b, g.fy);
}
template <int BYTES, class F>
__device__ __forceinline__ float filt3(const G3<BYTES>& g, F ch) {
    float a0 = lerpf(ch(g.t[0]), ch(g.t[1]), g.fx);
    float b0 = lerpf(ch(g.t[2]), ch(g.t[3]), g.fx);
    float a1 = lerpf(ch(g.t[4]), ch(g.t[5]), g.fx);
    float b1 = lerpf(ch(g.t[6]), ch(g.t[7]), g.fx);
    float s0 = lerpf(a0, b0, g.fy);
    float s1 = lerpf(a1, b1, g.fy);
    return lerpf(s0, s1, g.fz);
}

// Integer filtering of an 8-bit channel with 1/256 weights. With t in 0..255 and f = fk / 256, the float
// steps above are exact through two levels -- lerp(a, b, f) = (a (256 - fk) + b fk) / 256 needs 16 bits,
// the second level 24 -- and the third rounds once (one FMA on exact operands): filt2 = M / 2^16 and
// filt3 = RN(N / 2^24) for the integer sums M < 2^24, N < 2^32. The same sums in integer arithmetic --
// byte pairs by v_perm_b32, each level a v_dot2_u32_u16 against the weight pair (256 - fk, fk), the last
// level of filt3 in 24-bit multiplies -- then one conversion (RN(N) 2^-24 = RN(N / 2^24)) give the same
// bits in 16 instead of 23 operations per channel.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t dot2u(uint32_t a, uint32_t w, uint32_t c) {
    return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, w), c, false);
}
// (byte K of a, byte K of b) as a 16-bit pair (v_perm_b32 selector 12 gives a zero byte)
template <int K>
__device__ __forceinline__ uint32_t pair8(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_perm(b, a, (0x0cu << 24) | ((4u + (uint32_t)K) << 16) | (0x0cu << 8) | (uint32_t)K);
}
// (low half of a, low half of b) as a 16-bit pair
__device__ __forceinline__ uint32_t pair16(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x05040100u); }

// (byte K, byte 2 + K) of one word as a 16-bit pair: channel K of the two 2-byte texels of an apron row pair
template <int K>
__device__ __forceinline__ uint32_t pair8w(uint32_t w) {
    return __builtin_amdgcn_perm(0u, w, (0x0cu << 24) | ((2u + (uint32_t)K) << 16) | (0x0cu << 8) | (uint32_t)K);
}

template <int K, int BYTES>
__device__ __forceinline__ uint32_t isum2(const G2<BYTES>& g) {
    uint32_t l0, l1;
    if constexpr (BYTES == 2) {  // apron gather: t[0] and t[2] are whole row-pair words (ld_pair)
        l0 = dot2u(pair8w<K>(g.t[0]), g.wx, 0u);
        l1 = dot2u(pair8w<K>(g.t[2]), g.wx, 0u);
    } else {
        l0 = dot2u(pair8<K>(g.t[0], g.t[1]), g.wx, 0u);
        l1 = dot2u(pair8<K>(g.t[2], g.t[3]), g.wx, 0u);
    }
    return dot2u(pair16(l0, l1), g.wy, 0u);
}
template <int K, int BYTES>
__device__ __forceinline__ uint32_t isum3(const G3<BYTES>& g) {
    const uint32_t a0 = dot2u(pair8<K>(g.t[0], g.t[1]), g.wx, 0u);
    const uint32_t b0 = dot2u(pair8<K>(g.t[2], g.t[3]), g.wx, 0u);
    const uint32_t a1 = dot2u(pair8<K>(g.t[4], g.t[5]), g.wx, 0u);
    const uint32_t b1 = dot2u(pair8<K>(g.t[6], g.t[7]), g.wx, 0u);
    const uint32_t s0 = dot2u(pair16(a0, b0), g.wy, 0u);
    const uint32_t s1 = dot2u(pair16(a1, b1), g.wy, 0u);
    return __umul24(s0, g.wz & 0xffffu) + __umul24(s1, g.wz >> 16);
}

// Channel decoders.
template <int K> struct Byte {
    __device__ __forceinline__ float operator()(uint32_t v) const { return (float)((v >> (8 * K)) & 0xffu); }
};
template <int K> struct Half4 {  // channel K of 4 x FP16 packed in a uint2
    __device__ __forceinline__ float operator()(uint2 v) const {
        uint32_t w = (K < 2) ? v.x : v.y;
        uint16_t h = (uint16_t)((K & 1) ? (w >> 16) : (w & 0xffffu));
        return __half2float(__ushort_as_half(h));
    }
};
template <int K> struct Fp8 {  // channel K of 4 x FP8 e4m3 (OCP) packed in a uint32
    __device__ __forceinline__ float operator()(uint32_t v) const {
        return __builtin_amdgcn_cvt_f32_fp8((int)v, K);
    }
};

constexpr float kInv255 = 1.0f / 255.0f;

// Byte channel K filtered and scaled: filt(...) * s (s * 2^-16 and s * 2^-24 are exact scalings of s)
template <int K, int BYTES>
__device__ __forceinline__ float chan2(const G2<BYTES>& g, float s) {
    if (g.q8) return (float)isum2<K>(g) * (s * 0x1p-16f);
    return filt2(g, Byte<K>()) * s;
}
template <int K, int BYTES>
__device__ __forceinline__ float chan3(const G3<BYTES>& g, float s) {
    if (g.q8) return (float)isum3<K>(g) * (s * 0x1p-24f);
    return filt3(g, Byte<K>()) * s;
}

// ------------------------------------------------------------------ per-format texture reads
struct Weather {
    float coverage, type;
};

template <int FSET>
__device__ __forceinline__ Weather read_weather(const TexView& tv, float u, float v, int q8) {
    Weather w;
    if constexpr (kSet<FSET> == mm::FSET_REF) {
        G2<4> g = gather2<4, false>(tv, u, v, q8);
        if (q8) {
            const uint32_t c = isum2<0>(g), t = isum2<2>(g);
            w.coverage = (float)c * (kInv255 * 0x1p-16f);
            w.type = (float)t * (kInv255 * 0x1p-16f);
        } else {
            w.coverage = filt2(g, Byte<0>()) * kInv255;
            w.type = filt2(g, Byte<2>()) * kInv255;
        }
    } else {
        G2<2> g = gather2<2, true>(tv, u, v, q8);
        if (q8) {
            const uint32_t c = isum2<0>(g), t = isum2<1>(g);
            w.coverage = (float)c * (kInv255 * 0x1p-16f);
            w.type = (float)t * (kInv255 * 0x1p-16f);
        } else {
            w.coverage = filt2(g, Byte<0>()) * kInv255;
            w.type = filt2(g, Byte<1>()) * kInv255;
        }
    }
    return w;
}

// Twice the filtered curl texel (the 2 cv of comp:218's 2 cv - 1). Doubling commutes with rounding for
// these values (no overflow or underflow), so the 2 folds into the scale: 2 RN(x s) = RN(x (2 s)).
template <int FSET>
__device__ __forceinline__ V3 read_curl2(const TexView& tv, float u, float v, int q8) {
    if constexpr (kSet<FSET> == mm::FSET_PREC) {
        G2<8> g = gather2<8, true>(tv, u, v, q8);
        return mk(2.0f * filt2(g, Half4<0>()), 2.0f * filt2(g, Half4<1>()), 2.0f * filt2(g, Half4<2>()));
    } else {
        G2<4> g = gather2<4, kSet<FSET> != mm::FSET_REF>(tv, u, v, q8);
        constexpr float s16 = 2.0f * kInv255 * 0x1p-16f, s = 2.0f * kInv255;
        if (q8) {
            const uint32_t x = isum2<0>(g), y = isum2<1>(g), z = isum2<2>(g);
            return mk((float)x * s16, (float)y * s16, (float)z * s16);
        }
        return mk(filt2(g, Byte<0>()) * s, filt2(g, Byte<1>()) * s, filt2(g, Byte<2>()) * s);
    }
}

// Hi-res erosion 0.625 r + 0.25 g + 0.125 b of the filtered texel (comp:223-224).
template <int FSET>
__device__ __forceinline__ float read_hi_erosion(const TexView& tv, V3 p, int q8) {
    if constexpr (kSet<FSET> == mm::FSET_PREC) {
        G3<8> g = gather3<8, true>(tv, p.x, p.y, p.z, q8);
        float r = filt3(g, Half4<0>()), gg = filt3(g, Half4<1>()), b = filt3(g, Half4<2>());
        return ((0.625f * r) + (0.25f * gg)) + (0.125f * b);
    } else {
        G3<4> g = gather3<4, kSet<FSET> != mm::FSET_REF>(tv, p.x, p.y, p.z, q8);
        // the power-of-two weights fold into the channel scale: RN(x s) / 4 = RN(x (s / 4)) for these
        // normal-range values
        float r, gg4, b8;
        if (q8) {
            const uint32_t nr = isum3<0>(g), ng = isum3<1>(g), nb = isum3<2>(g);
            r = (float)nr * (kInv255 * 0x1p-24f);
            gg4 = (float)ng * (0.25f * kInv255 * 0x1p-24f);
            b8 = (float)nb * (0.125f * kInv255 * 0x1p-24f);
        } else {
            r = filt3(g, Byte<0>()) * kInv255;
            gg4 = filt3(g, Byte<1>()) * (0.25f * kInv255);
            b8 = filt3(g, Byte<2>()) * (0.125f * kInv255);
        }
        return ((0.625f * r) + gg4) + b8;
    }
}

// Low-res read in two halves: the base shape first (so the < 1e-4 early out of comp:243 skips the
// erosion decode), the erosion sum 0.625 g + 0.25 b + 0.125 a afterwards.
template <int FSET> struct LowG {
    G3<4> g;
};

template <int FSET>
__device__ __forceinline__ LowG<FSET> gather_low(const TexView& tv, V3 p, int q8) {
    LowG<FSET> r;
    r.g = gather3<4, kSet<FSET> != mm::FSET_REF>(tv, p.x, p.y, p.z, q8);
    return r;
}
template <int FSET>
__device__ __forceinline__ float low_shape(const LowG<FSET>& l, const TexView& tv) {
    if constexpr (kSet<FSET> == mm::FSET_PREC) return filt3(l.g, Fp8<0>()) * tv.scale[0];
    else return chan3<0>(l.g, kInv255);
}
template <int FSET>
__device__ __forceinline__ float low_erosion(const LowG<FSET>& l, const TexView& tv) {
    if constexpr (kSet<FSET> == mm::FSET_PREC) {
        float g = filt3(l.g, Fp8<1>()) * tv.scale[1], b = filt3(l.g, Fp8<2>()) * tv.scale[2],
              a = filt3(l.g, Fp8<3>()) * tv.scale[3];
        return ((0.625f * g) + (0.25f * b)) + (0.125f * a);
    } else {
        float g, b4, a8;  // power-of-two weights folded into the scale, as in read_hi_erosion
        if (l.g.q8) {
            const uint32_t ng = isum3<1>(l.g), nb = isum3<2>(l.g), na = isum3<3>(l.g);
            g = (float)ng * (kInv255 * 0x1p-24f);
            b4 = (float)nb * (0.25f * kInv255 * 0x1p-24f);
            a8 = (float)na * (0.125f * kInv255 * 0x1p-24f);
        } else {
            g = filt3(l.g, Byte<1>()) * kInv255;
            b4 = filt3(l.g, Byte<2>()) * (0.25f * kInv255);
            a8 = filt3(l.g, Byte<3>()) * (0.125f * kInv255);
        }
        return ((0.625f * g) + b4) + a8;
    }
}

// ------------------------------------------------------------------ shader functions
// getProjectedShellPoint (comp:180-182)
__device__ __forceinline__ V3 proj_shell(const MarchConsts& c, V3 pt, V3 center) {
    return (c.projRadius * normalize3(pt - center)) + center;
}
// getRelativeHeight (comp:186-188)
__device__ __forceinline__ float rel_height(const MarchConsts& c, V3 pt, V3 proj) {
    return clampf(length3(pt - proj) * c.invThickness, 0.0f, 1.0f);
}

// cloudLayerDensity (comp:193-204)
__device__ __forceinline__ float layer_density(float h, float type) {
    h = clampf(h, 0.0f, 1.0f);
    // remap(h, 0, b, 0, 1) = 0 + h (1 / b): the 0 + only turns a -0 into +0, and these factors meet
    // only products, max(0, .) and mixes whose zero signs never reach a nonzero profile (the caller
    // tests the profile against 0)
    const float r02 = h * (1.0f / 0.2f);
    float cumulus = fmaxf(0.0f, r02 * remapk(h, 0.7f, 1.0f / (0.9f - 0.7f), 1.0f, -1.0f));
    float strato = fmaxf(0.0f, r02 * remapk(h, 0.2f, 1.0f / (0.7f - 0.2f), 1.0f, -1.0f));
    float stratus = fmaxf(0.0f, (h * (1.0f / 0.1f)) * remapk(h, 0.2f, 1.0f / (0.3f - 0.2f), 1.0f, -1.0f));
    float d1 = mixf(stratus, strato, clampf(type * 2.0f, 0.0f, 1.0f));
    float d2 = mixf(strato, cumulus, clampf((type - 0.5f) * 2.0f, 0.0f, 1.0f));
    return mixf(d1, d2, type);
}

// wind offset (comp:414, 445)
__device__ __forceinline__ V3 wind_offset(const MarchConsts& c, float h) {
    // h is a clamped relative height in [0, 1], so h * 0 is a zero
    V3 w = mk(c.wind[0] + h * 0.1f, c.wind[1] + h * 0.05f, c.wind[2] + 0.0f);
    return (c.windStrength * w) * (c.time + h * 200.0f);
}

// The light-cone surrogate's density probes only feed the network (the recorder of its training
// rows takes the same probes), so they use the hardware square root, reciprocal square root,
// reciprocal, exp2 and log2 (about 1 ulp) without the rounding fix-ups of the march's own samples.
__device__ __forceinline__ V3 proj_shell_fast(const MarchConsts& c, V3 pt, V3 center) {
    const V3 d = pt - center;
    return (c.projRadius * (d * __builtin_amdgcn_rsqf(dot3(d, d)))) + center;
}
__device__ __forceinline__ float rel_height_fast(const MarchConsts& c, V3 pt, V3 proj) {
    const V3 d = pt - proj;
    return clampf(__builtin_amdgcn_sqrtf(dot3(d, d)) * c.invThickness, 0.0f, 1.0f);
}
__device__ __forceinline__ float remapClamped01_fast(float v, float a) {
    return v > a ? clampf((v - a) * __builtin_amdgcn_rcpf(1.0f - a), 0.0f, 1.0f) : 0.0f;
}

// wout (optional) receives the weather coverage and type the test read (left untouched when the
// height skips the read). EXACT = false: the surrogate probes' arithmetic (proj_shell_fast above).
template <int FSET, bool COUNT, bool EXACT = true>
__device__ __forceinline__ float cloud_test(const MarchConsts& c, const MarchTextures& tx, V3 pos, float h, V3 center,
                                            Tally& nd, Weather* wout = nullptr) {
    if (COUNT) nd++;
    // every layer profile is 0 at h = 0 and below 1e-7 for h >= 0.9 (the cumulus top; stratocumulus
    // and stratus end at 0.7 and 0.3), for any cloud type in [0, 1]: the density test below would
    // return 0 there whatever the textures hold, so the weather and low-res reads are skipped
    if (!(h > 0.0f && h < 0.9f)) return 0.0f;
    if (COUNT) nd += kTallyWeather;
    constexpr int q8 = kQ8<FSET>;
    V3 proj = EXACT ? proj_shell(c, pos, center) : proj_shell_fast(c, pos, center);
    Weather w = read_weather<FSET>(tx.weather, c.weatherScale * (proj.x - c.camPos[0]),
                                   c.weatherScale * (proj.z - c.camPos[2]), q8);
    if (wout) *wout = w;
    float layer = layer_density(h, w.type);
    // density = layer * remapClamped(R, ...) with the remap in [0, 1]: a zero layer density gives 0
    // whatever the low-res texel, so its gather is skipped (comp:240-243 evaluate to the same 0)
    if (layer == 0.0f) return 0.0f;
    if (COUNT) nd += kTallyLow;
    LowG<FSET> lg = gather_low<FSET>(tx.low, pos * c.lowScale, q8);
    float shape = low_shape<FSET>(lg, tx.low);
    // remapClamped(shape, 0.3, 1, 0, 1); its 0 + (.) only flips the sign of a zero, which the
    // < 0.0001 test below treats alike
    float density = layer * clampf((shape - 0.3f) * (1.0f / (1.0f - 0.3f)), 0.0f, 1.0f);
    if (density < 0.0001f) return 0.0f;
    // remapClamped(density, erosion, 1, 0, 1) with erosion in [0, 1] is exactly 1 for density > 1:
    // (density - erosion) and (1 - erosion) round monotonically, so their quotient is >= 1 (or +inf
    // for erosion = 1) and the clamp returns 1. The layer profile reaches 5, so dense cloud interiors
    // skip the erosion channels, the coverage pow and both remap quotients.
    if (density > 1.0f) return 1.0f;
    // heightBiasCoverage(relativeHeight, min(0.85, cloudInfo.r)) (comp:206-208, 245)
    float e = clampf(remapk(fminf(0.85f, w.coverage), 0.7f, 1.0f / (0.8f - 0.7f), 1.0f, 0.8f - 1.0f), 0.8f, 1.0f);
    float erosion = low_erosion<FSET>(lg, tx.low);
    if constexpr (!EXACT) return remapClamped01_fast(density, remapClamped01_fast(erosion, powg(h, e)));
    float coverage = pow_rt(h, e);
    erosion = remapClamped01(erosion, coverage);
    return remapClamped01(density, erosion);
}

template <int FSET, bool COUNT>
__device__ __forceinline__ float cloud_hires(const MarchConsts& c, const MarchTextures& tx, V3 pos, float curlStrength,
                                             float orig, float h, int& nh) {
    if (COUNT) nh++;
    constexpr int q8 = kQ8<FSET>;
    V3 cv2 = read_curl2<FSET>(tx.curl, c.curlScale * pos.x, c.curlScale * pos.z, q8);
    V3 curl = mk(cv2.x - 1.0f, cv2.y - 1.0f, cv2.z - 1.0f);
    pos = pos + ((1.9f * curlStrength) * curl);
    float e = read_hi_erosion<FSET>(tx.hi, pos * c.hiScale, q8);
    e = mixf(e, 1.0f - e, clampf(h * 10.0f, 0.0f, 1.0f));
    return remapClamped01(orig, 1.0f * e);
}

// hgPhase (comp:73-77)
__device__ __forceinline__ float hg_phase(float cosTheta, float g) {
    float g2 = g * g;
    float inv = 1.0f / powg(((1.0f - (2.0f * g) * cosTheta) + g2), 1.5f);
    return 0.07957747154594767f * ((1.0f - g2) * inv);
}

// getAtmosphereColorPhysical (comp:88-127)
__device__ V3 atmosphere(const MarchConsts& c, V3 dir, V3 sunDir) {
    const float PI = 3.14159265f;
    const float sunE = c.sunIntensity;
    float zenith = acosf(fmaxf(0.0f, dir.y));
    float inverse = 1.0f / (cosf(zenith) + 0.15f * powg(93.885f - ((zenith * 180.0f) / PI), -1.253f));
    float sR = 8.4e3f * inverse;
    float sM = 1.25e3f * inverse;
    float fex[3], Lin[3];
    float cosTheta = dot3(sunDir, dir);
    float rc = cosTheta * 0.5f + 0.5f;
    float rPhase = 0.05968310365946075f * (1.0f + rc * rc);
    float mPhase = hg_phase(cosTheta, c.mieG);
    float yDot = 1.0f - sunDir.y;
    yDot = yDot * (((yDot * yDot) * yDot) * yDot);
    float yc = clampf(yDot, 0.0f, 1.0f);
    V3 out;
    float* o = &out.x;
    for (int k = 0; k < 3; k++) {
        fex[k] = expg((c.betaR[k] * (-sR)) + (c.betaM[k] * sM));
        float betas = ((c.betaR[k] * rPhase) + (c.betaM[k] * mPhase)) / (c.betaR[k] + c.betaM[k]);
        float a = (betas * sunE) * (1.0f - fex[k]);
        Lin[k] = powg(a, 1.5f);
        float b = (betas * sunE) * fex[k];
        Lin[k] = Lin[k] * mixf(1.0f, powg(b, 0.5f), yc);
        float L0 = fex[k] * 0.1f;
        o[k] = (Lin[k] + L0) * 0.04f;
    }
    out.y = out.y + 0.0003f;
    out.z = out.z + 0.00075f;
    out.x = out.x + 0.0f;
    return out;
}

// IEEE division / square root in the intersection's arithmetic type.
__device__ __forceinline__ float div_(float a, float b) { return __fdiv_rn(a, b); }
__device__ __forceinline__ double div_(double a, double b) { return __ddiv_rn(a, b); }
__device__ __forceinline__ float sqrtx(float a) { return sqrt_rn(a); }
__device__ __forceinline__ double sqrtx(double a) { return __dsqrt_rn(a); }

// raySphereIntersection (comp:147-177) in arithmetic type I; a miss yields t = 0.
template <class I>
__device__ __forceinline__ I sphere_t(I rox, I roy, I roz, I rdx, I rdy, I rdz, I cx, I cy, I cz, I w, bool& valid) {
    rox = div_(rox - cx, w); roy = div_(roy - cy, w); roz = div_(roz - cz, w);
    I A = (rdx * rdx + rdy * rdy) + rdz * rdz;
    I B = I(2) * ((rdx * rox + rdy * roy) + rdz * roz);
    I C = ((rox * rox + roy * roy) + roz * roz) - I(0.25);
    I disc = (B * B) - ((I(4) * A) * C);
    valid = false;
    if (disc < I(0)) return I(0);
    I sq = sqrtx(disc);
    I t = div_(-sq - B, A) * I(0.5);
    if (t < I(0)) t = div_(sq - B, A) * I(0.5);
    if (!(t >= I(0))) return I(0);
    valid = true;
    I px = (rox + rdx * t) * w + cx, py = (roy + rdy * t) * w + cy, pz = (roz + rdz * t) * w + cz;
    I dx = px - rox, dy = py - roy, dz = pz - roz;
    return sqrtx((dx * dx + dy * dy) + dz * dz);
}

__device__ __forceinline__ float pow_chain4(float d) { return d * ((d * d) * d); }

// Bilinear RGBA8 night-sky fetch with generic (non power of two) REPEAT wrap.
__device__ V3 night_sky(const MarchConsts& c, const MarchTextures& tx, float u, float v) {
    if (!tx.hasNight) return mk(0.f, 0.f, 0.f);
    const TexView& tv = tx.night;
    const int q8 = c.subtexelBits;
    float f[2];
    int i0[2], i1[2];
    const float uv[2] = {u, v};
    const int n[2] = {tv.w, tv.h};
    for (int a = 0; a < 2; a++) {
        float x = uv[a] * (float)n[a] - 0.5f;
        if (q8) x = floorf(x * 256.0f + 0.5f) * (1.0f / 256.0f);
        float fl = floorf(x);
        f[a] = x - fl;
        int ii = (int)fl % n[a];
        if (ii < 0) ii += n[a];
        i0[a] = ii;
        i1[a] = (ii + 1 == n[a]) ? 0 : ii + 1;
    }
    const uint8_t* r0 = tv.data + (size_t)i0[1] * tv.stride * 4;
    const uint8_t* r1 = tv.data + (size_t)i1[1] * tv.stride * 4;
    G2<4> g;
    g.t[0] = ld_u32(r0 + i0[0] * 4); g.t[1] = ld_u32(r0 + i1[0] * 4);
    g.t[2] = ld_u32(r1 + i0[0] * 4); g.t[3] = ld_u32(r1 + i1[0] * 4);
    g.fx = f[0]; g.fy = f[1];
    return mk(filt2(g, Byte<0>()) * kInv255, filt2(g, Byte<1>()) * kInv255, filt2(g, Byte<2>()) * kInv255);
}

// ------------------------------------------------------------------ per-ray pieces
// Everything a ray needs before and after its march (comp:303-384, 484-499).
struct RaySetup {
    V3 rd, bg;
    float alpha;        // finalColor.a before the opacity factor
    float tStart, tEnd;
    float cosTheta, HG, beerBlend;
    bool night, march;  // march = false: the sky colour was already stored (comp:351-354)
};

__device__ __forceinline__ RaySetup ray_setup(const MarchConsts& c, const MarchTextures& tx, int px, int py) {
    RaySetup r;
    // ray generation (comp:303-322); IEEE division: the pixel's uv feeds the ray direction
    float uvx = __fdiv_rn((float)px, (float)c.width), uvy = __fdiv_rn((float)py, (float)c.height);
    float sx = uvx * 2.0f - 1.0f, sy = uvy * 2.0f - 1.0f;
    const V3 camPos = ld3(c.camPos);
    V3 p = (ld3(c.refPoint) + (((c.aspect * sx) * c.tanH) * ld3(c.camRight))) - ((sy * c.tanH) * ld3(c.camUp));
    r.rd = normalize3(p - camPos);
    const V3 rd = r.rd;
    const V3 sunDir = ld3(c.sunDir);

    // sun glow / alpha (comp:324-340)
    const float SUN_COS = 0.999956676946448443553574619906976478926848692873900859324f;
    float dotToSun = fmaxf(0.0f, dot3(sunDir, rd));
    float skyAmbient = dotToSun * 0.18f;
    skyAmbient = skyAmbient * (skyAmbient * skyAmbient);
    float sunDisk = smoothstepf(SUN_COS, SUN_COS + 0.00003f, dotToSun);
    dotToSun = pow_chain4(pow_chain4(pow_chain4(pow_chain4(dotToSun))));
    dotToSun = dotToSun * (dotToSun * dotToSun);
    r.night = c.sunDirY < 0.0f;
    if (r.night) {
        float d = dotToSun;
        dotToSun = d * ((((((d * d) * d) * d) * d) * d) * d);
    }
    sunDisk = fmaxf(0.0f, fmaxf(sunDisk, dotToSun));
    r.bg = mk(0.f, 0.f, 0.f);
    r.alpha = 0.0f;
    if (!r.night) {
        r.bg = atmosphere(c, rd, sunDir);
        r.alpha = fmaxf(skyAmbient, sunDisk);
    }
    r.march = !(rd.y < 0.0f);
    r.tStart = r.tEnd = 0.0f;
    r.cosTheta = r.HG = r.beerBlend = 0.0f;
    if (!r.march) return r;

    V3 center = camPos;
    center.y = c.earthCenter[1];
    bool vi, vo;
    if (c.flags & MM_MARCH_ISECT_FP64) {
        r.tStart = (float)sphere_t<double>(camPos.x, camPos.y, camPos.z, rd.x, rd.y, rd.z, center.x, center.y, center.z,
                                           (double)c.innerDiam, vi);
        r.tEnd = (float)sphere_t<double>(camPos.x, camPos.y, camPos.z, rd.x, rd.y, rd.z, center.x, center.y, center.z,
                                         (double)c.outerDiam, vo);
    } else {
        r.tStart = sphere_t<float>(camPos.x, camPos.y, camPos.z, rd.x, rd.y, rd.z, center.x, center.y, center.z,
                                   c.innerDiam, vi);
        r.tEnd = sphere_t<float>(camPos.x, camPos.y, camPos.z, rd.x, rd.y, rd.z, center.x, center.y, center.z,
                                 c.outerDiam, vo);
    }
    if (r.night) {  // comp:365-384
        V3 c0 = ld3(c.nightRot[0]), c1 = ld3(c.nightRot[1]), c2 = ld3(c.nightRot[2]);
        V3 rrd = ((c0 * rd.x) + (c1 * rd.y)) + (c2 * rd.z);
        V3 rro = ((c0 * camPos.x) + (c1 * camPos.y)) + (c2 * camPos.z);
        V3 point = (rrd * r.tEnd) + rro;
        V3 pr = proj_shell(c, point, center);
        float u = 0.00002f * (pr.x - camPos.x) + 0.35f;
        float v = 0.00002f * (pr.z - camPos.z) + 0.35f;
        V3 bg = night_sky(c, tx, u, v);
        bg = mk(bg.x * (sqrt_(bg.x) * 0.75f), bg.y * (sqrt_(bg.y) * 0.75f), bg.z * (sqrt_(bg.z) * 0.75f));
        bg = mk(powg(bg.x, 2.2f), powg(bg.y, 2.2f), powg(bg.z, 2.2f));
        bg = bg * 10.0f;
        float falloff = powg(rd.y, 6.0f);
        bg = bg * falloff;
        float pw = powg(rd.y, 0.03125f);
        bg = mk(mixf(0.3f * 0.05f, bg.x, pw), mixf(0.6f * 0.05f, bg.y, pw), mixf(4.0f * 0.05f, bg.z, pw));
        r.bg = bg + mk(sunDisk, sunDisk, sunDisk);
        r.alpha = sunDisk;
    }
    r.cosTheta = dot3(rd, sunDir);
    r.HG = fmaxf(hg_phase(r.cosTheta, 0.6f), 0.7f * hg_phase(r.cosTheta, 0.99f - 0.1f));
    r.beerBlend = (-r.cosTheta) * 0.5f + 0.5f;
    return r;
}

// compose (comp:484-499)
// The opacity accumulator after the horizon fade and the 0.999 cap (comp:484-486).
__device__ __forceinline__ float final_accum(const RaySetup& r, float accum) {
    accum *= smoothstepf(0.0f, 1.0f, fminf(1.0f, remapk(r.rd.y, 0.0f, 1.0f / 0.1f, 0.0f, 1.0f)));
    return fminf(accum, 0.999f);
}

__device__ __forceinline__ float4 compose(const MarchConsts& c, const RaySetup& r, float accum, float transmittance) {
    accum = final_accum(r, accum);
    const float I = c.sunIntensity;
    const float e = expg(-transmittance);
    V3 amb;
    if (!r.night) amb = (r.bg * 0.08f) * e;
    else {
        float pw = powg(r.rd.y, 0.03125f);
        amb = (mk(0.3f * 0.05f, 0.6f * 0.05f, 4.0f * 0.05f) * pw * 0.08f) * e;
    }
    const float direct = I * fmaxf(0.0f, transmittance);
    V3 cloud = mk(c.sunColor[0] * (direct + amb.x), c.sunColor[1] * (direct + amb.y), c.sunColor[2] * (direct + amb.z));
    float4 res;
    res.x = mixf(r.bg.x, cloud.x, accum);
    res.y = mixf(r.bg.y, cloud.y, accum);
    res.z = mixf(r.bg.z, cloud.z, accum);
    res.w = r.alpha * fmaxf(1.0f - accum, 0.0f);
    return res;
}

// Density along the sun: the six-sample light cone (comp:440-453).
template <int FSET, bool COUNT>
__device__ __forceinline__ float cone_dal(const MarchConsts& c, const MarchTextures& tx, V3 currentPos, float stepSize,
                                          Tally& nd, int& nh) {
    V3 center = ld3(c.camPos);
    center.y = c.earthCenter[1];
    float dal = 0.0f;
    // two independent samples in flight per lane (unrolling all six is slower: code size)
#pragma unroll 2
    for (int i = 0; i < 6; i++) {
        V3 lsPos = currentPos + ((3.0f * stepSize) * ld3(c.coneOffsets[i]));
        V3 lsProj = proj_shell(c, lsPos, center);
        float lsHeight = rel_height(c, lsPos, lsProj);
        V3 lwo = wind_offset(c, lsHeight);
        float ls = cloud_test<FSET, COUNT>(c, tx, lsPos + lwo, lsHeight, center, nd);
        if (ls > 0.0f) dal += cloud_hires<FSET, COUNT>(c, tx, lsPos + lwo, stepSize, ls, lsHeight, nh);
    }
    return dal;
}

// Beer's law with the artistic floor, as a multiplier of the in-scatter term (comp:456-460).
__device__ __forceinline__ float beers_of(float dal, float beerBlend) {
    float beers = expg(-dal);
    float beersMod = fmaxf(beers, 0.7f * expg(-0.25f * dal));
    return mixf(beers, beersMod, beerBlend);
}

// Inputs of the light-cone surrogate for one job: hit state (relative height, low-res density,
// eroded density), weather coverage/type at the hit, sun elevation, three density probes along the
// sun direction (the cone's axial offsets 0.6, 2.5 and 6 step lengths x 3), step scale, and the
// position within the hi-res tile as sin/cos pairs. (Cheaper probes -- the low-res base shape alone,
// scaled by the layer profile -- cost a third as much and made the trained network's error on the
// 1080p frame with every cone replaced 3.5 times larger, 1.68 against 0.48 / 255.)
// The coverage and type are those the hit's own density test read (the weather map at the
// projection of the wind-offset hit position), carried in the job queue.
template <int FSET, bool COUNT>
__device__ __forceinline__ void cone_features(const MarchConsts& c, const MarchTextures& tx, V3 pos, float stepSize,
                                              float rh, float lo, float dens, float cov, float type, float* f,
                                              Tally& nd) {
    V3 center = ld3(c.camPos);
    center.y = c.earthCenter[1];
    const V3 p = pos + wind_offset(c, rh);
    f[0] = rh; f[1] = lo; f[2] = dens; f[3] = cov; f[4] = type; f[5] = c.sunDir[1];
#pragma unroll 1
    for (int i = 0; i < 3; i++) {
        V3 lsPos = pos + ((3.0f * stepSize) * ld3(c.probeOffsets[i]));
        V3 lsProj = proj_shell_fast(c, lsPos, center);
        float lsHeight = rel_height_fast(c, lsPos, lsProj);
        f[6 + i] = cloud_test<FSET, COUNT, false>(c, tx, lsPos + wind_offset(c, lsHeight), lsHeight, center, nd);
    }
    f[9] = stepSize * c.invThickness;
    const float k = 6.283185307179586f * c.hiScale;
    f[10] = __sinf(k * p.x); f[11] = __cosf(k * p.x);
    f[12] = __sinf(k * p.y); f[13] = __cosf(k * p.y);
    f[14] = __sinf(k * p.z); f[15] = __cosf(k * p.z);
}

// in-scatter x phase of a fine hit (comp:461-462, 464): the factor that multiplies Beer's law
__device__ __forceinline__ float inscatter_hg(float loDensity, float rHeight, float HG) {
    float e1 = clampf(remapk(rHeight, 0.3f, 1.0f / (0.85f - 0.3f), 0.5f, 2.0f - 0.5f), 0.5f, 2.0f);
    float inScatter = 0.09f + powg(loDensity, e1);
    float e2 = clampf(remapk(rHeight, 0.07f, 1.0f / (0.34f - 0.07f), 0.1f, 1.0f - 0.1f), 0.1f, 1.0f);
    inScatter *= powg(e2, 0.8f);
    return inScatter * HG;
}

// State of one ray's march loop (comp:386-482) between loop trips.
struct MarchState {
    float t, stepSize, accum;
    int misses, steps, mc0, mc1, iterations;
    bool noHits, active;
};

__device__ __forceinline__ void march_init(const MarchConsts& c, const RaySetup& r, MarchState& m) {
    m.t = r.tStart;
    m.stepSize = c.stepSize0;
    m.accum = 0.0f;
    m.misses = m.steps = m.mc0 = m.mc1 = m.iterations = 0;
    m.noHits = true;
    m.active = r.march && (m.t < r.tEnd);
}

// The samples of one loop trip at t (comp:406-431): position, projected point and relative height,
// the wind-offset density and, in fine mode where it is positive, the hi-res eroded density. It
// depends only on t, the mode and the step size, so several trips' samples can be taken at once.
struct TripSample {
    V3 pos;
    float rHeight, lo, dens;
    Weather w;  // the weather the density test read (kept for the light-cone surrogate's features)
    Tally nd;
    int nh;  // density / detail samples taken (counted once the trip commits)
};

template <int FSET, bool COUNT, bool KEEPW = false>
__device__ __forceinline__ TripSample trip_eval(const MarchConsts& c, const MarchTextures& tx, const RaySetup& r,
                                                bool noHits, float stepSize, float t) {
    TripSample s;
    s.nd = s.nh = 0;
    s.w.coverage = s.w.type = 0.0f;
    const V3 camPos = ld3(c.camPos);
    V3 center = camPos;
    center.y = c.earthCenter[1];
    s.pos = camPos + (t * r.rd);
    const V3 proj = proj_shell(c, s.pos, center);
    s.rHeight = rel_height(c, s.pos, proj);
    const V3 wo = wind_offset(c, s.rHeight);
    s.lo = cloud_test<FSET, COUNT>(c, tx, s.pos + wo, s.rHeight, center, s.nd, KEEPW ? &s.w : nullptr);
    s.dens = s.lo;
    if (s.lo > 0.0f && !noHits) s.dens = cloud_hires<FSET, COUNT>(c, tx, s.pos + wo, stepSize, s.lo, s.rHeight, s.nh);
    return s;
}

// The rest of the trip (comp:386-482) given its samples, including the loop increment. A fine-mode
// hit that survives the hi-res erosion returns a light-cone job (the in-scatter factor and the mix
// weight 1 - accum at this trip) instead of evaluating the cone in place; its result only enters the
// transmittance recurrence, never the loop's control flow.
// The checks that end a trip before it samples: the hard bound on loop trips (termination even if
// t += stepSize stops advancing t) and, in bounded mode, the coarse / fine step budgets.
template <int F>
__device__ __forceinline__ bool trip_begin(const MarchConsts& c, MarchState& m) {
    if (++m.iterations > MM_MAX_MARCH_ITERATIONS) { m.active = false; return false; }
    if constexpr (kBS<F>) {
        if (m.noHits) { if (++m.mc0 > c.maxCoarse) { m.active = false; return false; } }
        else { if (++m.mc1 > c.maxFine) { m.active = false; return false; } }
    } else if (c.flags & MM_MARCH_BOUNDED) {
        if (m.noHits) { if (c.maxCoarse >= 0 && ++m.mc0 > c.maxCoarse) { m.active = false; return false; } }
        else { if (c.maxFine >= 0 && ++m.mc1 > c.maxFine) { m.active = false; return false; } }
    }
    return true;
}

// After trip_begin: the state update of the trip with its samples (low-res density lo, eroded
// density dens). A job's mix weight 1 - accum is returned in jobW; the caller takes its in-scatter
// factor from the trip's own samples.
template <int F>
__device__ __forceinline__ bool trip_commit(const MarchConsts& c, const RaySetup& r, MarchState& m, float lo,
                                            float dens, float& jobW) {
    bool job = false;
    const bool coarse = m.noHits;
    const bool bounded = kBS<F> || (c.flags & MM_MARCH_BOUNDED) != 0;
    bool cont = false;  // `continue`: skip the loop tail but still run the increment
    if (lo > 0.0f) {
        m.misses = 0;
        if (m.noHits) {
            m.t -= m.stepSize;
            m.stepSize *= c.fineMul;
            m.noHits = false;
            cont = true;
        } else {
            if (dens < 0.0001f) {
                cont = true;
            } else {
                job = true;
                jobW = 1.0f - m.accum;
                m.accum += dens;
            }
        }
    } else if (!m.noHits) {
        m.misses++;
        if (m.misses >= c.missLimit) {
            m.noHits = true;
            m.stepSize = __fdiv_rn(m.stepSize, c.fineMul);  // IEEE: every later sample position depends on it
        }
    }
    if (!cont) {
        if (m.accum > c.opacityCutoff) { m.accum = 1.0f; m.active = false; return job; }
        if (!kBS<F> && c.maxSteps >= 0 && ++m.steps > c.maxSteps) { m.active = false; return job; }
        if (!bounded) {
            if (coarse) { if (c.maxCoarse >= 0 && ++m.mc0 > c.maxCoarse) { m.active = false; return job; } }
            else { if (c.maxFine >= 0 && ++m.mc1 > c.maxFine) { m.active = false; return job; } }
        }
    }
    m.t += m.stepSize;
    if (!(m.t < r.tEnd)) m.active = false;
    return job;
}

// One trip of the reference's for-loop: its samples, then its state update.
template <int FSET, bool COUNT, bool KEEPW = false>
__device__ __forceinline__ bool march_trip(const MarchConsts& c, const MarchTextures& tx, const RaySetup& r,
                                           MarchState& m, V3& jobPos, float& jobStep, float& jobX0, float& jobW,
                                           float& jobRh, float& jobLo, float& jobDens, Weather& jobWeather, Tally& nd,
                                           int& nh) {
    if (!trip_begin<FSET>(c, m)) return false;
    const TripSample s = trip_eval<FSET, COUNT, KEEPW>(c, tx, r, m.noHits, m.stepSize, m.t);
    const float step = m.stepSize;
    const bool job = trip_commit<FSET>(c, r, m, s.lo, s.dens, jobW);
    if (job) {
        jobX0 = inscatter_hg(s.lo, s.rHeight, r.HG);
        jobPos = s.pos;
        jobStep = step;
        jobRh = s.rHeight;
        jobLo = s.lo;
        jobDens = s.dens;
        jobWeather = s.w;
    }
    nd += s.nd;
    nh += s.nh;
    return job;
}

// Tile order: block b marches the tile in column b mod tilesX of logical row b / tilesX. The hardware
// deals workgroups round-robin to the 8 XCDs, so consecutive tiles run on different XCDs and every
// XCD gets an even share of every image region. (A contiguous band of tiles per XCD keeps each L2's
// footprint compact, but the band that holds the long rays above the horizon then finishes last:
// 1.86x slower on the 1080p frame.) Logical rows start at the horizon and go up the sky, then the
// rows below the horizon follow (rowStart, rowDir, set on the host): the longest rays start first.
__device__ __forceinline__ int tile_row(const MarchConsts& c, int r) {
    int p = c.rowStart + c.rowDir * r;
    if (p < 0) p += c.tilesY;
    else if (p >= c.tilesY) p -= c.tilesY;
    return p;
}

__device__ __forceinline__ void tile_pixel(const MarchConsts& c, int gi, int gj, int& px, int& py) {
    if (c.phase >= 0) {
        px = gi * 4 + (c.phase & 3);
        py = gj * 4 + (c.phase >> 2);
    } else {
        px = gi;
        py = gj;
    }
}

// Physical tile index (tile row x tilesX + column) of logical launch tile `logical` (rows in tile_row's
// horizon-first order).
__device__ __forceinline__ int phys_tile(const MarchConsts& c, int logical) {
    return tile_row(c, logical / c.tilesX) * c.tilesX + logical % c.tilesX;
}

// The physical tile block b marches: from the caller's tile order (a permutation of the physical
// tiles, longest first, built from the previous launch's costs) or the horizon-first launch order;
// -1 for a block past the tiles or an out-of-range order entry.
// Block 0 of a scheduled launch clears the record's bucket counts for the next order (the order
// launches before this one are done with them).
__device__ __forceinline__ void clear_order_counts(unsigned* sched) {
    for (int i = threadIdx.x; i < mm::kOrderBuckets; i += blockDim.x) sched[mm::kSchedCount + i] = 0u;
}

__device__ __forceinline__ int block_tile(const MarchConsts& c) {
    const int b = (int)blockIdx.x;
    if (b == 0 && c.tileSched) clear_order_counts(c.tileSched);
    if (b >= c.tiles) return -1;
    const int t = c.tileOrder ? c.tileOrder[b] : phys_tile(c, b);
    return (unsigned)t < (unsigned)c.tiles ? t : -1;
}

// Where a tile records its cost: a schedule record (or null) and the tile's slot in it.
struct CostRec {
    unsigned* sched;
    int index;
};

// The tile's cost for the next launch's order: the longest ray's loop trips plus its fine-mode trips
// (those also take the detail sample; counted in bounded mode) -- a ray's lanes all count its trips
// -- plus an eighth of the tile's light-cone jobs, which run in 64-job batches. Every lane of the
// wave takes part. (Measured at 1080p against the trips alone and against each tile's own duration in
// the previous launch, this order gave the shortest launches.)
__device__ __forceinline__ void record_tile_cost(CostRec rec, int trips, int tileJobs) {
    if (!rec.sched) return;  // uniform
    int v = trips;
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if (threadIdx.x == 0) rec.sched[mm::kSchedHeader + rec.index] = (unsigned)v + (unsigned)tileJobs / 8u;
}

// Adds a lane's density tally to counts[0] and, for a six-slot counter array, its weather and
// low-res reads to counts[4] and counts[5].
__device__ __forceinline__ void add_tally(const MarchConsts& c, unsigned long long* counts, Tally nd) {
    atomicAdd(&counts[0], nd & kTallyMask);
    if (c.countSlots >= 6) {
        atomicAdd(&counts[4], (nd >> 21) & kTallyMask);
        atomicAdd(&counts[5], (nd >> 42) & kTallyMask);
    }
}

// ------------------------------------------------------------------ lock-step kernel
// One wave per 64-pixel tile; every lane evaluates its light cones in place (the shader's structure).
template <int FSET, bool COUNT>
__global__ __launch_bounds__(64) void march_kernel(const MarchConsts c, const MarchTextures tx, float* __restrict__ out,
                                                   unsigned long long* __restrict__ counts) {
    const int tile = blockIdx.x;
    if (tile >= c.tiles) return;
    const int lane = threadIdx.x;
    const int gi = (tile % c.tilesX) * (1 << c.tileShift) + (lane & ((1 << c.tileShift) - 1));
    const int gj = tile_row(c, tile / c.tilesX) * (64 >> c.tileShift) + (lane >> c.tileShift);
    if (gi >= c.gridW || gj >= c.gridH) return;
    int px, py;
    tile_pixel(c, gi, gj, px, py);
    Tally nd = 0;
    int nh = 0, nj = 0, nc = 0;  // density, detail samples; cone jobs, jobs on the plain cone
    const RaySetup r = ray_setup(c, tx, px, py);
    float4* dst = reinterpret_cast<float4*>(out) + ((size_t)py * c.pitch + px);
    float2* st = c.stateOut ? reinterpret_cast<float2*>(c.stateOut) + ((size_t)py * c.pitch + px) : nullptr;
    if (!r.march) {
        *dst = make_float4(r.bg.x, r.bg.y, r.bg.z, r.alpha);
        if (st) *st = make_float2(0.0f, 1.0f);
        return;
    }
    MarchState m;
    march_init(c, r, m);
    float transmittance = 1.0f;
    while (m.active) {
        V3 jp;
        float js, jx0, jw, jrh, jlo, jd;
        Weather jwe;
        if (march_trip<FSET, COUNT>(c, tx, r, m, jp, js, jx0, jw, jrh, jlo, jd, jwe, nd, nh)) {
            float dal = cone_dal<FSET, COUNT>(c, tx, jp, js, nd, nh);
            transmittance = mixf(transmittance, jx0 * beers_of(dal, r.beerBlend), jw);
            if (COUNT) { nj++; nc++; }
        }
    }
    *dst = compose(c, r, m.accum, transmittance);
    if (st) *st = make_float2(final_accum(r, m.accum), transmittance);
    if (COUNT) {
        add_tally(c, counts, nd);
        atomicAdd(&counts[1], (unsigned long long)nh);
        atomicAdd(&counts[2], (unsigned long long)nj);
        atomicAdd(&counts[3], (unsigned long long)nc);
    }
}

// ------------------------------------------------------------------ cone-batched kernel
// One wave per 64-pixel tile. Lanes run their primary march in lock-step but hand light-cone work to a
// wave-local LDS queue (slots assigned with ballot + mbcnt in creation order); once 64 jobs are
// queued (or no lane is still marching) the wave evaluates a full batch, one job per lane, and each
// lane folds its results into its transmittance in creation order. Cone work therefore runs at
// full SIMD width instead of waiting on the few lanes that are inside a cloud, and every lane
// performs exactly the shader's operations in the shader's order.
constexpr int kQueueCap = 128;

// light-cone modes: MM_CONE_PLAIN, MM_CONE_SURROGATE, and the training-sample recorder
enum ConeMode { CONE_PLAIN = MM_CONE_PLAIN, CONE_SURROGATE = MM_CONE_SURROGATE, CONE_COLLECT = 2 };

// The plain queue is 4 KB per wave (LDS then allows the full 8 waves per SIMD); the surrogate and
// recorder modes also keep each job's surrogate features and a compaction list.
struct ConeQueuePlain {
    float x[kQueueCap], y[kQueueCap], z[kQueueCap], step[kQueueCap];
    float x0[kQueueCap], w[kQueueCap], res[kQueueCap];
    int owner[kQueueCap];
};
struct ConeQueueFull : ConeQueuePlain {
    float rh[kQueueCap], lo[kQueueCap], dens[kQueueCap], cov[kQueueCap], type[kQueueCap];
    int fallback[64];
    float dal[64];
};
template <int MODE> struct ConeQueueOf { typedef ConeQueueFull type; };
template <> struct ConeQueueOf<CONE_PLAIN> { typedef ConeQueuePlain type; };

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// Evaluates the first min(qn, 64) queued cone jobs, one per lane, leaving each job's term
// x0 * Beer(D) in q.res (and, in the recorder mode, appending training rows).
// myBlend is the calling lane's own Beer blend; a job's blend is its owner's, fetched by a lane
// shuffle rather than stored in the queue.
template <int FSET, bool COUNT, int MODE, class Q>
__device__ __forceinline__ void run_batch(const MarchConsts& c, const MarchTextures& tx, Q& q, const int nb,
                                          const float myBlend, Tally& nd, int& nh, int& nj, int& nc) {
    const int lane = threadIdx.x;
    const float blend = __shfl(myBlend, q.owner[lane]);  // every lane takes part; slots >= nb are unused
    if (COUNT) nj += nb;
    if constexpr (MODE == CONE_PLAIN) {
        if (COUNT) nc += nb;
        if (lane < nb) {
            float dal = cone_dal<FSET, COUNT>(c, tx, mk(q.x[lane], q.y[lane], q.z[lane]), q.step[lane], nd, nh);
            q.res[lane] = q.x0[lane] * beers_of(dal, blend);
        }
    } else if constexpr (MODE == CONE_COLLECT) {
        if (COUNT) nc += nb;
        // plain cone + record (features, D) of one job in c.collectKeep
        float f[surr::kIn];
        float dal = 0.0f;
        bool keep = false;
        if (lane < nb) {
            const V3 pos = mk(q.x[lane], q.y[lane], q.z[lane]);
            dal = cone_dal<FSET, COUNT>(c, tx, pos, q.step[lane], nd, nh);
            q.res[lane] = q.x0[lane] * beers_of(dal, blend);
            const uint32_t hsh = mix32(__float_as_uint(pos.x) ^ mix32(__float_as_uint(pos.y) ^
                                       mix32(__float_as_uint(pos.z) ^ c.collectSeed)));
            keep = (hsh % c.collectKeep) == 0u;
            if (keep)
                cone_features<FSET, COUNT>(c, tx, pos, q.step[lane], q.rh[lane], q.lo[lane], q.dens[lane], q.cov[lane],
                                           q.type[lane], f, nd);
        }
        const unsigned long long km = __ballot(keep);
        if (km) {
            // every lane takes part in the atomic (lane 0 adds the batch's count, the others 0), so no
            // divergent branch sits inside the march loop; lane 0's old value is the batch's base
            const unsigned mine = atomicAdd(c.collectCount, lane == 0 ? (unsigned)__popcll(km) : 0u);
            const unsigned base = __builtin_amdgcn_readlane(mine, 0);
            const unsigned slot = base + __builtin_amdgcn_mbcnt_hi((unsigned)(km >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)km, 0u));
            if (keep && slot < c.collectCap) {
                float* dst = c.collectBuf + (size_t)slot * (surr::kIn + 1);
                for (int k = 0; k < surr::kIn; k++) dst[k] = f[k];
                dst[surr::kIn] = dal;
            }
        }
    } else {
        // surrogate: features + MLP for every queued job, plain cone only where the job's weighted
        // error bound is not below the threshold (compacted onto the first lanes)
        float f[surr::kIn];
        for (int k = 0; k < surr::kIn; k++) f[k] = 0.0f;
        if (lane < nb)
            cone_features<FSET, COUNT>(c, tx, mk(q.x[lane], q.y[lane], q.z[lane]), q.step[lane], q.rh[lane],
                                       q.lo[lane], q.dens[lane], q.cov[lane], q.type[lane], f, nd);
        for (int k = 0; k < surr::kIn; k++) f[k] = (f[k] - c.surrStats[k]) * c.surrStats[16 + k];
        const surr::FwdOut o = surr::mlp_infer64_h(c.surrFrags, c.surrParams + surr::oB4, f, nb);
        // the job's share of the pixel: its cone result enters T with weight (1 - accum) * inScatter * HG
        const float bound = lane < nb ? surr::softplus(o.o1) * (q.w[lane] * fabsf(q.x0[lane])) : 0.0f;
        const bool fb = lane < nb && !(bound < c.surrThreshold);
        const unsigned long long fm = __ballot(fb);
        const int nf = __popcll(fm);
        if (fb) q.fallback[__builtin_amdgcn_mbcnt_hi((unsigned)(fm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)fm, 0u))] = lane;
        __syncthreads();
        if (COUNT) nc += nf;
        if (lane < nf) {
            const int j = q.fallback[lane];
            q.dal[j] = cone_dal<FSET, COUNT>(c, tx, mk(q.x[j], q.y[j], q.z[j]), q.step[j], nd, nh);
        }
        __syncthreads();
        if (lane < nb) {
            const float dal = fb ? q.dal[lane] : fmaxf(0.0f, -o.o0);
            q.res[lane] = q.x0[lane] * beers_of(dal, blend);
        }
    }
    __syncthreads();
}

// Drops the first nb (= 64 when qn > 64) jobs of the queue, moving the remainder to the front.
template <class Q>
__device__ __forceinline__ void queue_shift(Q& q, const int nb, const int rem) {
    const int lane = threadIdx.x;
    __syncthreads();
    if (lane < rem) {  // rem < 64 and only when nb == 64: source and destination slots are disjoint
        const int s = nb + lane;
        q.x[lane] = q.x[s]; q.y[lane] = q.y[s]; q.z[lane] = q.z[s]; q.step[lane] = q.step[s];
        q.x0[lane] = q.x0[s]; q.w[lane] = q.w[s]; q.owner[lane] = q.owner[s];
        if constexpr (sizeof(Q) == sizeof(ConeQueueFull)) {
            q.rh[lane] = q.rh[s]; q.lo[lane] = q.lo[s]; q.dens[lane] = q.dens[s];
            q.cov[lane] = q.cov[s]; q.type[lane] = q.type[s];
        }
    }
    __syncthreads();
}

// One wave marches one 64-pixel tile ((1 << tileShift) x (64 >> tileShift) pixels of the marched
// grid, lane-major in x). Lanes run their primary march in lock-step; light-cone jobs go to the LDS
// queue and are evaluated 64 at a time (run_batch), each lane folding its results in creation order.
template <int FSET, bool COUNT, int MODE>
__device__ __forceinline__ void march_tile(const MarchConsts& c, const MarchTextures& tx, const int tile,
                                           float* __restrict__ out, unsigned long long* __restrict__ counts,
                                           CostRec costRec) {
    __shared__ typename ConeQueueOf<MODE>::type q;
    const int lane = threadIdx.x;
    // tile: physical tile index (row x tilesX + column)
    const int gi = (tile % c.tilesX) * (1 << c.tileShift) + (lane & ((1 << c.tileShift) - 1));
    const int gj = (tile / c.tilesX) * (64 >> c.tileShift) + (lane >> c.tileShift);
    const bool valid = gi < c.gridW && gj < c.gridH;
    int px = 0, py = 0;
    if (valid) tile_pixel(c, gi, gj, px, py);
    Tally nd = 0;
    int nh = 0, nj = 0, nc = 0;  // density, detail samples; cone jobs, jobs on the plain cone
    RaySetup r;
    MarchState m;
    m.active = false;
    if (valid) {
        r = ray_setup(c, tx, px, py);
        if (!r.march)
            reinterpret_cast<float4*>(out)[(size_t)py * c.pitch + px] = make_float4(r.bg.x, r.bg.y, r.bg.z, r.alpha);
        march_init(c, r, m);
    }
    float transmittance = 1.0f;
    int qn = 0;        // queued jobs (wave-uniform)
    int tileJobs = 0;  // light-cone jobs of the tile (wave-uniform)
    for (;;) {
        // primary march until a full batch is queued or every ray has left its loop
        while (qn < 64 && __any(m.active)) {
            V3 jp;
            float js = 0.f, jx0 = 0.f, jw = 0.f, jrh = 0.f, jlo = 0.f, jd = 0.f;
            Weather jwe;
            jwe.coverage = jwe.type = 0.0f;
            bool job = false;
            if (m.active)
                job = march_trip<FSET, COUNT, MODE != CONE_PLAIN>(c, tx, r, m, jp, js, jx0, jw, jrh, jlo, jd, jwe, nd, nh);
            const unsigned long long mask = __ballot(job);
            if (job) {
                const int slot = qn + __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                q.x[slot] = jp.x; q.y[slot] = jp.y; q.z[slot] = jp.z; q.step[slot] = js;
                q.x0[slot] = jx0; q.w[slot] = jw; q.owner[slot] = lane;
                if constexpr (MODE != CONE_PLAIN) {
                    q.rh[slot] = jrh; q.lo[slot] = jlo; q.dens[slot] = jd;
                    q.cov[slot] = jwe.coverage; q.type[slot] = jwe.type;
                }
            }
            qn += __popcll(mask);
            tileJobs += __popcll(mask);
        }
        if (qn == 0) break;
        __syncthreads();
        const int nb = qn < 64 ? qn : 64;
        run_batch<FSET, COUNT, MODE>(c, tx, q, nb, valid ? r.beerBlend : 0.0f, nd, nh, nj, nc);
        for (int s = 0; s < nb; s++)
            if (q.owner[s] == lane) transmittance = mixf(transmittance, q.res[s], q.w[s]);
        queue_shift(q, nb, qn - nb);
        qn -= nb;
    }
    if (valid && r.march) reinterpret_cast<float4*>(out)[(size_t)py * c.pitch + px] = compose(c, r, m.accum, transmittance);
    record_tile_cost(costRec, valid ? m.iterations + m.mc1 : 0, tileJobs);
    if (COUNT) {
        add_tally(c, counts, nd);
        atomicAdd(&counts[1], (unsigned long long)nh);
        if (lane == 0) {
            atomicAdd(&counts[2], (unsigned long long)nj);
            atomicAdd(&counts[3], (unsigned long long)nc);
        }
    }
}

template <int FSET, bool COUNT, int MODE>
__global__ __launch_bounds__(64) void march_batched_kernel(const MarchConsts c, const MarchTextures tx,
                                                           float* __restrict__ out,
                                                           unsigned long long* __restrict__ counts) {
    const int tile = block_tile(c);
    if (tile < 0) return;  // uniform across the block
    march_tile<FSET, COUNT, MODE>(c, tx, tile, out, counts, CostRec{c.tileSched, tile});
}

// ---------------------------------------------------------------- ray-parallel march
// LPR lanes per ray (2 or 4), 64 / LPR rays per wave in an 8-wide tile. Every lane of a ray holds
// the ray's whole state. Each loop round, lane k of the ray samples trip k ahead: at t plus k times
// the step (the same float additions the loop makes), in the current mode. The ray's lanes then
// commit the trips in order, identically, from the exchanged samples, and stop at the first trip
// whose mode, step or t no longer match what was sampled (a mode switch or the end of the ray): the
// committed trips are exactly the loop's, so the image is the one-lane-per-ray image. It shortens
// a ray's critical path by up to LPR, which is what bounds a single 1/16 phase (two waves per SIMD,
// each as long as its longest ray). Light-cone jobs go to the wave's queue like the batched kernel's.
template <int FSET, bool COUNT, int LPR, int MODE = CONE_PLAIN>
__device__ __forceinline__ void march_rp_tile(const MarchConsts& c, const MarchTextures& tx, const int tile,
                                              float* __restrict__ out, unsigned long long* __restrict__ counts,
                                              CostRec costRec) {
    static_assert(LPR == 2 || LPR == 4 || LPR == 8, "lanes per ray");
    constexpr int TH = 64 / LPR / 8;  // tile: 8 x TH rays
    __shared__ typename ConeQueueOf<MODE>::type q;
    const int lane = threadIdx.x, ray = lane / LPR, k = lane % LPR, first = lane - k;
    // tile: physical tile index (row x tilesX + column)
    const int gi = (tile % c.tilesX) * 8 + (ray & 7);
    const int gj = (tile / c.tilesX) * TH + (ray >> 3);
    const bool valid = gi < c.gridW && gj < c.gridH;
    int px = 0, py = 0;
    if (valid) tile_pixel(c, gi, gj, px, py);
    Tally nd = 0;
    int nh = 0, nj = 0, nc = 0;
    RaySetup r;
    MarchState m;
    m.active = false;
    if (valid) {
        r = ray_setup(c, tx, px, py);
        if (!r.march && k == 0)
            reinterpret_cast<float4*>(out)[(size_t)py * c.pitch + px] = make_float4(r.bg.x, r.bg.y, r.bg.z, r.alpha);
        march_init(c, r, m);
    }
    float transmittance = 1.0f;
    int qn = 0;        // queued jobs (wave-uniform)
    int tileJobs = 0;  // light-cone jobs of the tile (wave-uniform)
    for (;;) {
        while (qn < 64 && __any(m.active)) {
            const bool noHits0 = m.noHits;
            const float step0 = m.stepSize;
            float tk = m.t;
            for (int j = 0; j < k; j++) tk += step0;
            TripSample s;
            s.pos = mk(0.f, 0.f, 0.f);
            s.rHeight = s.lo = s.dens = 0.0f;
            s.nd = s.nh = 0;
            if (m.active) s = trip_eval<FSET, COUNT, MODE != CONE_PLAIN>(c, tx, r, noHits0, step0, tk);
            float lo[LPR], dens[LPR];
#pragma unroll
            for (int j = 0; j < LPR; j++) {
                lo[j] = __shfl(s.lo, first + j);
                dens[j] = __shfl(s.dens, first + j);
            }
            bool mine = false;  // this lane's trip committed with a light-cone job
            float jw = 0.0f;
            float tj = m.t;
#pragma unroll
            for (int j = 0; j < LPR; j++) {
                if (!m.active) break;
                if (j > 0 && !(m.noHits == noHits0 && m.stepSize == step0 && m.t == tj)) break;
                if (!trip_begin<FSET>(c, m)) break;
                float w = 0.0f;
                const bool job = trip_commit<FSET>(c, r, m, lo[j], dens[j], w);
                if (j == k) {
                    if (COUNT) { nd += s.nd; nh += s.nh; }
                    mine = job;
                    jw = w;
                }
                tj = tj + step0;
            }
            const unsigned long long mask = __ballot(mine);
            if (mine) {
                const int slot = qn + __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                q.x[slot] = s.pos.x; q.y[slot] = s.pos.y; q.z[slot] = s.pos.z; q.step[slot] = step0;
                q.x0[slot] = inscatter_hg(s.lo, s.rHeight, r.HG); q.w[slot] = jw; q.owner[slot] = lane;
                if constexpr (MODE != CONE_PLAIN) {
                    q.rh[slot] = s.rHeight; q.lo[slot] = s.lo; q.dens[slot] = s.dens;
                    q.cov[slot] = s.w.coverage; q.type[slot] = s.w.type;
                }
            }
            qn += __popcll(mask);
            tileJobs += __popcll(mask);
        }
        if (qn == 0) break;
        __syncthreads();
        const int nb = qn < 64 ? qn : 64;
        run_batch<FSET, COUNT, MODE>(c, tx, q, nb, valid ? r.beerBlend : 0.0f, nd, nh, nj, nc);
        for (int j = 0; j < nb; j++)
            if (q.owner[j] / LPR == ray) transmittance = mixf(transmittance, q.res[j], q.w[j]);
        queue_shift(q, nb, qn - nb);
        qn -= nb;
    }
    if (valid && r.march && k == 0)
        reinterpret_cast<float4*>(out)[(size_t)py * c.pitch + px] = compose(c, r, m.accum, transmittance);
    record_tile_cost(costRec, valid ? m.iterations + m.mc1 : 0, tileJobs);
    if (COUNT) {
        add_tally(c, counts, nd);
        atomicAdd(&counts[1], (unsigned long long)nh);
        if (lane == 0) {
            atomicAdd(&counts[2], (unsigned long long)nj);
            atomicAdd(&counts[3], (unsigned long long)nc);
        }
    }
}

template <int FSET, bool COUNT, int LPR>
__global__ __launch_bounds__(64) void march_rp_kernel(const MarchConsts c, const MarchTextures tx,
                                                      float* __restrict__ out, unsigned long long* __restrict__ counts) {
    const int tile = block_tile(c);
    if (tile < 0) return;  // uniform across the block
    march_rp_tile<FSET, COUNT, LPR>(c, tx, tile, out, counts, CostRec{c.tileSched, tile});
}

#ifndef MM_SURROGATE_WAVES_PER_EU
#define MM_SURROGATE_WAVES_PER_EU 5
#endif
// The surrogate variant bounds its register allocation: the MLP's peak (accumulators, operands)
// otherwise sets the occupancy of the whole march. Its job queue (7 KB per wave with the features'
// inputs) already holds a CU to 22 waves, 5 per SIMD, so the cap is 5: at 6 the MLP spills to
// scratch for no extra wave (1080p, every cone replaced, two lanes per ray: 4 / 5 / 6 waves
// 2.27 / 2.15 / 2.22 ms).
template <int FSET, bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MM_SURROGATE_WAVES_PER_EU)))
void march_surrogate_kernel(const MarchConsts c, const MarchTextures tx, float* __restrict__ out,
                            unsigned long long* __restrict__ counts) {
    const int tile = block_tile(c);
    if (tile < 0) return;
    march_tile<FSET, COUNT, CONE_SURROGATE>(c, tx, tile, out, counts, CostRec{c.tileSched, tile});
}

// The surrogate-lit march with two lanes per ray (the plain march's default shape for full frames).
template <int FSET, bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MM_SURROGATE_WAVES_PER_EU)))
void march_rp_surrogate_kernel(const MarchConsts c, const MarchTextures tx, float* __restrict__ out,
                               unsigned long long* __restrict__ counts) {
    const int tile = block_tile(c);
    if (tile < 0) return;
    march_rp_tile<FSET, COUNT, 2, CONE_SURROGATE>(c, tx, tile, out, counts, CostRec{c.tileSched, tile});
}

// ------------------------------------------------------------------ cloud shadows (model.frag)
// One work-item per shaded point: the mesh pass's six-step low-res shadow march (model.frag:240-270)
// with that shader's own constants and cloud functions (model.frag:52-129). Remaps divide exactly
// like the shader and the filters round their products like the scalar restatement (the march
// fuses them), so the REF texture set reproduces the restatement bit for bit.
__device__ __forceinline__ float sh_remap(float v, float a, float b, float c, float d) {
    return c + ((__fdiv_rn(v - a, b - a)) * (d - c));
}
__device__ __forceinline__ float sh_remap_clamped(float v, float a, float b, float c, float d) {
    return clampf(c + ((__fdiv_rn(v - a, b - a)) * (d - c)), c, d);
}
__device__ __forceinline__ float sh_layer(float h, float type) {
    h = clampf(h, 0.0f, 1.0f);
    const float strato = fmaxf(0.0f, sh_remap(h, 0.0f, 0.2f, 0.0f, 1.0f) * sh_remap(h, 0.2f, 0.7f, 1.0f, 0.0f));
    const float stratus = fmaxf(0.0f, sh_remap(h, 0.0f, 0.1f, 0.0f, 1.0f) * sh_remap(h, 0.2f, 0.3f, 1.0f, 0.0f));
    const float d1 = mixf(stratus, strato, clampf(type * 2.0f, 0.0f, 1.0f));
    const float d2 = mixf(strato, stratus, clampf((type - 0.5f) * 2.0f, 0.0f, 1.0f));
    return mixf(d1, d2, type);
}

template <int FSET>
__device__ float sh_cloud_test(const mm::ShadowConsts& c, const TexView& wt, const TexView& lt, V3 pos, float h,
                               V3 center) {
    const V3 proj = (c.projRadius * normalize3(pos - center)) + center;
    const Weather w = read_weather<FSET>(wt, c.weatherScale * (proj.x - c.camPos[0]),
                                                c.weatherScale * (proj.z - c.camPos[2]), kQ8<FSET>);
    const float layer = sh_layer(h, w.type);
    const LowG<FSET> lg = gather_low<FSET>(lt, pos * c.lowScale, kQ8<FSET>);
    float density = layer * sh_remap_clamped(low_shape<FSET>(lg, lt), 0.3f, 1.0f, 0.0f, 1.0f);
    if (density < 0.0001f) return 0.0f;
    const float e = clampf(sh_remap(fminf(0.85f, w.coverage), 0.7f, 0.8f, 1.0f, 0.6f), 0.6f, 1.0f);
    // pow evaluated as exp2(e * log2(h)) in double and rounded once (the scalar restatement uses the
    // same formula): the shadow is not on the frame's critical path, and the single rounding keeps
    // the 1e-4 early out and the erosion remap from amplifying last-bit differences
    const float coverage = (float)exp2((double)e * log2((double)h));
    const float erosion = sh_remap_clamped(low_erosion<FSET>(lg, lt), coverage, 1.0f, 0.0f, 1.0f);
    return sh_remap_clamped(density, erosion, 1.0f, 0.0f, 1.0f);
}

template <int FSET>
__global__ __launch_bounds__(256) void cloud_shadow_kernel(const mm::ShadowConsts c, const TexView wt, const TexView lt,
                                                           const float* __restrict__ pts, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c.n) return;
    const V3 wc = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    const V3 sunW = ld3(c.sunW);
    const V3 camPos = ld3(c.camPos);
    V3 center = camPos;
    center.y = c.centerY;
    bool hit;
    float t = c.fp64 ? (float)sphere_t<double>(wc.x, wc.y, wc.z, sunW.x, sunW.y, sunW.z, center.x, center.y, center.z,
                                               (double)c.innerDiam, hit)
                     : sphere_t<float>(wc.x, wc.y, wc.z, sunW.x, sunW.y, sunW.z, center.x, center.y, center.z,
                                       c.innerDiam, hit);
    const V3 L = ld3(c.L);
    const V3 origin = wc * c.originScale;
    const V3 wind = ld3(c.wind);
    float accum = 0.0f;
    for (int s = 0; s < c.steps; s++) {
        const V3 pos = origin + L * t;
        const V3 proj = (c.projRadius * normalize3(pos - center)) + center;
        const float h = clampf(__fdiv_rn(length3(pos - proj), c.thickness), 0.0f, 1.0f);
        V3 wo = c.windStrength * (wind + mk(0.0f, 0.2f * h, 0.0f));
        wo = wo * (c.time + h * 200.0f);
        accum = fmaxf(sh_cloud_test<FSET>(c, wt, lt, pos + wo, h, center), accum);
        if (accum > 0.99f) {
            accum = 1.0f;
            break;
        }
        t += c.step;
    }
    out[i] = accum;
}

// ------------------------------------------------------------------ multi-view kernel
// One launch over the tiles of all views: block b takes tile b of the concatenated tile lists, finds
// its view by binary search over the views' first-tile indices, and marches the tile with that
// view's constants and weather map. The hardware dispatcher hands blocks to CUs as slots free up,
// which is the global work queue; no counter or persistent loop is needed.
template <int FSET, bool COUNT, int MODE, int LPR = 1>
__device__ __forceinline__ void views_group(const ViewDesc* __restrict__ views, int nViews, const MarchTextures& tx0,
                                            int total, unsigned long long* __restrict__ counts,
                                            const int* __restrict__ order, unsigned* __restrict__ sched) {
    // global tile g: the plan's longest-first order (from its previous launch's tile costs) or the
    // concatenated tile lists in launch order
    const int b = blockIdx.x;
    if (b == 0 && sched) clear_order_counts(sched);
    if (b >= total) return;  // uniform across the block
    const int g = order ? order[b] : b;
    if ((unsigned)g >= (unsigned)total) return;
    int lo = 0, hi = nViews - 1;  // last view whose first tile is <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].tileBase <= g) lo = mid;
        else hi = mid - 1;
    }
    const ViewDesc& v = views[lo];
    // a view list whose tile counts no longer add up to the launch's total (a plan updated without a
    // new launch shape) must not march a tile index past its view's own tiles
    if (g - v.tileBase >= v.c.tiles) return;
    MarchTextures tx = tx0;
    tx.weather = v.weather;
    const CostRec rec{sched, g};
    if constexpr (LPR > 1) march_rp_tile<FSET, COUNT, LPR>(v.c, tx, phys_tile(v.c, g - v.tileBase), v.out, counts, rec);
    else march_tile<FSET, COUNT, MODE>(v.c, tx, phys_tile(v.c, g - v.tileBase), v.out, counts, rec);
}

template <int FSET, bool COUNT, int MODE, int LPR>
__global__ __launch_bounds__(64)
void march_views_kernel(const ViewDesc* __restrict__ views, int nViews, const MarchTextures tx, int total,
                        unsigned long long* __restrict__ counts, const int* __restrict__ order,
                        unsigned* __restrict__ sched) {
    views_group<FSET, COUNT, MODE, LPR>(views, nViews, tx, total, counts, order, sched);
}

template <int FSET, bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MM_SURROGATE_WAVES_PER_EU)))
void march_views_surrogate_kernel(const ViewDesc* __restrict__ views, int nViews, const MarchTextures tx, int total,
                                  unsigned long long* __restrict__ counts, const int* __restrict__ order,
                                  unsigned* __restrict__ sched) {
    views_group<FSET, COUNT, CONE_SURROGATE>(views, nViews, tx, total, counts, order, sched);
}

template <int FSET, bool COUNT>
const void* views_kernel_ptr(int mode, int lanes) {
    if constexpr ((FSET & kFloatWeights) != 0) lanes = 1;  // one kernel shape for the float sampler
    if (mode == CONE_SURROGATE) {
        if constexpr (kSet<FSET> != mm::FSET_PREC) return (const void*)march_views_surrogate_kernel<FSET, COUNT>;
        return nullptr;
    }
    if constexpr ((FSET & kFloatWeights) == 0) {
        switch (lanes) {
            case 2: return (const void*)march_views_kernel<FSET, COUNT, CONE_PLAIN, 2>;
            case 4: return (const void*)march_views_kernel<FSET, COUNT, CONE_PLAIN, 4>;
            case 8: return (const void*)march_views_kernel<FSET, COUNT, CONE_PLAIN, 8>;
        }
    }
    return (const void*)march_views_kernel<FSET, COUNT, CONE_PLAIN, 1>;
}

template <int FSET>
const void* views_kernel_ptr_c(int mode, int lanes, bool count, bool floatWeights, bool boundedSteps) {
    if (floatWeights)
        return count ? views_kernel_ptr<FSET | kFloatWeights, true>(mode, lanes)
                     : views_kernel_ptr<FSET | kFloatWeights, false>(mode, lanes);
    if (boundedSteps)
        return count ? views_kernel_ptr<FSET | kBoundedSteps, true>(mode, lanes)
                     : views_kernel_ptr<FSET | kBoundedSteps, false>(mode, lanes);
    return count ? views_kernel_ptr<FSET, true>(mode, lanes) : views_kernel_ptr<FSET, false>(mode, lanes);
}

// The float sampler (subtexel_bits 0) is a checking mode: one kernel per cone mode, the lock-step one
// for plain cones (every kernel shape gives the same image and counts).
template <int FSET, bool COUNT>
int launch_float_weights(const MarchConsts& c, const MarchTextures& t, float* out, unsigned long long* counts,
                         hipStream_t s) {
    dim3 grid((unsigned)c.tiles), block(64);
    if (c.coneMode == CONE_PLAIN)
        hipLaunchKernelGGL((march_kernel<FSET, COUNT>), grid, block, 0, s, c, t, out, counts);
    else if constexpr (kSet<FSET> != mm::FSET_PREC) {
        if (c.coneMode == CONE_COLLECT)
            hipLaunchKernelGGL((march_batched_kernel<FSET, COUNT, CONE_COLLECT>), grid, block, 0, s, c, t, out, counts);
        else
            hipLaunchKernelGGL((march_surrogate_kernel<FSET, COUNT>), grid, block, 0, s, c, t, out, counts);
    } else {
        return MM_E_FORMAT;
    }
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

template <int FSET, bool COUNT>
int launch_t(const MarchConsts& c, const MarchTextures& t, float* out, unsigned long long* counts, hipStream_t s) {
    dim3 grid((unsigned)c.tiles), block(64);
    if (c.coneMode == CONE_PLAIN && (c.flags & MM_MARCH_NO_BATCH))
        hipLaunchKernelGGL((march_kernel<FSET, COUNT>), grid, block, 0, s, c, t, out, counts);
    else if (c.coneMode == CONE_PLAIN && c.lanesPerRay == 8)
        hipLaunchKernelGGL((march_rp_kernel<FSET, COUNT, 8>), grid, block, 0, s, c, t, out, counts);
    else if (c.coneMode == CONE_PLAIN && c.lanesPerRay == 4)
        hipLaunchKernelGGL((march_rp_kernel<FSET, COUNT, 4>), grid, block, 0, s, c, t, out, counts);
    else if (c.coneMode == CONE_PLAIN && c.lanesPerRay == 2)
        hipLaunchKernelGGL((march_rp_kernel<FSET, COUNT, 2>), grid, block, 0, s, c, t, out, counts);
    else if (c.coneMode == CONE_PLAIN)
        hipLaunchKernelGGL((march_batched_kernel<FSET, COUNT, CONE_PLAIN>), grid, block, 0, s, c, t, out, counts);
    else if constexpr (kSet<FSET> != mm::FSET_PREC) {
        if (c.coneMode == CONE_SURROGATE && c.lanesPerRay == 2)
            hipLaunchKernelGGL((march_rp_surrogate_kernel<FSET, COUNT>), grid, block, 0, s, c, t, out, counts);
        else if (c.coneMode == CONE_COLLECT)
            hipLaunchKernelGGL((march_batched_kernel<FSET, COUNT, CONE_COLLECT>), grid, block, 0, s, c, t, out, counts);
        else
            hipLaunchKernelGGL((march_surrogate_kernel<FSET, COUNT>), grid, block, 0, s, c, t, out, counts);
    } else {
        return MM_E_FORMAT;
    }
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

// ---------------------------------------------------------------- tile schedule
// Longest-first order of the physical tiles from their costs in the previous launch of the same shape,
// in three launches over blocks of kOrderChunk tiles: (1) each block's histogram of cost buckets
// added into the record's per-bucket counts; (2) one workgroup turns the counts into each bucket's
// first slot, from the largest bucket down; (3) each block reserves a run of slots per bucket with one
// global atomic and hands them to its tiles with LDS atomics. Every tile gets exactly one slot of
// [0, n), so the result is a permutation; the order within a bucket is whatever the atomics give
// (the image does not depend on the order). The march that follows clears the counts (its block 0)
// and records the new costs.
constexpr int kOrderBlock = 256, kOrderPer = 16, kOrderChunk = kOrderBlock * kOrderPer;
// The LDS histograms are kept in kOrderCopies interleaved copies (thread t counts into copy t % 8):
// neighbouring tiles have similar costs, and one LDS atomic whose lanes fall into one bucket then
// queues 8 deep instead of 64. Entry (bucket b, copy r) sits at b * kOrderCopies + r.
constexpr int kOrderCopies = 8;
using mm::kOrderBuckets;

// Cost buckets on a log scale with 32 steps per octave: costs below 32 are their own buckets, a larger
// cost 2^e (1 + m / 32 + ...) goes to 32 + 32 (e - 5) + m. Monotone in the cost, at most 895, and
// within 3 % of the cost everywhere; needs no bound on the costs. (Ordering a recorded 1080p frame's
// 64 800 tile durations by these buckets, a list-scheduling model of the 1536 wave slots reaches the
// same makespan as the exact longest-first order: the total-work bound.)
__device__ __forceinline__ unsigned cost_bucket(unsigned c) {
    if (c < 32u) return c;
    const unsigned e = 31u - (unsigned)__clz((int)c);  // 5 .. 31
    return 32u + (e - 5u) * 32u + ((c >> (e - 5u)) & 31u);
}
static_assert(32 + 26 * 32 + 31 < mm::kOrderBuckets, "cost buckets");

// This block's tiles' costs (kOrderPer per thread, coalesced) counted per bucket into cnt.
__device__ __forceinline__ void order_block_counts(const unsigned* __restrict__ cost, int n, unsigned* cnt,
                                                   unsigned (&bk)[kOrderPer]) {
    const int t = threadIdx.x, copy = t % kOrderCopies, base = blockIdx.x * kOrderChunk + t;
    for (int e = t; e < kOrderBuckets * kOrderCopies; e += kOrderBlock) cnt[e] = 0u;
    unsigned v[kOrderPer];
#pragma unroll
    for (int k = 0; k < kOrderPer; k++) {
        const int i = base + k * kOrderBlock;
        v[k] = i < n ? cost[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kOrderPer; k++) bk[k] = cost_bucket(v[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kOrderPer; k++)
        if (base + k * kOrderBlock < n) atomicAdd(&cnt[bk[k] * kOrderCopies + copy], 1u);
    __syncthreads();
}

__global__ __launch_bounds__(kOrderBlock) void order_count_kernel(unsigned* __restrict__ sched, int n) {
    __shared__ unsigned cnt[kOrderBuckets * kOrderCopies];
    unsigned bk[kOrderPer];
    order_block_counts(sched + mm::kSchedHeader, n, cnt, bk);
    for (int b = threadIdx.x; b < kOrderBuckets; b += kOrderBlock) {
        unsigned s = 0;
#pragma unroll
        for (int r = 0; r < kOrderCopies; r++) s += cnt[b * kOrderCopies + r];
        if (s) atomicAdd(&sched[mm::kSchedCount + b], s);
    }
}

// One workgroup of kOrderBuckets threads: thread t takes bucket kOrderBuckets - 1 - t; an inclusive
// wave scan of the counts, then the wave totals, give each bucket's first slot.
__global__ __launch_bounds__(kOrderBuckets) void order_scan_kernel(unsigned* __restrict__ sched) {
    __shared__ unsigned part[kOrderBuckets / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, b = kOrderBuckets - 1 - t;
    const unsigned mine = sched[mm::kSchedCount + b];
    unsigned incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned v = (unsigned)__shfl_up((int)incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) part[wave] = incl;
    __syncthreads();
    unsigned before = 0;
    for (int w = 0; w < wave; w++) before += part[w];
    sched[mm::kSchedCursor + b] = before + incl - mine;
}

__global__ __launch_bounds__(kOrderBlock) void order_place_kernel(unsigned* __restrict__ sched, int n,
                                                                   int* __restrict__ order) {
    __shared__ unsigned cnt[kOrderBuckets * kOrderCopies];
    unsigned bk[kOrderPer];
    order_block_counts(sched + mm::kSchedHeader, n, cnt, bk);
    // the block's run of slots per bucket, split among the copies
    for (int b = threadIdx.x; b < kOrderBuckets; b += kOrderBlock) {
        unsigned loc[kOrderCopies], s = 0;
#pragma unroll
        for (int r = 0; r < kOrderCopies; r++) {
            loc[r] = s;
            s += cnt[b * kOrderCopies + r];
        }
        const unsigned first = s ? atomicAdd(&sched[mm::kSchedCursor + b], s) : 0u;
#pragma unroll
        for (int r = 0; r < kOrderCopies; r++) cnt[b * kOrderCopies + r] = first + loc[r];
    }
    __syncthreads();
    const int t = threadIdx.x, copy = t % kOrderCopies, base = blockIdx.x * kOrderChunk + t;
#pragma unroll
    for (int k = 0; k < kOrderPer; k++) {
        const int i = base + k * kOrderBlock;
        if (i < n) {
            const unsigned slot = atomicAdd(&cnt[bk[k] * kOrderCopies + copy], 1u);
            if (slot < (unsigned)n) order[slot] = i;  // always, with counts from this record
        }
    }
}

}  // namespace

namespace mm {

int launch_shadow(int fset, const ShadowConsts& c, const TexView& weather, const TexView& low, const float* pts,
                  float* out, void* stream) {
    dim3 grid((unsigned)((c.n + 255) / 256)), block(256);
    hipStream_t s = (hipStream_t)stream;
    const bool fw = c.subtexelBits == 0;
    if (fset == FSET_REF && fw)
        hipLaunchKernelGGL(cloud_shadow_kernel<FSET_REF | kFloatWeights>, grid, block, 0, s, c, weather, low, pts, out);
    else if (fset == FSET_REF)
        hipLaunchKernelGGL(cloud_shadow_kernel<FSET_REF>, grid, block, 0, s, c, weather, low, pts, out);
    else if (fset == FSET_FAST && fw)
        hipLaunchKernelGGL(cloud_shadow_kernel<FSET_FAST | kFloatWeights>, grid, block, 0, s, c, weather, low, pts, out);
    else if (fset == FSET_FAST)
        hipLaunchKernelGGL(cloud_shadow_kernel<FSET_FAST>, grid, block, 0, s, c, weather, low, pts, out);
    else
        return MM_E_FORMAT;
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int launch_views(int fset, int mode, int lanesPerRay, bool floatWeights, bool boundedSteps, const ViewDesc* views,
                 int nViews, const MarchTextures& t, int total, unsigned long long* counts, const int* order,
                 unsigned* sched, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const bool cnt = counts != nullptr;
    const bool bs = boundedSteps;
    const void* k = fset == FSET_REF ? views_kernel_ptr_c<FSET_REF>(mode, lanesPerRay, cnt, floatWeights, bs)
                    : fset == FSET_FAST ? views_kernel_ptr_c<FSET_FAST>(mode, lanesPerRay, cnt, floatWeights, bs)
                                        : views_kernel_ptr_c<FSET_PREC>(mode, lanesPerRay, cnt, floatWeights, bs);
    if (!k) return MM_E_FORMAT;
    const unsigned grid = (unsigned)total;
    void* args[] = {(void*)&views, (void*)&nViews, (void*)&t, (void*)&total, (void*)&counts, (void*)&order,
                    (void*)&sched};
    if (hipLaunchKernel(k, dim3(grid), dim3(64), args, 0, s) != hipSuccess) return MM_E_DEVICE;
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int launch_tile_order(unsigned* sched, int n, int* order, void* stream) {
    if (n <= 0) return MM_OK;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((n + kOrderChunk - 1) / kOrderChunk));
    hipLaunchKernelGGL(order_count_kernel, grid, dim3(kOrderBlock), 0, s, sched, n);
    if (hipGetLastError() != hipSuccess) return MM_E_DEVICE;
    hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(kOrderBuckets), 0, s, sched);
    if (hipGetLastError() != hipSuccess) return MM_E_DEVICE;
    hipLaunchKernelGGL(order_place_kernel, grid, dim3(kOrderBlock), 0, s, sched, n, order);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int launch_march(int fset, const MarchConsts& c, const MarchTextures& t, float* out, unsigned long long* counts,
                 void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const bool count = counts != nullptr;
    if (c.subtexelBits == 0) {
        switch (fset) {
            case FSET_REF: return count ? launch_float_weights<FSET_REF | kFloatWeights, true>(c, t, out, counts, s)
                                        : launch_float_weights<FSET_REF | kFloatWeights, false>(c, t, out, counts, s);
            case FSET_FAST: return count ? launch_float_weights<FSET_FAST | kFloatWeights, true>(c, t, out, counts, s)
                                         : launch_float_weights<FSET_FAST | kFloatWeights, false>(c, t, out, counts, s);
            case FSET_PREC: return count ? launch_float_weights<FSET_PREC | kFloatWeights, true>(c, t, out, counts, s)
                                         : launch_float_weights<FSET_PREC | kFloatWeights, false>(c, t, out, counts, s);
        }
        return MM_E_FORMAT;
    }
    if (bounded_steps(c.flags, c.maxSteps, c.maxCoarse, c.maxFine)) {
        constexpr int B = kBoundedSteps;
        switch (fset) {
            case FSET_REF: return count ? launch_t<FSET_REF | B, true>(c, t, out, counts, s)
                                        : launch_t<FSET_REF | B, false>(c, t, out, counts, s);
            case FSET_FAST: return count ? launch_t<FSET_FAST | B, true>(c, t, out, counts, s)
                                         : launch_t<FSET_FAST | B, false>(c, t, out, counts, s);
            case FSET_PREC: return count ? launch_t<FSET_PREC | B, true>(c, t, out, counts, s)
                                         : launch_t<FSET_PREC | B, false>(c, t, out, counts, s);
        }
        return MM_E_FORMAT;
    }
    switch (fset) {
        case FSET_REF: return count ? launch_t<FSET_REF, true>(c, t, out, counts, s) : launch_t<FSET_REF, false>(c, t, out, counts, s);
        case FSET_FAST: return count ? launch_t<FSET_FAST, true>(c, t, out, counts, s) : launch_t<FSET_FAST, false>(c, t, out, counts, s);
        case FSET_PREC: return count ? launch_t<FSET_PREC, true>(c, t, out, counts, s) : launch_t<FSET_PREC, false>(c, t, out, counts, s);
    }
    return MM_E_FORMAT;
}

}  // namespace mm
