// Device side of the learned light-cone surrogate: a 16-64-64-64-2 MLP on gfx950 MFMA with FP32
// accumulation. Training runs BF16 (v_mfma_f32_32x32x16_bf16); inference -- the march's light-cone
// batches and batch prediction -- runs FP16 (v_mfma_f32_32x32x16_f16) with packed-FP16 activations.
//
// Layout. The forward pass computes Z^T = W^T X^T per layer: weights are the A operand (M = units),
// activations the B operand (N = 32 samples per tile). A 32x32 accumulator puts row 8g + 4h + q
// (g, q in 0..3) of column c in lane 32h + c, element 4g + q. Eight consecutive accumulator elements
// 8t..8t+7 of one lane therefore hold 8 rows of one sample, which is exactly one 16-wide k-step of
// the next layer's B operand provided that layer's k index is permuted the same way:
//     k-step kt, lane half h, element j  ->  unit (kt / 2) * 32 + 8 * (2 * (kt % 2) + j / 4) + 4 h + j % 4
// Weights are stored pre-permuted (mm::surrogate packing), one 16-byte fragment per lane and
// k-step, so activations stay in registers from layer to layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace surr {

constexpr int kIn = 16, kHidden = 64, kOut = 2;
// FP32 master parameter offsets: W1[16][64], b1[64], W2[64][64], b2, W3[64][64], b3, W4[64][2], b4[2]
constexpr int oW1 = 0, oB1 = oW1 + kIn * kHidden, oW2 = oB1 + kHidden, oB2 = oW2 + kHidden * kHidden,
              oW3 = oB2 + kHidden, oB3 = oW3 + kHidden * kHidden, oW4 = oB3 + kHidden, oB4 = oW4 + kHidden * kOut,
              kParams = oB4 + kOut;  // 9538
// Packed BF16 A-operand fragments (ushort), one uint4 (8 bf16) per (tile, kstep, lane):
//   forward  L1: 2 m-tiles x 1 kstep, L2/L3: 2 m-tiles x 4 ksteps, L4: 1 m-tile x 4 ksteps
//   backward L4 (dA3 = W4 dZ4): 2 m-tiles x 1 kstep (k = 2 outputs padded to 16)
//            L3, L2 (dA = W dZ): 2 m-tiles x 4 ksteps
constexpr int kFragF1 = 0, kFragF2 = kFragF1 + 2 * 1, kFragF3 = kFragF2 + 2 * 4, kFragF4 = kFragF3 + 2 * 4,
              kFragB4 = kFragF4 + 1 * 4, kFragB3 = kFragB4 + 2 * 1, kFragB2 = kFragB3 + 2 * 4,
              kFrags = kFragB2 + 2 * 4;  // 36 fragments x 64 lanes x 16 B = 36 KiB

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// Permuted unit index of the hidden-layer k dimension (see the header comment).
__host__ __device__ inline int kperm(int kt, int h, int j) { return (kt / 2) * 32 + 8 * (2 * (kt % 2) + j / 4) + 4 * h + j % 4; }
// Row of accumulator element r in lane half h.
__host__ __device__ inline int crow(int r, int h) { return 8 * (r / 4) + 4 * h + (r % 4); }

// Hidden-unit activation: hard SiLU (hard-swish, MobileNetV3), z * clamp(z / 6 + 1/2, 0, 1), the
// piecewise-quadratic form of SiLU z * sigmoid(z): three full-rate VALU operations instead of an
// exponential and a reciprocal, which at 192 activations per job cost as much as a plain light cone.
__device__ __forceinline__ float hswish(float z) {
    return z * __builtin_amdgcn_fmed3f(__builtin_fmaf(z, 1.0f / 6.0f, 0.5f), 0.0f, 1.0f);
}
// d hswish / dz: 0 below -3, 1 above 3, z / 3 + 1/2 between
__device__ __forceinline__ float hswish_grad(float z) {
    return z <= -3.0f ? 0.0f : (z >= 3.0f ? 1.0f : __builtin_fmaf(z, 1.0f / 3.0f, 0.5f));
}

__device__ __forceinline__ bf16x8 load_frag(const uint4* frags, int frag, int lane) {
    uint4 u = frags[frag * 64 + lane];
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ bf16x8 pack8(const float* v) {
    bf16x8 r;
    for (int j = 0; j < 8; j++) r[j] = (__bf16)v[j];
    return r;
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Forward pass of one 32-sample tile. `xb` is the B operand of layer 1 (this lane's 8 input
// features of sample lane % 32: features 8h..8h+7, already normalised). Writes the pre-activations
// of the three hidden layers when `zsave` is non-null (16 floats per m-tile per layer per lane) and
// returns the two outputs of sample lane % 32 in lanes of half 0 (rows 0 and 1 of the last tile).
struct FwdOut {
    float o0, o1;
};

__device__ __forceinline__ FwdOut mlp_forward_tile(const uint4* __restrict__ frags, const float* __restrict__ params,
                                                   bf16x8 xb, float* zsave /* [3][2][16] or null */) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    // layer 1: K = 16 (natural feature order)
    f32x16 acc[2];
    for (int mt = 0; mt < 2; mt++) {
        f32x16 z = {};
        acc[mt] = mfma(load_frag(frags, kFragF1 + mt, lane), xb, z);
    }
    bf16x8 b[4];
    const int bias[3] = {oB1, oB2, oB3};
    const int ffrag[2] = {kFragF2, kFragF3};
    for (int layer = 0; layer < 3; layer++) {
        // bias + activation, repack as the next layer's B operand (8 consecutive accumulators = one k-step)
        for (int mt = 0; mt < 2; mt++) {
            float v[16];
            for (int r = 0; r < 16; r++) {
                float zz = acc[mt][r] + params[bias[layer] + mt * 32 + crow(r, h)];
                if (zsave) zsave[(layer * 2 + mt) * 16 + r] = zz;
                v[r] = hswish(zz);
            }
            b[mt * 2 + 0] = pack8(v);
            b[mt * 2 + 1] = pack8(v + 8);
        }
        if (layer == 2) break;
        for (int mt = 0; mt < 2; mt++) {
            f32x16 z = {};
            for (int kt = 0; kt < 4; kt++) z = mfma(load_frag(frags, ffrag[layer] + mt * 4 + kt, lane), b[kt], z);
            acc[mt] = z;
        }
    }
    // output layer: 1 m-tile (rows 0, 1 used), K = 64
    f32x16 o = {};
    for (int kt = 0; kt < 4; kt++) o = mfma(load_frag(frags, kFragF4 + kt, lane), b[kt], o);
    FwdOut r;
    r.o0 = o[0] + params[oB4 + 0];  // row crow(0, 0) = 0 in lanes of half 0
    r.o1 = o[1] + params[oB4 + 1];  // row 1
    return r;
}

// A wave-uniform pointer made explicit (scalar registers), whatever the optimiser could prove.
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const uint64_t v = (uint64_t)p;
    // readfirstlane returns a signed int: keep both halves unsigned before widening
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T*)(((uint64_t)hi << 32) | (uint64_t)lo);
}


__device__ __forceinline__ float softplus(float x) { return x > 20.0f ? x : __logf(1.0f + __expf(x)); }

// ---------------------------------------------------------------- FP16 inference (the march's path)
// The light-cone batches and mm_surrogate_predict evaluate the trained network with
// v_mfma_f32_32x32x16_f16 and packed-FP16 activations; training keeps the BF16 kernels above.
//  * Hidden units: z (FP32 accumulator) rounded to FP16 (round to nearest even, v_cvt_pk_f16_f32),
//    u = clamp(fma(z, 1/6, 1/2), 0, 1) and z u, each rounded to FP16 (v_pk_fma_f16 with clamp,
//    v_pk_mul_f16): three instructions per two units, the result already the next layer's B operand.
//  * Biases: one more MFMA per m-tile, the bias fragment (k = 0: FP16 b, k = 1: FP16 (b - FP16 b))
//    against a constant B operand that is 1 at k = 0, 1 and 0 elsewhere, so the accumulator starts at
//    b to about 22 bits with no per-lane vector loads. The two output biases are uniform and added
//    after the last layer.
// Inference fragments (one uint4 of 8 FP16 per lane), A-operand layout as the BF16 ones:
//   L1: m-tile mt at kHF1 + 2 mt: [bias, W1 (k = 16 features)]
//   L2, L3: m-tile mt at kHF2/kHF3 + 5 mt: [bias, 4 k-steps (k permuted as kperm)]
//   L4: kHF4 + kt: 4 k-steps (rows 0 and 1 are the two outputs)
constexpr int kHF1 = 0, kHF2 = 4, kHF3 = 14, kHF4 = 24, kHFrags = 28;  // 28 KiB

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));

// Fragments through a global-address-space pointer: global loads with a scalar base and 32-bit lane
// offsets (a generic pointer would give flat loads and 64-bit address arithmetic).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) u32x4 guint4;

__device__ __forceinline__ f16x8 load_hfrag(guint4* __restrict__ frags, int frag, int lane) {
    const u32x4 u = frags[frag * 64 + lane];
    return __builtin_bit_cast(f16x8, u);
}

__device__ __forceinline__ f32x16 mfma_h(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// Hard SiLU of accumulator elements 8q .. 8q + 7 in packed FP16, as one B-operand k-step.
__device__ __forceinline__ f16x8 hswish8_h(const f32x16& z, int q) {
    const h2 k6 = {(_Float16)(1.0f / 6.0f), (_Float16)(1.0f / 6.0f)};
    const h2 half = {(_Float16)0.5f, (_Float16)0.5f};
    const h2 zero = {(_Float16)0.0f, (_Float16)0.0f}, one = {(_Float16)1.0f, (_Float16)1.0f};
    uint32_t w[4];
    for (int i = 0; i < 4; i++) {
        const f2 zz = {z[8 * q + 2 * i], z[8 * q + 2 * i + 1]};
        const h2 v = __builtin_convertvector(zz, h2);
        h2 u = __builtin_elementwise_fma(v, k6, half);
        u = __builtin_elementwise_min(__builtin_elementwise_max(u, zero), one);
        w[i] = __builtin_bit_cast(uint32_t, v * u);
    }
    return __builtin_bit_cast(f16x8, make_uint4(w[0], w[1], w[2], w[3]));
}

// Inference of one 32-sample tile: xb is this lane's B operand of layer 1 (8 normalised FP16 features
// of sample lane % 32, features 8h .. 8h + 7), ones the constant bias operand. Returns the two
// outputs of sample lane % 32 in lanes of half 0, without the output biases.
__device__ __forceinline__ FwdOut mlp_infer_tile_h(guint4* __restrict__ frags, f16x8 xb, f16x8 ones) {
    const int lane = threadIdx.x & 63;
    f16x8 b[4];
    for (int mt = 0; mt < 2; mt++) {
        const f16x8 wb = load_hfrag(frags, kHF1 + 2 * mt, lane), w = load_hfrag(frags, kHF1 + 2 * mt + 1, lane);
        f32x16 z = {};
        z = mfma_h(wb, ones, z);
        z = mfma_h(w, xb, z);
        b[2 * mt] = hswish8_h(z, 0);
        b[2 * mt + 1] = hswish8_h(z, 1);
    }
#pragma unroll 1
    for (int layer = 0; layer < 2; layer++) {
        const int base = layer ? kHF3 : kHF2;
        f16x8 nb[4];
        for (int mt = 0; mt < 2; mt++) {
            f16x8 w[5];
            for (int k = 0; k < 5; k++) w[k] = load_hfrag(frags, base + 5 * mt + k, lane);
            f32x16 z = {};
            z = mfma_h(w[0], ones, z);
            for (int kt = 0; kt < 4; kt++) z = mfma_h(w[1 + kt], b[kt], z);
            nb[2 * mt] = hswish8_h(z, 0);
            nb[2 * mt + 1] = hswish8_h(z, 1);
        }
        for (int k = 0; k < 4; k++) b[k] = nb[k];
    }
    f16x8 w[4];
    for (int kt = 0; kt < 4; kt++) w[kt] = load_hfrag(frags, kHF4 + kt, lane);
    f32x16 o = {};
    for (int kt = 0; kt < 4; kt++) o = mfma_h(w[kt], b[kt], o);
    __builtin_amdgcn_sched_barrier(0);
    FwdOut r;
    r.o0 = o[0];  // row crow(0, 0) = 0 in lanes of half 0
    r.o1 = o[1];  // row 1
    return r;
}

// FP16 inference for up to 64 samples, one per lane (x: this lane's normalised features; n
// wave-uniform, lanes 0..n-1 in use; n <= 32 skips the second tile). b4 holds the two output
// biases. Every lane of the wave must call it. Returns this lane's outputs.
__device__ __forceinline__ FwdOut mlp_infer64_h(const uint4* frags, const float* b4, const float* x, int n = 64) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    // opaque to loop-invariant code motion: keeps the fragment loads from being hoisted out of an
    // enclosing loop (where they would hold registers for the whole loop)
    frags = uniform_ptr(frags);
    asm volatile("" : "+s"(frags));
    guint4* gfrags = (guint4*)frags;
    uint32_t lw[4], hw[4];
    for (int i = 0; i < 4; i++) {
        const f2 a = {x[2 * i], x[2 * i + 1]}, b = {x[8 + 2 * i], x[8 + 2 * i + 1]};
        lw[i] = __builtin_bit_cast(uint32_t, __builtin_convertvector(a, h2));
        hw[i] = __builtin_bit_cast(uint32_t, __builtin_convertvector(b, h2));
    }
    // B operand of tile t: lane (h, c) needs features 8h .. 8h + 7 of sample 32 t + c
    uint32_t sl[4], sh[4];
    for (int i = 0; i < 4; i++) {
        sl[i] = __shfl_xor(lw[i], 32);
        sh[i] = __shfl_xor(hw[i], 32);
    }
    uint32_t t0[4], t1[4];
    for (int i = 0; i < 4; i++) {
        t0[i] = h ? sh[i] : lw[i];  // tile 0 (samples 0..31)
        t1[i] = h ? hw[i] : sl[i];  // tile 1 (samples 32..63)
    }
    const f16x8 b0 = __builtin_bit_cast(f16x8, make_uint4(t0[0], t0[1], t0[2], t0[3]));
    const f16x8 b1 = __builtin_bit_cast(f16x8, make_uint4(t1[0], t1[1], t1[2], t1[3]));
    const uint32_t one2 = 0x3c003c00u;  // FP16 1.0 at k = 0 and 1 (lane half 0, elements 0 and 1)
    const f16x8 ones = __builtin_bit_cast(f16x8, make_uint4(h ? 0u : one2, 0u, 0u, 0u));
    FwdOut t0o = mlp_infer_tile_h(gfrags, b0, ones);
    __builtin_amdgcn_sched_barrier(0);
    FwdOut t1o;
    t1o.o0 = t1o.o1 = 0.0f;
    if (n > 32) t1o = mlp_infer_tile_h(gfrags, b1, ones);
    const float s0 = __shfl_xor(t1o.o0, 32), s1 = __shfl_xor(t1o.o1, 32);
    FwdOut r;
    r.o0 = (h ? s0 : t0o.o0) + b4[0];
    r.o1 = (h ? s1 : t0o.o1) + b4[1];
    return r;
}

}  // namespace surr
