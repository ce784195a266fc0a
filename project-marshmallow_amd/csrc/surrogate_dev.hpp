// Device side of the learned light-cone surrogate: a 16-64-64-64-2 MLP evaluated with
// gfx950 BF16 MFMA (v_mfma_f32_32x32x16_bf16), FP32 accumulation, shared by the training kernels
// and by the march's light-cone batches.
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


// Bias of rows crow(r, h) of m-tile `off` as an accumulator initialiser.
__device__ __forceinline__ f32x16 bias16(const float* __restrict__ p, int off, int h) {
    f32x16 z;
    for (int r = 0; r < 16; r++) z[r] = p[off + crow(r, h)];
    return z;
}

// Inference forward of one 32-sample tile with a small register footprint: one m-tile at a time,
// biases folded into the accumulator initialiser. The four weight fragments of an m-tile are loaded
// together before its MFMA chain, so the chain waits for one load latency rather than four.
__device__ __forceinline__ FwdOut mlp_infer_tile(const uint4* __restrict__ frags, const float* __restrict__ params,
                                                 bf16x8 xb) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    bf16x8 b[4], nb[4];
    for (int mt = 0; mt < 2; mt++) {
        f32x16 z = bias16(params, oB1 + mt * 32, h);
        z = mfma(load_frag(frags, kFragF1 + mt, lane), xb, z);
        float v[16];
        for (int r = 0; r < 16; r++) v[r] = hswish(z[r]);
        b[mt * 2] = pack8(v);
        b[mt * 2 + 1] = pack8(v + 8);
    }
    const int bias[2] = {oB2, oB3};
    const int ffrag[2] = {kFragF2, kFragF3};
#pragma unroll 1
    for (int layer = 0; layer < 2; layer++) {
        // the two m-tiles' MFMA chains interleaved: each dependent MFMA has the other chain's in front
        bf16x8 w0[4], w1[4];
        for (int kt = 0; kt < 4; kt++) {
            w0[kt] = load_frag(frags, ffrag[layer] + kt, lane);
            w1[kt] = load_frag(frags, ffrag[layer] + 4 + kt, lane);
        }
        f32x16 zz[2] = {bias16(params, bias[layer], h), bias16(params, bias[layer] + 32, h)};
        for (int kt = 0; kt < 4; kt++) {
            zz[0] = mfma(w0[kt], b[kt], zz[0]);
            zz[1] = mfma(w1[kt], b[kt], zz[1]);
        }
        __builtin_amdgcn_sched_barrier(0);
        for (int mt = 0; mt < 2; mt++) {
            const f32x16 z = zz[mt];
            float v[16];
            for (int r = 0; r < 16; r++) v[r] = hswish(z[r]);
            nb[mt * 2] = pack8(v);
            nb[mt * 2 + 1] = pack8(v + 8);
        }
        for (int k = 0; k < 4; k++) b[k] = nb[k];
    }
    bf16x8 w[4];
    for (int kt = 0; kt < 4; kt++) w[kt] = load_frag(frags, kFragF4 + kt, lane);
    f32x16 o = {};
    for (int kt = 0; kt < 4; kt++) o = mfma(w[kt], b[kt], o);
    __builtin_amdgcn_sched_barrier(0);
    FwdOut r;
    r.o0 = o[0] + params[oB4 + 0];  // row crow(0, 0) = 0 in lanes of half 0
    r.o1 = o[1] + params[oB4 + 1];  // row 1
    return r;
}

// Inference for up to 64 samples (two tiles), one sample per lane. `x` are this lane's normalised
// features; `n` (wave-uniform) is the number of samples in use, lanes 0..n-1: with n <= 32 the second
// tile (samples 32..63) is not evaluated and lanes 32..63 return zeros. Every lane of the wave must
// call it (MFMA is wave-wide). Returns this lane's outputs.
__device__ __forceinline__ FwdOut mlp_forward64(const uint4* frags, const float* params, const float* x, int n = 64) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    // opaque to loop-invariant code motion: keeps the weight loads from being hoisted out of an
    // enclosing loop (where they would hold registers for the whole loop)
    frags = uniform_ptr(frags);
    params = uniform_ptr(params);
    asm volatile("" : "+s"(frags));
    asm volatile("" : "+s"(params));
    const bf16x8 lo = pack8(x), hi = pack8(x + 8);
    // B operand of tile t: lane (h, c) needs features 8h..8h+7 of sample 32t + c
    uint4 ul = __builtin_bit_cast(uint4, lo), uh = __builtin_bit_cast(uint4, hi), sl, sh;
    sl.x = __shfl_xor(ul.x, 32); sl.y = __shfl_xor(ul.y, 32); sl.z = __shfl_xor(ul.z, 32); sl.w = __shfl_xor(ul.w, 32);
    sh.x = __shfl_xor(uh.x, 32); sh.y = __shfl_xor(uh.y, 32); sh.z = __shfl_xor(uh.z, 32); sh.w = __shfl_xor(uh.w, 32);
    const bf16x8 xlo = __builtin_bit_cast(bf16x8, sl), xhi = __builtin_bit_cast(bf16x8, sh);
    const bf16x8 b0 = h ? xhi : lo;  // tile 0 (samples 0..31)
    const bf16x8 b1 = h ? hi : xlo;  // tile 1 (samples 32..63)
    FwdOut t0 = mlp_infer_tile(frags, params, b0);
    __builtin_amdgcn_sched_barrier(0);
    FwdOut t1;
    t1.o0 = t1.o1 = 0.0f;
    if (n > 32) t1 = mlp_infer_tile(frags, params, b1);  // wave-uniform: a half-empty batch skips its second tile
    // tile 0 results sit in lane c (= sample c); tile 1 results for sample 32 + c sit in lane c
    FwdOut r;
    const float s0 = __shfl_xor(t1.o0, 32), s1 = __shfl_xor(t1.o1, 32);
    r.o0 = h ? s0 : t0.o0;
    r.o1 = h ? s1 : t0.o1;
    return r;
}

__device__ __forceinline__ float softplus(float x) { return x > 20.0f ? x : __logf(1.0f + __expf(x)); }

}  // namespace surr
