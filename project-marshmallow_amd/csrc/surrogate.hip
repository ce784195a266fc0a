// Learned light-cone surrogate: parameter init, feature statistics, BF16 fragment packing, MFMA
// training step (forward, backward, gradient accumulation), Adam, and batch prediction.
//
// The model predicts, for one light-cone job of the march (comp:441-453), the log sun
// transmittance o0 = -D (D = density along the light) and an error bound e = softplus(o1) in
// transmittance units. Loss per sample (T = exp(-D)): (exp(o0) - T)^2 + (e - |exp(o0) - T|)^2, the
// second term with the residual held fixed.
#include <hip/hip_runtime.h>
#include <math.h>
#include "mm_internal.hpp"
#include "surrogate_dev.hpp"

using namespace surr;

namespace {

// Model block layout (bytes). All regions 256-byte aligned.
constexpr size_t kPadP = ((kParams * 4 + 255) / 256) * 256;
constexpr size_t offParams = 0, offGrads = kPadP, offM = 2 * kPadP, offV = 3 * kPadP, offStats = 4 * kPadP,
                 offFrags = offStats + 512, offHFrags = offFrags + (size_t)kFrags * 64 * 16,
                 kModelBytes = offHFrags + (size_t)kHFrags * 64 * 16;
// one packing thread per (fragment, lane): the BF16 training fragments, then the FP16 inference ones
constexpr int kPackThreads = (kFrags + kHFrags) * 64;

struct ModelPtrs {
    float *params, *grads, *m, *v, *stats;  // stats: mean[16], inv_std[16], step, loss_sum, loss_n
    uint4* frags;   // BF16 training fragments (kFrags)
    uint4* hfrags;  // FP16 inference fragments (kHFrags)
};

ModelPtrs ptrs(void* base) {
    uint8_t* b = (uint8_t*)base;
    return ModelPtrs{(float*)(b + offParams), (float*)(b + offGrads), (float*)(b + offM), (float*)(b + offV),
                     (float*)(b + offStats), (uint4*)(b + offFrags), (uint4*)(b + offHFrags)};
}

__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// Glorot-uniform weights, zero biases; Adam state cleared; stats = identity normalisation.
__global__ void init_kernel(float* params, float* m, float* v, float* stats, uint32_t seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams) {
        float w = 0.0f, lim = 0.0f;
        if (i < oB1) lim = sqrtf(6.0f / (kIn + kHidden));
        else if (i >= oW2 && i < oB2) lim = sqrtf(6.0f / (2 * kHidden));
        else if (i >= oW3 && i < oB3) lim = sqrtf(6.0f / (2 * kHidden));
        else if (i >= oW4 && i < oB4) lim = sqrtf(6.0f / (kHidden + kOut));
        if (lim > 0.0f) w = ((hash_u32(seed * 0x9E3779B9u + (uint32_t)i) >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * lim;
        params[i] = w;
        m[i] = 0.0f;
        v[i] = 0.0f;
    }
    if (i < 16) {
        stats[i] = 0.0f;
        stats[16 + i] = 1.0f;
    }
    if (i >= 32 && i < 64) stats[i] = 0.0f;
}

// Per-feature sums (double) over the sample set, then mean / inverse std.
__global__ void stats_sum_kernel(const float* __restrict__ samples, int n, int stride, double* acc) {
    __shared__ double s[2][kIn][64];
    const int lane = threadIdx.x;
    double sum[kIn], sq[kIn];
    for (int k = 0; k < kIn; k++) sum[k] = sq[k] = 0.0;
    for (long i = (long)blockIdx.x * 64 + lane; i < n; i += (long)gridDim.x * 64) {
        for (int k = 0; k < kIn; k++) {
            double x = samples[i * stride + k];
            sum[k] += x;
            sq[k] += x * x;
        }
    }
    for (int k = 0; k < kIn; k++) { s[0][k][lane] = sum[k]; s[1][k][lane] = sq[k]; }
    __syncthreads();
    if (lane < 2 * kIn) {
        const int w = lane / kIn, k = lane % kIn;
        double t = 0.0;
        for (int j = 0; j < 64; j++) t += s[w][k][j];
        atomicAdd(&acc[w * kIn + k], t);
    }
}

__global__ void stats_final_kernel(const double* acc, int n, float* stats) {
    const int k = threadIdx.x;
    if (k < kIn) {
        double mean = acc[k] / n, var = acc[kIn + k] / n - mean * mean;
        stats[k] = (float)mean;
        stats[16 + k] = (float)(1.0 / sqrt(var > 1e-12 ? var : 1e-12));
    }
}

// FP16 inference fragment element: fragment f, lane half h, column c, element j (k = 8 h + j).
// The inference weights are the training pass's BF16 weights: FP16 holds a BF16 value exactly when
// its magnitude lies in [2^-17, 65504] (8 significant bits; FP16 subnormals keep 8 bits down to
// 2^-17), so the FP16 fragments carry the same weights and only the activations gain precision.
__device__ float bf16_value(float w) { return (float)(__bf16)w; }

__device__ float hfrag_value(const float* __restrict__ p, int f, int h, int c, int j) {
    const int k = 8 * h + j;
    auto bias = [&](int off, int u) {  // k = 0: FP16 b, k = 1: the FP16 rounding of the remainder
        const float b = p[off + u], hi = (float)(_Float16)b;
        return k == 0 ? hi : (k == 1 ? b - hi : 0.0f);
    };
    if (f < kHF2) {
        const int mt = f / 2, u = mt * 32 + c;
        return (f % 2 == 0) ? bias(oB1, u) : bf16_value(p[oW1 + k * kHidden + u]);
    }
    if (f < kHF4) {
        const int base = f < kHF3 ? kHF2 : kHF3, off = f < kHF3 ? oW2 : oW3, boff = f < kHF3 ? oB2 : oB3;
        const int mt = (f - base) / 5, which = (f - base) % 5, u = mt * 32 + c;
        return which == 0 ? bias(boff, u) : bf16_value(p[off + kperm(which - 1, h, j) * kHidden + u]);
    }
    const int kt = f - kHF4;
    return c < kOut ? bf16_value(p[oW4 + kperm(kt, h, j) * kOut + c]) : 0.0f;
}

// BF16 A-operand fragments (see surrogate_dev.hpp), then the FP16 inference fragments. One thread per
// (fragment, lane).
__global__ void pack_kernel(const float* __restrict__ p, uint4* __restrict__ frags, uint4* __restrict__ hfrags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kPackThreads) return;
    if (t >= kFrags * 64) {
        const int ti = t - kFrags * 64, f = ti / 64, lane = ti % 64;
        f16x8 r;
        for (int j = 0; j < 8; j++) r[j] = (_Float16)hfrag_value(p, f, lane >> 5, lane & 31, j);
        hfrags[ti] = __builtin_bit_cast(uint4, r);
        return;
    }
    const int frag = t / 64, lane = t % 64, h = lane >> 5, c = lane & 31;
    float v[8];
    for (int j = 0; j < 8; j++) {
        float w = 0.0f;
        if (frag < kFragF2) {                        // forward L1: W1[k][u], k = 8h + j, u = mt*32 + c
            const int mt = frag - kFragF1;
            w = p[oW1 + (8 * h + j) * kHidden + mt * 32 + c];
        } else if (frag < kFragF4) {                 // forward L2/L3: W[k = kperm][u]
            const int base = frag < kFragF3 ? kFragF2 : kFragF3, off = frag < kFragF3 ? oW2 : oW3;
            const int mt = (frag - base) / 4, kt = (frag - base) % 4;
            w = p[off + kperm(kt, h, j) * kHidden + mt * 32 + c];
        } else if (frag < kFragB4) {                 // forward L4: W4[k = kperm][u = c] for u < 2
            const int kt = frag - kFragF4;
            w = c < kOut ? p[oW4 + kperm(kt, h, j) * kOut + c] : 0.0f;
        } else if (frag < kFragB3) {                 // backward L4: A[u3 = mt*32 + c][k = out 8h + j]
            const int mt = frag - kFragB4, k = 8 * h + j;
            w = k < kOut ? p[oW4 + (mt * 32 + c) * kOut + k] : 0.0f;
        } else {                                     // backward L3/L2: A[u_in = mt*32 + c][k = kperm(u_out)]
            const int base = frag < kFragB2 ? kFragB3 : kFragB2, off = frag < kFragB2 ? oW3 : oW2;
            const int mt = (frag - base) / 4, kt = (frag - base) % 4;
            w = p[off + (mt * 32 + c) * kHidden + kperm(kt, h, j)];
        }
        v[j] = w;
    }
    bf16x8 b = pack8(v);
    frags[t] = __builtin_bit_cast(uint4, b);
}

// One Adam step's gradients over a batch: block = one wave = 64 random samples (two 32-sample
// tiles). Forward in registers (pre-activations parked in LDS), backward dA = W dZ with the same
// fragment layout, weight gradients dW^T = dZ^T a^T from LDS-transposed tiles, accumulated into
// the global FP32 gradient buffer with atomics.
struct TrainLds {
    float z[2][3][2][16][64];          // [tile][layer][m-tile][acc][lane] pre-activations
    __bf16 dz[64][40];                 // dZ^T rows = units, 32 samples (+pad)
    __bf16 act[64][40];                // a^T rows = input units/features, 32 samples (+pad)
    float x[64][kIn];                  // normalised inputs of the 64 samples
    float dout[64][2];
};

__global__ __launch_bounds__(64) void train_kernel(const float* __restrict__ samples, int n, int stride,
                                                   const float* __restrict__ params, const float* __restrict__ stats,
                                                   const uint4* __restrict__ frags, float* __restrict__ grads,
                                                   float* __restrict__ loss_acc, uint32_t seed, int step,
                                                   int sequential, float invB) {
    __shared__ TrainLds L;
    const int lane = threadIdx.x, h = lane >> 5, c = lane & 31;
    // rows: a seeded hash draw (training) or rows blockIdx * 64 + lane < n (exact mean gradient)
    uint32_t idx;
    bool valid = true;
    if (sequential) {
        idx = blockIdx.x * 64u + lane;
        valid = idx < (uint32_t)n;
        if (!valid) idx = 0;
    } else {
        idx = hash_u32(seed ^ hash_u32((uint32_t)step * 0x632BE5ABu + blockIdx.x * 64u + lane)) % (uint32_t)n;
    }
    float x[kIn];
    for (int k = 0; k < kIn; k++) {
        x[k] = (samples[(size_t)idx * stride + k] - stats[k]) * stats[16 + k];
        L.x[lane][k] = x[k];
    }
    const float D = samples[(size_t)idx * stride + kIn];
    const float T = __expf(-D);
    const float rowW = valid ? 1.0f : 0.0f;
    __syncthreads();

    float lossv = 0.0f;
    for (int tile = 0; tile < 2; tile++) {
        // B operand of layer 1 for this tile: features 8h..8h+7 of sample 32*tile + c
        float xb[8];
        for (int j = 0; j < 8; j++) xb[j] = L.x[32 * tile + c][8 * h + j];
        float zs[96];
        FwdOut o = mlp_forward_tile(frags, params, pack8(xb), zs);
        for (int l = 0; l < 3; l++)
            for (int mt = 0; mt < 2; mt++)
                for (int r = 0; r < 16; r++) L.z[tile][l][mt][r][lane] = zs[(l * 2 + mt) * 16 + r];
        // loss gradient for sample 32*tile + c (its outputs are held by lanes of half 0)
        const float Ts = __shfl(T, 32 * tile + c);  // all lanes shuffle; the target lives in lane s
        const float ws = __shfl(rowW, 32 * tile + c);
        if (h == 0) {
            const int s = 32 * tile + c;
            const float p = __expf(o.o0), r = p - Ts, ar = fabsf(r);
            const float e = softplus(o.o1), sig = 1.0f / (1.0f + __expf(-o.o1));
            lossv += ws * (r * r + (e - ar) * (e - ar));
            L.dout[s][0] = ws * (2.0f * r * p);
            L.dout[s][1] = ws * (2.0f * (e - ar) * sig);
        }
    }
    __syncthreads();

    for (int tile = 0; tile < 2; tile++) {
        // ---- layer 4: dZ4^T (2 rows) as B operand of dA3 = W4 dZ4 (k = outputs 8h + j)
        float dz4[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (h == 0) {
            dz4[0] = L.dout[32 * tile + c][0] * invB;
            dz4[1] = L.dout[32 * tile + c][1] * invB;
        }
        // dW4^T[u][k] = sum_s dZ4[u][s] a3[k][s], db4: accumulate directly (2 outputs x 64 units)
        // a3 = hswish(z3) of this lane's sample c for units crow(r, h) + 32 mt
        {
            // write a3^T and dz4^T tiles to LDS
            for (int mt = 0; mt < 2; mt++)
                for (int r = 0; r < 16; r++)
                    L.act[mt * 32 + crow(r, h)][c] = (__bf16)hswish(L.z[tile][2][mt][r][lane]);
            if (h == 0) {
                L.dz[0][c] = (__bf16)dz4[0];
                L.dz[1][c] = (__bf16)dz4[1];
            }
            __syncthreads();
            // dW4: M = 2 outputs (rows of dz, padded to 32 with zero rows), N = 64 units, K = 32 samples
            for (int nt = 0; nt < 2; nt++) {
                f32x16 acc = {};
                for (int kt = 0; kt < 2; kt++) {
                    float a8[8], b8[8];
                    for (int j = 0; j < 8; j++) {
                        a8[j] = c < kOut ? (float)L.dz[c][16 * kt + 8 * h + j] : 0.0f;
                        b8[j] = (float)L.act[nt * 32 + c][16 * kt + 8 * h + j];
                    }
                    acc = mfma(pack8(a8), pack8(b8), acc);
                }
                for (int r = 0; r < 16; r++) {
                    const int u = crow(r, h), k = nt * 32 + c;
                    if (u < kOut) atomicAdd(&grads[oW4 + k * kOut + u], acc[r]);
                }
            }
            if (lane < kOut) {
                float sb = 0.0f;
                for (int s = 0; s < 32; s++) sb += (float)L.dz[lane][s];
                atomicAdd(&grads[oB4 + lane], sb);
            }
            __syncthreads();
        }
        // dA3 = W4 dZ4 -> dZ3 = dA3 * hswish'(z3)
        f32x16 dz[2];
        {
            const bf16x8 b = pack8(dz4);
            for (int mt = 0; mt < 2; mt++) {
                f32x16 z = {};
                z = mfma(load_frag(frags, kFragB4 + mt, lane), b, z);
                for (int r = 0; r < 16; r++) z[r] *= hswish_grad(L.z[tile][2][mt][r][lane]);
                dz[mt] = z;
            }
        }
        // ---- layers 3, 2, 1: weight gradients from (dZ_l, a_{l-1}); propagate to l-1
        const int wOff[3] = {oW3, oW2, oW1}, bOff[3] = {oB3, oB2, oB1}, bfrag[2] = {kFragB3, kFragB2};
        for (int li = 0; li < 3; li++) {
            const int layer = 2 - li;  // z index of this layer's pre-activation: 2 = layer 3
            const int kin = layer == 0 ? kIn : kHidden;
            // LDS: dz^T rows = units of this layer; act^T rows = inputs of this layer
            for (int mt = 0; mt < 2; mt++)
                for (int r = 0; r < 16; r++) L.dz[mt * 32 + crow(r, h)][c] = (__bf16)dz[mt][r];
            if (layer > 0) {
                for (int mt = 0; mt < 2; mt++)
                    for (int r = 0; r < 16; r++)
                        L.act[mt * 32 + crow(r, h)][c] = (__bf16)hswish(L.z[tile][layer - 1][mt][r][lane]);
            } else {
                for (int k = 0; k < kIn; k++)
                    if (h == 0) L.act[k][c] = (__bf16)L.x[32 * tile + c][k];
            }
            __syncthreads();
            // dW^T[u][k] : M = 64 units (2 m-tiles), N = kin inputs (1 or 2 n-tiles), K = 32 samples
            for (int mt = 0; mt < 2; mt++) {
                for (int nt = 0; nt < (kin + 31) / 32; nt++) {
                    f32x16 acc = {};
                    for (int kt = 0; kt < 2; kt++) {
                        float a8[8], b8[8];
                        for (int j = 0; j < 8; j++) {
                            a8[j] = (float)L.dz[mt * 32 + c][16 * kt + 8 * h + j];
                            const int k = nt * 32 + c;
                            b8[j] = k < kin ? (float)L.act[k][16 * kt + 8 * h + j] : 0.0f;
                        }
                        acc = mfma(pack8(a8), pack8(b8), acc);
                    }
                    for (int r = 0; r < 16; r++) {
                        const int u = mt * 32 + crow(r, h), k = nt * 32 + c;
                        if (k < kin) atomicAdd(&grads[wOff[li] + k * kHidden + u], acc[r]);
                    }
                }
            }
            {
                float sb = 0.0f;
                for (int s = 0; s < 32; s++) sb += (float)L.dz[lane][s];
                atomicAdd(&grads[bOff[li] + lane], sb);
            }
            __syncthreads();
            if (layer == 0) break;
            // dA_{l-1} = W_l dZ_l (B operand = dZ registers, k permuted), times hswish'(z_{l-1})
            bf16x8 b[4];
            for (int mt = 0; mt < 2; mt++) {
                float v[16];
                for (int r = 0; r < 16; r++) v[r] = dz[mt][r];
                b[mt * 2 + 0] = pack8(v);
                b[mt * 2 + 1] = pack8(v + 8);
            }
            for (int mt = 0; mt < 2; mt++) {
                f32x16 z = {};
                for (int kt = 0; kt < 4; kt++) z = mfma(load_frag(frags, bfrag[li] + mt * 4 + kt, lane), b[kt], z);
                for (int r = 0; r < 16; r++) z[r] *= hswish_grad(L.z[tile][layer - 1][mt][r][lane]);
                dz[mt] = z;
            }
        }
    }
    // loss: lanes of half 0 carry the per-sample terms
    for (int off = 32; off > 0; off >>= 1) lossv += __shfl_xor(lossv, off);
    if (lane == 0) atomicAdd(loss_acc, lossv);
}

__global__ void adam_kernel(float* params, float* grads, float* m, float* v, float lr, float b1, float b2, float eps,
                            int t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float g = grads[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)t)), vh = vi / (1.0f - powf(b2, (float)t));
    params[i] -= lr * mh / (sqrtf(vh) + eps);
    grads[i] = 0.0f;
}

__global__ void zero_kernel(float* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Batch inference with the march's FP16 path: 64 samples per wave; out[i] = (o0, softplus(o1)).
__global__ __launch_bounds__(64) void predict_kernel(const float* __restrict__ x, int n, int stride,
                                                     const float* __restrict__ params, const float* __restrict__ stats,
                                                     const uint4* __restrict__ hfrags, float* __restrict__ out) {
    const int lane = threadIdx.x;
    const long i = (long)blockIdx.x * 64 + lane;
    float f[kIn];
    for (int k = 0; k < kIn; k++) f[k] = i < n ? (x[i * stride + k] - stats[k]) * stats[16 + k] : 0.0f;
    const long left = (long)n - (long)blockIdx.x * 64;
    FwdOut o = mlp_infer64_h(hfrags, params + oB4, f, left < 64 ? (int)left : 64);
    if (i < n) {
        out[i * 2 + 0] = o.o0;
        out[i * 2 + 1] = softplus(o.o1);
    }
}

inline unsigned blocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace mm {

size_t surrogate_model_bytes() { return kModelBytes; }

int surrogate_init(void* model, uint32_t seed, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    ModelPtrs p = ptrs(model);
    hipLaunchKernelGGL(init_kernel, dim3(blocks(kParams, 256)), dim3(256), 0, s, p.params, p.m, p.v, p.stats, seed);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks(kParams, 256)), dim3(256), 0, s, p.grads, kParams);
    hipLaunchKernelGGL(pack_kernel, dim3(blocks(kPackThreads, 256)), dim3(256), 0, s, p.params, p.frags, p.hfrags);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int surrogate_fit_stats(void* model, const float* samples, int n, int stride, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    ModelPtrs p = ptrs(model);
    double* acc = nullptr;
    if (hipMallocAsync((void**)&acc, sizeof(double) * 2 * kIn, s) != hipSuccess) return MM_E_DEVICE;
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, (float*)acc, 4 * kIn);
    hipLaunchKernelGGL(stats_sum_kernel, dim3(256), dim3(64), 0, s, samples, n, stride, acc);
    hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(64), 0, s, acc, n, p.stats);
    (void)hipFreeAsync(acc, s);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int surrogate_train(void* model, const float* samples, int n, int stride, int batch, int steps, float lr,
                    uint32_t seed, float* loss_out, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    ModelPtrs p = ptrs(model);
    float* step_counter = p.stats + 32;
    int t0 = 0;
    if (hipMemcpyAsync(&t0, step_counter, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return MM_E_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MM_E_DEVICE;
    const int waves = batch / 64;
    for (int i = 0; i < steps; i++) {
        float* lacc = loss_out ? loss_out + i : p.stats + 33;
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, lacc, 1);
        hipLaunchKernelGGL(train_kernel, dim3(waves), dim3(64), 0, s, samples, n, stride, p.params, p.stats, p.frags,
                           p.grads, lacc, seed, t0 + i, 0, 1.0f / (float)batch);
        hipLaunchKernelGGL(adam_kernel, dim3(blocks(kParams, 256)), dim3(256), 0, s, p.params, p.grads, p.m, p.v, lr,
                           0.9f, 0.999f, 1e-8f, t0 + i + 1);
        hipLaunchKernelGGL(pack_kernel, dim3(blocks(kPackThreads, 256)), dim3(256), 0, s, p.params, p.frags, p.hfrags);
    }
    int t1 = t0 + steps;
    if (hipMemcpyAsync(step_counter, &t1, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) return MM_E_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MM_E_DEVICE;
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int surrogate_load(void* model, const float* params, const float* stats, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    ModelPtrs p = ptrs(model);
    if (params && hipMemcpyAsync(p.params, params, sizeof(float) * kParams, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return MM_E_DEVICE;
    if (stats && hipMemcpyAsync(p.stats, stats, sizeof(float) * 2 * kIn, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return MM_E_DEVICE;
    hipLaunchKernelGGL(pack_kernel, dim3(blocks(kPackThreads, 256)), dim3(256), 0, s, p.params, p.frags, p.hfrags);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int surrogate_gradient(const void* model, const float* samples, int n, int stride, float* grad_out, float* loss_out,
                       void* stream) {
    hipStream_t s = (hipStream_t)stream;
    ModelPtrs p = ptrs(const_cast<void*>(model));
    hipLaunchKernelGGL(zero_kernel, dim3(blocks(kParams, 256)), dim3(256), 0, s, grad_out, kParams);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, loss_out, 1);
    hipLaunchKernelGGL(train_kernel, dim3(blocks(n, 64)), dim3(64), 0, s, samples, n, stride, p.params, p.stats, p.frags,
                       grad_out, loss_out, 0u, 0, 1, 1.0f / (float)n);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

int surrogate_predict(const void* model, const float* x, int n, int stride, float* out, void* stream) {
    ModelPtrs p = ptrs(const_cast<void*>(model));
    hipLaunchKernelGGL(predict_kernel, dim3(blocks(n, 64)), dim3(64), 0, (hipStream_t)stream, x, n, stride, p.params,
                       p.stats, p.hfrags, out);
    return hipGetLastError() == hipSuccess ? MM_OK : MM_E_DEVICE;
}

const float* surrogate_params(const void* model) { return ptrs(const_cast<void*>(model)).params; }
const float* surrogate_stats(const void* model) { return ptrs(const_cast<void*>(model)).stats; }
const uint4* surrogate_frags(const void* model) { return ptrs(const_cast<void*>(model)).frags; }
const uint4* surrogate_infer_frags(const void* model) { return ptrs(const_cast<void*>(model)).hfrags; }

}  // namespace mm
