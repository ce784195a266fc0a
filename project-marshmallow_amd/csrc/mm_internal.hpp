// Internal declarations shared by the C-ABI layer and the HIP launchers.
#pragma once
#include <mutex>
#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime.h>
#include "../../include/mm.h"

namespace mm {

// Per-dispatch constants derived once on the host from the uniform blocks, exactly as the shader
// derives them per invocation (compute-clouds.comp:289-401). Passed by value as a kernel argument.
struct MarchConsts {
    float camPos[3];
    float camLook[3], camRight[3], camUp[3];
    float refPoint[3];           // cameraPos - camLook (comp:318)
    float aspect, tanH;          // cameraParams.xy
    float sunDir[3];             // normalize(directionBasis[1].xyz) (comp:324)
    float coneOffsets[6][3];     // basis * cone table (comp:392-401)
    float earthCenter[3];        // (cam.x, center_y, cam.z) (comp:357-358)
    float innerDiam, outerDiam, thickness, invThickness, projRadius;
    float stepSize0;             // coarse_step_frac * thickness (comp:390)
    float fineMul;
    float betaR[3], betaM[3];
    float sunIntensity;
    float sunColor[3];
    float sunDirY;               // sun.direction.y (day/night switch, comp:336, 344, 365)
    float mieG;
    float wind[3], time, windStrength;
    float weatherScale, lowScale, curlScale, hiScale;
    float opacityCutoff;
    float nightRot[3][3];        // fromAngleAxis(normalize(1,0,1), sun.direction.y * 0.5) columns
    int missLimit, maxSteps, maxCoarse, maxFine;
    int flags;
    int subtexelBits;
    int width, height, pitch;    // image size used for uv and bounds, row pitch in pixels
    int phase;                   // -1 = all pixels
    int gridW, gridH;            // marched pixel grid
    int tileShift;                // tiles of (1 << tileShift) x (64 >> tileShift) pixels: 3, 4 or 5
    int lanesPerRay;              // 1, or 2 / 4 / 8 for the ray-parallel kernel (tiles of 8 x (8 / lanesPerRay) rays)
    int tilesX, tilesY;           // tiles of the marched grid
    int tiles;                    // tilesX * tilesY: one block per tile
    int rowStart, rowDir;         // launch order of tile rows: logical row r is row (rowStart + rowDir r) mod tilesY
    float probeOffsets[3][3];     // basis * (0, k, 0), k = 0.6, 2.5, 6: surrogate density probes
    // light-cone mode: MM_CONE_PLAIN, MM_CONE_SURROGATE, or 2 = plain + record surrogate training samples
    int coneMode;
    float surrThreshold;
    const float* surrParams;      // FP32 parameters of the surrogate MLP
    const float* surrStats;       // feature mean[16], inverse std[16]
    const uint4* surrFrags;       // FP16 MFMA inference fragments (surr::kHF*)
    int countSlots;               // 4, or 6: counts[4], counts[5] also receive the weather / low-res reads
    float* stateOut;              // lock-step kernel only, or null: per pixel (final opacity accumulator,
                                  // light-energy accumulator), 2 floats at the image's pixel index
    float* collectBuf;            // [collectCap][17]: 16 features + density along the light
    unsigned* collectCount;
    unsigned collectCap, collectKeep, collectSeed;
    // tile schedule (mm_march_clouds_scheduled), both null otherwise: the physical tile each block
    // marches (a permutation of 0 .. tiles - 1) and the schedule record (kSchedHeader words of order
    // state, then each physical tile's cost for the next launch)
    const int* tileOrder;
    unsigned* tileSched;
};

// A tile schedule's device record: kOrderBuckets per-bucket tile counts, kOrderBuckets per-bucket
// next slots, then one cost per tile (from kSchedHeader on).
constexpr int kOrderBuckets = 1024;
constexpr int kSchedCount = 0, kSchedCursor = kOrderBuckets, kSchedHeader = 2 * kOrderBuckets;

// Device view of one texture.
struct TexView {
    const uint8_t* data;
    int w, h, d;                 // logical dims
    int stride;                  // texels per row including the apron (w or w + 1)
    int sliceStride;             // texels per slice including the apron
    float scale[4];
    int lw, lh;                  // log2 of w and h for power-of-two dims (else 0)
};

enum FormatSet { FSET_REF = 0, FSET_FAST = 1, FSET_PREC = 2 };

struct MarchTextures {
    TexView weather, curl, low, hi, night;
    int hasNight;
};

// Constants of one reprojection pass (reproject.comp:91-153), derived on the host.
struct ReprojConsts {
    float camPos[3], camRight[3], camUp[3], refPoint[3];
    float aspect, tanH;
    float prevView[16];
    float centerY, innerDiam;
    int fp64;
    int width, height, pitch;
    int skipPhase;                // -1, or the 1/16 pixel phase left to a march that follows
};
int launch_reproject(const ReprojConsts& c, const float* src, float* dst, void* stream);

// A second stream of this device for work beside the caller's stream. side_fork (under side_lock)
// returns it ordered after the work already on `main`, or nullptr (then use `main`); side_join makes
// `main` wait for what was enqueued on it.
std::unique_lock<std::mutex> side_lock();
hipStream_t side_fork(hipStream_t main);
int side_join(hipStream_t main, hipStream_t side);

// Constants of one cloud-shadow pass (model.frag:240-283), derived on the host.
struct ShadowConsts {
    float L[3];                   // view-space sun direction, flipped when L.y < -0.05
    float sunW[3];                // sun.directionBasis[1].xyz (world, not normalised)
    float camPos[3];
    float wind[3], time;
    float centerY, innerDiam, thickness, projRadius, step, originScale, windStrength, weatherScale, lowScale;
    int steps, subtexelBits, fp64, n;
};
int launch_shadow(int fset, const ShadowConsts& c, const TexView& weather, const TexView& low, const float* pts,
                  float* out, void* stream);

// Constants of one post-chain pass (god-ray.frag, radialBlur.frag, tonemap.frag), derived on the host.
struct PostConsts {
    int width, height, pitch, bits, samples, below;
    float sunX, sunY;             // the sun in NDC: ((proj * view) * location).xy / w
    float k;                      // (1 / samples) * density
    float sw;                     // 1 / samples
    float decay, godExposure, blurScale, blurGain;
    float gc[3];                  // sun.color.rgb * sun.intensity
    float exposure, invGamma, whitemap, vignette[3];
};
int launch_post(const PostConsts& c, const float* image, float* scratch, float* hdr, uint8_t* rgba8, void* stream);

// One view of a multi-view launch (device array, built on the host).
struct ViewDesc {
    MarchConsts c;
    TexView weather;
    float* out;
    int tileBase;                 // global index of this view's first tile
};

// The march's step checks reduce to the two budget tests (the kernels' specialised loop) when the
// march is bounded with no counted-step limit and both budgets set.
inline bool bounded_steps(uint32_t flags, int maxSteps, int maxCoarse, int maxFine) {
    return (flags & MM_MARCH_BOUNDED) != 0 && maxSteps < 0 && maxCoarse >= 0 && maxFine >= 0;
}

// order: null (launch order) or a permutation of the plan's global tiles; sched: null or the plan's
// schedule record (kSchedHeader words, then one cost slot per global tile for the next launch's order).
int launch_views(int fset, int mode, int lanesPerRay, bool floatWeights, bool boundedSteps, const ViewDesc* views,
                 int nViews, const MarchTextures& t, int total, unsigned long long* counts, const int* order,
                 unsigned* sched, void* stream);

size_t surrogate_model_bytes();
int surrogate_init(void* model, uint32_t seed, void* stream);
int surrogate_fit_stats(void* model, const float* samples, int n, int stride, void* stream);
int surrogate_train(void* model, const float* samples, int n, int stride, int batch, int steps, float lr,
                    uint32_t seed, float* loss_out, void* stream);
int surrogate_predict(const void* model, const float* x, int n, int stride, float* out, void* stream);
int surrogate_load(void* model, const float* params, const float* stats, void* stream);
int surrogate_gradient(const void* model, const float* samples, int n, int stride, float* grad_out, float* loss_out,
                       void* stream);
const float* surrogate_params(const void* model);
const float* surrogate_stats(const void* model);
const uint4* surrogate_frags(const void* model);
const uint4* surrogate_infer_frags(const void* model);  // FP16 inference fragments (surr::kHF*)

int launch_march(int fset, const MarchConsts& c, const MarchTextures& t, float* out, unsigned long long* counts,
                 void* stream);
// order <- the n physical tiles of a schedule record sorted by cost, largest first (bucketed: 1024 cost
// buckets), in three launches on `stream`. The record's counts must be clear (every scheduled march
// clears them for the next order).
int launch_tile_order(unsigned* sched, int n, int* order, void* stream);

// Noise generation launchers.
// allow_side = false keeps every generator on `stream` (no second stream is created or used)
int launch_generate(const mm_noise_params* p, const mm_noise_set* out, void* workspace, void* stream,
                    bool allow_side = true);
size_t noise_workspace_bytes();
size_t static_scratch_bytes();
int static_scratch(void** ptr);
int launch_pack(int kind, const mm_texture* src, const mm_texture* dst, void* stream);

bool texture_dims_ok(const mm_texture* t, int dims);

}  // namespace mm
