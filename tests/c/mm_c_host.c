/* A plain C host of the C ABI (include/mm.h): the way the reference's C++ application would call the
 * library, with no Python or PyTorch in the process.
 *
 *   mm_c_host check           calls that need no GPU: version, error strings, texture sizes, presets,
 *                             the uniform builder and argument validation; exits 0 when all hold
 *   mm_c_host run W H OUT     on a GPU: allocates the FAST texture set with hipMalloc, generates the
 *                             noise (seed 1, weather seed 1, coverage 0.5), marches the W x H scenario
 *                             frame on a stream, and writes the RGBA32F image to OUT
 */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mm.h"

#define EXPECT(cond)                                                     \
    do {                                                                 \
        if (!(cond)) {                                                   \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            return 1;                                                    \
        }                                                                \
    } while (0)

static int check(void) {
    EXPECT(mm_abi_version() == MM_ABI_VERSION);
    EXPECT(strcmp(mm_strerror(MM_OK), "ok") == 0);
    EXPECT(mm_strerror(MM_E_IO) != NULL);
    EXPECT(mm_texture_bytes(128, 128, 128, MM_FMT_RGBA8, MM_LAYOUT_LINEAR) == (size_t)128 * 128 * 128 * 4);
    EXPECT(mm_texture_bytes(32, 32, 32, MM_FMT_RGBA8, MM_LAYOUT_APRON) == (size_t)33 * 33 * 33 * 4);
    EXPECT(sizeof(mm_camera_ubo) == 160 && sizeof(mm_sun_ubo) == 128 && sizeof(mm_sky_ubo) == 64);
    mm_march_params p;
    mm_march_params_reference(&p);
    EXPECT(p.size == sizeof(mm_march_params));
    mm_scene s;
    mm_scene_scenario(&s, 200.0f, 20.0f, 60.0f, 16.0f / 9.0f, 25.0f, 0.0f);
    mm_camera_ubo cam;
    mm_sun_ubo sun;
    mm_sky_ubo sky;
    EXPECT(mm_build_ubos(&s, &cam, &sun, &sky) == MM_OK);
    EXPECT(sun.direction[1] > 0.0f);  /* 25 degrees above the horizon */
    /* validation happens before any device work */
    EXPECT(mm_march_clouds(NULL, &sun, &sky, NULL, NULL, NULL, 64, 64, 64, 0, NULL, NULL) == MM_E_NULL);
    int32_t w, h;
    EXPECT(mm_image_info("/nonexistent/file.tga", &w, &h) == MM_E_IO);
    mm_post_params pp;
    mm_post_params_reference(&pp);
    EXPECT(pp.god_ray_samples == 8 && pp.subtexel_bits == 8);
    return 0;
}

static int texture(mm_texture* t, int w, int h, int d, int fmt, int layout) {
    memset(t, 0, sizeof(*t));
    t->width = w;
    t->height = h;
    t->depth = d;
    t->format = fmt;
    t->layout = layout;
    t->bytes = mm_texture_bytes(w, h, d, fmt, layout);
    for (int k = 0; k < 4; k++) t->scale[k] = 1.0f;
    return hipMalloc(&t->data, t->bytes) == hipSuccess ? 0 : 1;
}

static int run(int W, int H, const char* path) {
    hipStream_t stream;
    EXPECT(hipStreamCreate(&stream) == hipSuccess);
    mm_noise_set set;
    EXPECT(!texture(&set.weather, 512, 512, 1, MM_FMT_RG8, MM_LAYOUT_APRON));
    EXPECT(!texture(&set.curl, 128, 128, 1, MM_FMT_RGBA8, MM_LAYOUT_APRON));
    EXPECT(!texture(&set.low_res, 128, 128, 128, MM_FMT_RGBA8, MM_LAYOUT_APRON));
    EXPECT(!texture(&set.hi_res, 32, 32, 32, MM_FMT_RGBA8, MM_LAYOUT_APRON));
    mm_noise_params np;
    mm_noise_params_default(&np);
    np.seed = 1;
    np.weather_seed = 1;
    np.coverage = 0.5f;
    int e = mm_generate_noise(&np, &set, NULL, stream);
    if (e) { fprintf(stderr, "mm_generate_noise: %s\n", mm_strerror(e)); return 1; }
    mm_scene s;
    mm_scene_scenario(&s, 200.0f, 20.0f, 60.0f, (float)W / (float)H, 25.0f, 0.0f);
    mm_camera_ubo cam;
    mm_sun_ubo sun;
    mm_sky_ubo sky;
    EXPECT(mm_build_ubos(&s, &cam, &sun, &sky) == MM_OK);
    mm_march_params p;
    mm_march_params_scenario(&p);
    float* out;
    const size_t bytes = (size_t)W * H * 4 * sizeof(float);
    EXPECT(hipMalloc((void**)&out, bytes) == hipSuccess);
    e = mm_march_clouds(&cam, &sun, &sky, &set, NULL, out, W, H, W, -1, &p, stream);
    if (e) { fprintf(stderr, "mm_march_clouds: %s\n", mm_strerror(e)); return 1; }
    EXPECT(hipStreamSynchronize(stream) == hipSuccess);
    float* host = (float*)malloc(bytes);
    EXPECT(host && hipMemcpy(host, out, bytes, hipMemcpyDeviceToHost) == hipSuccess);
    FILE* f = fopen(path, "wb");
    EXPECT(f && fwrite(host, 1, bytes, f) == bytes);
    fclose(f);
    free(host);
    hipFree(out);
    hipFree(set.weather.data);
    hipFree(set.curl.data);
    hipFree(set.low_res.data);
    hipFree(set.hi_res.data);
    hipStreamDestroy(stream);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && strcmp(argv[1], "check") == 0) return check();
    if (argc >= 5 && strcmp(argv[1], "run") == 0) return run(atoi(argv[2]), atoi(argv[3]), argv[4]);
    fprintf(stderr, "usage: %s check | run W H OUT\n", argv[0]);
    return 2;
}
