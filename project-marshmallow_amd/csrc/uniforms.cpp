// Caller side of the boundary: derives the three uniform blocks of the cloud dispatch from a scene
// description, restating what VulkanApplication::updateUniformBuffer (VulkanApplication.cpp:351-395)
// computes through Camera (camera.cpp:27-39, 218-238; camera.h:15, 72) and SkyManager
// (SkyManager.cpp:15-70, 108-116). Float/double mixing follows the reference's expressions
// (e.g. `2.0 * PI * (elevation - 0.5)` is a double product narrowed to float).
#include <math.h>
#include <string.h>
#include "mm_internal.hpp"

namespace {

const float kDeg2Rad = 0.01745f;   // camera.h:15 DEG2RAD
const float kPi = 3.14159265f;     // SkyManager.cpp:6
const float kE = 2.718281828459f;  // SkyManager.cpp:7

struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 cross(F3 a, F3 b) {
    return f3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline F3 normalize(F3 a) {
    float s = 1.0f / sqrtf(dot(a, a));
    return f3(a.x * s, a.y * s, a.z * s);
}
inline float clampf_(float t, float lo, float hi) { return fmaxf(lo, fminf(hi, t)); }  // SkyManager.cpp:11-13

bool finite3(const float* v) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }

}  // namespace

extern "C" void mm_scene_default(mm_scene* s) {
    if (!s) return;
    memset(s, 0, sizeof(*s));
    s->camera_pos[0] = 0.f; s->camera_pos[1] = 1.f; s->camera_pos[2] = 1.f;  // VulkanApplication.cpp:76
    s->pitch = 0.f;
    s->yaw = 0.f;
    s->fov_deg = 45.f;
    s->aspect = 1920.0f / 1080.0f;        // camera.h:30
    s->sun_elevation = kPi / 4.f;          // SkyManager.cpp:74-75
    s->sun_azimuth = kPi / 8.f;
    s->time = 0.f;
    s->wind[0] = 1.f; s->wind[1] = 0.05f; s->wind[2] = 1.f;  // SkyManager.cpp:88
    s->rayleigh = 2.f;                     // SkyManager.cpp:97
    s->mie = 0.005f;                       // SkyManager.cpp:95
    s->mie_directional = 0.8f;             // SkyManager.cpp:96
    s->turbidity = 10.f;                   // never initialised by the reference (SkyManager.h:55)
    s->phase = 0;
}

extern "C" void mm_scene_scenario(mm_scene* s, float altitude, float pitch_deg, float fov_deg, float aspect,
                                  float sun_elevation_deg, float time) {
    if (!s) return;
    mm_scene_default(s);
    s->camera_pos[0] = 0.f;
    s->camera_pos[1] = altitude;
    s->camera_pos[2] = 0.f;
    // The shader looks along -forward (comp:312, 318); forward = (cos yaw cos pitch, sin pitch,
    // sin yaw cos pitch) (camera.cpp:223). Looking up by pitch_deg towards -z needs forward =
    // (0, -sin, cos): pitch = -pitch_deg (converted with the reference's DEG2RAD), yaw = pi/2.
    s->pitch = -pitch_deg * kDeg2Rad;
    s->yaw = 1.57079632679489662f;
    s->fov_deg = fov_deg;
    s->aspect = aspect;
    // SkyManager parametrisation (SkyManager.cpp:46-54): azimuth 0.25 gives phi = -pi/2 and
    // dir = (0, -sin theta, -cos theta); theta = 2 pi (e - 0.5) = -elevation puts the sun at
    // `sun_elevation_deg` above the horizon towards -z.
    s->sun_elevation = 0.5f - sun_elevation_deg / 360.0f;
    s->sun_azimuth = 0.25f;
    s->time = time;
}

extern "C" int mm_build_ubos(const mm_scene* s, mm_camera_ubo* cam, mm_sun_ubo* sun, mm_sky_ubo* sky) {
    if (!s || !cam || !sun || !sky) return MM_E_NULL;
    if (!finite3(s->camera_pos) || !isfinite(s->pitch) || !isfinite(s->yaw) || !isfinite(s->fov_deg) ||
        !isfinite(s->aspect) || !isfinite(s->sun_elevation) || !isfinite(s->sun_azimuth) || !isfinite(s->time) ||
        !finite3(s->wind))
        return MM_E_PARAM;
    if (s->fov_deg <= 0.f || s->fov_deg >= 180.f || s->aspect <= 0.f) return MM_E_PARAM;
    if (s->phase < 0 || s->phase > 15) return MM_E_PHASE;

    // ---- Camera (camera.cpp:218-238 forward/right/up; camera.cpp:27-39 view) ----
    const float cp = cosf(s->pitch), sp = sinf(s->pitch), cy = cosf(s->yaw), sy = sinf(s->yaw);
    F3 fwd = f3(cy * cp, sp, sy * cp);
    F3 wUp = (1.0f - fabsf(dot(fwd, f3(0, 1, 0))) < 0.00001f) ? f3(0, 0, 1) : f3(0, 1, 0);
    F3 right = normalize(cross(fwd, wUp));
    F3 up = normalize(cross(right, fwd));
    const float* P = s->camera_pos;
    memset(cam, 0, sizeof(*cam));
    float* V = cam->view;  // column-major: V[col * 4 + row]
    V[0] = right.x; V[1] = up.x; V[2] = fwd.x; V[3] = 0.f;
    V[4] = right.y; V[5] = up.y; V[6] = fwd.y; V[7] = 0.f;
    V[8] = right.z; V[9] = up.z; V[10] = fwd.z; V[11] = 0.f;
    // view * translate(-pos): column 3 = V0 * -px + V1 * -py + V2 * -pz + V3 * 1 (glm mat4 product order)
    for (int r = 0; r < 3; r++) V[12 + r] = ((V[r] * -P[0] + V[4 + r] * -P[1]) + V[8 + r] * -P[2]) + 0.f * 1.f;
    V[15] = 1.f;
    // glm::perspective(DEG2RAD * fov, aspect, 0.1, 1000) (right-handed, -1..1 depth), then [1][1] *= -1
    // (VulkanApplication.cpp:361-362).
    const float zn = 0.1f, zf = 1000.0f;
    const float th = tanf(kDeg2Rad * s->fov_deg / 2.0f);
    float* Pm = cam->proj;
    Pm[0] = 1.0f / (s->aspect * th);
    Pm[5] = -(1.0f / th);
    Pm[10] = -(zf + zn) / (zf - zn);
    Pm[11] = -1.0f;
    Pm[14] = -(2.0f * zf * zn) / (zf - zn);
    cam->cameraPosition[0] = P[0]; cam->cameraPosition[1] = P[1]; cam->cameraPosition[2] = P[2];
    cam->cameraPosition[3] = 1.0f;
    cam->cameraParams[0] = s->aspect;
    cam->cameraParams[1] = tanf(0.5f * kDeg2Rad * s->fov_deg);  // camera.h:72 getHTanFov

    // ---- SkyManager::calcSunPosition (SkyManager.cpp:46-70) ----
    memset(sun, 0, sizeof(*sun));
    const float theta = (float)(2.0 * (double)kPi * ((double)s->sun_elevation - 0.5));
    const float phi = (float)(2.0 * (double)kPi * ((double)s->sun_azimuth - 0.5));
    F3 dir = f3(cosf(phi), sinf(phi) * sinf(theta), sinf(phi) * cosf(theta));
    if (dot(dir, dir) == 0.0f) return MM_E_PARAM;
    const bool below = dir.y < 0.0f;
    sun->direction[0] = dir.x; sun->direction[1] = dir.y; sun->direction[2] = dir.z; sun->direction[3] = 0.f;
    const float sd = 400000.0f;  // SUN_DISTANCE
    sun->location[0] = sd * dir.x; sun->location[1] = sd * dir.y; sun->location[2] = sd * dir.z;
    sun->location[3] = 1.0f;
    if (below)
        for (int k = 0; k < 4; k++) sun->location[k] *= -1.0f;
    F3 rv = (fabsf(dir.y) < 0.001f) ? ((dir.z < 0) ? f3(0, 1, 0) : f3(0, -1, 0)) : f3(0, 0, 1);
    F3 dirN = normalize(cross(dir, rv));
    F3 dirB = normalize(cross(dir, dirN));
    float* B = sun->directionBasis;
    B[0] = dirN.x; B[1] = dirN.y; B[2] = dirN.z; B[3] = 0.f;
    B[4] = dir.x; B[5] = dir.y; B[6] = dir.z; B[7] = 0.f;
    B[8] = dirB.x; B[9] = dirB.y; B[10] = dirB.z; B[11] = 0.f;
    B[12] = 0.f; B[13] = 0.f; B[14] = 0.f; B[15] = 1.f;
    if (below)
        for (int k = 0; k < 16; k++) B[k] *= -1.0f;

    // ---- calcSunIntensity (SkyManager.cpp:28-34) ----
    const float zc = clampf_(dir.y, -1.f, 1.f);
    sun->intensity = 1000.0f * fmaxf(0.f, 1.f - powf(kE, -((1.6110731557f - acosf(zc)) / 1.5f)));
    if (below) sun->intensity = 2.0f;

    // ---- calcSunColor (SkyManager.cpp:15-26); alpha carries the pixel phase ----
    if (below) {
        sun->color[0] = 0.8f; sun->color[1] = 0.9f; sun->color[2] = 1.0f;
    } else {
        const float sunset[3] = {2.f, (float)0.33922, 0.0431f};
        float t = clampf_(dir.y * 13.f, 0.f, 1.f);
        for (int k = 0; k < 3; k++) sun->color[k] = (1.f - t) * sunset[k] + t * 1.0f;
    }
    sun->color[3] = (float)s->phase;

    // ---- calcSkyBetaR / calcSkyBetaV (SkyManager.cpp:36-44) ----
    memset(sky, 0, sizeof(*sky));
    const float RT[3] = {(float)5.804542996261093E-6, (float)1.3562911419845635E-5, (float)3.0265902468824876E-5};
    const float MC[3] = {1.839991851443397f, 2.779802391966052f, 4.079047954386109f};
    const float sunFade = 1.0f - clampf_((float)(1.0 - exp((double)sun->location[1] / 450000.0)), 0.0f, 1.0f);
    for (int k = 0; k < 3; k++) sky->betaR[k] = RT[k] * (s->rayleigh - 1.f + sunFade);
    const float c = (0.2f * s->turbidity) * 10E-18f;
    for (int k = 0; k < 3; k++) sky->betaV[k] = ((0.434f * c) * MC[k]) * s->mie;
    sky->wind[0] = s->wind[0]; sky->wind[1] = s->wind[1]; sky->wind[2] = s->wind[2];
    sky->wind[3] = s->time;
    sky->mie_directional = s->mie_directional;
    return MM_OK;
}
