// Synthetic depth frames for tests and benchmarks: closed-form ray casting of
// an axis-aligned room seen from inside, spheres and axis-aligned boxes, with
// an optional depth-dependent Gaussian noise and random dropouts.
//
// The reference has no depth renderer (its inputs are point files and analytic
// point sets, cpp/tests/analytic_scenes.hpp); this produces the depth images
// the depth-frame entry points consume. Pixel convention as the back-projection
// (crates/tartanair/src/lib.rs:323-326): pixel (u, v) looks along
// R_wc * ((u - cx) / fx, (v - cy) / fy, 1), and the stored value is planar
// depth (camera z) in metres, 0 where nothing is hit or the pixel drops out.

#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/osn_tsdf_hip.h"
#include "../../include/osn_scene_hip.h"

namespace {

struct SceneDev {
    float room_lo[3], room_hi[3];
    int32_t n_spheres, n_boxes;
    float sph[OSN_SCENE_MAX_PRIMS][4];     // centre xyz, radius
    float box[OSN_SCENE_MAX_PRIMS][6];     // lo xyz, hi xyz
};

struct CamDev {
    float R[9];
    float t[3];
    float fx, fy, cx, cy;
    int32_t W, H;
    float noise_k, dropout;
    unsigned long long seed;
    uint32_t frame;
};

__device__ inline uint32_t mix32(uint32_t x) {
    // lowbias32 (Wellons)
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ inline float u01(uint32_t h) { return ((h >> 8) + 0.5f) * (1.0f / 16777216.0f); }

__global__ void k_render(SceneDev s, CamDev c, float* out) {
    const int u = blockIdx.x * 16 + (threadIdx.x & 15);
    const int v = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (u >= c.W || v >= c.H) return;
    // ray in camera coords with z = 1, so the hit parameter is planar depth
    const float dc0 = ((float)u - c.cx) / c.fx;
    const float dc1 = ((float)v - c.cy) / c.fy;
    float d[3], o[3];
    for (int i = 0; i < 3; ++i) {
        d[i] = c.R[i * 3 + 0] * dc0 + c.R[i * 3 + 1] * dc1 + c.R[i * 3 + 2];
        o[i] = c.t[i];
    }
    // room, from inside: the nearest exit through any wall
    float best = 3.0e38f;
    for (int i = 0; i < 3; ++i) {
        if (d[i] > 0.0f) best = fminf(best, (s.room_hi[i] - o[i]) / d[i]);
        else if (d[i] < 0.0f) best = fminf(best, (s.room_lo[i] - o[i]) / d[i]);
    }
    for (int k = 0; k < s.n_spheres; ++k) {
        const float ox = o[0] - s.sph[k][0], oy = o[1] - s.sph[k][1], oz = o[2] - s.sph[k][2];
        const float a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const float b = ox * d[0] + oy * d[1] + oz * d[2];
        const float cc = ox * ox + oy * oy + oz * oz - s.sph[k][3] * s.sph[k][3];
        const float disc = b * b - a * cc;
        if (disc < 0.0f) continue;
        const float sq = sqrtf(disc);
        float tt = (-b - sq) / a;
        if (tt <= 1e-4f) tt = (-b + sq) / a;
        if (tt > 1e-4f) best = fminf(best, tt);
    }
    for (int k = 0; k < s.n_boxes; ++k) {
        float t0 = -3.0e38f, t1 = 3.0e38f;
        bool miss = false;
        for (int i = 0; i < 3; ++i) {
            if (d[i] == 0.0f) {
                if (o[i] < s.box[k][i] || o[i] > s.box[k][3 + i]) miss = true;
                continue;
            }
            float a = (s.box[k][i] - o[i]) / d[i];
            float b = (s.box[k][3 + i] - o[i]) / d[i];
            if (a > b) { const float tmp = a; a = b; b = tmp; }
            t0 = fmaxf(t0, a);
            t1 = fminf(t1, b);
        }
        if (miss || t0 > t1) continue;
        const float tt = t0 > 1e-4f ? t0 : t1;
        if (tt > 1e-4f) best = fminf(best, tt);
    }
    float z = best < 1.0e30f ? best : 0.0f;
    if (z > 0.0f && (c.noise_k > 0.0f || c.dropout > 0.0f)) {
        const uint32_t pix = (uint32_t)(v * c.W + u);
        const uint32_t h0 = mix32(pix ^ mix32((uint32_t)c.seed ^ mix32(c.frame * 0x9E3779B9u + (uint32_t)(c.seed >> 32))));
        const uint32_t h1 = mix32(h0 + 0x68E31DA4u);
        const uint32_t h2 = mix32(h1 + 0xB5297A4Du);
        if (c.dropout > 0.0f && u01(h2) < c.dropout) {
            z = 0.0f;
        } else if (c.noise_k > 0.0f) {
            // Box-Muller
            const float r = sqrtf(-2.0f * logf(u01(h0)));
            const float g = r * cosf(6.28318530718f * u01(h1));
            z = z + c.noise_k * z * z * g;
            if (!(z > 0.0f)) z = 0.0f;
        }
    }
    out[(size_t)v * c.W + u] = z;
}

}  // namespace

extern "C" int32_t osn_scene_render_depth(const OsnScene* scene, const OsnTsdfHipFrame* frame,
                                          float noise_k, float dropout, uint64_t seed,
                                          uint32_t frame_index, float* d_depth, void* stream) {
    if (!scene || !frame || !d_depth || frame->width <= 0 || frame->height <= 0) return 1;
    if (scene->n_spheres < 0 || scene->n_spheres > OSN_SCENE_MAX_PRIMS || scene->n_boxes < 0 ||
        scene->n_boxes > OSN_SCENE_MAX_PRIMS)
        return 1;
    SceneDev s{};
    std::memcpy(s.room_lo, scene->room_lo, sizeof(s.room_lo));
    std::memcpy(s.room_hi, scene->room_hi, sizeof(s.room_hi));
    s.n_spheres = scene->n_spheres;
    s.n_boxes = scene->n_boxes;
    std::memcpy(s.sph, scene->spheres, sizeof(float) * 4 * scene->n_spheres);
    std::memcpy(s.box, scene->boxes, sizeof(float) * 6 * scene->n_boxes);
    CamDev c{};
    std::memcpy(c.R, frame->R_wc, sizeof(c.R));
    std::memcpy(c.t, frame->t_wc, sizeof(c.t));
    c.fx = frame->fx; c.fy = frame->fy; c.cx = frame->cx; c.cy = frame->cy;
    c.W = frame->width; c.H = frame->height;
    c.noise_k = noise_k;
    c.dropout = dropout;
    c.seed = seed;
    c.frame = frame_index;
    dim3 grid((c.W + 15) / 16, (c.H + 15) / 16);
    k_render<<<grid, 256, 0, (hipStream_t)stream>>>(s, c, d_depth);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
