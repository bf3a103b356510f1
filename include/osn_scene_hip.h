/*
 * Synthetic depth-frame renderer shipped with the gfx950 TSDF library, for
 * tests and benchmarks. No reference counterpart: the reference's inputs are
 * analytic point sets (cpp/tests/analytic_scenes.hpp:19-46) and point files;
 * this renders the depth images the depth-frame entry points consume.
 */
#ifndef OSN_SCENE_HIP_H
#define OSN_SCENE_HIP_H

#include <stdint.h>

#include "osn_tsdf_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OSN_SCENE_MAX_PRIMS 64

/* An axis-aligned room seen from inside, plus spheres and axis-aligned boxes. */
typedef struct {
    float room_lo[3], room_hi[3];
    int32_t n_spheres, n_boxes;
    float spheres[OSN_SCENE_MAX_PRIMS][4]; /* centre xyz, radius */
    float boxes[OSN_SCENE_MAX_PRIMS][6];   /* lo xyz, hi xyz */
} OsnScene;

/* Render one planar-depth frame (metres, 0 = no return) into d_depth
 * (width*height floats, device), on `stream` (a hipStream_t, NULL = default).
 * Noise: z += N(0, (noise_k * z^2)^2), then each pixel is dropped to 0 with
 * probability `dropout`; both seeded by (seed, frame_index, pixel).
 * Returns 0, 1 on bad arguments, 2 on launch failure. Asynchronous. */
int32_t osn_scene_render_depth(const OsnScene* scene, const OsnTsdfHipFrame* frame,
                               float noise_k, float dropout, uint64_t seed,
                               uint32_t frame_index, float* d_depth, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OSN_SCENE_HIP_H */
