/* TartanAir V2 ingest for the depth-frame path: pose composition on the host
 * in FP64, depth decoding, and the frame descriptor the device walks.
 *
 * Conventions restated from the reference's loader:
 *   - a pose is one text line of 7 numbers, tx ty tz qx qy qz qw, NED body axes
 *     (x forward, y right, z down) (crates/tartanair/src/lib.rs:170-177,
 *     :379-402);
 *   - the camera frame is x right, y down, z forward, so camera-to-world is
 *     R(q) @ NED_R_CAM with NED_R_CAM = [[0,0,1],[1,0,0],[0,1,0]]
 *     (lib.rs:58,179-190; tools/tartanair_points.py:21-27): body = (Zc, Xc, Yc);
 *   - R(q) normalises q first (lib.rs:193-214);
 *   - a depth PNG is 8-bit RGBA whose bytes taken as B, G, R, A are one
 *     little-endian float32, planar depth in metres (lib.rs:104-112,
 *     tools/tartanair.py:56-69);
 *   - measured image constants 640 x 640, fx = fy = cx = cy = 320 (lib.rs:43-53).
 *
 * Return codes as in osn_io.h: 0 ok; 1 bad arguments; 2 cannot open; 3 bad
 * line (not 7 numbers); 5 caller buffer too small.
 */
#ifndef OSN_TARTANAIR_H
#define OSN_TARTANAIR_H

#include <stdint.h>

#include "osn_tsdf_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OSN_TA_WIDTH 640
#define OSN_TA_HEIGHT 640
#define OSN_TA_FX 320.0
#define OSN_TA_FY 320.0
#define OSN_TA_CX 320.0
#define OSN_TA_CY 320.0

/* Camera-to-world rotation of a NED pose (row-major, FP64). */
int32_t osn_ta_pose_rotation(const double pose[7], double R_wc[9]);

/* The frame descriptor of a NED pose: FP64 composition, FP32 copies for the
 * device, camera centre = the pose's translation. */
int32_t osn_ta_frame(const double pose[7], int32_t width, int32_t height, double fx, double fy,
                     double cx, double cy, float max_range_m, OsnTsdfHipFrame* out);

/* Poses of a trajectory file (blank lines skipped) into caller memory of
 * cap * 7 doubles; *n gets the count found. */
int32_t osn_ta_load_poses(const char* path, double* poses, int32_t cap, int32_t* n);

/* Planar depth from decoded RGBA samples (4 bytes per pixel, file order). */
void osn_ta_depth_from_rgba(const uint8_t* rgba, int64_t n_pixels, float* depth);

#ifdef __cplusplus
}
#endif

#endif /* OSN_TARTANAIR_H */
