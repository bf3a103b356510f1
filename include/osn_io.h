/* On-disk formats and synthetic point clouds of the reference harness.
 *
 * OSNP point cloud (cpp/bench/bench_arms.cu:148-170, tools/tartanair_points.py:
 * 11-17,95-110), little-endian:
 *     char[4] "OSNP"; int32 version = 1; int32 n; float32 cam[3];
 *     float32 points[3 * n]                 (world xyz, interleaved)
 * OSN1 raw mesh dump (bench_arms.cu:506-523, tools/gen_geometry.py:28):
 *     uint32 magic = 0x314E534F ("OSN1"); uint32 n_vertices;
 *     float32 posnor[6 * n_vertices]        (position, normal per vertex, as
 *                                            extraction writes them)
 * Synthetic clouds: the sphere and plane generators of the reference's
 * analytic scenes (cpp/tests/analytic_scenes.hpp:18-44), same FP32 expressions,
 * so the same parameters give bit-identical points.
 *
 * Return codes: 0 ok; 1 bad arguments; 2 cannot open / write; 3 bad header
 * (magic, version, count); 4 truncated file; 5 caller buffer too small.
 */
#ifndef OSN_IO_H
#define OSN_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Header only: point count and camera centre. */
int32_t osn_osnp_info(const char* path, int32_t* n_points, float cam[3]);
/* Points into caller memory of cap_points * 3 floats. */
int32_t osn_osnp_read(const char* path, float* points, int32_t cap_points, int32_t* n_points,
                      float cam[3]);
int32_t osn_osnp_write(const char* path, const float* points, int32_t n_points,
                       const float cam[3]);

int32_t osn_osn1_info(const char* path, int32_t* n_vertices);
int32_t osn_osn1_read(const char* path, float* posnor, int32_t cap_vertices,
                      int32_t* n_vertices);
int32_t osn_osn1_write(const char* path, const float* posnor, int32_t n_vertices);

/* Sphere of `radius` about the origin: n_theta * n_phi points (3 floats each). */
void osn_points_sphere(float radius, int32_t n_theta, int32_t n_phi, float* out);
/* z = z0 plane patch over [-half_extent, half_extent]^2: n * n points. */
void osn_points_plane(float z0, float half_extent, int32_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif /* OSN_IO_H */
