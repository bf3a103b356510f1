/* Mesh-comparison gate: area-weighted surface sampling, point-to-triangle
 * symmetric distances, Chamfer / Hausdorff statistics and the pass/fail gate.
 *
 * Restates the reference's mesh-metrics crate (crates/mesh-metrics/src/lib.rs),
 * which its correctness gates use to compare an arm's triangle soup with a
 * golden mesh: same seeded xorshift sampler, same closest-point-on-triangle
 * distance, same statistics and default tolerances, so the same inputs give
 * the same numbers. Host code (no GPU), multi-threaded over the samples.
 *
 * Meshes are triangle soups: 3 consecutive vertices per triangle, `stride`
 * floats per vertex with x, y, z first (3 for bare positions, 6 for the
 * position + normal records osn_tsdf_hip_extract_mesh writes).
 */
#ifndef OSN_MESH_METRICS_H
#define OSN_MESH_METRICS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Distances from one sample set to the other mesh (lib.rs:315-325). */
typedef struct {
    double mean, rms, p95, p99;
    double max;    /* one-sided Hausdorff */
    int64_t n;
} OsnDistanceStats;

/* Symmetric agreement (lib.rs:356-368). */
typedef struct {
    OsnDistanceStats a_to_b, b_to_a;
    double mean_surface;   /* (a_to_b.mean + b_to_a.mean) / 2: the Chamfer mean */
    double hausdorff;      /* max of the two maxima */
    double hausdorff_p99;  /* max of the two 99th percentiles */
} OsnMeshAgreement;

/* Gate thresholds (lib.rs:394-416); osn_mesh_tolerance_for_voxel gives the
 * reference defaults: mean 0.1 voxel, p99 one voxel, vertex count 2 %. */
typedef struct {
    double mean_surface_m;
    double hausdorff_p99_m;
    double vertex_count_rel;
} OsnMeshTolerance;

enum {
    OSN_GATE_FAIL_MEAN = 1,
    OSN_GATE_FAIL_P99 = 2,
    OSN_GATE_FAIL_COUNT = 4,
};

typedef struct {
    OsnMeshAgreement agreement;
    int64_t n_vertices_a, n_vertices_b;
    double vertex_count_rel;
    int32_t passed;    /* 1 when no threshold is exceeded */
    int32_t failures;  /* OSN_GATE_FAIL_* bits */
} OsnGateResult;

OsnMeshTolerance osn_mesh_tolerance_for_voxel(double voxel_m);

/* Seeded area-weighted sampling of n_samples surface points (lib.rs:102-150)
 * into out_xyz (3 doubles each). A mesh of zero total area yields its vertices
 * instead (at most n_vertices points). *n_out receives the count. 0 ok, 1 bad
 * arguments. */
int32_t osn_mesh_sample_surface(const float* xyz, int64_t n_vertices, int32_t stride,
                                int64_t n_samples, uint64_t seed, double* out_xyz,
                                int64_t out_cap, int64_t* n_out);

/* Squared distance from each query point (3 doubles each) to the nearest
 * triangle of the mesh, through the spatial hash of lib.rs:230-300 (`cell`
 * sizes the hash and changes speed only). threads <= 0: all host threads. */
int32_t osn_mesh_nearest_sq(const float* xyz, int64_t n_vertices, int32_t stride,
                            const double* queries, int64_t n_queries, double cell,
                            int32_t threads, double* out_sq);

/* compare() of lib.rs:370-392: n_samples per side, sides seeded with seed and
 * seed ^ 0x9E3779B97F4A7C15. */
int32_t osn_mesh_compare(const float* xyz_a, int64_t n_vertices_a, int32_t stride_a,
                         const float* xyz_b, int64_t n_vertices_b, int32_t stride_b,
                         int64_t n_samples, double cell, uint64_t seed, int32_t threads,
                         OsnMeshAgreement* out);

/* gate() of lib.rs:427-463 (seed 0xC0FFEE). */
int32_t osn_mesh_gate(const float* xyz_a, int64_t n_vertices_a, int32_t stride_a,
                      const float* xyz_b, int64_t n_vertices_b, int32_t stride_b,
                      OsnMeshTolerance tol, int64_t n_samples, double cell, int32_t threads,
                      OsnGateResult* out);

#ifdef __cplusplus
}
#endif

#endif /* OSN_MESH_METRICS_H */
