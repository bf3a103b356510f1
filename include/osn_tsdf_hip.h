/*
 * C ABI of the gfx950 (MI355X) arm of the hash-blocked TSDF volume.
 *
 * The first group mirrors the reference's device-pointer ABI one for one,
 * under an osn_tsdf_hip_ prefix instead of osn_tsdf_cuda_: same arguments,
 * same meaning, same return codes. Each declaration cites the reference
 * declaration it replaces (cpp/include/osn_tsdf/c_api.h). The second group
 * adds what the reference has no counterpart for: depth-image integration
 * with a projective update and FP16 storage, device-resident extraction,
 * probe statistics and stage timing, and many-scene batches.
 *
 * Conventions (c_api.h:14-19, 88-106; cpp/src/c_api_cuda.cu:188-246):
 *   - positions are DEVICE pointers to 3*n interleaved world-space floats;
 *   - trunc_m <= 0 selects 4 * voxel_size_m; weight_cap <= 0 disables the cap;
 *   - radius_m <= 0 disables the far cull; the camera centre is required;
 *   - int32 status: 0 ok, 1 bad arguments (null volume or pointer, n <= 0),
 *     2 device launch failure (the volume is then marked invalid), 3 the
 *     operation does not apply to this volume's storage or table;
 *   - pool or probe exhaustion is not an error: it increments drop_count;
 *   - calls are asynchronous on the volume's own stream except the getters,
 *     extract_mesh (host buffers) and reset, which synchronise;
 *   - a volume is not safe to use from two host threads at once.
 */
#ifndef OSN_TSDF_HIP_H
#define OSN_TSDF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OsnTsdfHipVolume OsnTsdfHipVolume;

/* Layout-identical to OsnTsdfDeviceView (c_api.h:61-78, 104 bytes), so an
 * external arm written against the reference view can run its own kernels
 * over this library's memory. Only volumes created by osn_tsdf_hip_create
 * (reference hash, weighted-sum storage) can hand one out. */
#ifndef OSN_TSDF_DEVICE_VIEW_DEFINED
#define OSN_TSDF_DEVICE_VIEW_DEFINED
typedef struct {
    void* table;          /* hash_mask+1 entries of {int64 key, int32 block_idx, int32 pad} */
    void* block_count;    /* int32 */
    void* drop_count;     /* uint64 */
    void* block_coord;    /* int32[3] per pool slot */
    void* tsdf;           /* float, weighted sum */
    void* weight;         /* float */
    void* r;
    void* g;
    void* b;
    uint32_t hash_mask;
    uint32_t scratch_base;  /* first of 2^20 sentinel slots past the live table */
    int32_t pool_capacity;
    int32_t block_dim;
    float voxel_size_m;
    float trunc_m;
    float weight_cap;
} OsnTsdfDeviceView;
#endif

/* ---- reference-mirroring entry points ----------------------------------- */

/* c_api.h:80-81 osn_tsdf_cuda_create. NULL on bad parameters or when device
 * allocation fails. Table = next_pow2(2 * pool) slots (cpp/src/volume.cpp:48). */
OsnTsdfHipVolume* osn_tsdf_hip_create(float voxel_size_m, int32_t pool_capacity_blocks,
                                      float trunc_m, float weight_cap);
/* c_api.h:82 osn_tsdf_cuda_destroy */
void osn_tsdf_hip_destroy(OsnTsdfHipVolume* v);
/* c_api.h:83 osn_tsdf_cuda_valid: 1 when usable, 0 otherwise (or NULL) */
int32_t osn_tsdf_hip_valid(const OsnTsdfHipVolume* v);
/* c_api.h:86 osn_tsdf_cuda_device_view. 1 on NULL arguments, 3 when the volume
 * uses a storage or table the reference view cannot describe. */
int32_t osn_tsdf_hip_device_view(const OsnTsdfHipVolume* v, OsnTsdfDeviceView* out);
/* c_api.h:89-91 osn_tsdf_cuda_allocate_blocks */
int32_t osn_tsdf_hip_allocate_blocks(OsnTsdfHipVolume* v, const float* d_positions,
                                     int32_t n_points, float cam_x, float cam_y, float cam_z,
                                     float radius_m);
/* c_api.h:92-94 osn_tsdf_cuda_update_voxels (scatter sums, weighted-sum storage only).
 * The first call allocates the update's scratch (about 1 KiB per pool block, and
 * entries for up to 1.86 M points at the default band); larger calls run in parts. */
int32_t osn_tsdf_hip_update_voxels(OsnTsdfHipVolume* v, const float* d_positions,
                                   int32_t n_points, float cam_x, float cam_y, float cam_z,
                                   float radius_m);
/* c_api.h:97-98 osn_tsdf_cuda_extract_mesh: HOST buffers, 6 floats per
 * vertex, rgb may be NULL. Vertex count, or -1 when vert_cap is too small
 * (nothing written) or on error. The copy goes through 8 MiB of pinned host
 * memory per volume (taken at the first call) and a process-wide pool of up
 * to 7 copy threads. */
int32_t osn_tsdf_hip_extract_mesh(OsnTsdfHipVolume* v, float* buffer_posnor, uint8_t* buffer_rgb,
                                  int32_t vert_cap, float min_weight, float iso);
/* c_api.h:106 osn_tsdf_cuda_reset */
int32_t osn_tsdf_hip_reset(OsnTsdfHipVolume* v);
/* c_api.h:110-112 osn_tsdf_cuda_allocate_counting_cas: compare-exchange
 * ATTEMPTS are added to drop_count (only meaningful when nothing drops). */
int32_t osn_tsdf_hip_allocate_counting_cas(OsnTsdfHipVolume* v, const float* d_positions,
                                           int32_t n_points, float cam_x, float cam_y,
                                           float cam_z, float radius_m);
/* c_api.h:115-117 osn_tsdf_cuda_allocate_no_cas: probe and read, never insert */
int32_t osn_tsdf_hip_allocate_no_cas(OsnTsdfHipVolume* v, const float* d_positions,
                                     int32_t n_points, float cam_x, float cam_y, float cam_z,
                                     float radius_m);
/* c_api.h:120-122 osn_tsdf_cuda_allocate_slot_index. The reference variant
 * exists to price the spin on an unpublished block index; this library never
 * spins (indices are assigned after all keys are in), so the call is the
 * ordinary allocate and the result is identical. */
int32_t osn_tsdf_hip_allocate_slot_index(OsnTsdfHipVolume* v, const float* d_positions,
                                         int32_t n_points, float cam_x, float cam_y,
                                         float cam_z, float radius_m);
/* c_api.h:124 osn_tsdf_cuda_block_count (synchronises) */
int32_t osn_tsdf_hip_block_count(const OsnTsdfHipVolume* v);
/* c_api.h:125 osn_tsdf_cuda_drop_count (synchronises) */
uint64_t osn_tsdf_hip_drop_count(const OsnTsdfHipVolume* v);
/* c_api.h:126 osn_tsdf_cuda_synchronize */
void osn_tsdf_hip_synchronize(const OsnTsdfHipVolume* v);
/* cpp/include/osn_tsdf/volume_cuda.hpp:55 allocate_blocks_probe_depth and
 * cpp/src/volume_cuda.cu:271-300: allocate while recording, per distinct key
 * sent to the table, how many slots it inspected. lane_steps = sum over lanes;
 * wave_steps = sum over wave batches of 64 x the deepest lane (the lockstep
 * cost); hist (may be NULL) receives 64 bins, bin i = depth i+1, bin 63 = 64+.
 * Statistics are cumulative since create/reset. Synchronises. */
int32_t osn_tsdf_hip_allocate_probe_depth(OsnTsdfHipVolume* v, const float* d_positions,
                                          int32_t n_points, float cam_x, float cam_y,
                                          float cam_z, float radius_m, uint64_t* lane_steps,
                                          uint64_t* wave_steps, uint64_t* hist);

/* ---- additions: no reference counterpart -------------------------------- */

enum {
    OSN_TSDF_HIP_STORE_SUMS = 0, /* reference: FP32 weighted sums tsdf, weight, r, g, b */
    OSN_TSDF_HIP_STORE_F32 = 1,  /* projective: FP32 mean tsdf + FP32 weight, 8 B/voxel */
    OSN_TSDF_HIP_STORE_F16 = 2,  /* projective: FP16 mean tsdf/trunc + u16 weight, 4 B/voxel */
    OSN_TSDF_HIP_STORE_F64 = 3   /* projective FP64 oracle: 16 B/voxel */
};
enum {
    OSN_TSDF_HIP_HASH_REFERENCE = 0, /* Teschner mix, power-of-two mask (types.hpp:150-155) */
    OSN_TSDF_HIP_HASH_MIX = 1        /* 64-bit mix + multiply-shift range: any capacity */
};

typedef struct {
    float voxel_size_m;
    int32_t pool_capacity_blocks;
    float trunc_m;          /* <= 0: 4 * voxel */
    float weight_cap;       /* a voxel stops changing once its weight reaches the cap
                               (types.hpp:99-101); <= 0: no cap (projective storage
                               freezes at the u16 limit 65535) */
    int32_t storage;        /* OSN_TSDF_HIP_STORE_* */
    uint32_t table_slots;   /* 0: next_pow2(2 * pool) */
    int32_t hash_kind;      /* OSN_TSDF_HIP_HASH_*; REFERENCE needs a power of two */
    uint32_t probe_limit;   /* slots a lane inspects before giving up; 0: table_slots */
    int32_t lds_set;        /* per-workgroup LDS key set: power of two <= 1024 (larger is clamped), 0 off, -1 default */
    int32_t update_threads; /* projective update workgroup size (128/256/512), 0 default */
    int32_t update_grid;    /* projective update workgroups, 0 default */
    int32_t alloc_grid;     /* allocate workgroups (each loops over pixel tiles), 0 default */
} OsnTsdfHipConfig;

/* NULL on bad parameters or allocation failure. */
OsnTsdfHipVolume* osn_tsdf_hip_create_ex(const OsnTsdfHipConfig* cfg);

/* One depth frame: planar depth in metres (0 = invalid), pinhole intrinsics,
 * and the camera pose. Pose composition happens in FP64 on the host; the FP32
 * world-from-camera rotation and camera centre are what the device uses. The
 * FP64 copies are read only by the FP64 oracle storage. */
typedef struct {
    int32_t width, height;
    float fx, fy, cx, cy;
    float R_wc[9];        /* world-from-camera rotation, row-major */
    float t_wc[3];        /* camera centre in world */
    double R_wc64[9];
    double t_wc64[3];
    float max_range_m;    /* drop pixels deeper than this (> 0 enables) */
    float radius_m;       /* band-walk far cull as in allocate_blocks (> 0 enables) */
} OsnTsdfHipFrame;

/* Allocate the blocks of a depth frame: each valid pixel is back-projected in
 * FP32 (x = (u - cx) * z / fx, ...) and band-walked exactly as
 * allocate_blocks would walk that point; also records the frame's touched
 * blocks for update_depth. d_depth is a device pointer to width*height floats. */
int32_t osn_tsdf_hip_allocate_depth(OsnTsdfHipVolume* v, const float* d_depth,
                                    const OsnTsdfHipFrame* frame);
/* Projective update of the blocks the last allocate_depth touched (projective
 * storage only: 3 otherwise). */
int32_t osn_tsdf_hip_update_depth(OsnTsdfHipVolume* v, const float* d_depth,
                                  const OsnTsdfHipFrame* frame);
/* allocate_depth then update_depth. */
int32_t osn_tsdf_hip_integrate_depth(OsnTsdfHipVolume* v, const float* d_depth,
                                     const OsnTsdfHipFrame* frame);
/* Back-project every pixel to world (FP32, the expressions allocate_depth
 * uses) into d_points_out (3 floats per pixel, NaN for invalid pixels). */
int32_t osn_tsdf_hip_backproject(const float* d_depth, const OsnTsdfHipFrame* frame,
                                 float* d_points_out);
/* Insert block coordinates directly (int32 x,y,z per key, device pointer):
 * the allocate path without the band walk, for allocator stress tests. */
int32_t osn_tsdf_hip_insert_blocks(OsnTsdfHipVolume* v, const int32_t* d_coords, int32_t n);
/* insert_blocks with the allocate measurement modes: 0 plain, 1 count
 * compare-exchange attempts (osn_tsdf_hip_cas_attempts), 2 probe-depth
 * statistics (osn_tsdf_hip_probe_stats). */
int32_t osn_tsdf_hip_insert_blocks_mode(OsnTsdfHipVolume* v, const int32_t* d_coords, int32_t n,
                                        int32_t mode);
/* Total compare-exchange attempts counted since create/reset (synchronises). */
int32_t osn_tsdf_hip_cas_attempts(const OsnTsdfHipVolume* v, uint64_t* out);
/* Uncontended 64-bit compare-exchange ceiling of the device: 2^log2_slots
 * CASes, one per slot in a scattered order, averaged over `reps` timed runs;
 * milliseconds per run into *ms. Allocates and frees its own buffer. */
int32_t osn_tsdf_hip_cas_ceiling(int32_t log2_slots, int32_t reps, double* ms);

/* Distinct blocks the last allocate call visited (synchronises). */
int32_t osn_tsdf_hip_touched_count(const OsnTsdfHipVolume* v);
/* Packed keys of pool blocks [0, min(cap, block_count)) in pool order into a
 * host buffer; returns the count written (synchronises). */
int32_t osn_tsdf_hip_block_keys(const OsnTsdfHipVolume* v, int64_t* host_keys, int32_t cap);
/* Voxel data of pool blocks [first, first+n) to host, 512 per block in
 * (z, y, x) order: tsdf (weighted sum for SUMS storage, mean otherwise) and
 * weight, as float. Returns 0, or 1 on bad range. Synchronises. */
int32_t osn_tsdf_hip_read_voxels(const OsnTsdfHipVolume* v, int32_t first, int32_t n,
                                 float* host_tsdf, float* host_weight);
/* FP64 storage only: means and weights as double. */
int32_t osn_tsdf_hip_read_voxels_f64(const OsnTsdfHipVolume* v, int32_t first, int32_t n,
                                     double* host_tsdf, double* host_weight);
/* Extraction into DEVICE buffers, asynchronous: the vertex count (or -1 on
 * overflow, nothing written) lands in *d_count (device int32). The output
 * order is block (pool) order, voxel (z, y, x) order, tetrahedron order. */
int32_t osn_tsdf_hip_extract_mesh_device(OsnTsdfHipVolume* v, float* d_posnor, uint8_t* d_rgb,
                                         int32_t vert_cap, float min_weight, float iso,
                                         int32_t* d_count);
/* osn_tsdf_hip_extract_mesh_device without any device allocation. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The per-block scratch (neighbour indices, vertex
 * counts, offsets, chunk sums, the scan's arrival counter) lives in one
 * library-wide device array of OSN_TSDF_HIP_NOALLOC_SCRATCH_BYTES on each
 * device, instead of the volume's own, grown on demand; neighbour indices are
 * looked up afresh on every call. Same arguments and results as the
 * allocating call, and byte-identical output; returns 4 (nothing launched)
 * when the volume's pool exceeds OSN_TSDF_HIP_NOALLOC_MAX_BLOCKS blocks.
 * Every call on one device shares the array: calls that overlap on that
 * device (different streams, no synchronisation between them) race. */
#define OSN_TSDF_HIP_NOALLOC_MAX_BLOCKS (1 << 21)
#define OSN_TSDF_HIP_NOALLOC_SCRATCH_BYTES 83894784 /* 40 B per block + chunk sums + counter */
int32_t osn_tsdf_hip_extract_mesh_device_noalloc(OsnTsdfHipVolume* v, float* d_posnor,
                                                 uint8_t* d_rgb, int32_t vert_cap,
                                                 float min_weight, float iso,
                                                 int32_t* d_count);
/* Cumulative probe statistics (see allocate_probe_depth) and the number of
 * distinct keys sent to the table; any pointer may be NULL. Synchronises. */
int32_t osn_tsdf_hip_probe_stats(const OsnTsdfHipVolume* v, uint64_t* hist64,
                                 uint64_t* lane_steps, uint64_t* wave_steps, uint64_t* lookups);

/* Stage timing with device events on the volume's stream. Stages:
 * 0 allocate (band walk and insert), 1 index assignment, 2 update,
 * 3 extract. read_timing synchronises, returns per-stage milliseconds summed
 * since the last read and per-stage call counts, then clears them. */
int32_t osn_tsdf_hip_set_timing(OsnTsdfHipVolume* v, int32_t enable);
int32_t osn_tsdf_hip_read_timing(OsnTsdfHipVolume* v, double* ms4, int64_t* calls4);
/* The volume's stream (a hipStream_t), for callers that order their own work
 * against it. */
void* osn_tsdf_hip_stream(const OsnTsdfHipVolume* v);

/* ---- batches: many independent scenes per launch ------------------------- */

typedef struct OsnTsdfHipBatch OsnTsdfHipBatch;

/* Group volumes (all projective, all the same storage) to be integrated
 * together. The batch runs on its own stream; while it exists the volumes
 * must not be integrated through the single-volume calls. NULL on error. */
OsnTsdfHipBatch* osn_tsdf_hip_batch_create(OsnTsdfHipVolume* const* volumes, int32_t n_volumes);  /* 1..64 volumes */
void osn_tsdf_hip_batch_destroy(OsnTsdfHipBatch* b);
/* Upload the schedule: step k of scene s integrates d_depths[k*n + s] with
 * frames[k*n + s]; a NULL depth pointer leaves the scene idle in that step.
 * Images may differ in size between scenes. Resets the step cursor to 0. */
int32_t osn_tsdf_hip_batch_set_schedule(OsnTsdfHipBatch* b, int32_t n_steps,
                                        const float* const* d_depths,
                                        const OsnTsdfHipFrame* frames);
/* Integrate the next n_steps steps (allocate, assign, projective update for
 * every active scene: three launches per step). With use_graph the per-step
 * chain is captured once in a hipGraph and replayed, the step index living
 * in device memory. Asynchronous; 1 when the schedule would be exceeded. */
int32_t osn_tsdf_hip_batch_run(OsnTsdfHipBatch* b, int32_t n_steps, int32_t use_graph);
void osn_tsdf_hip_batch_synchronize(const OsnTsdfHipBatch* b);
/* Batch launch shape: allocate workgroups, update workgroups and update
 * workgroup size (0 keeps the default). Invalidates a captured graph. */
int32_t osn_tsdf_hip_batch_tune(OsnTsdfHipBatch* b, int32_t alloc_grid, int32_t update_grid,
                                int32_t update_threads, int32_t lds_set);
/* Mesh extraction inside batched steps: after every `every`-th step (step
 * numbers count from the schedule's start; 0 turns extraction off) scene s's
 * triangle soup goes to d_posnor[s] (6 floats per vertex, capacity
 * vert_caps[s] vertices) and its vertex count, or -1 on overflow, to
 * d_counts[s], all device memory owned by the caller. The count, scan and
 * emit passes are captured with the step (a second graph for the steps that
 * extract) and read block counts on the device, so nothing synchronises.
 * Extends osn_tsdf_cuda_extract_mesh (c_api.h:96-98) to the batched, device-
 * resident case; same min_weight/iso meaning. 0 ok, 1 bad arguments, 2 device
 * error. */
int32_t osn_tsdf_hip_batch_set_extract(OsnTsdfHipBatch* b, float* const* d_posnor,
                                       const int32_t* vert_caps, int32_t* d_counts,
                                       float min_weight, float iso, int32_t every);

/* Instrumented batched allocation, the batched form of the reference's
 * measurement variants (cpp/src/volume_cuda.cu:59-61,81 counting CAS,
 * :109-154,271-300 probe depth): mode 0 normal, 1 counts every compare-and-
 * swap attempt (into cas_attempts and, as the reference does, drop_count),
 * 2 records each member volume's probe-depth histogram, lane and wave step
 * totals and distinct-key lookups (read with osn_tsdf_hip_probe_stats; the
 * counters accumulate). Results are unchanged. Invalidates a captured graph.
 * 0 ok, 1 bad arguments, 2 device error. */
int32_t osn_tsdf_hip_batch_set_alloc_mode(OsnTsdfHipBatch* b, int32_t mode);

/* Stage timing for batched steps, as osn_tsdf_hip_read_timing (stages 0-2,
 * and 3 for the extraction after a step; with use_graph the whole step is
 * reported as stage 0). */
int32_t osn_tsdf_hip_batch_set_timing(OsnTsdfHipBatch* b, int32_t enable);
int32_t osn_tsdf_hip_batch_read_timing(OsnTsdfHipBatch* b, double* ms4, int64_t* calls4);

#ifdef __cplusplus
}
#endif

#endif /* OSN_TSDF_HIP_H */
