// Host-side launch interface between the C ABI layer and the kernels.
#pragma once

#include <hip/hip_ext.h>

#include "upd_params.hpp"

namespace osn_hip {

// Where a batch of points comes from: world-space points (the reference
// shape), or a depth image back-projected on the fly.
struct PointSrc {
    const float* pos;      // 3*n interleaved, device
    int32_t n;
    const float* depth;    // W*H planar depth in metres, device; 0 = invalid
    float* tiled;          // when set, the depth is also copied here in 8x4 tiles
    int32_t W, H;
    float fx, fy, cx, cy;
    float R[9];            // world-from-camera rotation, row-major (FP32)
    float max_range;       // drop pixels with depth > max_range (> 0 enables)
    float cam[3];          // camera centre in world (the band walk's origin)
    float radius_sq;       // band-walk far cull, 0 disables
};

// Which call this is: the stamp marks blocks touched by it, the parity picks
// the list-counter set it appends to.
struct CallInfo {
    int32_t stamp;
    int32_t parity;
};

enum AllocMode : int32_t {
    kAllocNormal = 0,
    kAllocCountCas = 1,  // tally every compare-exchange attempt
    kAllocProbe = 2,     // probe-depth histogram and lane/wave step totals
    kAllocNoCas = 3,     // probe and read only, never insert
};

struct AllocOpts {
    int32_t mode = kAllocNormal;
    bool track_touched = false;
    int32_t lds_set = 1024;   // per-workgroup LDS key-set entries (0: every batch to the table)
    int32_t grid = 0;         // workgroups; 0 = one per tile up to a cap
};

// One scene of a batched step: its volume, its frame (allocation source and
// projection camera) and whether it has a frame this step.
struct SceneStep {
    PointSrc src;
    ProjCam cam;
    int32_t active;
    int32_t tiles;      // depth work units of this scene's image (the batch's depth_sub)
};

// Depth images are allocated in work units of sub x sub tiles of 16x16
// pixels, one workgroup each. Batches use kBatchDepthSub, unless the whole
// batch would then have fewer than kBatchMinUnits units (a single scene or a
// few small ones): a unit per 16x16 tile then, so the launch still fills the
// chip.
constexpr int kBatchDepthSub = 2;
constexpr int kBatchMinUnits = 2048;
__host__ __device__ inline int32_t depth_units(int32_t W, int32_t H, int32_t sub) {
    const int32_t span = 16 * sub;
    return ((W + span - 1) / span) * ((H + span - 1) / span);
}

constexpr int kMaxBatchScenes = 64;  // volumes per batch

// A batched step over many scenes, read from device memory so that a captured
// graph can replay it: scene s of step k is steps[k * n_scenes + s], with k
// read from *step_ctr when step_ctr is set, else `step`. When step_advance is
// set, the step's update kernel (which runs after every reader of the
// counter in the step) adds one to it for the next replay.
struct BatchDev {
    const VolDev* vols;
    const SceneStep* steps;
    const int32_t* tile_prefix;  // n_scenes + 1, exclusive prefix of tiles
    const int32_t* step_ctr;
    int32_t* step_advance;
    const int32_t* call_base;    // per scene: calls the volume had made before step 0
    UpdParams* upd;              // per scene: the update's parameters, written each step
    int32_t step;
    int32_t n_scenes;
    int32_t depth_sub;           // work units of depth_sub x depth_sub 16x16 tiles
};

// Call number c of a volume (0-based) stamps its blocks with c + 1 and
// appends to list set c & 1; step k of a batch is call call_base[s] + k.
__device__ inline CallInfo batch_call(const BatchDev& b, int sc, int k) {
    const int32_t c = b.call_base[sc] + k;
    return CallInfo{c + 1, c & 1};
}

// Extraction scans the per-block vertex counts in chunks of this many blocks
// (chunk_sums holds one total per chunk).
constexpr int kScanChunk = 1024;

struct ExtractArgs {
    float min_weight;
    float iso;
    int32_t n_blocks;      // blocks to walk (<= block_count), -1 = read block_count
    int32_t* block_vcount; // per block vertex counts
    int32_t* block_voff;   // per block exclusive offsets (emit reads)
    int32_t* block_nb;     // per block the pool indices of its 8 neighbours (count writes)
    int32_t* chunk_sums;   // per 1024-block chunk sums
    int32_t* scan_flag;    // arrival counter for the chunk scan
    float* posnor;         // device output, 6 floats per vertex
    uint8_t* rgb;          // device output, 3 bytes per vertex, may be null
    int32_t cap;           // output capacity in vertices
    int32_t* count_out;    // device int32: vertex count or -1 on overflow
    int32_t nb_cache;      // 1: block_nb is the volume's own neighbour cache (reuse the
                           // indices of blocks below Counters.nb_blocks, then advance it);
                           // 0: look every neighbour up and leave nb_blocks alone
    int32_t clear_flag;    // 1: the count pass zeroes scan_flag before the scan reads it
};

// Index assignment works in tiles of kSortTile new keys: level 0 sorts each
// tile, level l >= 1 merges pairs of sorted runs of kSortTile << (l - 1) keys,
// one kSortTile-output chunk per work item.
constexpr int kSortTile = 4096;
constexpr int kAssignMaxGrid = 256;

// Optional events of one kernel launch. When set, the launch records them
// itself (hipExtLaunchKernel): they mark its own dispatch start and end, not
// separate stream commands around it, so their interval is the kernel's.
struct LaunchEvents {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};

template <typename... KArgs, typename... Args>
inline void launch_kernel(void (*k)(KArgs...), dim3 grid, dim3 block, hipStream_t s,
                          const LaunchEvents& ev, Args... args) {
    if (ev.start)
        hipExtLaunchKernelGGL(k, grid, block, 0, s, ev.start, ev.stop, 0, args...);
    else
        hipLaunchKernelGGL(k, grid, block, 0, s, args...);
}

hipError_t launch_alloc(const VolDev& v, const PointSrc& src, const AllocOpts& o,
                        const int32_t* coords, CallInfo call, hipStream_t s);
hipError_t launch_alloc_batch(const BatchDev& b, int32_t total_tiles, const AllocOpts& o,
                              hipStream_t s, const LaunchEvents& ev = {});
hipError_t launch_backproject(const PointSrc& src, float* out, hipStream_t s);
hipError_t launch_tile_depth(const float* depth, int32_t W, int32_t H, float* tiled, hipStream_t s);
hipError_t launch_assign(const VolDev& v, CallInfo call, hipStream_t s);
hipError_t launch_assign_batch(const BatchDev& b, uint32_t max_capacity, hipStream_t s,
                               const LaunchEvents& ev = {});
// Scatter update (reference semantics) in two passes. The visit pass sums
// each workgroup's visits per voxel in LDS and writes the distinct voxels as
// entries grouped by pool block into the workgroup's own region of `entries`;
// each (workgroup, block) group becomes one segment of that block. The apply
// pass takes every touched block in one workgroup, sums its segments in LDS
// and adds the totals to the voxel sums with plain loads and stores, the
// weight cap applied once per voxel. Visits past a full LDS set, a full
// per-workgroup block set or a block's kScatterSegs segments are added by
// the visit pass itself with global atomics, before the apply pass runs.
constexpr int kScatterThreads = 256;
// LDS voxel slots per workgroup: 3,072 (4 workgroups per CU) for sparse
// clouds, whose workgroups see up to ~1,600 distinct voxels; 2,048 (5 per CU)
// for dense ones, where more workgroups per CU pay and the set stays below
// half full.
constexpr int kScatterSet = 3072;
constexpr int kScatterSetDense = 2048;
// Calls of at most kScatterSplitPoints points take one sample per lane (the
// visit pass's SPLIT layout): a workgroup then holds kScatterThreads visits,
// and its set kScatterSetSplit slots.
constexpr int kScatterSplitPoints = 1 << 17;
constexpr int kScatterSetSplit = 1024;
// Points per touched block at which a call counts as dense.
constexpr int kScatterDensePoints = 100;
constexpr int kScatterBlocks = 256;  // distinct blocks per workgroup (one per thread)
constexpr int kScatterSegs = 128;    // segments kept per block and call
constexpr int kScatterApplyGrid = 2048;
// Touched blocks are listed in kScatterLists lists by pool index mod
// kScatterLists, so that first touches spread their atomics over as many
// counters (kScatterCtrStride ints apart: one counter per 128 B).
constexpr int kScatterLists = 64;
constexpr int kScatterCtrStride = 32;
constexpr int kScatterCtrInts = 2 * kScatterLists * kScatterCtrStride;

struct ScatterSeg {
    uint32_t off;  // first entry in `entries`
    uint32_t cnt;
};

// A workgroup's distinct voxels number at most min(kScatterSet, its points x
// samples per point): the stride of the regions in `entries`.
__host__ __device__ inline int32_t scatter_region(int32_t steps) {
    const int64_t visits = (int64_t)kScatterThreads * (2 * (int64_t)steps + 1);
    return visits < kScatterSet ? (int32_t)visits : kScatterSet;
}
__host__ __device__ constexpr int32_t scatter_region_split() { return kScatterThreads; }
// Visit-pass workgroups of a call of n points: one per kScatterThreads points,
// or (split) ceil(samples / waves) per 64 points.
__host__ __device__ inline int64_t scatter_visit_grid(int32_t n, int32_t steps, bool split) {
    if (!split) return ((int64_t)n + kScatterThreads - 1) / kScatterThreads;
    const int64_t waves = kScatterThreads / 64, samples = 2 * (int64_t)steps + 1;
    return ((int64_t)n + 63) / 64 * ((samples + waves - 1) / waves);
}

struct ScatterScratch {
    uint2* entries;          // x = voxel in block | visits << 9, y = tsdf sum (float bits)
    int32_t region;          // entries per workgroup region (scatter_region)
    ScatterSeg* segs;        // kScatterSegs per pool block
    int32_t* seg_count;      // per pool block, zero between calls
    int32_t* touched;        // pool blocks with segments this call: list j at j * list_cap
    int32_t* touched_count;  // per parity and list (kScatterCtrStride apart); the visit
                             // pass zeroes the other parity's
    int32_t list_cap;        // ceil(pool capacity / kScatterLists)
    int32_t* touched_out;    // device view of a host int: the apply pass leaves the call's
                             // touched-block count there (may be null)
    int32_t parity;
};

// dense: the visit pass takes the smaller LDS set (kScatterSetDense);
// split: one sample per lane (calls of at most kScatterSplitPoints points).
hipError_t launch_scatter_update(const VolDev& v, const PointSrc& src, const ScatterScratch& sc,
                                 bool dense, bool split, hipStream_t s);
hipError_t launch_proj_update(const VolDev& v, const ProjCam& cam, CallInfo call, int32_t grid,
                              int32_t block_threads, hipStream_t s);
// max_w_max: the largest weight cap of the batch's volumes (FP16 records
// with caps up to 64 take their averaging reciprocals from a table)
hipError_t launch_proj_update_batch(const BatchDev& b, int32_t storage, int32_t grid,
                                    int32_t block_threads, hipStream_t s,
                                    const LaunchEvents& ev = {}, int32_t max_w_max = 65535);
hipError_t launch_extract(const VolDev& v, const ExtractArgs& a, bool emit, hipStream_t s);
// Count + scan (emit false) or emit for every volume of a batch, one launch
// each: blockIdx.y is the volume, grid_per_scene one-wave workgroups walk its
// blocks (fewer when the pools are smaller).
constexpr int32_t kExtractGridPerScene = 16384;
// XCDs of the chip: workgroups are dealt to them round-robin by id, each XCD
// with its own L2.
constexpr int kXcds = 8;
hipError_t launch_extract_batch(const VolDev* vols, const ExtractArgs* args, int32_t n_scenes,
                                int32_t storage, uint32_t max_pool_capacity,
                                int32_t grid_per_scene, bool emit, hipStream_t s);
hipError_t launch_fill_table(HashEntry* table, uint32_t n, hipStream_t s);
hipError_t launch_cas_ceiling(unsigned long long* slots, uint32_t mask, int32_t n_ops,
                              hipStream_t s);

__host__ __device__ constexpr int32_t ceil_log2(int32_t x) {
    int32_t l = 0;
    while ((int32_t(1) << l) < x) ++l;
    return l;
}

// A call inserts at most `capacity` keys, so at most this many tiles.
__host__ __device__ inline int32_t assign_tiles_max(uint32_t capacity) {
    return (int32_t)((capacity + kSortTile - 1) / kSortTile);
}

// assign_done holds, per parity, one counter per (level, run): level l of a
// call with T tiles has ceil(T / 2^l) runs; the stride per level is the
// volume's maximum tile count.
__host__ __device__ inline int32_t assign_done_per_parity(uint32_t capacity) {
    const int32_t t = assign_tiles_max(capacity);
    return (ceil_log2(t) + 1) * t;
}

}  // namespace osn_hip
