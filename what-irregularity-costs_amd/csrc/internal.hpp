// Host-side launch interface between the C ABI layer and the kernels.
#pragma once

#include "common.hpp"

namespace osn_hip {

// Where a batch of points comes from: world-space points (the reference
// shape), or a depth image back-projected on the fly.
struct PointSrc {
    const float* pos;      // 3*n interleaved, device
    int32_t n;
    const float* depth;    // W*H planar depth in metres, device; 0 = invalid
    int32_t W, H;
    float fx, fy, cx, cy;
    float R[9];            // world-from-camera rotation, row-major (FP32)
    float max_range;       // drop pixels with depth > max_range (> 0 enables)
    float cam[3];          // camera centre in world (the band walk's origin)
    float radius_sq;       // band-walk far cull, 0 disables
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
    int32_t lds_set = 1024;   // per-workgroup pre-hash entries (0 disables)
};

// Camera for the projective update: world -> camera rotation and the camera
// centre, FP32, plus intrinsics and the depth image.
struct ProjCam {
    float Rcw[9];
    float c[3];
    float fx, fy, cx, cy;
    int32_t W, H;
    const float* depth;
    double Rcw64[9];   // FP64 pose, read only by the FP64 oracle kernel
    double c64[3];
};

struct ExtractArgs {
    float min_weight;
    float iso;
    int32_t n_blocks;      // blocks to walk (<= block_count), -1 = read block_count
    int32_t* block_vcount; // per block vertex counts
    int32_t* block_voff;   // per block exclusive offsets (emit reads)
    int32_t* chunk_sums;   // per 1024-block chunk sums
    int32_t* scan_flag;    // arrival counter for the chunk scan
    float* posnor;         // device output, 6 floats per vertex
    uint8_t* rgb;          // device output, 3 bytes per vertex, may be null
    int32_t cap;           // output capacity in vertices
    int32_t* count_out;    // device int32: vertex count or -1 on overflow
};

constexpr int kSortTile = 4096;
constexpr int kMaxSortLevels = 24;

hipError_t launch_begin_call(const VolDev& v, hipStream_t s);
hipError_t launch_alloc(const VolDev& v, const PointSrc& src, const AllocOpts& o,
                        const int32_t* coords, hipStream_t s);
hipError_t launch_backproject(const PointSrc& src, float* out, hipStream_t s);
hipError_t launch_assign(const VolDev& v, hipStream_t s);
hipError_t launch_scatter_update(const VolDev& v, const PointSrc& src, hipStream_t s);
hipError_t launch_proj_update(const VolDev& v, const ProjCam& cam, int32_t grid,
                              int32_t block_threads, hipStream_t s);
hipError_t launch_extract(const VolDev& v, const ExtractArgs& a, bool emit, hipStream_t s);
hipError_t launch_fill_table(HashEntry* table, uint32_t n, hipStream_t s);
hipError_t launch_zero_words(uint32_t* p, size_t n_words, hipStream_t s);

int32_t sort_flag_stride(uint32_t capacity);

}  // namespace osn_hip
