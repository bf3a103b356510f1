// Per-scene parameters of the projective update, built once per frame
// (kernel arguments for one volume; one device array per batched step, read by
// the update through scalar loads).
#pragma once

#include "common.hpp"

namespace osn_hip {

// The projective update reads depth from a copy in 8x4-pixel tiles: tile
// (tx, ty) holds 32 floats, row-major within the tile, at
// (ty * ceil(W / 8) + tx) * 32, so one tile is one 128-byte cache line and the
// pixels a wave gathers for a block fall in about half as many lines as in
// row-major order. The copy covers the image rounded up to whole tiles (padding
// 0). Frames are limited to kMaxFrameDim per side, so every index fits 32 bits
// and the index arithmetic fits 24-bit multiplies.
constexpr int32_t kMaxFrameDim = 1 << 14;
__host__ __device__ inline int32_t tiled_width(int32_t W) { return (W + 7) & ~7; }
__host__ __device__ inline int32_t tiled_height(int32_t H) { return (H + 3) & ~3; }
__host__ __device__ inline int32_t tiled_stride_k(int32_t W) { return 4 * tiled_width(W) - 32; }
// Index of pixel (u, w) with k = tiled_stride_k(W): for u = 8a + b, w = 4c + d
// it is c * 4 * tiled_width + 32 a + 8 d + b.
__device__ inline uint32_t tiled_index(uint32_t u, uint32_t w, uint32_t k) {
    return __umul24(w >> 2, k) + (w << 3) + (__umul24(u >> 3, 24u) + u);
}

// Camera depth range of an observable voxel centre, metres. Within it the
// reciprocals of a run of four depths come from one division (batch
// inversion: the product's reciprocal times partial products), never
// overflowing or underflowing.
constexpr float kZMin = 1.0f / 1024.0f;
constexpr float kZMax = 16384.0f;

// Camera for the projective update: world -> camera rotation and the camera
// centre, FP32, plus intrinsics and the depth image.
struct ProjCam {
    float Rcw[9];
    float c[3];
    float fx, fy, cx, cy;
    int32_t W, H;
    const float* depth;  // the depth image in 8x4-pixel tiles
    double Rcw64[9];   // FP64 pose, read only by the FP64 oracle kernel
    double c64[3];
};

struct UpdParams {
    const HashEntry* table;
    const int32_t* block_coord;
    char* vox;
    const TouchEntry* touched;
    const float* depth;  // 8x4-pixel tiles
    float Rcw[9];
    float c[3];
    float g[3][3];      // g[F][j]: step of form F (A, B, Z) along voxel axis j
    float fx, fy, cxh, cyh;
    int32_t W, H;
    float voxel, trunc, inv_trunc, wmax;
    int32_t wcap;       // wmax as an integer, 1..65535
    int32_t n;          // touched blocks
    int32_t tile_k;     // tiled_stride_k(W)
    double Rcw64[9];    // FP64 oracle kernel only
    double c64[3];
};

// The pixel of a voxel is a ratio of two linear forms of its index,
// A = fx p.x + (cx + 0.5) p.z and B = fy p.y + (cy + 0.5) p.z, over Z = p.z;
// g holds their steps along the three voxel axes (e_j = voxel * R_cw[:, j]).
// FP32, every product rounded on its own (built with -ffp-contract=off).
__device__ inline UpdParams make_params(const VolDev& v, const ProjCam& k, int parity) {
    UpdParams p;
    p.table = v.table;
    p.block_coord = v.block_coord;
    p.vox = (char*)v.vox;
    p.touched = v.touched;
    p.depth = k.depth;
    for (int i = 0; i < 9; ++i) { p.Rcw[i] = k.Rcw[i]; p.Rcw64[i] = k.Rcw64[i]; }
    for (int i = 0; i < 3; ++i) { p.c[i] = k.c[i]; p.c64[i] = k.c64[i]; }
    p.fx = k.fx; p.fy = k.fy;
    p.cxh = k.cx + 0.5f;
    p.cyh = k.cy + 0.5f;
    for (int j = 0; j < 3; ++j) {
        const float ex = v.voxel * k.Rcw[0 * 3 + j];
        const float ey = v.voxel * k.Rcw[1 * 3 + j];
        const float ez = v.voxel * k.Rcw[2 * 3 + j];
        p.g[0][j] = p.fx * ex + p.cxh * ez;
        p.g[1][j] = p.fy * ey + p.cyh * ez;
        p.g[2][j] = ez;
    }
    p.W = k.W; p.H = k.H;
    p.voxel = v.voxel;
    p.trunc = v.trunc;
    p.inv_trunc = 1.0f / v.trunc;
    p.wmax = (float)v.w_max;
    p.wcap = v.w_max;
    p.n = v.ctr->lists[parity].touched_count;
    p.tile_k = tiled_stride_k(k.W);
    return p;
}

}  // namespace osn_hip
