// Per-scene parameters of the projective update, built once per frame
// (kernel arguments for one volume; one device array per batched step, read by
// the update through scalar loads).
#pragma once

#include "common.hpp"

namespace osn_hip {

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

struct UpdParams {
    const HashEntry* table;
    const int32_t* block_coord;
    char* vox;
    const TouchEntry* touched;
    const float* depth;
    float Rcw[9];
    float c[3];
    float g[3][3];      // g[F][j]: step of form F (A, B, Z) along voxel axis j
    float fx, fy, cxh, cyh;
    int32_t W, H;
    float voxel, trunc, inv_trunc, wmax;
    int32_t n;          // touched blocks
    int32_t pad;
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
    p.n = v.ctr->lists[parity].touched_count;
    p.pad = 0;
    return p;
}

}  // namespace osn_hip
