// Projective per-voxel update over the blocks an allocate call touched.
//
// No reference counterpart: the reference integrates by scattering along each
// point's view ray (cpp/src/volume.cpp:200-299). This update instead visits
// every voxel of every touched block once, projects its centre into the depth
// image and folds the observation into a running weighted average:
//
//   p   = R_cw * ((v + 0.5) * voxel - c)          camera coordinates (FP32),
//         evaluated as row = (d0 + ly*ey) + lz*ez, p = row + lx*ex
//   iz  = 1 / p.z                                  (skip unless p.z > 0)
//   pix = floor(fx * (p.x * iz) + cx + 0.5), floor(fy * (p.y * iz) + cy + 0.5)
//   d   = depth[pix]                               (skip unless inside, d > 0)
//   sdf = d - p.z                                  (skip if sdf < -trunc)
//   x   = min(sdf * (1 / trunc), 1)
//   t  <- (t * w + x) / (w + 1),  w <- min(w + 1, w_max)
//
// One wave per 8^3 block: each load of a lane is 16 contiguous bytes of the
// block's voxel record (4 voxels of {half, u16}, 2 of {float, float}, 1 of
// {double, double}), consecutive lanes read consecutive 16-byte pieces, and
// each voxel is owned by exactly one lane: no atomics, deterministic. The
// voxels of one load share (ly, lz), so their row term is computed once.
// The discrete choices (pixel, gates) are made by the same FP32 expression
// sequence in every storage mode, including the FP64 oracle kernel, which
// then evaluates the signed distance and the average in FP64 from the FP64
// pose. Compiled with -ffp-contract=off; the FP32 average uses an explicit
// fused multiply-add.

#include <hip/hip_fp16.h>

#include "internal.hpp"

namespace osn_hip {
namespace {

// Everything the update reads for one scene, gathered once per launch (kernel
// arguments in the single-volume kernel, LDS in the batched one) so that the
// per-voxel loop never re-reads descriptors from global memory.
#define OSN_GLOBAL __attribute__((address_space(1)))
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

struct UpdParams {
    // global address space, so loads through pointers kept in LDS stay
    // global (not flat) loads
    const HashEntry OSN_GLOBAL* table;
    const int32_t OSN_GLOBAL* block_coord;
    char OSN_GLOBAL* vox;
    const TouchEntry OSN_GLOBAL* touched;
    const float OSN_GLOBAL* depth;
    float Rcw[9];
    float c[3];
    float fx, fy, cx, cy;
    int32_t W, H;
    float voxel, trunc, inv_trunc, wmax;
    int32_t n;          // touched blocks
    int32_t pad;
    double Rcw64[9];    // FP64 oracle kernel only
    double c64[3];
};

__device__ inline UpdParams make_params(const VolDev& v, const ProjCam& k, int parity) {
    UpdParams p;
    p.table = (const HashEntry OSN_GLOBAL*)v.table;
    p.block_coord = (const int32_t OSN_GLOBAL*)v.block_coord;
    p.vox = (char OSN_GLOBAL*)v.vox;
    p.touched = (const TouchEntry OSN_GLOBAL*)v.touched;
    p.depth = (const float OSN_GLOBAL*)k.depth;
    for (int i = 0; i < 9; ++i) { p.Rcw[i] = k.Rcw[i]; p.Rcw64[i] = k.Rcw64[i]; }
    for (int i = 0; i < 3; ++i) { p.c[i] = k.c[i]; p.c64[i] = k.c64[i]; }
    p.fx = k.fx; p.fy = k.fy; p.cx = k.cx; p.cy = k.cy;
    p.W = k.W; p.H = k.H;
    p.voxel = v.voxel;
    p.trunc = v.trunc;
    p.inv_trunc = 1.0f / v.trunc;
    p.wmax = (float)v.w_max;
    p.n = v.ctr->lists[parity].touched_count;
    p.pad = 0;
    return p;
}

struct BlockFrame {
    float d0[3];     // camera coords of the block's first voxel centre
    float e[3][3];   // e[j] = voxel * R_cw[:, j]
};

__device__ inline void block_frame(const UpdParams& k, int32_t bx, int32_t by, int32_t bz,
                                   BlockFrame& f) {
    const float voxel = k.voxel;
    const float h = 0.5f * voxel;
    const float ox = (float)(bx * 8) * voxel + h - k.c[0];
    const float oy = (float)(by * 8) * voxel + h - k.c[1];
    const float oz = (float)(bz * 8) * voxel + h - k.c[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        f.d0[i] = k.Rcw[i * 3 + 0] * ox + k.Rcw[i * 3 + 1] * oy + k.Rcw[i * 3 + 2] * oz;
#pragma unroll
        for (int j = 0; j < 3; ++j) f.e[j][i] = voxel * k.Rcw[i * 3 + j];
    }
}

// FP32 decision for the voxel at camera point p: true when it gets an
// observation, with the pixel's depth and the FP32 sdf. Evaluated without
// branches (an out-of-image voxel reads pixel 0 and is then rejected), so a
// wave never splits on it.
__device__ inline bool observe(const UpdParams& k, float p0, float p1, float p2, float& sdf_out,
                               float& depth_out) {
    const float iz = 1.0f / p2;
    const float uf = k.fx * (p0 * iz) + k.cx + 0.5f;
    const float vf = k.fy * (p1 * iz) + k.cy + 0.5f;
    const bool inb = (p2 > 0.0f) & (uf >= 0.0f) & (vf >= 0.0f) & (uf < (float)k.W) &
                     (vf < (float)k.H);
    const int u = inb ? (int)uf : 0, w = inb ? (int)vf : 0;  // uf, vf >= 0: truncation is floor
    const float d = k.depth[w * k.W + u];
    const float sdf = d - p2;
    sdf_out = sdf;
    depth_out = d;
    return inb & (d > 0.0f) & !(sdf < -k.trunc);
}

__device__ inline float avg_step(float t, float w, float x) {
    // (t * w + x) / (w + 1); the reciprocal's 1-ulp error is far inside the
    // storage rounding this feeds
    return __builtin_fmaf(t, w, x) * __builtin_amdgcn_rcpf(w + 1.0f);
}

template <int STORE>
struct Layout;
template <>
struct Layout<kStoreF16> { static constexpr int kVpl = 4; };  // voxels per 16-byte load
template <>
struct Layout<kStoreF32> { static constexpr int kVpl = 2; };
template <>
struct Layout<kStoreF64> { static constexpr int kVpl = 1; };

// One wave updates the whole block of touch entry `item` (uniform over the
// wave). Load i of lane l covers voxels [i*64*V + l*V, +V).
template <int STORE>
__device__ inline void update_block(const UpdParams& k, int item, int lane) {
    constexpr int V = Layout<STORE>::kVpl;
    constexpr int L = kBlockVoxels / (64 * V);
    // {int64 key, uint32 slot, int32 bi} in one 16-byte load
    const u32x4 e = *reinterpret_cast<const u32x4 OSN_GLOBAL*>(k.touched + item);
    int32_t bi = (int32_t)e.w;
    if (bi == kIdxPending) bi = k.table[e.z].block_idx;  // inserted by this call
    if (bi < 0) return;
    int32_t bx, by, bz;
    unpack_key((int64_t)(((uint64_t)e.y << 32) | e.x), bx, by, bz);
    BlockFrame f;
    block_frame(k, bx, by, bz, f);

#pragma unroll
    for (int i = 0; i < L; ++i) {
        const int vi0 = i * 64 * V + lane * V;
        const int lx0 = vi0 & 7, ly = (vi0 >> 3) & 7, lz = vi0 >> 6;
        const float fly = (float)ly, flz = (float)lz;
        float row[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) row[c] = (f.d0[c] + fly * f.e[1][c]) + flz * f.e[2][c];

        if constexpr (STORE == kStoreF16) {
            u32x4 OSN_GLOBAL* rec = reinterpret_cast<u32x4 OSN_GLOBAL*>(
                k.vox + ((size_t)bi * kBlockVoxels + vi0) * 4);
            const u32x4 q = *rec;
            uint32_t word[4] = {q.x, q.y, q.z, q.w};
            bool dirty = false;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float flx = (float)(lx0 + j);
                float sdf, d;
                const bool ok = observe(k, row[0] + flx * f.e[0][0], row[1] + flx * f.e[0][1],
                                        row[2] + flx * f.e[0][2], sdf, d);
                const float x = fminf(sdf * k.inv_trunc, 1.0f);
                const float tt = __half2float(__ushort_as_half((unsigned short)(word[j] & 0xffffu)));
                const float w = (float)(word[j] >> 16);
                const float tn = avg_step(tt, w, x);
                const float wn = fminf(w + 1.0f, k.wmax);
                const uint32_t nw = (uint32_t)__half_as_ushort(__float2half_rn(tn)) | ((uint32_t)wn << 16);
                word[j] = ok ? nw : word[j];
                dirty |= ok;
            }
            if (dirty) *rec = u32x4{word[0], word[1], word[2], word[3]};
        } else if constexpr (STORE == kStoreF32) {
            f32x4 OSN_GLOBAL* rec = reinterpret_cast<f32x4 OSN_GLOBAL*>(
                k.vox + ((size_t)bi * kBlockVoxels + vi0) * 8);
            const f32x4 q = *rec;
            float tv[2] = {q.x, q.z}, wv[2] = {q.y, q.w};
            bool dirty = false;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const float flx = (float)(lx0 + j);
                float sdf, d;
                const bool ok = observe(k, row[0] + flx * f.e[0][0], row[1] + flx * f.e[0][1],
                                        row[2] + flx * f.e[0][2], sdf, d);
                const float x = fminf(sdf * k.inv_trunc, 1.0f);
                const float tn = avg_step(tv[j], wv[j], x);
                const float wn = fminf(wv[j] + 1.0f, k.wmax);
                tv[j] = ok ? tn : tv[j];
                wv[j] = ok ? wn : wv[j];
                dirty |= ok;
            }
            if (dirty) *rec = f32x4{tv[0], wv[0], tv[1], wv[1]};
        } else {
            f64x2 OSN_GLOBAL* rec = reinterpret_cast<f64x2 OSN_GLOBAL*>(
                k.vox + ((size_t)bi * kBlockVoxels + vi0) * 16);
            const float flx = (float)lx0;
            float sdf, d;
            if (observe(k, row[0] + flx * f.e[0][0], row[1] + flx * f.e[0][1],
                        row[2] + flx * f.e[0][2], sdf, d)) {
                f64x2 q = *rec;
                const double vs = (double)k.voxel;
                const double wx = ((double)(bx * 8 + lx0) + 0.5) * vs - k.c64[0];
                const double wy = ((double)(by * 8 + ly) + 0.5) * vs - k.c64[1];
                const double wz = ((double)(bz * 8 + lz) + 0.5) * vs - k.c64[2];
                const double z = k.Rcw64[6] * wx + k.Rcw64[7] * wy + k.Rcw64[8] * wz;
                double x = ((double)d - z) / (double)k.trunc;
                x = x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x);
                const double w = q.y;
                q.x = (q.x * w + x) / (w + 1.0);
                q.y = (w + 1.0 < (double)k.wmax) ? w + 1.0 : (double)k.wmax;
                *rec = q;
            }
        }
    }
}

__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

template <int STORE>
__global__ __launch_bounds__(512) void k_proj_update(VolDev v, ProjCam cam, CallInfo call) {
    const int waves = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const UpdParams k = make_params(v, cam, call.parity);
    for (int item = blockIdx.x * waves + wave_id(); item < k.n; item += gridDim.x * waves)
        update_block<STORE>(k, item, lane);
}

// All scenes of a batched step in one launch: the work list is the
// concatenation of every active scene's touched blocks. Each workgroup gathers
// the scenes' parameters and the prefix of their touched counts into LDS. The
// list is cut into one contiguous chunk per XCD (workgroups are dealt to the
// 8 XCDs round-robin by id), so each XCD's L2 serves the depth images of only
// the few scenes in its chunk.
constexpr int kXcds = 8;

template <int STORE>
__global__ __launch_bounds__(512) void k_proj_update_batch(BatchDev b) {
    __shared__ UpdParams s_p[kMaxBatchScenes];
    __shared__ int s_prefix[kMaxBatchScenes + 1];
    const int waves = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const int k = b.step_ctr ? *b.step_ctr : b.step;
    const int ns = b.n_scenes;
    if ((int)threadIdx.x < ns) {
        const int sc = threadIdx.x;
        const SceneStep& st = b.steps[k * ns + sc];
        s_p[sc] = make_params(b.vols[sc], st.cam, batch_call(b, sc, k).parity);
        if (!st.active) s_p[sc].n = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int sc = 0; sc < ns; ++sc) { s_prefix[sc] = acc; acc += s_p[sc].n; }
        s_prefix[ns] = acc;
    }
    __syncthreads();
    const int n = s_prefix[ns];
    const int xcd = blockIdx.x % kXcds;
    const int per_xcd = gridDim.x / kXcds;  // the launcher keeps the grid a multiple of 8
    const int lo = (int)((long long)n * xcd / kXcds);
    const int hi = (int)((long long)n * (xcd + 1) / kXcds);
    int sc = 0;
    for (int item = lo + (int)(blockIdx.x / kXcds) * waves + wave_id(); item < hi;
         item += per_xcd * waves) {
        while (s_prefix[sc + 1] <= item) ++sc;  // items ascend, so the scene only moves forward
        const UpdParams& p = s_p[sc];
        update_block<STORE>(p, item - s_prefix[sc], lane);
    }
}

}  // namespace

hipError_t launch_proj_update(const VolDev& v, const ProjCam& cam, CallInfo call, int32_t grid,
                              int32_t block_threads, hipStream_t s) {
    if (grid <= 0) grid = 2048;
    if (block_threads != 256 && block_threads != 128 && block_threads != 512 &&
        block_threads != 64)
        return hipErrorInvalidValue;
    switch (v.storage) {
        case kStoreF16: k_proj_update<kStoreF16><<<grid, block_threads, 0, s>>>(v, cam, call); break;
        case kStoreF32: k_proj_update<kStoreF32><<<grid, block_threads, 0, s>>>(v, cam, call); break;
        case kStoreF64: k_proj_update<kStoreF64><<<grid, block_threads, 0, s>>>(v, cam, call); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_proj_update_batch(const BatchDev& b, int32_t storage, int32_t grid,
                                    int32_t block_threads, hipStream_t s) {
    if (b.n_scenes > kMaxBatchScenes) return hipErrorInvalidValue;
    if (grid <= 0) grid = 4096;
    grid = (grid + kXcds - 1) / kXcds * kXcds;
    if (block_threads != 256 && block_threads != 128 && block_threads != 512)
        return hipErrorInvalidValue;
    if (block_threads < b.n_scenes) block_threads = 512;  // one thread per scene gathers
    switch (storage) {
        case kStoreF16: k_proj_update_batch<kStoreF16><<<grid, block_threads, 0, s>>>(b); break;
        case kStoreF32: k_proj_update_batch<kStoreF32><<<grid, block_threads, 0, s>>>(b); break;
        case kStoreF64: k_proj_update_batch<kStoreF64><<<grid, block_threads, 0, s>>>(b); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace osn_hip
