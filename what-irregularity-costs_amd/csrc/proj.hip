// Projective per-voxel update over the blocks an allocate call touched.
//
// No reference counterpart: the reference integrates by scattering along each
// point's view ray (cpp/src/volume.cpp:200-299). This update instead visits
// every voxel of every touched block once, projects its centre into the depth
// image and folds the observation into a running weighted average:
//
//   p   = R_cw * ((v + 0.5) * voxel - c)          camera coordinates (FP32)
//   iz  = 1 / p.z                                  (skip unless p.z > 0)
//   pix = floor(fx * (p.x * iz) + cx + 0.5), floor(fy * (p.y * iz) + cy + 0.5)
//   d   = depth[pix]                               (skip unless inside, d > 0)
//   sdf = d - p.z                                  (skip if sdf < -trunc)
//   x   = min(sdf * (1 / trunc), 1)
//   t  <- (t * w + x) / (w + 1),  w <- min(w + 1, w_max)
//
// One workgroup slice per 8^3 block, each thread owning 16 contiguous bytes of
// the block's voxel record (4 voxels of {half, u16}, 2 of {float, float}, 1 of
// {double, double}), so every load and store is a coalesced 16-byte access and
// each voxel is owned by exactly one thread: no atomics, deterministic.
// The discrete choices (pixel, gates) are made by the same FP32 expression
// sequence in every storage mode, including the FP64 oracle kernel, which
// then evaluates the signed distance and the average in FP64 from the FP64
// pose. Compiled with -ffp-contract=off; the FP32 average uses an explicit
// fused multiply-add.

#include <hip/hip_fp16.h>

#include "internal.hpp"

namespace osn_hip {
namespace {

struct BlockFrame {
    float d0[3];     // camera coords of the block's first voxel centre
    float e[3][3];   // e[j] = voxel * R_cw[:, j]
};

__device__ inline void block_frame(const ProjCam& k, float voxel, int32_t bx, int32_t by,
                                   int32_t bz, BlockFrame& f) {
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

// FP32 decision for one voxel. Returns false when the voxel gets no
// observation; otherwise the pixel's depth and the FP32 sdf.
__device__ inline bool observe(const ProjCam& k, const BlockFrame& f, float trunc, int lx, int ly,
                               int lz, float& sdf_out, float& depth_out) {
    const float fl0 = (float)lx, fl1 = (float)ly, fl2 = (float)lz;
    const float p0 = f.d0[0] + fl0 * f.e[0][0] + fl1 * f.e[1][0] + fl2 * f.e[2][0];
    const float p1 = f.d0[1] + fl0 * f.e[0][1] + fl1 * f.e[1][1] + fl2 * f.e[2][1];
    const float p2 = f.d0[2] + fl0 * f.e[0][2] + fl1 * f.e[1][2] + fl2 * f.e[2][2];
    if (!(p2 > 0.0f)) return false;
    const float iz = 1.0f / p2;
    const float uf = k.fx * (p0 * iz) + k.cx + 0.5f;
    const float vf = k.fy * (p1 * iz) + k.cy + 0.5f;
    if (!(uf >= 0.0f) || !(vf >= 0.0f) || !(uf < (float)k.W) || !(vf < (float)k.H)) return false;
    const int u = (int)uf, w = (int)vf;  // uf, vf >= 0: truncation is floor
    const float d = k.depth[(size_t)w * k.W + u];
    if (!(d > 0.0f)) return false;
    const float sdf = d - p2;
    if (sdf < -trunc) return false;
    sdf_out = sdf;
    depth_out = d;
    return true;
}

__device__ inline float avg_step(float t, float w, float x) {
    // (t * w + x) / (w + 1); the reciprocal's 1-ulp error is far inside the
    // storage rounding this feeds
    return __builtin_fmaf(t, w, x) * __builtin_amdgcn_rcpf(w + 1.0f);
}

template <int STORE>
struct Layout;
template <>
struct Layout<kStoreF16> { static constexpr int kVpt = 4; };
template <>
struct Layout<kStoreF32> { static constexpr int kVpt = 2; };
template <>
struct Layout<kStoreF64> { static constexpr int kVpt = 1; };

template <int STORE>
__global__ __launch_bounds__(512) void k_proj_update(VolDev v, ProjCam k) {
    constexpr int kVpt = Layout<STORE>::kVpt;
    constexpr int kTpb = kBlockVoxels / kVpt;  // threads per block record
    const int per_wg = blockDim.x / kTpb;        // block records per workgroup pass
    const int sub = threadIdx.x / kTpb;
    const int t = threadIdx.x % kTpb;
    const int n = v.ctr->touched_count;
    const float inv_trunc = 1.0f / v.trunc;
    const float wmax = (float)v.w_max;

    for (int item = blockIdx.x * per_wg + sub; item < n; item += gridDim.x * per_wg) {
        const uint32_t slot = v.touched[item];
        const int32_t bi = v.table[slot].block_idx;
        if (bi < 0) continue;
        const int32_t bx = v.block_coord[bi * 3 + 0];
        const int32_t by = v.block_coord[bi * 3 + 1];
        const int32_t bz = v.block_coord[bi * 3 + 2];
        BlockFrame f;
        block_frame(k, v.voxel, bx, by, bz, f);
        const int vi0 = t * kVpt;

        if constexpr (STORE == kStoreF16) {
            uint4* rec = reinterpret_cast<uint4*>(static_cast<uint32_t*>(v.vox) +
                                                  (size_t)bi * kBlockVoxels) + t;
            uint4 q = *rec;
            uint32_t word[4] = {q.x, q.y, q.z, q.w};
            bool dirty = false;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int vi = vi0 + j;
                float sdf, d;
                if (!observe(k, f, v.trunc, vi & 7, (vi >> 3) & 7, vi >> 6, sdf, d)) continue;
                const float x = fminf(sdf * inv_trunc, 1.0f);
                const float tt = __half2float(__ushort_as_half((unsigned short)(word[j] & 0xffffu)));
                const float w = (float)(word[j] >> 16);
                const float tn = avg_step(tt, w, x);
                const float wn = fminf(w + 1.0f, wmax);
                word[j] = (uint32_t)__half_as_ushort(__float2half_rn(tn)) |
                          ((uint32_t)wn << 16);
                dirty = true;
            }
            if (dirty) *rec = make_uint4(word[0], word[1], word[2], word[3]);
        } else if constexpr (STORE == kStoreF32) {
            float4* rec = reinterpret_cast<float4*>(static_cast<float*>(v.vox) +
                                                    (size_t)bi * kBlockVoxels * 2) + t;
            float4 q = *rec;
            float tv[2] = {q.x, q.z}, wv[2] = {q.y, q.w};
            bool dirty = false;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int vi = vi0 + j;
                float sdf, d;
                if (!observe(k, f, v.trunc, vi & 7, (vi >> 3) & 7, vi >> 6, sdf, d)) continue;
                const float x = fminf(sdf * inv_trunc, 1.0f);
                tv[j] = avg_step(tv[j], wv[j], x);
                wv[j] = fminf(wv[j] + 1.0f, wmax);
                dirty = true;
            }
            if (dirty) *rec = make_float4(tv[0], wv[0], tv[1], wv[1]);
        } else {
            double2* rec = reinterpret_cast<double2*>(static_cast<double*>(v.vox) +
                                                      (size_t)bi * kBlockVoxels * 2) + t;
            const int vi = vi0;
            const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
            float sdf, d;
            if (observe(k, f, v.trunc, lx, ly, lz, sdf, d)) {
                double2 q = *rec;
                const double vs = (double)v.voxel;
                const double wx = ((double)(bx * 8 + lx) + 0.5) * vs - k.c64[0];
                const double wy = ((double)(by * 8 + ly) + 0.5) * vs - k.c64[1];
                const double wz = ((double)(bz * 8 + lz) + 0.5) * vs - k.c64[2];
                const double z = k.Rcw64[6] * wx + k.Rcw64[7] * wy + k.Rcw64[8] * wz;
                double x = ((double)d - z) / (double)v.trunc;
                x = x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x);
                const double w = q.y;
                q.x = (q.x * w + x) / (w + 1.0);
                q.y = (w + 1.0 < (double)v.w_max) ? w + 1.0 : (double)v.w_max;
                *rec = q;
            }
        }
    }
}

}  // namespace

hipError_t launch_proj_update(const VolDev& v, const ProjCam& cam, int32_t grid,
                              int32_t block_threads, hipStream_t s) {
    if (grid <= 0) grid = 2048;
    if (block_threads != 256 && block_threads != 128 && block_threads != 512)
        return hipErrorInvalidValue;
    switch (v.storage) {
        case kStoreF16:
            if (block_threads < 128) return hipErrorInvalidValue;
            k_proj_update<kStoreF16><<<grid, block_threads, 0, s>>>(v, cam);
            break;
        case kStoreF32:
            if (block_threads < 256) return hipErrorInvalidValue;
            k_proj_update<kStoreF32><<<grid, block_threads, 0, s>>>(v, cam);
            break;
        case kStoreF64:
            if (block_threads < 512) block_threads = 512;
            k_proj_update<kStoreF64><<<grid, block_threads, 0, s>>>(v, cam);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace osn_hip
