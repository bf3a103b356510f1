// Projective per-voxel update over the blocks an allocate call touched.
//
// No reference counterpart: the reference integrates by scattering along each
// point's view ray (cpp/src/volume.cpp:200-299). This update instead visits
// every voxel of every touched block once, projects its centre into the depth
// image and folds the observation into a running weighted average:
//
//   p   = R_cw * ((v + 0.5) * voxel - c)                camera coordinates
//   pix = floor(fx * p.x / p.z + cx + 0.5), floor(fy * p.y / p.z + cy + 0.5)
//   d   = depth[pix]                    (skip unless p.z > 0, inside, d > 0)
//   sdf = d - p.z                       (skip if sdf < -trunc)
//   x   = min(sdf * (1 / trunc), 1)
//   t  <- (t * w + x) / (w + 1),  w <- min(w + 1, w_max)
//
// FP32 evaluation. The pixel is a ratio of two linear forms of the voxel
// index, A = fx p.x + (cx + 0.5) p.z and B = fy p.y + (cy + 0.5) p.z, over
// Z = p.z, so each form is its block-origin value plus a per-lane offset, then
// one add per step along x:
//   per frame   (A_j, B_j, Z_j) for the voxel axes j (e_j = voxel * R_cw[:, j])
//   per block   (A0, B0, Z0) from d0 = R_cw * (block origin centre - c)
//   per lane    off = (ly * F_y + lz * F_z) + lx0 * F_x  for its run of 4
//               voxels along x starting at lx0 (0 or 4)
//   per voxel   F = F0 + off, then F += F_x;  iz = 1 / Z; uf = A * iz, vf = B * iz
// with every product rounded on its own (-ffp-contract=off).
//
// One wave per 8^3 block, each lane owning two runs of 4 consecutive voxels
// along x (lz and lz + 4), so the lane's records are 16 contiguous bytes (FP16
// {half, u16}), 32 (FP32 {float, float}) or 64 (FP64); each voxel has one
// owner: no atomics, deterministic. Every storage mode, the FP64 oracle kernel
// included, makes the discrete choices (pixel, gates) with this same FP32
// sequence; the FP64 kernel then evaluates the signed distance and the
// average in FP64 from the FP64 pose. The FP32 average uses an explicit fused
// multiply-add.

#include <hip/hip_fp16.h>

#include "internal.hpp"

namespace osn_hip {
namespace {

// Pointers of the parameters are dereferenced through the global address
// space, so every access is a global (not flat) load or store; a batch reads
// its per-scene parameter array through the constant address space (scalar
// loads).
#define OSN_GLOBAL __attribute__((address_space(1)))
#define OSN_CONST __attribute__((address_space(4)))
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ inline T OSN_GLOBAL* glob(T* p) {
    return (T OSN_GLOBAL*)p;
}

// A lane's two runs: voxels lane * 4 + 256 * r + [0, 4) of the block, r = 0, 1.
struct LaneRuns {
    float off[2][3];  // off[r][F]
    int32_t lx0, ly, lz0;
};

template <typename Params>  // UpdParams in any address space
__device__ inline LaneRuns lane_runs(const Params& k, int lane) {
    LaneRuns L;
    const int vi = lane * 4;
    L.lx0 = vi & 7;
    L.ly = (vi >> 3) & 7;
    L.lz0 = vi >> 6;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q)
            L.off[r][q] = ((float)L.ly * k.g[q][1] + (float)(L.lz0 + 4 * r) * k.g[q][2]) +
                          (float)L.lx0 * k.g[q][0];
    return L;
}

// Block-origin values (A0, B0, Z0) of block (bx, by, bz).
template <typename Params>
__device__ inline void block_forms(const Params& k, int32_t bx, int32_t by, int32_t bz,
                                   float f0[3]) {
    const float voxel = k.voxel;
    const float h = 0.5f * voxel;
    const float ox = (float)(bx * 8) * voxel + h - k.c[0];
    const float oy = (float)(by * 8) * voxel + h - k.c[1];
    const float oz = (float)(bz * 8) * voxel + h - k.c[2];
    float d0[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        d0[i] = k.Rcw[i * 3 + 0] * ox + k.Rcw[i * 3 + 1] * oy + k.Rcw[i * 3 + 2] * oz;
    f0[0] = k.fx * d0[0] + k.cxh * d0[2];
    f0[1] = k.fy * d0[1] + k.cyh * d0[2];
    f0[2] = d0[2];
}

__device__ inline float avg_step(float t, float w, float x) {
    // (t * w + x) / (w + 1); the reciprocal's 1-ulp error is far inside the
    // storage rounding this feeds
    return __builtin_fmaf(t, w, x) * __builtin_amdgcn_rcpf(w + 1.0f);
}

// A touch entry {int64 key, uint32 slot, int32 bi} is fetched one dword per
// lane (lanes 0-3), so the value stays a vector register while the load is in
// flight: the next block's entry is requested a whole block ahead, and the
// wave waits for it only when it reads the dwords out (touch_entry).
template <typename Params>
__device__ inline uint32_t load_touch(const Params& k, int item, int lane) {
    return reinterpret_cast<const uint32_t OSN_GLOBAL*>(glob(k.touched + item))[lane & 3];
}

__device__ inline u32x4 touch_entry(uint32_t part) {
    return u32x4{(uint32_t)__builtin_amdgcn_readlane((int)part, 0),
                 (uint32_t)__builtin_amdgcn_readlane((int)part, 1),
                 (uint32_t)__builtin_amdgcn_readlane((int)part, 2),
                 (uint32_t)__builtin_amdgcn_readlane((int)part, 3)};
}

// One wave updates the whole block of touch entry e (uniform over the wave).
// The update is latency-bound before it is bandwidth-bound, so everything a
// block needs from memory is requested before any of it is used: the lane's
// two voxel records, then the depth of all 8 of its voxels (pixels chosen
// without branches: an out-of-image voxel reads pixel 0 and is then rejected).
template <int STORE, typename Params>
__device__ inline void update_block(const Params& k, const LaneRuns& L, uint32_t part, int lane) {
    const u32x4 e = touch_entry(part);
    int32_t bi = (int32_t)e.w;
    if (bi == kIdxPending) bi = glob(k.table)[e.z].block_idx;  // inserted by this call
    if (bi < 0) return;
    int32_t bx, by, bz;
    unpack_key((int64_t)(((uint64_t)e.y << 32) | e.x), bx, by, bz);
    const size_t v0 = (size_t)bi * kBlockVoxels + lane * 4;  // first voxel of run 0

    u32x4 h16[2];
    f32x4 h32[2][2];
    if constexpr (STORE == kStoreF16) {
        const u32x4 OSN_GLOBAL* rec = reinterpret_cast<const u32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 4);
        h16[0] = rec[0];
        h16[1] = rec[64];  // run 1 starts 256 voxels (1 KiB) later
    } else if constexpr (STORE == kStoreF32) {
        const f32x4 OSN_GLOBAL* rec = reinterpret_cast<const f32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 8);
        h32[0][0] = rec[0];
        h32[0][1] = rec[1];
        h32[1][0] = rec[128];
        h32[1][1] = rec[129];
    }

    float f0[3];
    block_forms(k, bx, by, bz, f0);
    float Zv[8];
    uint32_t pix[8];
    bool inb[8];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        float A = f0[0] + L.off[r][0], B = f0[1] + L.off[r][1], Z = f0[2] + L.off[r][2];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float iz = 1.0f / Z;
            const float uf = A * iz;
            const float vf = B * iz;
            const bool in = (Z > 0.0f) & (uf >= 0.0f) & (vf >= 0.0f) & (uf < (float)k.W) &
                            (vf < (float)k.H);
            // uf, vf >= 0: truncation is floor; W, H < 2^24, so the index fits 24x24 bits
            const uint32_t u = in ? (uint32_t)uf : 0u, w = in ? (uint32_t)vf : 0u;
            pix[r * 4 + j] = __umul24(w, (uint32_t)k.W) + u;
            inb[r * 4 + j] = in;
            Zv[r * 4 + j] = Z;
            A = A + k.g[0][0];
            B = B + k.g[1][0];
            Z = Z + k.g[2][0];
        }
    }
    float d[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) d[i] = glob(k.depth)[pix[i]];
    // The records are read before the gates decide whether they change; left
    // alone, the compiler would move a record load behind the gates (into the
    // store branch) and expose its latency. This empty statement, which emits
    // no instruction, needs them here, after the gathers were issued.
    if constexpr (STORE == kStoreF16) asm volatile("" ::"v"(h16[0]), "v"(h16[1]));
    if constexpr (STORE == kStoreF32)
        asm volatile("" ::"v"(h32[0][0]), "v"(h32[0][1]), "v"(h32[1][0]), "v"(h32[1][1]));
    bool ok[8];
    float sdf[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        sdf[i] = d[i] - Zv[i];
        ok[i] = inb[i] & (d[i] > 0.0f) & !(sdf[i] < -k.trunc);
    }

#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const bool* o = ok + r * 4;
        const float* sd = sdf + r * 4;
        if constexpr (STORE == kStoreF16) {
            if (o[0] | o[1] | o[2] | o[3]) {
                uint32_t word[4] = {h16[r].x, h16[r].y, h16[r].z, h16[r].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float x = fminf(sd[j] * k.inv_trunc, 1.0f);
                    const float tt = __half2float(__ushort_as_half((unsigned short)(word[j] & 0xffffu)));
                    const float w = (float)(word[j] >> 16);
                    const float tn = avg_step(tt, w, x);
                    const float wn = fminf(w + 1.0f, k.wmax);
                    const uint32_t nw = __builtin_amdgcn_perm(
                        (uint32_t)wn, (uint32_t)__half_as_ushort(__float2half_rn(tn)), 0x05040100u);
                    word[j] = o[j] ? nw : word[j];
                }
                u32x4 OSN_GLOBAL* rec = reinterpret_cast<u32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 4);
                rec[r * 64] = u32x4{word[0], word[1], word[2], word[3]};
            }
        } else if constexpr (STORE == kStoreF32) {
            if (o[0] | o[1] | o[2] | o[3]) {
                float tv[4] = {h32[r][0].x, h32[r][0].z, h32[r][1].x, h32[r][1].z};
                float wv[4] = {h32[r][0].y, h32[r][0].w, h32[r][1].y, h32[r][1].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float x = fminf(sd[j] * k.inv_trunc, 1.0f);
                    const float tn = avg_step(tv[j], wv[j], x);
                    const float wn = fminf(wv[j] + 1.0f, k.wmax);
                    tv[j] = o[j] ? tn : tv[j];
                    wv[j] = o[j] ? wn : wv[j];
                }
                f32x4 OSN_GLOBAL* rec = reinterpret_cast<f32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 8);
                if (o[0] | o[1]) rec[r * 128 + 0] = f32x4{tv[0], wv[0], tv[1], wv[1]};
                if (o[2] | o[3]) rec[r * 128 + 1] = f32x4{tv[2], wv[2], tv[3], wv[3]};
            }
        } else {
            const int ly = L.ly, lz = L.lz0 + 4 * r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!o[j]) continue;
                f64x2 OSN_GLOBAL* rec = reinterpret_cast<f64x2 OSN_GLOBAL*>(
                    glob(k.vox) + (v0 + r * 256 + j) * 16);
                f64x2 q = *rec;
                const double vs = (double)k.voxel;
                const double wx = ((double)(bx * 8 + L.lx0 + j) + 0.5) * vs - k.c64[0];
                const double wy = ((double)(by * 8 + ly) + 0.5) * vs - k.c64[1];
                const double wz = ((double)(bz * 8 + lz) + 0.5) * vs - k.c64[2];
                const double z = k.Rcw64[6] * wx + k.Rcw64[7] * wy + k.Rcw64[8] * wz;
                double x = ((double)d[r * 4 + j] - z) / (double)k.trunc;
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
    const LaneRuns L = lane_runs(k, lane);
    const int stride = gridDim.x * waves;
    int item = blockIdx.x * waves + wave_id();
    uint32_t e = item < k.n ? load_touch(k, item, lane) : 0u;
    while (item < k.n) {  // the next block's entry is requested a block ahead
        const int next = item + stride;
        const uint32_t en = next < k.n ? load_touch(k, next, lane) : 0u;
        update_block<STORE>(k, L, e, lane);
        e = en;
        item = next;
    }
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
    __shared__ int s_prefix[kMaxBatchScenes + 1];
    const int waves = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const int ns = b.n_scenes;
    const UpdParams OSN_CONST* P = (const UpdParams OSN_CONST*)b.upd;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int sc = 0; sc < ns; ++sc) { s_prefix[sc] = acc; acc += P[sc].n; }
        s_prefix[ns] = acc;
    }
    __syncthreads();
    const int n = s_prefix[ns];
    const int xcd = blockIdx.x % kXcds;
    const int per_xcd = gridDim.x / kXcds;  // the launcher keeps the grid a multiple of 8
    const int lo = (int)((long long)n * xcd / kXcds);
    const int hi = (int)((long long)n * (xcd + 1) / kXcds);
    const int stride = per_xcd * waves;
    int item = lo + (int)(blockIdx.x / kXcds) * waves + wave_id();
    int sc = 0, lanes_sc = -1;
    while (item < hi && s_prefix[sc + 1] <= item) ++sc;
    uint32_t e = item < hi ? load_touch(P[sc], item - s_prefix[sc], lane) : 0u;
    LaneRuns L{};
    while (item < hi) {  // the next block's entry is requested a block ahead
        const int next = item + stride;
        int scn = sc;  // items ascend, so the scene only moves forward
        while (next < hi && s_prefix[scn + 1] <= next) ++scn;
        const uint32_t en = next < hi ? load_touch(P[scn], next - s_prefix[scn], lane) : 0u;
        if (sc != lanes_sc) {  // the lane offsets depend on the scene's pose
            L = lane_runs(P[sc], lane);
            lanes_sc = sc;
        }
        update_block<STORE>(P[sc], L, e, lane);
        e = en;
        item = next;
        sc = scn;
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
