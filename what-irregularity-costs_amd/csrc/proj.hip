// Projective per-voxel update over the blocks an allocate call touched.
//
// No reference counterpart: the reference integrates by scattering along each
// point's view ray (cpp/src/volume.cpp:200-299). This update instead visits
// every voxel of every touched block once, projects its centre into the depth
// image and folds the observation into a running weighted average:
//
//   p   = R_cw * ((v + 0.5) * voxel - c)                camera coordinates
//   pix = floor(fx * p.x / p.z + cx + 0.5), floor(fy * p.y / p.z + cy + 0.5)
//   d   = depth[pix]                    (skip unless 2^-10 <= p.z <= 2^14,
//                                        inside, d > 0)
//                                       (read from a copy of the image in
//                                       8x4-pixel tiles, one cache line each)
//   sdf = d - p.z                       (skip if sdf < -trunc)
//   x   = min(sdf * (1 / trunc), 1)
//   skip if w >= w_max                  (frozen at the weight cap, as the
//                                        reference's cap freezes a voxel:
//                                        cpp/include/osn_tsdf/types.hpp:99-101)
//   t  <- (t * w + x) / (w + 1),  w <- w + 1
//
// so t is the mean of the voxel's first min(n, w_max) observations, the
// reference's sum / w (cpp/src/volume.cpp:322-324) with unit weights.
//
// FP32 evaluation. The pixel is a ratio of two linear forms of the voxel
// index, A = fx p.x + (cx + 0.5) p.z and B = fy p.y + (cy + 0.5) p.z, over
// Z = p.z, so each form is its block-origin value plus a per-lane offset, then
// one add per step along x:
//   per frame   (A_j, B_j, Z_j) for the voxel axes j (e_j = voxel * R_cw[:, j])
//   per block   (A0, B0, Z0) from d0 = R_cw * (block origin centre - c)
//   per lane    off = (ly * F_y + lz * F_z) + lx0 * F_x  for its run of 4
//               voxels along x starting at lx0 (0 or 4)
//   per voxel   F = F0 + off, then F += F_x;  uf = A * iz, vf = B * iz
// with every product rounded on its own (-ffp-contract=off), where iz comes
// from the run's batch inversion (run_reciprocals) of its four Z values.
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

// Reciprocals of the four depths of a run by batch inversion: clamp each to
// [kZMin, kZMax] (a voxel outside that range is rejected anyway), form the
// running products, divide once (correctly rounded), and peel the reciprocals
// off with the partial products. Six multiplies and one division instead of
// four divisions; a few ulp from 1 / z, every step rounded on its own. z gets
// the clamped depths: Z is in range exactly when z == Z (NaN included).
__device__ inline void run_reciprocals(const float Z[4], float iz[4], float z[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) z[j] = fminf(fmaxf(Z[j], kZMin), kZMax);
    const float p1 = z[0] * z[1];
    const float p2 = p1 * z[2];
    const float p3 = p2 * z[3];
    const float r3 = 1.0f / p3;
    iz[3] = r3 * p2;
    const float r2 = r3 * z[3];
    iz[2] = r2 * p1;
    const float r1 = r2 * z[2];
    iz[1] = r1 * z[0];
    iz[0] = r1 * z[1];
}

// Weight caps up to this many frames: the reciprocals 1 / (w + 1) of every
// weight a record is updated from (w < cap) come from a table, correctly
// rounded, instead of the transcendental unit.
constexpr int kRcpTable = 64;

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

// A touched block, uniform over the wave: its pool index (negative: dropped,
// nothing to update), block coordinates and block-origin forms (A0, B0, Z0).
struct BlockRef {
    int32_t bi;
    int32_t bx, by, bz;
    float f0[3];
};

template <typename Params>
__device__ inline BlockRef resolve_block(const Params& k, uint32_t part) {
    const u32x4 e = touch_entry(part);
    BlockRef b;
    b.bi = (int32_t)e.w;
    if (b.bi == kIdxPending) b.bi = glob(k.table)[e.z].block_idx;  // inserted by this call
    unpack_key((int64_t)(((uint64_t)e.y << 32) | e.x), b.bx, b.by, b.bz);
    // the forms are computed here, while the previous block's gathers are in
    // flight, and kept in scalar registers
    float f0[3];
    block_forms(k, b.bx, b.by, b.bz, f0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        b.f0[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(f0[i])));
    return b;
}

// The lane's voxel records of a block, read a whole block ahead of their
// update: FP16 one 16-byte record per run, FP32 two. FP64 records are read
// voxel by voxel after the gates.
template <int STORE>
struct Records {};
template <>
struct Records<kStoreF16> {
    u32x4 h[2];
};
template <>
struct Records<kStoreF32> {
    f32x4 h[2][2];
};

// The loads are issued for every block, a dropped or absent one reading pool
// block 0 (results unused), so the number of loads in flight is the same on
// every path and the wait for the gathers can leave these outstanding.
template <int STORE, typename Params>
__device__ inline Records<STORE> load_records(const Params& k, const BlockRef& b, int lane) {
    Records<STORE> R{};
    const size_t v0 = (size_t)(b.bi < 0 ? 0 : b.bi) * kBlockVoxels + lane * 4;
    if constexpr (STORE == kStoreF16) {
        const u32x4 OSN_GLOBAL* rec = reinterpret_cast<const u32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 4);
        // non-temporal: a record is read once per frame, and the caches are
        // better left to the depth images the whole chip gathers from
        R.h[0] = __builtin_nontemporal_load(&rec[0]);
        R.h[1] = __builtin_nontemporal_load(&rec[64]);  // run 1 starts 256 voxels (1 KiB) later
    } else if constexpr (STORE == kStoreF32) {
        const f32x4 OSN_GLOBAL* rec = reinterpret_cast<const f32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 8);
        R.h[0][0] = __builtin_nontemporal_load(&rec[0]);
        R.h[0][1] = __builtin_nontemporal_load(&rec[1]);
        R.h[1][0] = __builtin_nontemporal_load(&rec[128]);
        R.h[1][1] = __builtin_nontemporal_load(&rec[129]);
    }
    return R;
}

// One wave updates the whole block b (uniform over the wave) from its records
// R, requested one block earlier. The update is latency-bound before it is
// bandwidth-bound, so loads are kept in flight across blocks: the depth of all
// 8 of the lane's voxels is gathered first (pixels chosen without branches: an
// out-of-image voxel reads pixel 0 and is then rejected), then ahead() runs,
// which resolves the next block and requests its records, and only then does
// the wave wait, for the gathers alone (loads complete in issue order).
template <int STORE, bool TAB, typename Params, typename Ahead>
__device__ inline void update_block(const Params& k, const LaneRuns& L, const BlockRef& b,
                                    const Records<STORE>& R, int lane, const float* rtab,
                                    Ahead&& ahead) {
    if (b.bi < 0) {
        ahead();
        return;
    }
    const int32_t bx = b.bx, by = b.by, bz = b.bz;
    const size_t v0 = (size_t)b.bi * kBlockVoxels + lane * 4;  // first voxel of run 0

    const float* f0 = b.f0;
    float Zv[8];
    uint32_t pix[8];
    bool inb[8];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        float A[4], B[4], Z[4], iz[4], zc[4];
        A[0] = f0[0] + L.off[r][0];
        B[0] = f0[1] + L.off[r][1];
        Z[0] = f0[2] + L.off[r][2];
#pragma unroll
        for (int j = 1; j < 4; ++j) {
            A[j] = A[j - 1] + k.g[0][0];
            B[j] = B[j - 1] + k.g[1][0];
            Z[j] = Z[j - 1] + k.g[2][0];
        }
        run_reciprocals(Z, iz, zc);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float uf = A[j] * iz[j];
            const float vf = B[j] * iz[j];
            const bool in = (zc[j] == Z[j]) & (uf >= 0.0f) & (vf >= 0.0f) &
                            (uf < (float)k.W) & (vf < (float)k.H);
            // uf, vf >= 0: truncation is floor; the depth is read from the
            // 8x4-tiled copy of the image
            const uint32_t u = in ? (uint32_t)uf : 0u, w = in ? (uint32_t)vf : 0u;
            pix[r * 4 + j] = tiled_index(u, w, (uint32_t)k.tile_k);
            inb[r * 4 + j] = in;
            Zv[r * 4 + j] = Z[j];
        }
    }
    float d[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) d[i] = glob(k.depth)[pix[i]];
    ahead();
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
            uint32_t word[4] = {R.h[r].x, R.h[r].y, R.h[r].z, R.h[r].w};
            // a voxel whose weight has reached the cap is frozen
            bool up[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) up[j] = o[j] & ((word[j] >> 16) < (uint32_t)k.wcap);
            if (up[0] | up[1] | up[2] | up[3]) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float x = fminf(sd[j] * k.inv_trunc, 1.0f);
                    const float tt = __half2float(__ushort_as_half((unsigned short)(word[j] & 0xffffu)));
                    const uint32_t wi = word[j] >> 16;
                    float tn;
                    if constexpr (TAB) {
                        // wi < w_max <= kRcpTable when the record is updated
                        tn = __builtin_fmaf(tt, (float)wi, x) *
                             rtab[wi < kRcpTable - 1 ? wi : kRcpTable - 1];
                    } else {
                        tn = avg_step(tt, (float)wi, x);
                    }
                    // the weight w + 1 (below the cap, so no carry out of the
                    // high half), the new mean in the low half
                    const uint32_t nw = ((word[j] + 0x10000u) & 0xffff0000u) |
                                        (uint32_t)__half_as_ushort(__float2half_rn(tn));
                    word[j] = up[j] ? nw : word[j];
                }
                u32x4 OSN_GLOBAL* rec = reinterpret_cast<u32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 4);
                rec[r * 64] = u32x4{word[0], word[1], word[2], word[3]};
            }
        } else if constexpr (STORE == kStoreF32) {
            float tv[4] = {R.h[r][0].x, R.h[r][0].z, R.h[r][1].x, R.h[r][1].z};
            float wv[4] = {R.h[r][0].y, R.h[r][0].w, R.h[r][1].y, R.h[r][1].w};
            bool up[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) up[j] = o[j] & (wv[j] < k.wmax);
            if (up[0] | up[1] | up[2] | up[3]) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float x = fminf(sd[j] * k.inv_trunc, 1.0f);
                    const float tn = avg_step(tv[j], wv[j], x);
                    tv[j] = up[j] ? tn : tv[j];
                    wv[j] = up[j] ? wv[j] + 1.0f : wv[j];
                }
                f32x4 OSN_GLOBAL* rec = reinterpret_cast<f32x4 OSN_GLOBAL*>(glob(k.vox) + v0 * 8);
                if (up[0] | up[1]) rec[r * 128 + 0] = f32x4{tv[0], wv[0], tv[1], wv[1]};
                if (up[2] | up[3]) rec[r * 128 + 1] = f32x4{tv[2], wv[2], tv[3], wv[3]};
            }
        } else {
            const int ly = L.ly, lz = L.lz0 + 4 * r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!o[j]) continue;
                f64x2 OSN_GLOBAL* rec = reinterpret_cast<f64x2 OSN_GLOBAL*>(
                    glob(k.vox) + (v0 + r * 256 + j) * 16);
                f64x2 q = *rec;
                if (!(q.y < (double)k.wmax)) continue;  // frozen at the cap
                const double vs = (double)k.voxel;
                const double wx = ((double)(bx * 8 + L.lx0 + j) + 0.5) * vs - k.c64[0];
                const double wy = ((double)(by * 8 + ly) + 0.5) * vs - k.c64[1];
                const double wz = ((double)(bz * 8 + lz) + 0.5) * vs - k.c64[2];
                const double z = k.Rcw64[6] * wx + k.Rcw64[7] * wy + k.Rcw64[8] * wz;
                double x = ((double)d[r * 4 + j] - z) / (double)k.trunc;
                x = x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x);
                const double w = q.y;
                q.x = (q.x * w + x) / (w + 1.0);
                q.y = w + 1.0;
                *rec = q;
            }
        }
    }
}

__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

template <int STORE, bool TAB>
__global__ __launch_bounds__(512) void k_proj_update(VolDev v, ProjCam cam, CallInfo call) {
    [[maybe_unused]] __shared__ float s_rcp[TAB ? kRcpTable : 1];
    if constexpr (TAB) {
        if (threadIdx.x < kRcpTable) s_rcp[threadIdx.x] = 1.0f / (float)(threadIdx.x + 1);
        __syncthreads();
    }
    const int waves = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const UpdParams k = make_params(v, cam, call.parity);
    const LaneRuns L = lane_runs(k, lane);
    const int stride = gridDim.x * waves;
    int item = blockIdx.x * waves + wave_id();
    if (item >= k.n) return;
    // Pipeline: while block i is updated, the entry of block i + 2 and the
    // records of block i + 1 are in flight. A pass updates cur from `use` while
    // it fills `fill` and `ent_fill`; passes alternate the two register sets,
    // so nothing in flight is ever copied (a copy would wait for the load).
    BlockRef cur = resolve_block(k, load_touch(k, item, lane));
    Records<STORE> ra = load_records<STORE>(k, cur, lane), rb{};
    int next = item + stride;
    uint32_t ea = load_touch(k, next < k.n ? next : 0, lane), eb = 0u;
    auto pass = [&](const Records<STORE>& use, Records<STORE>& fill, uint32_t ent,
                    uint32_t& ent_fill) {
        const int next2 = next + stride;
        BlockRef nb{-1, 0, 0, 0, {0.0f, 0.0f, 0.0f}};
        update_block<STORE, TAB>(k, L, cur, use, lane, s_rcp, [&] {
            if (next < k.n) nb = resolve_block(k, ent);
            ent_fill = load_touch(k, next2 < k.n ? next2 : 0, lane);  // issued on every path
            fill = load_records<STORE>(k, nb, lane);
        });
        cur = nb;
        const bool more = next < k.n;
        next = next2;
        return more;
    };
    while (pass(ra, rb, ea, eb) && pass(rb, ra, eb, ea)) {
    }
}

// All scenes of a batched step in one launch: the work list is the
// concatenation of every active scene's touched blocks. Each wave holds the
// prefix of the scenes' touched counts in a register (lane s: the first item
// of scene s), so the scene of an item is one ballot. The list is cut into one
// contiguous chunk per XCD (workgroups are dealt to the 8 XCDs round-robin by
// id), so each XCD's L2 serves the depth images of only the few scenes in its
// chunk.
static_assert(kMaxBatchScenes <= 64, "one lane per scene");

template <int STORE, bool TAB>
__global__ __launch_bounds__(512) void k_proj_update_batch(BatchDev b) {
    [[maybe_unused]] __shared__ float s_rcp[TAB ? kRcpTable : 1];
    if constexpr (TAB) {
        if (threadIdx.x < kRcpTable) s_rcp[threadIdx.x] = 1.0f / (float)(threadIdx.x + 1);
        __syncthreads();
    }
    // the step's allocate and assign kernels, the only readers of the step
    // counter, have finished: advance it for the next replay
    if (b.step_advance && blockIdx.x == 0 && threadIdx.x == 0) *glob(b.step_advance) += 1;
    const int waves = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const UpdParams OSN_CONST* P = (const UpdParams OSN_CONST*)b.upd;
    const int cnt = lane < b.n_scenes ? P[lane].n : 0;
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    // lanes past the last scene hold n, above every item
    const int first = incl - cnt;
    const int n = __builtin_amdgcn_readlane(incl, 63);
    auto scene_of = [&](int x) { return (int)__popcll(__ballot(first <= x)) - 1; };
    auto first_of = [&](int sc) { return __builtin_amdgcn_readlane(first, sc); };
    const int xcd = blockIdx.x % kXcds;
    const int per_xcd = gridDim.x / kXcds;  // the launcher keeps the grid a multiple of 8
    const int lo = (int)((long long)n * xcd / kXcds);
    const int hi = (int)((long long)n * (xcd + 1) / kXcds);
    const int stride = per_xcd * waves;
    const int item = lo + (int)(blockIdx.x / kXcds) * waves + wave_id();
    if (item >= hi) return;
    // Pipeline as in k_proj_update.
    int sc = scene_of(item);
    BlockRef cur = resolve_block(P[sc], load_touch(P[sc], item - first_of(sc), lane));
    Records<STORE> ra = load_records<STORE>(P[sc], cur, lane), rb{};
    int next = item + stride;
    int scn = next < hi ? scene_of(next) : sc;
    // entry 0 of this scene stands in for an absent next entry (never used)
    uint32_t ea = load_touch(P[scn], next < hi ? next - first_of(scn) : 0, lane), eb = 0u;
    int lanes_sc = -1;
    LaneRuns L{};
    auto pass = [&](const Records<STORE>& use, Records<STORE>& fill, uint32_t ent,
                    uint32_t& ent_fill) {
        const int next2 = next + stride;
        const int scn2 = next2 < hi ? scene_of(next2) : scn;
        if (sc != lanes_sc) {  // the lane offsets depend on the scene's pose
            L = lane_runs(P[sc], lane);
            lanes_sc = sc;
        }
        BlockRef nb{-1, 0, 0, 0, {0.0f, 0.0f, 0.0f}};
        update_block<STORE, TAB>(P[sc], L, cur, use, lane, s_rcp, [&] {
            if (next < hi) nb = resolve_block(P[scn], ent);
            ent_fill = load_touch(P[scn2], next2 < hi ? next2 - first_of(scn2) : 0, lane);
            fill = load_records<STORE>(P[scn], nb, lane);
        });
        cur = nb;
        sc = scn;
        scn = scn2;
        const bool more = next < hi;
        next = next2;
        return more;
    };
    while (pass(ra, rb, ea, eb) && pass(rb, ra, eb, ea)) {
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
        case kStoreF16:
            if (v.w_max <= kRcpTable)
                k_proj_update<kStoreF16, true><<<grid, block_threads, 0, s>>>(v, cam, call);
            else
                k_proj_update<kStoreF16, false><<<grid, block_threads, 0, s>>>(v, cam, call);
            break;
        case kStoreF32: k_proj_update<kStoreF32, false><<<grid, block_threads, 0, s>>>(v, cam, call); break;
        case kStoreF64: k_proj_update<kStoreF64, false><<<grid, block_threads, 0, s>>>(v, cam, call); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_proj_update_batch(const BatchDev& b, int32_t storage, int32_t grid,
                                    int32_t block_threads, hipStream_t s,
                                    const LaunchEvents& ev, int32_t max_w_max) {
    if (b.n_scenes > kMaxBatchScenes) return hipErrorInvalidValue;
    if (grid <= 0) grid = 4096;
    grid = (grid + kXcds - 1) / kXcds * kXcds;
    if (block_threads != 256 && block_threads != 128 && block_threads != 512)
        return hipErrorInvalidValue;
    switch (storage) {
        case kStoreF16:
            if (max_w_max <= kRcpTable)
                launch_kernel(k_proj_update_batch<kStoreF16, true>, dim3(grid), dim3(block_threads), s, ev, b);
            else
                launch_kernel(k_proj_update_batch<kStoreF16, false>, dim3(grid), dim3(block_threads), s, ev, b);
            break;
        case kStoreF32: launch_kernel(k_proj_update_batch<kStoreF32, false>, dim3(grid), dim3(block_threads), s, ev, b); break;
        case kStoreF64: launch_kernel(k_proj_update_batch<kStoreF64, false>, dim3(grid), dim3(block_threads), s, ev, b); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace osn_hip
