// Marching-tetrahedra extraction as count -> scan -> emit, with no atomics on
// vertex placement: every vertex lands at (block offset + voxel offset within
// the block + tetrahedron order), so the same volume state always yields the
// same bytes.
//
// Per cube semantics follow cpp/src/marching_tets.cpp:84-243: a cube per voxel
// spanning (v..v+1)^3, corners sampled at voxel centres, any invalid corner
// (absent block, w < min_weight or w <= 0) skips the cube, 6 tetrahedra sharing
// the 0-7 diagonal, position-ordered interpolation, one face normal per
// triangle. The reference launches one thread per voxel and performs 8 hash
// lookups per voxel (cpp/src/marching_tets_cuda.cu:73-167); here a workgroup
// stages the block's 9x9x9 corner samples in LDS and classifies from there.
// The count pass looks the 7 neighbour blocks up once per block and keeps
// their indices for the emit pass, which assembles each round's vertices in
// LDS and writes them out as one contiguous, coalesced run.
//
// Compiled with -ffp-contract=off so vertex arithmetic matches the reference
// CPU extractor's rounding.

#include <hip/hip_fp16.h>

#include "internal.hpp"

namespace osn_hip {
namespace {

constexpr int kExThreads = 256;     // emit workgroup
constexpr int kCountThreads = 64;   // count workgroup: one wave per block, many blocks per CU
constexpr int kEmitWavesPerSimd = 6;  // emit's register bound (measured faster than 4 at 99 VGPRs)
constexpr int kCube = 9 * 9 * 9;
constexpr int kScanChunk = 1024;


constexpr int kStageFloats = 4096;  // emit staging: 682 vertices of 6 floats (16 KiB)

// The 9x9x9 corner samples of one block; colour planes only where the
// storage keeps colour (the reference's weighted sums).
template <int STORE>
struct Cube {
    float val[kCube];     // mean tsdf, NaN when the corner is not a valid sample
    int32_t nb[8];        // pool index of the block at offset (x, y, z) bits
};
template <>
struct Cube<kStoreSums> {
    float val[kCube];
    int32_t nb[8];
    float col[3][kCube];  // mean colour
};

__device__ inline int cube_index(int x, int y, int z) { return (z * 9 + y) * 9 + x; }

// Stage block bi's corner samples. The count pass (LOOKUP) finds the 7
// neighbour blocks in the table and records their indices in nb_out[bi * 8];
// the emit pass reads them back from there.
template <int STORE, bool LOOKUP>
__device__ void load_cube(const VolDev& v, int32_t bi, float min_w, Cube<STORE>& c,
                          int32_t* nb_out) {
    const int tid = threadIdx.x;
    if (tid < 8) {
        int32_t nbi = bi;
        if (LOOKUP) {
            if (tid != 0) {
                const int32_t bx = v.block_coord[bi * 3 + 0] + (tid & 1);
                const int32_t by = v.block_coord[bi * 3 + 1] + ((tid >> 1) & 1);
                const int32_t bz = v.block_coord[bi * 3 + 2] + ((tid >> 2) & 1);
                nbi = find_index(v, pack_coord(bx, by, bz));
            }
            nb_out[(size_t)bi * 8 + tid] = nbi;
        } else {
            nbi = nb_out[(size_t)bi * 8 + tid];
        }
        c.nb[tid] = nbi;
    }
    __syncthreads();
    for (int i = tid; i < kCube; i += blockDim.x) {
        const int x = i % 9, y = (i / 9) % 9, z = i / 81;
        const int o = (x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2);
        const int32_t nbi = c.nb[o];
        float t = __builtin_nanf("");
        float cr = 0.f, cg = 0.f, cb = 0.f;
        if (nbi >= 0) {
            const size_t vi = (size_t)nbi * kBlockVoxels + (((z & 7) * 8 + (y & 7)) * 8 + (x & 7));
            if (STORE == kStoreSums) {
                const float w = v.weight[vi];
                if (!(w < min_w || w <= 0.0f)) {
                    const float inv_w = 1.0f / w;
                    t = v.tsdf[vi] * inv_w;
                    cr = v.r[vi] * inv_w;
                    cg = v.g[vi] * inv_w;
                    cb = v.b[vi] * inv_w;
                }
            } else if (STORE == kStoreF32) {
                const float2 q = static_cast<const float2*>(v.vox)[vi];
                if (!(q.y < min_w || q.y <= 0.0f)) { t = q.x; cr = cg = cb = 128.0f; }
            } else if (STORE == kStoreF16) {
                const uint32_t q = static_cast<const uint32_t*>(v.vox)[vi];
                const float w = (float)(q >> 16);
                if (!(w < min_w || w <= 0.0f)) {
                    t = __half2float(__ushort_as_half((unsigned short)(q & 0xffffu)));
                    cr = cg = cb = 128.0f;
                }
            } else {
                const double2 q = static_cast<const double2*>(v.vox)[vi];
                const float w = (float)q.y;
                if (!(w < min_w || w <= 0.0f)) { t = (float)q.x; cr = cg = cb = 128.0f; }
            }
        }
        c.val[i] = t;
        if constexpr (STORE == kStoreSums) {
            c.col[0][i] = cr; c.col[1][i] = cg; c.col[2][i] = cb;
        }
    }
    __syncthreads();
}

// The 6 tetrahedra (corner ids, 4 bits each, lowest first):
// {0,5,1,7} {0,1,3,7} {0,3,2,7} {0,6,4,7} {0,4,5,7} {0,2,6,7}.
// Kept as immediates: an array indexed by a lane-varying t would live in
// scratch memory.
__device__ inline uint32_t tet_corners(int t) {
    return t < 4 ? (uint32_t)(0x7460723073107150ull >> (16 * t)) & 0xffffu
                 : (0x76207540u >> (16 * (t - 4))) & 0xffffu;
}

// Crossing code of tetrahedron t from the cube's corner mask (bit k set when
// corner k is below iso), and the vertices it emits: none, one triangle, or
// two from a quad.
__device__ inline int tet_code(uint32_t mask, int t) {
    const uint32_t tc = tet_corners(t);
    return (int)(((mask >> (tc & 15u)) & 1u) | (((mask >> ((tc >> 4) & 15u)) & 1u) << 1) |
                 (((mask >> ((tc >> 8) & 15u)) & 1u) << 2) | (((mask >> (tc >> 12)) & 1u) << 3));
}

__device__ inline int tet_vertices(int code) {
    return (code == 0 || code == 15) ? 0 : ((code == 3 || code == 5 || code == 6 ||
                                            code == 9 || code == 10 || code == 12) ? 6 : 3);
}

// Vertex count of voxel (lx, ly, lz)'s cube (0 when a corner is invalid) and
// its corner mask.
template <int STORE>
__device__ inline int cube_count(const Cube<STORE>& c, int lx, int ly, int lz, float iso,
                                 uint32_t& mask) {
    mask = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float x = c.val[cube_index(lx + (k & 1), ly + ((k >> 1) & 1), lz + ((k >> 2) & 1))];
        if (__builtin_isnan(x)) return 0;
        mask |= (x < iso ? 1u : 0u) << k;
    }
    int n = 0;
#pragma unroll
    for (int t = 0; t < 6; ++t) n += tet_vertices(tet_code(mask, t));
    return n;
}

struct P3 { float x, y, z; };
struct V6 { float p[3]; float c[3]; };

template <bool COL>
__device__ inline V6 interp(float iso, P3 pa, P3 pb, float va, float vb, P3 ca, P3 cb) {
    const bool sw = (pb.x < pa.x) || (pb.x == pa.x && pb.y < pa.y) ||
                    (pb.x == pa.x && pb.y == pa.y && pb.z < pa.z);
    const P3 p0 = sw ? pb : pa, p1 = sw ? pa : pb;
    const P3 c0 = sw ? cb : ca, c1 = sw ? ca : cb;
    const float v0 = sw ? vb : va;
    const float v1 = sw ? va : vb;
    const float den = v1 - v0;
    float t = fabsf(den) < 1e-12f ? 0.5f : (iso - v0) / den;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    V6 o;
    o.p[0] = p0.x + t * (p1.x - p0.x);
    o.p[1] = p0.y + t * (p1.y - p0.y);
    o.p[2] = p0.z + t * (p1.z - p0.z);
    o.c[0] = COL ? c0.x + t * (c1.x - c0.x) : 128.0f;
    o.c[1] = COL ? c0.y + t * (c1.y - c0.y) : 128.0f;
    o.c[2] = COL ? c0.z + t * (c1.z - c0.z) : 128.0f;
    return o;
}

// Triangle (a, b, c) with its face normal: positions and normals to posnor
// from float index at_pn * 6 (global memory or the LDS staging area), colours
// to rgb from byte index at_rgb * 3 when rgb is set.
__device__ inline void write_tri(const V6& a, const V6& b, const V6& c, float* posnor,
                                 int64_t at_pn, uint8_t* rgb, int64_t at_rgb) {
    const float e1[3] = {b.p[0] - a.p[0], b.p[1] - a.p[1], b.p[2] - a.p[2]};
    const float e2[3] = {c.p[0] - a.p[0], c.p[1] - a.p[1], c.p[2] - a.p[2]};
    float nx = e1[1] * e2[2] - e1[2] * e2[1];
    float ny = e1[2] * e2[0] - e1[0] * e2[2];
    float nz = e1[0] * e2[1] - e1[1] * e2[0];
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    if (len > 1e-20f) { nx /= len; ny /= len; nz /= len; }
    const V6* vs[3] = {&a, &b, &c};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float* d = posnor + (size_t)(at_pn + k) * 6;
        d[0] = vs[k]->p[0]; d[1] = vs[k]->p[1]; d[2] = vs[k]->p[2];
        d[3] = nx; d[4] = ny; d[5] = nz;
        if (rgb) {
            uint8_t* q = rgb + (size_t)(at_rgb + k) * 3;
            q[0] = (uint8_t)fminf(fmaxf(vs[k]->c[0], 0.0f), 255.0f);
            q[1] = (uint8_t)fminf(fmaxf(vs[k]->c[1], 0.0f), 255.0f);
            q[2] = (uint8_t)fminf(fmaxf(vs[k]->c[2], 0.0f), 255.0f);
        }
    }
}

// The triangles of tetrahedron t of voxel (lx, ly, lz)'s cube (corner mask
// `mask`), written from vertex index at (posnor) and at_rgb (rgb). Corner
// positions, values and colours are formed per corner id, never held in an
// array indexed by a lane-varying value.
template <int STORE>
__device__ void emit_tet(const Cube<STORE>& c, const VolDev& v, int32_t bx, int32_t by,
                         int32_t bz, int lx, int ly, int lz, float iso, uint32_t mask, int t,
                         float* posnor, int64_t at, uint8_t* rgb, int64_t at_rgb) {
    constexpr bool kCol = STORE == kStoreSums;
    const int32_t vx = bx * 8 + lx, vy = by * 8 + ly, vz = bz * 8 + lz;
    auto pos = [&](int k) {
        return P3{((float)(vx + (k & 1)) + 0.5f) * v.voxel,
                  ((float)(vy + ((k >> 1) & 1)) + 0.5f) * v.voxel,
                  ((float)(vz + ((k >> 2) & 1)) + 0.5f) * v.voxel};
    };
    auto ci = [&](int k) { return cube_index(lx + (k & 1), ly + ((k >> 1) & 1), lz + ((k >> 2) & 1)); };
    auto val = [&](int k) { return c.val[ci(k)]; };
    auto col = [&](int k) {
        if constexpr (kCol) return P3{c.col[0][ci(k)], c.col[1][ci(k)], c.col[2][ci(k)]};
        else return P3{128.0f, 128.0f, 128.0f};
    };
    const uint32_t tc = tet_corners(t);
    const int i0 = (int)(tc & 15u), i1 = (int)((tc >> 4) & 15u), i2 = (int)((tc >> 8) & 15u),
              i3 = (int)(tc >> 12);
    int a, b, cc, d;
    bool quad;
    switch (tet_code(mask, t)) {
        case 0x01: case 0x0E: a = i0; b = i1; cc = i2; d = i3; quad = false; break;
        case 0x02: case 0x0D: a = i1; b = i0; cc = i2; d = i3; quad = false; break;
        case 0x04: case 0x0B: a = i2; b = i0; cc = i1; d = i3; quad = false; break;
        case 0x08: case 0x07: a = i3; b = i0; cc = i1; d = i2; quad = false; break;
        case 0x03: a = i0; b = i1; cc = i2; d = i3; quad = true; break;
        case 0x0C: a = i2; b = i3; cc = i0; d = i1; quad = true; break;
        case 0x05: a = i0; b = i2; cc = i1; d = i3; quad = true; break;
        case 0x0A: a = i1; b = i3; cc = i0; d = i2; quad = true; break;
        case 0x06: a = i1; b = i2; cc = i0; d = i3; quad = true; break;
        case 0x09: a = i0; b = i3; cc = i1; d = i2; quad = true; break;
        default: return;
    }
    const P3 pa = pos(a), pb = pos(b), pc = pos(cc), pd = pos(d);
    const float va = val(a), vb = val(b), vc = val(cc), vd = val(d);
    const P3 ca = col(a), cb = col(b), c3 = col(cc), cd = col(d);
    if (!quad) {
        const V6 e0 = interp<kCol>(iso, pa, pb, va, vb, ca, cb);
        const V6 e1 = interp<kCol>(iso, pa, pc, va, vc, ca, c3);
        const V6 e2 = interp<kCol>(iso, pa, pd, va, vd, ca, cd);
        write_tri(e0, e1, e2, posnor, at, rgb, at_rgb);
    } else {
        const V6 ac = interp<kCol>(iso, pa, pc, va, vc, ca, c3);
        const V6 ad = interp<kCol>(iso, pa, pd, va, vd, ca, cd);
        const V6 bd = interp<kCol>(iso, pb, pd, vb, vd, cb, cd);
        const V6 bc = interp<kCol>(iso, pb, pc, vb, vc, cb, c3);
        write_tri(ac, ad, bd, posnor, at, rgb, at_rgb);
        write_tri(ac, bd, bc, posnor, at + 3, rgb, at_rgb + 3);
    }
}

// Every tetrahedron of one cube, in order.
template <int STORE>
__device__ void emit_cube(const Cube<STORE>& c, const VolDev& v, int32_t bx, int32_t by,
                          int32_t bz, int lx, int ly, int lz, float iso, uint32_t mask,
                          float* posnor, int64_t at, uint8_t* rgb, int64_t at_rgb) {
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int n = tet_vertices(tet_code(mask, t));
        if (!n) continue;
        emit_tet(c, v, bx, by, bz, lx, ly, lz, iso, mask, t, posnor, at, rgb, at_rgb);
        at += n;
        at_rgb += n;
    }
}

__device__ inline int32_t n_blocks_of(const VolDev& v, const ExtractArgs& a) {
    return a.n_blocks >= 0 ? a.n_blocks : v.ctr->block_count;
}

// Exclusive scan of one value per thread across the workgroup; returns the
// prefix and writes the workgroup total.
__device__ inline int wg_exclusive_scan(int x, int* s_wave, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int before = 0, all = 0;
    const int n_waves = blockDim.x >> 6;
    for (int w = 0; w < n_waves; ++w) {
        const int s = s_wave[w];
        if (w < wave) before += s;
        all += s;
    }
    __syncthreads();
    total = all;
    return before + incl - x;
}

template <int STORE>
__global__ __launch_bounds__(kCountThreads) void k_extract_count(VolDev v, ExtractArgs a) {
    __shared__ Cube<STORE> c;
    __shared__ int s_wave[kCountThreads / 64];
    const int32_t n = n_blocks_of(v, a);
    for (int32_t b = blockIdx.x; b < n; b += gridDim.x) {
        load_cube<STORE, true>(v, b, a.min_weight, c, a.block_nb);
        int cnt = 0;
        for (int vi = threadIdx.x; vi < kBlockVoxels; vi += kCountThreads) {
            uint32_t mask;
            cnt += cube_count(c, vi & 7, (vi >> 3) & 7, vi >> 6, a.iso, mask);
        }
        int total;
        wg_exclusive_scan(cnt, s_wave, total);
        if (threadIdx.x == 0) a.block_vcount[b] = total;
        __syncthreads();
    }
}

// Two-level exclusive scan of the per-block counts. Each workgroup scans a
// 1024-block chunk; the last to finish scans the chunk totals, publishes the
// vertex count (or -1 when it exceeds the capacity) and resets the arrival
// counter. Workgroups never wait on each other.
__global__ __launch_bounds__(kScanChunk) void k_extract_scan(VolDev v, ExtractArgs a) {
    __shared__ int s_wave[kScanChunk / 64];
    __shared__ int s_last;
    const int32_t n = n_blocks_of(v, a);
    const int n_chunks = (n + kScanChunk - 1) / kScanChunk;
    if (n == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *a.count_out = 0;
            v.ctr->mesh_vertices = 0;
        }
        return;
    }
    if ((int)blockIdx.x >= n_chunks) return;
    const int b = blockIdx.x * kScanChunk + threadIdx.x;
    const int x = b < n ? a.block_vcount[b] : 0;
    int total;
    const int ex = wg_exclusive_scan(x, s_wave, total);
    if (b < n) a.block_voff[b] = ex;
    if (threadIdx.x == 0) a.chunk_sums[blockIdx.x] = total;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(a.scan_flag, 1) == n_chunks - 1);
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    // last workgroup: exclusive scan of chunk totals in place, 64-bit total
    long long carry = 0;
    for (int base = 0; base < n_chunks; base += kScanChunk) {
        const int i = base + threadIdx.x;
        const int y = i < n_chunks ? a.chunk_sums[i] : 0;
        int t;
        const int e = wg_exclusive_scan(y, s_wave, t);
        if (i < n_chunks) a.chunk_sums[i] = (int)(carry + e);
        carry += t;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int32_t out = carry > (long long)a.cap ? -1 : (int32_t)carry;
        *a.count_out = out;
        v.ctr->mesh_vertices = out;
        *a.scan_flag = 0;
    }
}

// Emit: per block, two rounds of 256 voxels. A round's vertices are one
// contiguous run of the output (its offset from the scans); the workgroup
// writes them into LDS, then copies the run out 8 bytes per lane, so the
// stores are coalesced. A round too large for the staging area writes its
// vertices straight to the output.
template <int STORE>
__global__ __launch_bounds__(kExThreads, kEmitWavesPerSimd) void k_extract_emit(VolDev v, ExtractArgs a) {
    __shared__ Cube<STORE> c;
    __shared__ int s_wave[kExThreads / 64];
    __shared__ float2 s_stage[kStageFloats / 2];
    __shared__ uint32_t s_tets[kExThreads * 6];
    if (*a.count_out < 0) return;  // overflow: write nothing
    const int32_t n = n_blocks_of(v, a);
    for (int32_t b = blockIdx.x; b < n; b += gridDim.x) {
        if (a.block_vcount[b] == 0) continue;
        load_cube<STORE, false>(v, b, a.min_weight, c, a.block_nb);
        const int32_t bx = v.block_coord[b * 3 + 0];
        const int32_t by = v.block_coord[b * 3 + 1];
        const int32_t bz = v.block_coord[b * 3 + 2];
        int64_t base = (int64_t)a.chunk_sums[b / kScanChunk] + a.block_voff[b];
        for (int r = 0; r < kBlockVoxels / kExThreads; ++r) {
            const int vi = r * kExThreads + threadIdx.x;
            const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
            uint32_t mask;
            const int cnt = cube_count(c, lx, ly, lz, a.iso, mask);
            int nt = 0;         // tetrahedra with a crossing
            uint32_t tv = 0;    // 2 bits per tetrahedron: its vertex count / 3
            if (cnt) {
#pragma unroll
                for (int t = 0; t < 6; ++t) {
                    const int nv = tet_vertices(tet_code(mask, t));
                    tv |= (uint32_t)(nv / 3) << (2 * t);
                    nt += nv != 0;
                }
            }
            // one scan of both counts: vertices (low 16 bits, <= 256 x 36)
            // and tetrahedra (high bits, <= 256 x 6)
            int packed_total;
            const int packed_off = wg_exclusive_scan(cnt | (nt << 16), s_wave, packed_total);
            const int off = packed_off & 0xffff, total = packed_total & 0xffff;
            const int toff = packed_off >> 16, n_tets = packed_total >> 16;
            if (total * 6 <= kStageFloats) {
                // List the round's crossing tetrahedra with their vertex
                // offsets, then emit one tetrahedron per lane: a lane per
                // voxel would leave most lanes of a wave idle while a few
                // emit many triangles.
                int vo = off, k = toff;
#pragma unroll
                for (int t = 0; t < 6; ++t) {
                    const int nv = (int)((tv >> (2 * t)) & 3u) * 3;
                    if (nv) {
                        s_tets[k++] = ((uint32_t)vo << 12) | ((uint32_t)t << 9) | (uint32_t)vi;
                        vo += nv;
                    }
                }
                __syncthreads();
                for (int i = threadIdx.x; i < n_tets; i += kExThreads) {
                    const uint32_t it = s_tets[i];
                    const int vj = (int)(it & 511u), t = (int)((it >> 9) & 7u), at = (int)(it >> 12);
                    const int jx = vj & 7, jy = (vj >> 3) & 7, jz = vj >> 6;
                    uint32_t jm;
                    cube_count(c, jx, jy, jz, a.iso, jm);
                    emit_tet(c, v, bx, by, bz, jx, jy, jz, a.iso, jm, t,
                             reinterpret_cast<float*>(s_stage), at, a.rgb, base + at);
                }
                __syncthreads();
                float* out = a.posnor + base * 6;
                if ((reinterpret_cast<uintptr_t>(out) & 7u) == 0) {
                    float2* dst = reinterpret_cast<float2*>(out);
                    for (int i = threadIdx.x; i < total * 3; i += kExThreads) dst[i] = s_stage[i];
                } else {  // a caller's buffer need only be float-aligned
                    const float* src = reinterpret_cast<const float*>(s_stage);
                    for (int i = threadIdx.x; i < total * 6; i += kExThreads) out[i] = src[i];
                }
                __syncthreads();
            } else if (cnt) {
                emit_cube(c, v, bx, by, bz, lx, ly, lz, a.iso, mask, a.posnor, base + off, a.rgb,
                          base + off);
            }
            base += total;
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_extract(const VolDev& v, const ExtractArgs& a, bool emit, hipStream_t s) {
    // grid: one workgroup per block when the count is known on the host,
    // otherwise a grid-stride over the pool sized for the whole chip
    const int32_t grid_blocks = a.n_blocks >= 0 ? (a.n_blocks > 0 ? a.n_blocks : 1)
                                                : (v.pool_capacity < 8192 ? v.pool_capacity : 8192);
    // the one-wave count workgroups: four times the chip's worth when the
    // block count is read on the device
    const int32_t count_blocks = a.n_blocks >= 0 ? grid_blocks
                                                 : (v.pool_capacity < 32768 ? v.pool_capacity : 32768);
    const int32_t n_for_scan = a.n_blocks >= 0 ? a.n_blocks : v.pool_capacity;
    const int32_t scan_grid = (n_for_scan + kScanChunk - 1) / kScanChunk > 0
                                  ? (n_for_scan + kScanChunk - 1) / kScanChunk : 1;
    if (!emit) {
        switch (v.storage) {
            case kStoreSums: k_extract_count<kStoreSums><<<count_blocks, kCountThreads, 0, s>>>(v, a); break;
            case kStoreF32: k_extract_count<kStoreF32><<<count_blocks, kCountThreads, 0, s>>>(v, a); break;
            case kStoreF16: k_extract_count<kStoreF16><<<count_blocks, kCountThreads, 0, s>>>(v, a); break;
            case kStoreF64: k_extract_count<kStoreF64><<<count_blocks, kCountThreads, 0, s>>>(v, a); break;
            default: return hipErrorInvalidValue;
        }
        k_extract_scan<<<scan_grid, kScanChunk, 0, s>>>(v, a);
    } else {
        switch (v.storage) {
            case kStoreSums: k_extract_emit<kStoreSums><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
            case kStoreF32: k_extract_emit<kStoreF32><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
            case kStoreF16: k_extract_emit<kStoreF16><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
            case kStoreF64: k_extract_emit<kStoreF64><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace osn_hip
