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
// performs the 7 neighbour-block lookups once, stages the block's 9x9x9
// corner samples in LDS and classifies from there.
//
// Compiled with -ffp-contract=off so vertex arithmetic matches the reference
// CPU extractor's rounding.

#include <hip/hip_fp16.h>

#include "internal.hpp"

namespace osn_hip {
namespace {

constexpr int kExThreads = 256;
constexpr int kCube = 9 * 9 * 9;
constexpr int kScanChunk = 1024;

__constant__ int c_tets[6][4] = {
    {0, 5, 1, 7}, {0, 1, 3, 7}, {0, 3, 2, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 2, 6, 7},
};

struct Cube {
    float val[kCube];     // mean tsdf, NaN when the corner is not a valid sample
    float col[3][kCube];  // mean colour (sums mode only)
    int32_t nb[8];        // pool index of the block at offset (x, y, z) bits
};

__device__ inline int cube_index(int x, int y, int z) { return (z * 9 + y) * 9 + x; }

template <int STORE>
__device__ void load_cube(const VolDev& v, int32_t bi, float min_w, Cube& c) {
    const int tid = threadIdx.x;
    if (tid < 8) {
        int32_t nbi = bi;
        if (tid != 0) {
            const int32_t bx = v.block_coord[bi * 3 + 0] + (tid & 1);
            const int32_t by = v.block_coord[bi * 3 + 1] + ((tid >> 1) & 1);
            const int32_t bz = v.block_coord[bi * 3 + 2] + ((tid >> 2) & 1);
            nbi = find_index(v, pack_coord(bx, by, bz));
        }
        c.nb[tid] = nbi;
    }
    __syncthreads();
    for (int i = tid; i < kCube; i += kExThreads) {
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
        c.col[0][i] = cr; c.col[1][i] = cg; c.col[2][i] = cb;
    }
    __syncthreads();
}

// Vertex count of voxel (lx, ly, lz)'s cube; fills the corner values.
__device__ inline int cube_count(const Cube& c, int lx, int ly, int lz, float iso, float val[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        val[k] = c.val[cube_index(lx + (k & 1), ly + ((k >> 1) & 1), lz + ((k >> 2) & 1))];
        if (__builtin_isnan(val[k])) return 0;
    }
    int n = 0;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int code = (val[c_tets[t][0]] < iso ? 1 : 0) | (val[c_tets[t][1]] < iso ? 2 : 0) |
                         (val[c_tets[t][2]] < iso ? 4 : 0) | (val[c_tets[t][3]] < iso ? 8 : 0);
        n += (code == 0 || code == 15) ? 0 : ((code == 3 || code == 5 || code == 6 ||
                                              code == 9 || code == 10 || code == 12) ? 6 : 3);
    }
    return n;
}

struct V6 { float p[3]; float c[3]; };

__device__ inline V6 interp(float iso, const float* pa, const float* pb, float va, float vb,
                            const float* ca, const float* cb) {
    const bool sw = (pb[0] < pa[0]) || (pb[0] == pa[0] && pb[1] < pa[1]) ||
                    (pb[0] == pa[0] && pb[1] == pa[1] && pb[2] < pa[2]);
    const float* p0 = sw ? pb : pa;
    const float* p1 = sw ? pa : pb;
    const float* c0 = sw ? cb : ca;
    const float* c1 = sw ? ca : cb;
    const float v0 = sw ? vb : va;
    const float v1 = sw ? va : vb;
    const float den = v1 - v0;
    float t = fabsf(den) < 1e-12f ? 0.5f : (iso - v0) / den;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    V6 o;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o.p[k] = p0[k] + t * (p1[k] - p0[k]);
        o.c[k] = c0[k] + t * (c1[k] - c0[k]);
    }
    return o;
}

__device__ inline void write_tri(const V6& a, const V6& b, const V6& c, float* posnor,
                                 uint8_t* rgb, int64_t at) {
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
        float* d = posnor + (size_t)(at + k) * 6;
        d[0] = vs[k]->p[0]; d[1] = vs[k]->p[1]; d[2] = vs[k]->p[2];
        d[3] = nx; d[4] = ny; d[5] = nz;
        if (rgb) {
            uint8_t* q = rgb + (size_t)(at + k) * 3;
            q[0] = (uint8_t)fminf(fmaxf(vs[k]->c[0], 0.0f), 255.0f);
            q[1] = (uint8_t)fminf(fmaxf(vs[k]->c[1], 0.0f), 255.0f);
            q[2] = (uint8_t)fminf(fmaxf(vs[k]->c[2], 0.0f), 255.0f);
        }
    }
}

__device__ void emit_cube(const Cube& c, const VolDev& v, int32_t bx, int32_t by, int32_t bz,
                          int lx, int ly, int lz, float iso, const float val[8], float* posnor,
                          uint8_t* rgb, int64_t at) {
    float pos[8][3], col[8][3];
    const int32_t vx = bx * 8 + lx, vy = by * 8 + ly, vz = bz * 8 + lz;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        pos[k][0] = ((float)(vx + ox) + 0.5f) * v.voxel;
        pos[k][1] = ((float)(vy + oy) + 0.5f) * v.voxel;
        pos[k][2] = ((float)(vz + oz) + 0.5f) * v.voxel;
        const int ci = cube_index(lx + ox, ly + oy, lz + oz);
        col[k][0] = c.col[0][ci]; col[k][1] = c.col[1][ci]; col[k][2] = c.col[2][ci];
    }
    for (int t = 0; t < 6; ++t) {
        const int i0 = c_tets[t][0], i1 = c_tets[t][1], i2 = c_tets[t][2], i3 = c_tets[t][3];
        const int code = (val[i0] < iso ? 1 : 0) | (val[i1] < iso ? 2 : 0) |
                         (val[i2] < iso ? 4 : 0) | (val[i3] < iso ? 8 : 0);
        int a, b, cc, d;
        bool quad;
        switch (code) {
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
            default: continue;
        }
        if (!quad) {
            const V6 e0 = interp(iso, pos[a], pos[b], val[a], val[b], col[a], col[b]);
            const V6 e1 = interp(iso, pos[a], pos[cc], val[a], val[cc], col[a], col[cc]);
            const V6 e2 = interp(iso, pos[a], pos[d], val[a], val[d], col[a], col[d]);
            write_tri(e0, e1, e2, posnor, rgb, at);
            at += 3;
        } else {
            const V6 ac = interp(iso, pos[a], pos[cc], val[a], val[cc], col[a], col[cc]);
            const V6 ad = interp(iso, pos[a], pos[d], val[a], val[d], col[a], col[d]);
            const V6 bd = interp(iso, pos[b], pos[d], val[b], val[d], col[b], col[d]);
            const V6 bc = interp(iso, pos[b], pos[cc], val[b], val[cc], col[b], col[cc]);
            write_tri(ac, ad, bd, posnor, rgb, at);
            write_tri(ac, bd, bc, posnor, rgb, at + 3);
            at += 6;
        }
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
__global__ __launch_bounds__(kExThreads) void k_extract_count(VolDev v, ExtractArgs a) {
    __shared__ Cube c;
    __shared__ int s_wave[kExThreads / 64];
    const int32_t n = n_blocks_of(v, a);
    for (int32_t b = blockIdx.x; b < n; b += gridDim.x) {
        load_cube<STORE>(v, b, a.min_weight, c);
        int cnt = 0;
        for (int vi = threadIdx.x; vi < kBlockVoxels; vi += kExThreads) {
            float val[8];
            cnt += cube_count(c, vi & 7, (vi >> 3) & 7, vi >> 6, a.iso, val);
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

template <int STORE>
__global__ __launch_bounds__(kExThreads) void k_extract_emit(VolDev v, ExtractArgs a) {
    __shared__ Cube c;
    __shared__ int s_wave[kExThreads / 64];
    if (*a.count_out < 0) return;  // overflow: write nothing
    const int32_t n = n_blocks_of(v, a);
    for (int32_t b = blockIdx.x; b < n; b += gridDim.x) {
        if (a.block_vcount[b] == 0) continue;
        load_cube<STORE>(v, b, a.min_weight, c);
        const int32_t bx = v.block_coord[b * 3 + 0];
        const int32_t by = v.block_coord[b * 3 + 1];
        const int32_t bz = v.block_coord[b * 3 + 2];
        int64_t base = (int64_t)a.chunk_sums[b / kScanChunk] + a.block_voff[b];
        for (int r = 0; r < kBlockVoxels / kExThreads; ++r) {
            const int vi = r * kExThreads + threadIdx.x;
            const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
            float val[8];
            const int cnt = cube_count(c, lx, ly, lz, a.iso, val);
            int total;
            const int off = wg_exclusive_scan(cnt, s_wave, total);
            if (cnt) emit_cube(c, v, bx, by, bz, lx, ly, lz, a.iso, val, a.posnor, a.rgb, base + off);
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
    const int32_t n_for_scan = a.n_blocks >= 0 ? a.n_blocks : v.pool_capacity;
    const int32_t scan_grid = (n_for_scan + kScanChunk - 1) / kScanChunk > 0
                                  ? (n_for_scan + kScanChunk - 1) / kScanChunk : 1;
    if (!emit) {
        switch (v.storage) {
            case kStoreSums: k_extract_count<kStoreSums><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
            case kStoreF32: k_extract_count<kStoreF32><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
            case kStoreF16: k_extract_count<kStoreF16><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
            case kStoreF64: k_extract_count<kStoreF64><<<grid_blocks, kExThreads, 0, s>>>(v, a); break;
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
