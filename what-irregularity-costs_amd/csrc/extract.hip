// Marching-tetrahedra extraction as count -> scan -> emit, with no atomics on
// vertex placement: a block's vertices start at the scan of the per-block
// counts and its triangles follow in (voxel, tetrahedron, quad half) order, so
// the same volume state always yields the same bytes.
//
// Per cube semantics follow cpp/src/marching_tets.cpp:84-243: a cube per voxel
// spanning (v..v+1)^3, corners sampled at voxel centres, any invalid corner
// (absent block, w < min_weight or w <= 0) skips the cube, 6 tetrahedra sharing
// the 0-7 diagonal, position-ordered interpolation, one face normal per
// triangle. The reference launches one thread per voxel and performs 8 hash
// lookups per voxel (cpp/src/marching_tets_cuda.cu:73-167). Here one wave
// handles a block (the count pass two): it stages the block's 9x9x9 corner
// samples in LDS, each lane classifies a row of 8 cubes through a mask table,
// and the emit pass lists the block's triangles and emits one per lane in
// groups of 64, assembled in LDS and written out as one contiguous run. The
// count pass looks the 7 neighbour blocks up and keeps their indices, for the
// emit pass and for later count passes (pool indices never move until a
// reset). A batch runs each pass as one launch for all its volumes.
//
// Compiled with -ffp-contract=off so vertex arithmetic matches the reference
// CPU extractor's rounding.

#include <hip/hip_fp16.h>

#include "internal.hpp"

namespace osn_hip {
namespace {

// Device memory as the global address space. Pointers read from argument
// arrays are otherwise generic, and generic (flat) accesses count against
// the LDS wait counter too: every wait for an LDS read would then also wait
// for the output stores in flight.
template <typename T>
__device__ inline __attribute__((address_space(1))) T* global(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

constexpr int kWaveThreads = 64;     // count and emit: one wave per block, many per CU
// emit's register bound: 6 waves per SIMD with FP16 corners (80 VGPRs, what
// its LDS allows), 4 with FP32 (128), 2 with the colour planes of the
// weighted sums
template <int STORE>
constexpr int emit_waves_per_simd() { return STORE == kStoreSums ? 2 : (STORE == kStoreF16 ? 6 : 4); }
constexpr int kCube = 9 * 9 * 9;
constexpr int kScanGrid = 256;      // scan workgroups (grid-stride over chunks)
constexpr int kSerialChunks = 64;   // up to this many chunks one workgroup scans alone
constexpr int kStripeChunks = 16;   // chunks it loads and scans at a time
constexpr uint32_t kScanGridBatch = 32;  // the same per volume of a batch
constexpr int kSlotCap = 768;         // triangles listed per pass over a block
constexpr int kStageVerts = 3 * kWaveThreads;  // one group: a triangle per lane

// The 9x9x9 corner samples of one block; colour planes only where the
// storage keeps colour (the reference's weighted sums). The FP16 storage's
// mean tsdf is a half, kept as one in LDS (exact; NaN marks an invalid
// corner as for the others).
template <int STORE>
struct Cube {
    float val[kCube];     // mean tsdf, NaN when the corner is not a valid sample
    int32_t nb[8];        // pool index of the block at offset (x, y, z) bits
};
template <>
struct Cube<kStoreF16> {
    unsigned short val[kCube];  // half bits
    int32_t nb[8];
};
template <>
struct Cube<kStoreSums> {
    float val[kCube];
    int32_t nb[8];
    float col[3][kCube];  // mean colour
};

// A voxel record as the packed storages keep it.
template <int STORE> struct RawVox;
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <> struct RawVox<kStoreF32> { using T = f32x2; };    // tsdf, weight
template <> struct RawVox<kStoreF16> { using T = uint32_t; }; // half tsdf | u16 weight << 16
template <> struct RawVox<kStoreF64> { using T = f64x2; };    // tsdf, weight

__device__ inline float corner(float x) { return x; }
__device__ inline float corner(unsigned short h) { return __half2float(__ushort_as_half(h)); }

__device__ inline int cube_index(int x, int y, int z) { return (z * 9 + y) * 9 + x; }

// The FP16 storage's u16 weight w passes !(w < min_w || w <= 0) exactly when
// w >= this floor (w is an integer; a NaN min_w rejects only w = 0).
__device__ inline uint32_t f16_weight_floor(float min_w) {
    if (!(min_w > 1.0f)) return 1u;
    if (min_w > 65535.0f) return 65536u;
    return (uint32_t)ceilf(min_w);
}

// Stage block bi's corner samples. The count pass (LOOKUP) finds the 7
// neighbour blocks in the table and records their indices in nb_out[bi * 8];
// the emit pass reads them back from there. Pool indices never move until a
// reset, so the count pass reuses the indices it found in earlier calls for
// blocks below `cached` and looks up only neighbours that were absent then.
template <int STORE, bool LOOKUP>
__device__ void load_cube(const VolDev& v, int32_t bi, float min_w, Cube<STORE>& c,
                          int32_t* nb_out, int32_t cached = 0) {
    const int tid = threadIdx.x;
    if (tid < 8) {
        int32_t nbi = bi;
        if (LOOKUP) {
            const int32_t known = bi < cached ? global(nb_out)[(size_t)bi * 8 + tid] : -1;
            if (known >= 0) {
                nbi = known;
            } else if (tid != 0) {
                const int32_t bx = global(v.block_coord)[bi * 3 + 0] + (tid & 1);
                const int32_t by = global(v.block_coord)[bi * 3 + 1] + ((tid >> 1) & 1);
                const int32_t bz = global(v.block_coord)[bi * 3 + 2] + ((tid >> 2) & 1);
                nbi = find_index(v, pack_coord(bx, by, bz));
            }
            global(nb_out)[(size_t)bi * 8 + tid] = nbi;
        } else {
            nbi = global(nb_out)[(size_t)bi * 8 + tid];
        }
        c.nb[tid] = nbi;
    }
    __syncthreads();
    if constexpr (STORE != kStoreSums) {
        // Rows of 9 corners (y, z fixed): row `tid` and, for tid < 17, row
        // tid + 64. A row's corners x < 8 lie in one block, x = 8 in its +x
        // neighbour: two base addresses, immediate offsets. Every lane issues
        // all its loads before using any; a corner of an absent block reads
        // its own block (in bounds) and is then marked invalid.
        using T = typename RawVox<STORE>::T;
        const auto vox = global(static_cast<const T*>(v.vox));
        constexpr int kRows = 81;
        T q[2][9];
        bool ok_lo[2], ok_hi[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = tid + h * kWaveThreads;
            const int y = r % 9, z = r / 9;
            const int o = ((y >> 3) << 1) | ((z >> 3) << 2);
            const int32_t nlo = c.nb[o & 7], nhi = c.nb[(o | 1) & 7];
            ok_lo[h] = nlo >= 0;
            ok_hi[h] = nhi >= 0;
            const int row = ((z & 7) * 8 + (y & 7)) * 8;
            const auto plo = vox + (size_t)(ok_lo[h] ? nlo : bi) * kBlockVoxels + row;
            const auto phi = vox + (size_t)(ok_hi[h] ? nhi : bi) * kBlockVoxels + row;
            if (h == 0 || r < kRows) {
#pragma unroll
                for (int x = 0; x < 8; ++x) q[h][x] = plo[x];
                q[h][8] = phi[0];
            }
        }
        [[maybe_unused]] const uint32_t wfloor = f16_weight_floor(min_w);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = tid + h * kWaveThreads;
            if (h == 1 && r >= kRows) break;
#pragma unroll
            for (int x = 0; x < 9; ++x) {
                const bool present = x < 8 ? ok_lo[h] : ok_hi[h];
                if constexpr (STORE == kStoreF16) {
                    c.val[r * 9 + x] = present && (q[h][x] >> 16) >= wfloor
                                           ? (unsigned short)(q[h][x] & 0xffffu)
                                           : (unsigned short)0x7e00u;
                } else {
                    const float w = (float)q[h][x].y;
                    c.val[r * 9 + x] = present && !(w < min_w || w <= 0.0f) ? (float)q[h][x].x
                                                                           : __builtin_nanf("");
                }
            }
        }
    } else {
        for (int i = tid; i < kCube; i += kWaveThreads) {
            const int x = i % 9, y = (i / 9) % 9, z = i / 81;
            const int o = (x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2);
            const int32_t nbi = c.nb[o];
            float t = __builtin_nanf("");
            float cr = 0.f, cg = 0.f, cb = 0.f;
            if (nbi >= 0) {
                const size_t vi = (size_t)nbi * kBlockVoxels + (((z & 7) * 8 + (y & 7)) * 8 + (x & 7));
                const float w = v.weight[vi];
                if (!(w < min_w || w <= 0.0f)) {
                    const float inv_w = 1.0f / w;
                    t = v.tsdf[vi] * inv_w;
                    cr = v.r[vi] * inv_w;
                    cg = v.g[vi] * inv_w;
                    cb = v.b[vi] * inv_w;
                }
            }
            c.val[i] = t;
            c.col[0][i] = cr; c.col[1][i] = cg; c.col[2][i] = cb;
        }
    }
    __syncthreads();
}

// The count pass stages two blocks at once (the packed storages): their
// neighbour indices together (lanes 0-7 and 8-15), then every corner load of
// both in flight before the first wait, so two blocks cost one chain of two
// dependent round trips. b1 < 0: one block.
template <int STORE>
__device__ void load_cube_pair(const VolDev& v, int32_t b0, int32_t b1, float min_w,
                               Cube<STORE>* c, int32_t* nb_out, int32_t cached) {
    static_assert(STORE != kStoreSums, "packed storages only");
    const int tid = threadIdx.x;
    if (tid < 16) {
        const int k = tid & 7, w = tid >> 3;
        const int32_t bi = w ? b1 : b0;
        if (bi >= 0) {
            int32_t nbi = bi;
            const int32_t known = bi < cached ? global(nb_out)[(size_t)bi * 8 + k] : -1;
            if (known >= 0) {
                nbi = known;
            } else if (k != 0) {
                const int32_t bx = global(v.block_coord)[bi * 3 + 0] + (k & 1);
                const int32_t by = global(v.block_coord)[bi * 3 + 1] + ((k >> 1) & 1);
                const int32_t bz = global(v.block_coord)[bi * 3 + 2] + ((k >> 2) & 1);
                nbi = find_index(v, pack_coord(bx, by, bz));
            }
            global(nb_out)[(size_t)bi * 8 + k] = nbi;
            c[w].nb[k] = nbi;
        }
    }
    __syncthreads();
    using T = typename RawVox<STORE>::T;
    const auto vox = global(static_cast<const T*>(v.vox));
    constexpr int kRows = 81;
    const int nw = b1 >= 0 ? 2 : 1;
    T q[2][2][9];
    bool ok_lo[2][2], ok_hi[2][2];
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        if (w >= nw) break;
        const int32_t bi = w ? b1 : b0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = tid + h * kWaveThreads;
            const int y = r % 9, z = r / 9;
            const int o = ((y >> 3) << 1) | ((z >> 3) << 2);
            const int32_t nlo = c[w].nb[o & 7], nhi = c[w].nb[(o | 1) & 7];
            ok_lo[w][h] = nlo >= 0;
            ok_hi[w][h] = nhi >= 0;
            const int row = ((z & 7) * 8 + (y & 7)) * 8;
            const auto plo = vox + (size_t)(ok_lo[w][h] ? nlo : bi) * kBlockVoxels + row;
            const auto phi = vox + (size_t)(ok_hi[w][h] ? nhi : bi) * kBlockVoxels + row;
            if (h == 0 || r < kRows) {
#pragma unroll
                for (int x = 0; x < 8; ++x) q[w][h][x] = plo[x];
                q[w][h][8] = phi[0];
            }
        }
    }
    [[maybe_unused]] const uint32_t wfloor = f16_weight_floor(min_w);
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        if (w >= nw) break;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = tid + h * kWaveThreads;
            if (h == 1 && r >= kRows) break;
#pragma unroll
            for (int x = 0; x < 9; ++x) {
                const bool present = x < 8 ? ok_lo[w][h] : ok_hi[w][h];
                if constexpr (STORE == kStoreF16) {
                    c[w].val[r * 9 + x] = present && (q[w][h][x] >> 16) >= wfloor
                                              ? (unsigned short)(q[w][h][x] & 0xffffu)
                                              : (unsigned short)0x7e00u;
                } else {
                    const float wt = (float)q[w][h][x].y;
                    c[w].val[r * 9 + x] = present && !(wt < min_w || wt <= 0.0f)
                                              ? (float)q[w][h][x].x
                                              : __builtin_nanf("");
                }
            }
        }
    }
    __syncthreads();
}

// The 6 tetrahedra of a cube, corner ids as the reference lists them
// (cpp/src/marching_tets.cpp:37-41); every one holds the 0-7 diagonal.
constexpr int kTetList[6][4] = {{0, 5, 1, 7}, {0, 1, 3, 7}, {0, 3, 2, 7},
                                {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 2, 6, 7}};

// The same tetrahedra as corner bit sets, one byte each, lowest first.
constexpr unsigned long long make_tet_bits() {
    unsigned long long b = 0;
    for (int t = 0; t < 6; ++t)
        for (int j = 0; j < 4; ++j) b |= 1ull << (8 * t + kTetList[t][j]);
    return b;
}
constexpr unsigned long long kTetBits = make_tet_bits();
static_assert(kTetBits == 0xC5B1D18D8BA3ull, "tetrahedra bit sets");

// Classification of a cube from its corner mask (bit k set when corner k is
// below iso): per tetrahedron t its triangles in bits 2t..2t+1 (0 none, 1 one,
// 2 a quad's two) and the cube's triangles (vertices / 3) in bits 12-15. A
// tetrahedron with 1 or 3 corners below iso cuts one triangle, with 2 a quad.
constexpr uint16_t classify_mask(uint32_t mask) {
    uint32_t e = 0, n3 = 0;
    for (int t = 0; t < 6; ++t) {
        uint32_t m = mask & (uint32_t)((kTetBits >> (8 * t)) & 0xffu), pc = 0;
        for (; m; m &= m - 1u) ++pc;
        const uint32_t f = pc == 2u ? 2u : (pc & 1u);
        e |= f << (2 * t);
        n3 += f;
    }
    return (uint16_t)(e | (n3 << 12));
}

struct MaskTable {
    uint16_t e[256];
};
constexpr MaskTable make_mask_table() {
    MaskTable t{};
    for (uint32_t m = 0; m < 256; ++m) t.e[m] = classify_mask(m);
    return t;
}
// Built by the compiler; every workgroup copies it into LDS (8 bytes per
// lane) and reads it there for every block it walks.
__constant__ MaskTable kMaskTable = make_mask_table();

__device__ inline void build_table(uint16_t* table) {
    static_assert(sizeof(MaskTable) == kWaveThreads * 8, "one 8-byte piece per lane");
    reinterpret_cast<uint2*>(table)[threadIdx.x] =
        reinterpret_cast<const uint2*>(kMaskTable.e)[threadIdx.x];
    __syncthreads();
}

// The row of cubes (0..7, ly, lz) of the staged block: per cube its
// classification in e[x] (0 when a corner is not a valid sample). Corner
// columns x and x + 1 are shared by neighbouring cubes, so the row reads 36
// corners, not 64. Returns the row's triangles (vertices / 3).
template <int STORE>
__device__ inline uint32_t classify_row(const Cube<STORE>& c, const uint16_t* table, int ly,
                                        int lz, float iso, uint32_t (&e)[8]) {
    // FP16 corners compare as halves: a half x is below iso exactly when it
    // is below iso rounded up to a half (no half lies in between)
    [[maybe_unused]] const __half iso_h = __float2half_ru(iso);
    auto below = [&](auto x) -> bool {
        if constexpr (STORE == kStoreF16) return __hlt(__ushort_as_half(x), iso_h);
        else return x < iso;
    };
    auto invalid = [&](auto x) -> bool {
        if constexpr (STORE == kStoreF16) return __hisnan(__ushort_as_half(x));
        else return __builtin_isnan(x);
    };
    uint32_t prev_bits = 0, sum = 0;
    bool prev_ok = false;
#pragma unroll
    for (int x = 0; x < 9; ++x) {
        // corners (x, ly + dy, lz + dz) at mask bit 2 dy + 4 dz: the cube's
        // corner ids with dx = 0
        const auto p = c.val[cube_index(x, ly, lz)];
        const auto q = c.val[cube_index(x, ly + 1, lz)];
        const auto r = c.val[cube_index(x, ly, lz + 1)];
        const auto s = c.val[cube_index(x, ly + 1, lz + 1)];
        const bool ok = !(invalid(p) || invalid(q) || invalid(r) || invalid(s));
        const uint32_t bits = (below(p) ? 1u : 0u) | (below(q) ? 4u : 0u) | (below(r) ? 16u : 0u) |
                              (below(s) ? 64u : 0u);
        if (x > 0) {
            const uint32_t ent = (prev_ok && ok) ? table[prev_bits | (bits << 1)] : 0u;
            e[x - 1] = ent;
            sum += ent >> 12;
        }
        prev_bits = bits;
        prev_ok = ok;
    }
    return sum;
}

__device__ inline uint32_t wave_inclusive_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

struct P3 { float x, y, z; };
struct V6 { float p[3]; float c[3]; };

// A group's vertices as assembled in LDS: positions per vertex, the face
// normal once per triangle (vertex offsets are multiples of 3).
struct Stage {
    float pos[kStageVerts * 3];
    float nor[kStageVerts];
};

// Triangle (a, b, c) with its face normal into the staging area from vertex
// index at, colours to rgb from byte index at_rgb * 3 when rgb is set.
__device__ inline void write_tri(const V6& a, const V6& b, const V6& c, Stage& st, int at,
                                 uint8_t* rgb, int64_t at_rgb) {
    const float e1[3] = {b.p[0] - a.p[0], b.p[1] - a.p[1], b.p[2] - a.p[2]};
    const float e2[3] = {c.p[0] - a.p[0], c.p[1] - a.p[1], c.p[2] - a.p[2]};
    float nx = e1[1] * e2[2] - e1[2] * e2[1];
    float ny = e1[2] * e2[0] - e1[0] * e2[2];
    float nz = e1[0] * e2[1] - e1[1] * e2[0];
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    if (len > 1e-20f) { nx /= len; ny /= len; nz /= len; }
    st.nor[at] = nx; st.nor[at + 1] = ny; st.nor[at + 2] = nz;
    const V6* vs[3] = {&a, &b, &c};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float* d = st.pos + (at + k) * 3;
        d[0] = vs[k]->p[0]; d[1] = vs[k]->p[1]; d[2] = vs[k]->p[2];
        if (rgb) {
            const auto q = global(rgb) + (size_t)(at_rgb + k) * 3;
            q[0] = (uint8_t)fminf(fmaxf(vs[k]->c[0], 0.0f), 255.0f);
            q[1] = (uint8_t)fminf(fmaxf(vs[k]->c[1], 0.0f), 255.0f);
            q[2] = (uint8_t)fminf(fmaxf(vs[k]->c[2], 0.0f), 255.0f);
        }
    }
}

// Output float c (0-5: position, normal) of staged vertex i.
__device__ inline float staged(const Stage& st, int i, int c) {
    return c < 3 ? st.pos[i * 3 + c] : st.nor[i - i % 3 + (c - 3)];
}

// Every tetrahedron holds the cube's main diagonal, and its corners ordered by
// position (x, then y, then z: the order in which the reference takes an
// edge's endpoints, cpp/src/marching_tets.cpp:59-80) form a monotone path
// 0 -> e_a1 -> e_a1 + e_a2 -> 7 that sets one axis per step. So an edge's
// endpoints in position order are its corners in path order (rank 0-3), and
// along the edge an axis either stays or goes from the lower to the upper
// coordinate. The tables below are built by the compiler from the tetrahedra
// and the reference's case switch (cpp/src/marching_tets.cpp:157-183).
constexpr int corner_lds_offset(int k) { return (k & 1) + 9 * ((k >> 1) & 1) + 81 * (k >> 2); }
constexpr int lex_rank_key(int k) { return ((k & 1) << 2) | (k & 2) | ((k >> 2) & 1); }

struct EmitTables {
    // (flip, code, half) -> the triangle's three edges (vertex order), each
    // as its endpoint ranks r0 < r1 in bits 4j..4j+3 (r0 low); bit 15 set
    // for a triangle that exists. code bit r: the rank-r corner is below iso;
    // flip: the tetrahedron lists its rank-2 corner second.
    uint16_t tri[64];
    // per tetrahedron: cube-array offsets of its rank-1 and rank-2 corners
    // (bits 0-6, 7-13), flip (bit 14), the rank at which each axis becomes
    // the upper coordinate (x: bits 16-17, y: 18-19, z: 20-21)
    uint32_t tet[8];
};

constexpr EmitTables make_emit_tables() {
    EmitTables T{};
    for (int flip = 0; flip < 2; ++flip) {
        // rank of the tetrahedron's listed corner j
        const int rank[4] = {0, flip ? 2 : 1, flip ? 1 : 2, 3};
        for (int code = 0; code < 16; ++code) {
            int listed = 0;  // the crossing code in listed-corner order
            for (int j = 0; j < 4; ++j) listed |= ((code >> rank[j]) & 1) << j;
            // roles (a, b, c, d) as listed corners; single: triangle (ab, ac,
            // ad); quad: (ac, ad, bd) then (ac, bd, bc)
            int a = -1, b = 0, c = 0, d = 0;
            bool quad = false;
            switch (listed) {
                case 0x01: case 0x0E: a = 0; b = 1; c = 2; d = 3; break;
                case 0x02: case 0x0D: a = 1; b = 0; c = 2; d = 3; break;
                case 0x04: case 0x0B: a = 2; b = 0; c = 1; d = 3; break;
                case 0x08: case 0x07: a = 3; b = 0; c = 1; d = 2; break;
                case 0x03: a = 0; b = 1; c = 2; d = 3; quad = true; break;
                case 0x0C: a = 2; b = 3; c = 0; d = 1; quad = true; break;
                case 0x05: a = 0; b = 2; c = 1; d = 3; quad = true; break;
                case 0x0A: a = 1; b = 3; c = 0; d = 2; quad = true; break;
                case 0x06: a = 1; b = 2; c = 0; d = 3; quad = true; break;
                case 0x09: a = 0; b = 3; c = 1; d = 2; quad = true; break;
                default: break;
            }
            for (int half = 0; half < 2; ++half) {
                uint16_t e = 0;
                if (a >= 0 && (quad || half == 0)) {
                    int ed[3][2] = {{a, b}, {a, c}, {a, d}};
                    if (quad && half == 0) {
                        ed[0][0] = a; ed[0][1] = c; ed[1][0] = a; ed[1][1] = d;
                        ed[2][0] = b; ed[2][1] = d;
                    } else if (quad) {
                        ed[0][0] = a; ed[0][1] = c; ed[1][0] = b; ed[1][1] = d;
                        ed[2][0] = b; ed[2][1] = c;
                    }
                    e = 0x8000u;
                    for (int j = 0; j < 3; ++j) {
                        int r0 = rank[ed[j][0]], r1 = rank[ed[j][1]];
                        if (r1 < r0) { const int s = r0; r0 = r1; r1 = s; }
                        e = (uint16_t)(e | ((r0 | (r1 << 2)) << (4 * j)));
                    }
                }
                T.tri[(flip * 16 + code) * 2 + half] = e;
            }
        }
    }
    for (int t = 0; t < 6; ++t) {
        const int A = kTetList[t][1], B = kTetList[t][2];
        const bool flip = lex_rank_key(B) < lex_rank_key(A);
        const int L1 = flip ? B : A, L2 = flip ? A : B;
        uint32_t w = (uint32_t)corner_lds_offset(L1) | ((uint32_t)corner_lds_offset(L2) << 7) |
                     ((flip ? 1u : 0u) << 14);
        for (int ax = 0; ax < 3; ++ax) {
            const int s = ((L1 >> ax) & 1) ? 1 : (((L2 >> ax) & 1) ? 2 : 3);
            w |= (uint32_t)s << (16 + 2 * ax);
        }
        T.tet[t] = w;
    }
    return T;
}
static_assert(sizeof(EmitTables) % 4 == 0 && sizeof(EmitTables) / 4 <= 64, "one dword per lane");
__constant__ EmitTables kEmitTables = make_emit_tables();

__device__ inline void load_emit_tables(EmitTables& t) {
    if (threadIdx.x < sizeof(EmitTables) / 4)
        reinterpret_cast<uint32_t*>(&t)[threadIdx.x] =
            reinterpret_cast<const uint32_t*>(&kEmitTables)[threadIdx.x];
}

// Triangle `half` (0, or 1 for a quad's second) of tetrahedron t of voxel
// (lx, ly, lz)'s cube, staged from vertex index at, colours written from
// at_rgb. Every lane takes the same path: its tetrahedron's four corner
// values read once in rank order, one table entry for the case, three edge
// interpolations and one triangle; values, positions and colours are picked
// by rank with selects (no array indexed by a lane-varying value).
// Selects by a lane-varying rank, taking values (a conditional over lvalues
// would select their addresses, and the compiler keeps addressed locals in
// scratch memory).
__device__ inline float pick_lo(int r, float a, float b, float c) {  // r in 0-2
    return r == 0 ? a : (r == 1 ? b : c);
}
__device__ inline float pick_hi(int r, float b, float c, float d) {  // r in 1-3
    return r == 3 ? d : (r == 2 ? c : b);
}
__device__ inline int pick_lo_i(int r, int a, int b, int c) { return r == 0 ? a : (r == 1 ? b : c); }
__device__ inline int pick_hi_i(int r, int b, int c, int d) { return r == 3 ? d : (r == 2 ? c : b); }
__device__ inline float axis_point(int r0, int r1, int s, float a0, float a1, float d, float w) {
    const float p0 = r0 >= s ? a1 : a0;
    return p0 + w * ((r0 < s && r1 >= s) ? d : 0.0f);
}

template <int STORE>
__device__ void emit_slot(const Cube<STORE>& c, const EmitTables& et, const VolDev& v,
                          int32_t bx, int32_t by, int32_t bz, int lx, int ly, int lz, float iso,
                          int t, bool half, Stage& st, int at, uint8_t* rgb, int64_t at_rgb) {
    constexpr bool kCol = STORE == kStoreSums;
    const uint32_t tw = et.tet[t];
    const int o1 = (int)(tw & 127u), o2 = (int)((tw >> 7) & 127u);
    const int cb = cube_index(lx, ly, lz);
    const float V0 = corner(c.val[cb]), V1 = corner(c.val[cb + o1]);
    const float V2 = corner(c.val[cb + o2]), V3 = corner(c.val[cb + 91]);
    const uint32_t code = (V0 < iso ? 1u : 0u) | (V1 < iso ? 2u : 0u) | (V2 < iso ? 4u : 0u) |
                          (V3 < iso ? 8u : 0u);
    const uint32_t ent = et.tri[(((tw >> 10) & 16u) | code) * 2u + (half ? 1u : 0u)];
    if (!(ent & 0x8000u)) return;
    const int32_t vx = bx * 8 + lx, vy = by * 8 + ly, vz = bz * 8 + lz;
    // each axis' two corner coordinates, as the reference forms them
    const float x0 = ((float)vx + 0.5f) * v.voxel, x1 = ((float)(vx + 1) + 0.5f) * v.voxel;
    const float y0 = ((float)vy + 0.5f) * v.voxel, y1 = ((float)(vy + 1) + 0.5f) * v.voxel;
    const float z0 = ((float)vz + 0.5f) * v.voxel, z1 = ((float)(vz + 1) + 0.5f) * v.voxel;
    // an edge that moves along an axis goes from its lower to its upper
    // coordinate: p1 - p0 is then this difference, otherwise 0
    const float dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    const int sx = (int)((tw >> 16) & 3u), sy = (int)((tw >> 18) & 3u), sz = (int)((tw >> 20) & 3u);
    auto interp = [&](int j) {
        const int r0 = (int)((ent >> (4 * j)) & 3u), r1 = (int)((ent >> (4 * j + 2)) & 3u);
        // r0 < r1: r0 is 0-2, r1 is 1-3
        const float v0 = pick_lo(r0, V0, V1, V2);
        const float v1 = pick_hi(r1, V1, V2, V3);
        const float den = v1 - v0;
        float w = fabsf(den) < 1e-12f ? 0.5f : (iso - v0) / den;
        w = fminf(fmaxf(w, 0.0f), 1.0f);
        // p0 + w * (p1 - p0) per axis; w * 0 adds a zero, as in the reference
        V6 o;
        o.p[0] = axis_point(r0, r1, sx, x0, x1, dx, w);
        o.p[1] = axis_point(r0, r1, sy, y0, y1, dy, w);
        o.p[2] = axis_point(r0, r1, sz, z0, z1, dz, w);
        if constexpr (kCol) {
            const int i0 = cb + pick_lo_i(r0, 0, o1, o2);
            const int i1 = cb + pick_hi_i(r1, o1, o2, 91);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float c0 = c.col[q][i0], c1 = c.col[q][i1];
                o.c[q] = c0 + w * (c1 - c0);
            }
        } else {
            o.c[0] = o.c[1] = o.c[2] = 128.0f;
        }
        return o;
    };
    const V6 e0 = interp(0), e1 = interp(1), e2 = interp(2);
    write_tri(e0, e1, e2, st, at, rgb, at_rgb);
}

__device__ inline int32_t n_blocks_of(const VolDev& v, const ExtractArgs& a) {
    return a.n_blocks >= 0 ? a.n_blocks : global(v.ctr)->block_count;
}

// Blocks [lo, hi) of workgroup w of G (G >= 8, a multiple of 8 as the
// launchers round it) and its first index and stride in steps of `per`
// blocks. Workgroups are dealt to the 8 XCDs round-robin by id, so XCD
// w % 8 takes the (w % 8)-th eighth of the pool, in pool order: a block's +z
// and +y neighbours (usually a few indices on in key order) and its +x
// neighbour (a few hundred on) are read by workgroups of the same XCD close
// in time, from its L2, instead of from HBM by another XCD.
struct XcdSpan {
    int32_t lo, hi, first, stride;
};
__device__ inline XcdSpan xcd_span(int32_t w, int32_t G, int32_t n, int32_t per) {
    const int32_t xcd = w % kXcds, mine = (G - xcd + kXcds - 1) / kXcds;
    XcdSpan sp;
    sp.lo = (int32_t)((int64_t)n * xcd / kXcds);
    sp.hi = (int32_t)((int64_t)n * (xcd + 1) / kXcds);
    sp.first = sp.lo + per * (w / kXcds);
    sp.stride = per * mine;
    return sp;
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

// Count: one wave per block, lane = row (ly, lz) of 8 cubes; workgroup w of
// G walks its XCD's span of the pool (two consecutive blocks at a time for
// the packed storages).
template <int STORE>
__device__ void count_blocks(const VolDev& v, const ExtractArgs& a, int32_t w, int32_t G) {
    [[maybe_unused]] __shared__ Cube<STORE> c;
    __shared__ __attribute__((aligned(16))) uint16_t s_table[256];
    build_table(s_table);
    const int lane = threadIdx.x;
    const int32_t n = n_blocks_of(v, a);
    const int32_t cached = a.nb_cache ? global(v.ctr)->nb_blocks : 0;
    if (a.clear_flag && w == 0 && lane == 0) *global(a.scan_flag) = 0;
    if constexpr (STORE != kStoreSums) {
        // blocks b and b + 1 together
        __shared__ Cube<STORE> c2[2];
        const XcdSpan sp = xcd_span(w, G, n, 2);
        for (int32_t b = sp.first; b < sp.hi; b += sp.stride) {
            const int32_t b1 = b + 1 < sp.hi ? b + 1 : -1;
            load_cube_pair<STORE>(v, b, b1, a.min_weight, c2, a.block_nb, cached);
            uint32_t e[8];
            // two triangle counts (<= 6144 each) in one word, one reduction
            uint32_t n3 = classify_row(c2[0], s_table, lane & 7, lane >> 3, a.iso, e);
            if (b1 >= 0) n3 |= classify_row(c2[1], s_table, lane & 7, lane >> 3, a.iso, e) << 16;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) n3 += __shfl_xor(n3, o);
            if (lane == 0) {
                global(a.block_vcount)[b] = 3 * (int32_t)(n3 & 0xffffu);
                if (b1 >= 0) global(a.block_vcount)[b1] = 3 * (int32_t)(n3 >> 16);
            }
            __syncthreads();  // the cubes are read before the next pair replaces them
        }
    } else {
        const XcdSpan sp = xcd_span(w, G, n, 1);
        for (int32_t b = sp.first; b < sp.hi; b += sp.stride) {
            load_cube<STORE, true>(v, b, a.min_weight, c, a.block_nb, cached);
            uint32_t e[8];
            uint32_t n3 = classify_row(c, s_table, lane & 7, lane >> 3, a.iso, e);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) n3 += __shfl_xor(n3, o);
            if (lane == 0) global(a.block_vcount)[b] = 3 * (int32_t)n3;
            __syncthreads();  // the cube is read before the next block replaces it
        }
    }
}

template <int STORE>
__global__ __launch_bounds__(kWaveThreads) void k_extract_count(VolDev v, ExtractArgs a) {
    count_blocks<STORE>(v, a, blockIdx.x, gridDim.x);
}

// A batch of volumes: blockIdx.y is the volume.
template <int STORE>
__global__ __launch_bounds__(kWaveThreads) void k_extract_count_batch(
    const VolDev* __restrict__ vols, const ExtractArgs* __restrict__ args) {
    count_blocks<STORE>(vols[blockIdx.y], args[blockIdx.y], blockIdx.x, gridDim.x);
}

// Exclusive scan of the per-block counts; publishes the vertex count (or -1
// when it exceeds the capacity). Up to kSerialChunks chunks of 1024 blocks
// one workgroup scans them all. Above that, workgroups scan 1024-block chunks
// (grid-stride); whoever finishes the last chunk scans the chunk totals and
// resets the arrival counter. Workgroups never wait on each other.
__device__ void scan_chunks(const VolDev& v, const ExtractArgs& a) {
    __shared__ int s_wave[kScanChunk / 64];
    __shared__ int s_last;
    const int32_t n = n_blocks_of(v, a);
    const int n_chunks = (n + kScanChunk - 1) / kScanChunk;
    // the count pass that just ran holds every block's neighbour indices
    if (a.nb_cache && blockIdx.x == 0 && threadIdx.x == 0) global(v.ctr)->nb_blocks = n;
    if (n == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *a.count_out = 0;
            v.ctr->mesh_vertices = 0;
        }
        return;
    }
    if (n_chunks <= kSerialChunks) {
        // Up to 64 chunks (65,536 blocks): workgroup 0 alone, 16 chunks at a
        // time. Thread t holds the counts of blocks t + 1024 j of the group's
        // chunks (one coalesced load per chunk, all in flight together); the
        // chunks are scanned across the threads at once, and a block's offset
        // is the earlier chunks' totals plus the scan within its chunk.
        if (blockIdx.x != 0) return;
        __shared__ int s_tot[kStripeChunks * (kScanChunk / 64)];
        const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
        long long carry = 0;  // vertices of the chunks before the current one
        for (int g0 = 0; g0 < n_chunks; g0 += kStripeChunks) {
            int c[kStripeChunks], incl[kStripeChunks];
#pragma unroll
            for (int j = 0; j < kStripeChunks; ++j) {
                const int b = (g0 + j) * kScanChunk + t;
                c[j] = (g0 + j < n_chunks && b < n) ? global(a.block_vcount)[b] : 0;
            }
#pragma unroll
            for (int j = 0; j < kStripeChunks; ++j) {
                incl[j] = c[j];
                for (int o = 1; o < 64; o <<= 1) {
                    const int y = __shfl_up(incl[j], o);
                    if (lane >= o) incl[j] += y;
                }
                if (lane == 63) s_tot[j * (kScanChunk / 64) + wave] = incl[j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kStripeChunks; ++j) {
                int before = 0, tot = 0;
                for (int w = 0; w < kScanChunk / 64; ++w) {
                    const int x = s_tot[j * (kScanChunk / 64) + w];
                    if (w < wave) before += x;
                    tot += x;
                }
                const int b = (g0 + j) * kScanChunk + t;
                if (g0 + j < n_chunks && b < n)
                    global(a.block_voff)[b] = (int)(carry + before + incl[j] - c[j]);
                carry += tot;
            }
            __syncthreads();  // s_tot is rewritten by the next group
        }
        if (t < n_chunks) a.chunk_sums[t] = 0;
        if (t == 0) {
            const int32_t out = carry > (long long)a.cap ? -1 : (int32_t)carry;
            *a.count_out = out;
            v.ctr->mesh_vertices = out;
        }
        return;
    }
    if (threadIdx.x == 0) s_last = 0;  // workgroups with no chunk
    __syncthreads();
    for (int chunk = blockIdx.x; chunk < n_chunks; chunk += gridDim.x) {
        const int b = chunk * kScanChunk + threadIdx.x;
        const int x = b < n ? a.block_vcount[b] : 0;
        int total;
        const int ex = wg_exclusive_scan(x, s_wave, total);
        if (b < n) a.block_voff[b] = ex;
        if (threadIdx.x == 0) a.chunk_sums[chunk] = total;
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) s_last = (atomicAdd(a.scan_flag, 1) == n_chunks - 1);
        __syncthreads();
        // the last arrival is this workgroup's last chunk: its later chunks
        // would not have arrived yet
        if (s_last) break;
    }
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

__global__ __launch_bounds__(kScanChunk) void k_extract_scan(VolDev v, ExtractArgs a) {
    scan_chunks(v, a);
}

__global__ __launch_bounds__(kScanChunk) void k_extract_scan_batch(
    const VolDev* __restrict__ vols, const ExtractArgs* __restrict__ args) {
    scan_chunks(vols[blockIdx.y], args[blockIdx.y]);
}

// Emit: one wave per block, blocks dealt to workgroups round-robin (the
// blocks' triangle counts vary along the pool, and an interleaved deal keeps
// the XCDs evenly loaded where contiguous spans did not), lane = row (ly, lz)
// of 8 cubes. The block's
// triangles are listed in LDS in output order (voxel, tetrahedron, half), so
// triangle j's vertices are 3j.. of the block's run, and emitted one per lane
// in groups of 64; each group's 192 vertices are staged and copied out as one
// contiguous run.
template <int STORE>
__device__ void emit_blocks(const VolDev& v, const ExtractArgs& a, int32_t w, int32_t G) {
    __shared__ Cube<STORE> c;
    __shared__ __attribute__((aligned(16))) uint16_t s_table[256];
    __shared__ uint16_t s_slots[kSlotCap + kWaveThreads];
    __shared__ Stage s_stage;
    __shared__ EmitTables s_et;
    if (*global(a.count_out) < 0) return;  // overflow: write nothing
    load_emit_tables(s_et);
    build_table(s_table);  // (its barrier also covers s_et)
    const int lane = threadIdx.x;
    const int32_t n = n_blocks_of(v, a);
    for (int32_t b = w; b < n; b += G) {
        if (global(a.block_vcount)[b] == 0) continue;
        load_cube<STORE, false>(v, b, a.min_weight, c, a.block_nb);
        const int32_t bx = global(v.block_coord)[b * 3 + 0];
        const int32_t by = global(v.block_coord)[b * 3 + 1];
        const int32_t bz = global(v.block_coord)[b * 3 + 2];
        const int64_t base = (int64_t)global(a.chunk_sums)[b / kScanChunk] + global(a.block_voff)[b];
        uint32_t e[8];
        const uint32_t row = classify_row(c, s_table, lane & 7, lane >> 3, a.iso, e);
        const uint32_t incl = wave_inclusive_scan(row);
        const int n_tri = __builtin_amdgcn_readlane((int)incl, 63);
        const int before = (int)(incl - row);
        // passes of up to kSlotCap listed triangles (one pass unless the
        // block's surface is very dense)
        for (int g0 = 0; g0 < n_tri; g0 += kSlotCap) {
            // entry: voxel (bits 0-8), tetrahedron (9-11), half (12)
            int k = before;
            if (n_tri <= kSlotCap) {
                // one pass: no range checks, no branches; an item with
                // nothing to list writes this lane's dummy entry instead
#pragma unroll
                for (int x = 0; x < 8; ++x) {
                    const uint32_t ent = e[x];
#pragma nounroll
                    for (int t = 0; t < 6; ++t) {
                        const uint32_t f = (ent >> (2 * t)) & 3u;
                        const uint16_t item = (uint16_t)((lane * 8 + x) | (t << 9));
                        s_slots[f ? k : kSlotCap + lane] = item;
                        s_slots[f == 2u ? k + 1 : kSlotCap + lane] = (uint16_t)(item | (1u << 12));
                        k += (int)f;
                    }
                }
            } else
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const uint32_t ent = e[x];
                if (ent & 0xfffu) {
#pragma nounroll
                    for (int t = 0; t < 6; ++t) {
                        const uint32_t f = (ent >> (2 * t)) & 3u;
                        const uint16_t item = (uint16_t)((lane * 8 + x) | (t << 9));
                        if (f && (unsigned)(k - g0) < (unsigned)kSlotCap) s_slots[k - g0] = item;
                        if (f == 2u && (unsigned)(k + 1 - g0) < (unsigned)kSlotCap)
                            s_slots[k + 1 - g0] = (uint16_t)(item | (1u << 12));
                        k += (int)f;
                    }
                }
            }
            __syncthreads();
            const int n_pass = min(kSlotCap, n_tri - g0);
            for (int g = 0; g < n_pass; g += kWaveThreads) {
                const int i = g + lane;
                const int cnt = min(kWaveThreads, n_pass - g);
                const int64_t v_lo = base + 3 * (int64_t)(g0 + g);  // the group's first vertex
                if (i < n_pass) {
                    const uint32_t it = s_slots[i];
                    const int vi = (int)(it & 511u), t = (int)((it >> 9) & 7u);
                    emit_slot(c, s_et, v, bx, by, bz, vi & 7, (vi >> 3) & 7, vi >> 6, a.iso, t,
                              ((it >> 12) & 1u) != 0u, s_stage, 3 * lane, a.rgb, v_lo + 3 * lane);
                }
                const int n_verts = 3 * cnt;
                __syncthreads();
                const auto out = global(a.posnor) + v_lo * 6;
                if ((reinterpret_cast<uintptr_t>(out) & 7u) == 0) {
                    // one vertex per lane: three float2 stores, so each
                    // wave-wide store covers every third 8 bytes of one run
                    const auto dst = (__attribute__((address_space(1))) f32x2*)out;
                    for (int j = lane; j < n_verts; j += kWaveThreads) {
                        const float* p = s_stage.pos + j * 3;
                        const float* q = s_stage.nor + (j - j % 3);
                        dst[j * 3 + 0] = f32x2{p[0], p[1]};
                        dst[j * 3 + 1] = f32x2{p[2], q[0]};
                        dst[j * 3 + 2] = f32x2{q[1], q[2]};
                    }
                } else {  // a caller's buffer need only be float-aligned
                    for (int j = lane; j < n_verts * 6; j += kWaveThreads)
                        out[j] = staged(s_stage, j / 6, j % 6);
                }
                __syncthreads();
            }
        }
    }
}

template <int STORE>
__global__ __launch_bounds__(kWaveThreads, emit_waves_per_simd<STORE>()) void k_extract_emit(VolDev v, ExtractArgs a) {
    emit_blocks<STORE>(v, a, blockIdx.x, gridDim.x);
}

template <int STORE>
__global__ __launch_bounds__(kWaveThreads, emit_waves_per_simd<STORE>()) void k_extract_emit_batch(
    const VolDev* __restrict__ vols, const ExtractArgs* __restrict__ args) {
    emit_blocks<STORE>(vols[blockIdx.y], args[blockIdx.y], blockIdx.x, gridDim.x);
}

}  // namespace

hipError_t launch_extract(const VolDev& v, const ExtractArgs& a, bool emit, hipStream_t s) {
    // one-wave workgroups: one per block when the count is known on the host,
    // otherwise a grid-stride over the pool, four times the chip's worth
    auto xcd_round = [](int32_t g) { return (g + kXcds - 1) / kXcds * kXcds; };
    const int32_t grid_blocks = xcd_round(
        a.n_blocks >= 0 ? (a.n_blocks > 0 ? a.n_blocks : 1)
                        : (v.pool_capacity < 32768 ? v.pool_capacity : 32768));
    // count: two blocks per workgroup for the packed storages
    const int32_t count_grid =
        v.storage == kStoreSums ? grid_blocks : xcd_round((grid_blocks + 1) / 2);
    const int32_t n_for_scan = a.n_blocks >= 0 ? a.n_blocks : v.pool_capacity;
    const int32_t n_chunks = (n_for_scan + kScanChunk - 1) / kScanChunk;
    const int32_t scan_grid = n_chunks < 1 ? 1 : (n_chunks < kScanGrid ? n_chunks : kScanGrid);
    if (!emit) {
        switch (v.storage) {
            case kStoreSums: k_extract_count<kStoreSums><<<grid_blocks, kWaveThreads, 0, s>>>(v, a); break;
            case kStoreF32: k_extract_count<kStoreF32><<<count_grid, kWaveThreads, 0, s>>>(v, a); break;
            case kStoreF16: k_extract_count<kStoreF16><<<count_grid, kWaveThreads, 0, s>>>(v, a); break;
            case kStoreF64: k_extract_count<kStoreF64><<<count_grid, kWaveThreads, 0, s>>>(v, a); break;
            default: return hipErrorInvalidValue;
        }
        k_extract_scan<<<scan_grid, kScanChunk, 0, s>>>(v, a);
    } else {
        switch (v.storage) {
            case kStoreSums: k_extract_emit<kStoreSums><<<grid_blocks, kWaveThreads, 0, s>>>(v, a); break;
            case kStoreF32: k_extract_emit<kStoreF32><<<grid_blocks, kWaveThreads, 0, s>>>(v, a); break;
            case kStoreF16: k_extract_emit<kStoreF16><<<grid_blocks, kWaveThreads, 0, s>>>(v, a); break;
            case kStoreF64: k_extract_emit<kStoreF64><<<grid_blocks, kWaveThreads, 0, s>>>(v, a); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_extract_batch(const VolDev* vols, const ExtractArgs* args, int32_t n_scenes,
                                int32_t storage, uint32_t max_pool_capacity,
                                int32_t grid_per_scene, bool emit, hipStream_t s) {
    if (n_scenes <= 0 || grid_per_scene <= 0) return hipErrorInvalidValue;
    const uint32_t cap = max_pool_capacity > 0 ? max_pool_capacity : 1;
    auto xcd_round = [](uint32_t g) { return (g + kXcds - 1) / kXcds * kXcds; };
    const dim3 grid(xcd_round(cap < (uint32_t)grid_per_scene ? cap : (uint32_t)grid_per_scene),
                    n_scenes);
    // count: two blocks per workgroup
    const dim3 count_grid(xcd_round((grid.x + 1) / 2), n_scenes);
    const uint32_t n_chunks = (cap + kScanChunk - 1) / kScanChunk;
    const dim3 scan_grid(n_chunks < kScanGridBatch ? n_chunks : kScanGridBatch, n_scenes);
    if (!emit) {
        switch (storage) {
            case kStoreF32: k_extract_count_batch<kStoreF32><<<count_grid, kWaveThreads, 0, s>>>(vols, args); break;
            case kStoreF16: k_extract_count_batch<kStoreF16><<<count_grid, kWaveThreads, 0, s>>>(vols, args); break;
            case kStoreF64: k_extract_count_batch<kStoreF64><<<count_grid, kWaveThreads, 0, s>>>(vols, args); break;
            default: return hipErrorInvalidValue;
        }
        k_extract_scan_batch<<<scan_grid, kScanChunk, 0, s>>>(vols, args);
    } else {
        switch (storage) {
            case kStoreF32: k_extract_emit_batch<kStoreF32><<<grid, kWaveThreads, 0, s>>>(vols, args); break;
            case kStoreF16: k_extract_emit_batch<kStoreF16><<<grid, kWaveThreads, 0, s>>>(vols, args); break;
            case kStoreF64: k_extract_emit_batch<kStoreF64><<<grid, kWaveThreads, 0, s>>>(vols, args); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace osn_hip
