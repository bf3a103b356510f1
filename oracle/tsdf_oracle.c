/*
 * tsdf_oracle.c -- CPU restatement of the hash-blocked TSDF path, used ONLY as
 * the checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only to check or time against
 * it. Nothing in the product (the HIP library and its host wrapper) links,
 * loads or calls it; the product fails loudly when its HIP build is missing.
 *
 * What it restates (citations relative to the reference checkout):
 *   key codec, hash, floor-div ........ cpp/include/osn_tsdf/types.hpp:50-69,150-165
 *   table sizing / init ............... cpp/src/volume.cpp:14-19,41-61
 *   lookup / find-or-insert ........... cpp/src/volume.cpp:71-145
 *   allocate (band walk) .............. cpp/src/volume.cpp:147-198
 *   update (scatter sums) ............. cpp/src/volume.cpp:200-299
 *   sample (mean on read) ............. cpp/src/volume.cpp:306-329
 *   marching tetrahedra ............... cpp/src/marching_tets.cpp:37-243
 *   analytic point scenes ............. cpp/tests/analytic_scenes.hpp:19-46
 *   back-projection convention ........ crates/tartanair/src/lib.rs:307-343
 *                                       (pixel index u, not u+0.5)
 *
 * It is single-threaded and sequential, so its results are deterministic: the
 * reference arm's atomics make the last ulp of a tsdf sum and the weight-cap
 * cut order-dependent, this restatement does not. Integer work (keys, block
 * counts, weights with unit inputs) is exact and is compared bit-for-bit.
 *
 * Parts with no reference counterpart (the projective update, FP64 and
 * FP16-storage variants, the FP32 back-projection used by the fused device
 * kernel) are this build's own definitions; their pins are listed in the test
 * files that exercise them.
 *
 * Build: -O2 -ffp-contract=off, no -ffast-math, no -march=native, so that
 * float expressions round exactly as the reference's g++ x86-64 build does
 * (SSE, no FMA).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BLOCK_DIM 8
#define ORC_BLOCK_VOXELS 512
#define ORC_EMPTY_KEY ((int64_t)-1)
#define ORC_COORD_BIAS ((int64_t)1 << 20)
#define ORC_COORD_BITS 21

/* ORC_PLANT selects a deliberately wrong running average; only the tests that
 * show the average's closed-form pin catches such mistakes build it (into a
 * separate library). 0 is the oracle. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

/* ---- a1/a2: key codec, hash, floor division ------------------------------ */

/* types.hpp:64-69 */
int64_t orc_pack_coord(int32_t x, int32_t y, int32_t z) {
    const int64_t bx = (int64_t)x + ORC_COORD_BIAS;
    const int64_t by = (int64_t)y + ORC_COORD_BIAS;
    const int64_t bz = (int64_t)z + ORC_COORD_BIAS;
    return (bx << (2 * ORC_COORD_BITS)) | (by << ORC_COORD_BITS) | bz;
}

void orc_unpack_key(int64_t key, int32_t out[3]) {
    const int64_t m = ((int64_t)1 << ORC_COORD_BITS) - 1;
    out[0] = (int32_t)(((key >> (2 * ORC_COORD_BITS)) & m) - ORC_COORD_BIAS);
    out[1] = (int32_t)(((key >> ORC_COORD_BITS) & m) - ORC_COORD_BIAS);
    out[2] = (int32_t)((key & m) - ORC_COORD_BIAS);
}

/* types.hpp:150-155 (Teschner multiply-xor, masked to a power of two) */
uint32_t orc_hash_block(int32_t x, int32_t y, int32_t z, uint32_t mask) {
    const uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^
                       (uint32_t)z * 83492791u;
    return h & mask;
}

/* types.hpp:162-165 */
int32_t orc_floor_div(int32_t a, int32_t b) {
    const int32_t q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

/* volume.cpp:14-19 */
static uint32_t next_pow2(uint32_t v) {
    if (v < 2) return 2;
    --v;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}

/* ---- a3: volume storage ---------------------------------------------------- */

typedef struct {
    int64_t key;
    int32_t block_idx;
    int32_t pad;
} OrcEntry;

typedef struct OrcVolume {
    float voxel;
    float trunc;
    float weight_cap;
    int32_t pool;
    uint32_t mask;
    OrcEntry* table;
    int32_t block_count;
    uint64_t drop_count;
    /* weighted sums, SoA, pool*512 each (volume.hpp:99-116) */
    float* tsdf;
    float* weight;
    float* r;
    float* g;
    float* b;
    int32_t* coord; /* pool*3 */
    /* 1 = tsdf/weight hold means (loaded from a projective device state), the
     * sample step then skips the divide. 0 = reference weighted sums. */
    int32_t mean_mode;
} OrcVolume;

void orc_destroy(OrcVolume* v) {
    if (!v) return;
    free(v->table); free(v->tsdf); free(v->weight);
    free(v->r); free(v->g); free(v->b); free(v->coord);
    free(v);
}

/* volume.cpp:41-61 and types.hpp:97-106 (trunc <= 0 selects 4 voxels). */
OrcVolume* orc_create(float voxel, int32_t pool, float trunc, float weight_cap) {
    if (!(voxel > 0.0f) || pool <= 0) return NULL;
    OrcVolume* v = (OrcVolume*)calloc(1, sizeof(OrcVolume));
    if (!v) return NULL;
    v->voxel = voxel;
    v->trunc = trunc > 0.0f ? trunc : 4.0f * voxel;
    v->weight_cap = weight_cap;
    v->pool = pool;
    const uint32_t size = next_pow2((uint32_t)pool * 2u);
    v->mask = size - 1;
    v->table = (OrcEntry*)malloc((size_t)size * sizeof(OrcEntry));
    const size_t n = (size_t)pool * ORC_BLOCK_VOXELS;
    v->tsdf = (float*)calloc(n, sizeof(float));
    v->weight = (float*)calloc(n, sizeof(float));
    v->r = (float*)calloc(n, sizeof(float));
    v->g = (float*)calloc(n, sizeof(float));
    v->b = (float*)calloc(n, sizeof(float));
    v->coord = (int32_t*)calloc((size_t)pool * 3, sizeof(int32_t));
    if (!v->table || !v->tsdf || !v->weight || !v->r || !v->g || !v->b || !v->coord) {
        orc_destroy(v);
        return NULL;
    }
    for (uint32_t i = 0; i < size; ++i) {
        v->table[i].key = ORC_EMPTY_KEY;
        v->table[i].block_idx = -1;
        v->table[i].pad = 0;
    }
    return v;
}

int32_t orc_block_count(const OrcVolume* v) { return v ? v->block_count : 0; }
uint64_t orc_drop_count(const OrcVolume* v) { return v ? v->drop_count : 0; }
uint32_t orc_table_size(const OrcVolume* v) { return v ? v->mask + 1 : 0; }
float orc_trunc(const OrcVolume* v) { return v ? v->trunc : 0.0f; }

/* ---- a6: lookup, a5: find-or-insert (volume.cpp:71-145) ------------------- */

static int32_t find_block(const OrcVolume* v, int32_t bx, int32_t by, int32_t bz) {
    const int64_t want = orc_pack_coord(bx, by, bz);
    const uint32_t slot = orc_hash_block(bx, by, bz, v->mask);
    const uint32_t size = v->mask + 1;
    for (uint32_t p = 0; p < size; ++p) {
        const OrcEntry* e = &v->table[(slot + p) & v->mask];
        if (e->key == ORC_EMPTY_KEY) return -1;
        if (e->key == want) return e->block_idx;
    }
    return -1;
}

static int32_t find_or_insert(OrcVolume* v, int32_t bx, int32_t by, int32_t bz) {
    const int64_t want = orc_pack_coord(bx, by, bz);
    const uint32_t slot = orc_hash_block(bx, by, bz, v->mask);
    const uint32_t size = v->mask + 1;
    for (uint32_t p = 0; p < size; ++p) {
        OrcEntry* e = &v->table[(slot + p) & v->mask];
        if (e->key == want) return e->block_idx;
        if (e->key == ORC_EMPTY_KEY) {
            const int32_t idx = v->block_count;
            if (idx >= v->pool) {
                /* pool exhausted: slot stays empty, the visit is counted */
                v->drop_count++;
                return -1;
            }
            e->key = want;
            e->block_idx = idx;
            v->block_count = idx + 1;
            v->coord[idx * 3 + 0] = bx;
            v->coord[idx * 3 + 1] = by;
            v->coord[idx * 3 + 2] = bz;
            return idx;
        }
    }
    v->drop_count++;
    return -1;
}

/* ---- a4: band walk ---------------------------------------------------------
 * volume.cpp:147-198 (allocate) and :200-299 (update) walk the same band with
 * the same float expressions. The callback receives the voxel and the signed
 * distance at its centre; occluded samples (sdf < -trunc) are not delivered.
 */
typedef void (*visit_fn)(void* ctx, int32_t i, int32_t vx, int32_t vy, int32_t vz, float sdf);

static void walk_points(const OrcVolume* v, const float* pos, int32_t n, const float cam[3],
                        float radius_m, visit_fn fn, void* ctx) {
    const float inv_voxel = 1.0f / v->voxel;
    const float trunc = v->trunc;
    const float radius_sq = radius_m > 0.0f ? radius_m * radius_m : 0.0f;
    const int32_t csteps = (int32_t)ceilf(trunc * inv_voxel);
    const int32_t steps = csteps > 1 ? csteps : 1;
    for (int32_t i = 0; i < n; ++i) {
        const float px = pos[i * 3 + 0], py = pos[i * 3 + 1], pz = pos[i * 3 + 2];
        const float dx = px - cam[0], dy = py - cam[1], dz = pz - cam[2];
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (radius_sq > 0.0f && d2 > radius_sq) continue;
        const float dist = sqrtf(d2);
        if (!(dist > 1e-6f)) continue;
        const float ux = dx / dist, uy = dy / dist, uz = dz / dist;
        for (int32_t s = -steps; s <= steps; ++s) {
            const float t = (float)s * v->voxel;
            const int32_t vx = (int32_t)floorf((px + ux * t) * inv_voxel);
            const int32_t vy = (int32_t)floorf((py + uy * t) * inv_voxel);
            const int32_t vz = (int32_t)floorf((pz + uz * t) * inv_voxel);
            const float cx = ((float)vx + 0.5f) * v->voxel;
            const float cy = ((float)vy + 0.5f) * v->voxel;
            const float cz = ((float)vz + 0.5f) * v->voxel;
            const float ex = cx - cam[0], ey = cy - cam[1], ez = cz - cam[2];
            const float sdf = dist - sqrtf(ex * ex + ey * ey + ez * ez);
            if (sdf < -trunc) continue;
            fn(ctx, i, vx, vy, vz, sdf);
        }
    }
}

static void alloc_visit(void* ctx, int32_t i, int32_t vx, int32_t vy, int32_t vz, float sdf) {
    (void)i; (void)sdf;
    OrcVolume* v = (OrcVolume*)ctx;
    find_or_insert(v, orc_floor_div(vx, ORC_BLOCK_DIM), orc_floor_div(vy, ORC_BLOCK_DIM),
                   orc_floor_div(vz, ORC_BLOCK_DIM));
}

/* Returns 0 ok, 1 bad arguments (c_api.cpp:310-320). */
int32_t orc_allocate_blocks(OrcVolume* v, const float* pos, int32_t n, float cx, float cy,
                            float cz, float radius_m) {
    if (!v || !pos || n <= 0) return 1;
    const float cam[3] = {cx, cy, cz};
    walk_points(v, pos, n, cam, radius_m, alloc_visit, v);
    return 0;
}

typedef struct {
    OrcVolume* v;
    const uint8_t* colors;
    const float* weights;
    float inv_trunc;
} UpdateCtx;

/* volume.cpp:240-296. Sequential, so the weight cap cut is exact. */
static void update_visit(void* ctxp, int32_t i, int32_t vx, int32_t vy, int32_t vz, float sdf) {
    UpdateCtx* c = (UpdateCtx*)ctxp;
    OrcVolume* v = c->v;
    const int32_t bx = orc_floor_div(vx, ORC_BLOCK_DIM);
    const int32_t by = orc_floor_div(vy, ORC_BLOCK_DIM);
    const int32_t bz = orc_floor_div(vz, ORC_BLOCK_DIM);
    const int32_t bi = find_block(v, bx, by, bz);
    if (bi < 0) return;
    float sdf_n = sdf * c->inv_trunc;
    sdf_n = sdf_n < -1.0f ? -1.0f : (sdf_n > 1.0f ? 1.0f : sdf_n);
    const int32_t lx = vx - bx * ORC_BLOCK_DIM;
    const int32_t ly = vy - by * ORC_BLOCK_DIM;
    const int32_t lz = vz - bz * ORC_BLOCK_DIM;
    const size_t idx = (size_t)bi * ORC_BLOCK_VOXELS + (size_t)((lz * 8 + ly) * 8 + lx);
    if (v->weight_cap > 0.0f && v->weight[idx] >= v->weight_cap) return;
    const float w_in = c->weights ? c->weights[i] : 1.0f;
    float cr = 128.0f, cg = 128.0f, cb = 128.0f;
    if (c->colors) {
        cr = (float)c->colors[i * 3 + 0];
        cg = (float)c->colors[i * 3 + 1];
        cb = (float)c->colors[i * 3 + 2];
    }
    v->weight[idx] += w_in;
    v->tsdf[idx] += sdf_n * w_in;
    v->r[idx] += cr * w_in;
    v->g[idx] += cg * w_in;
    v->b[idx] += cb * w_in;
}

int32_t orc_update_voxels(OrcVolume* v, const float* pos, const uint8_t* colors,
                          const float* weights, int32_t n, float cx, float cy, float cz,
                          float radius_m) {
    if (!v || !pos || n <= 0) return 1;
    const float cam[3] = {cx, cy, cz};
    UpdateCtx c = {v, colors, weights, 1.0f / v->trunc};
    walk_points(v, pos, n, cam, radius_m, update_visit, &c);
    return 0;
}

int32_t orc_integrate(OrcVolume* v, const float* pos, const uint8_t* colors,
                      const float* weights, int32_t n, float cx, float cy, float cz,
                      float radius_m) {
    const int32_t rc = orc_allocate_blocks(v, pos, n, cx, cy, cz, radius_m);
    if (rc) return rc;
    return orc_update_voxels(v, pos, colors, weights, n, cx, cy, cz, radius_m);
}

/* Count of visits (samples that pass the cull) a batch generates. */
typedef struct { int64_t n; } CountCtx;
static void count_visit(void* ctx, int32_t i, int32_t vx, int32_t vy, int32_t vz, float sdf) {
    (void)i; (void)vx; (void)vy; (void)vz; (void)sdf;
    ((CountCtx*)ctx)->n++;
}
int64_t orc_count_visits(const OrcVolume* v, const float* pos, int32_t n, float cx, float cy,
                         float cz, float radius_m) {
    if (!v || !pos || n <= 0) return 0;
    const float cam[3] = {cx, cy, cz};
    CountCtx c = {0};
    walk_points(v, pos, n, cam, radius_m, count_visit, &c);
    return c.n;
}

/* Distinct block keys the batch's band walk visits, in first-visit order,
 * restricted to blocks present in the volume (dropped blocks are absent).
 * This is the "touched" set a frame's projective update covers. Returns the
 * count, or -1 when `cap` is too small. */
typedef struct {
    const OrcVolume* v;
    int64_t* out;
    int32_t cap, n;
    int64_t* set;
    uint32_t mask;
    int overflow;
} TouchCtx;

static void touch_visit(void* ctxp, int32_t i, int32_t vx, int32_t vy, int32_t vz, float sdf) {
    (void)i; (void)sdf;
    TouchCtx* c = (TouchCtx*)ctxp;
    const int32_t bx = orc_floor_div(vx, 8), by = orc_floor_div(vy, 8), bz = orc_floor_div(vz, 8);
    if (find_block(c->v, bx, by, bz) < 0) return;
    const int64_t key = orc_pack_coord(bx, by, bz);
    uint32_t s = (uint32_t)(((uint64_t)key * 0x9E3779B97F4A7C15ull) >> 32) & c->mask;
    for (;;) {
        if (c->set[s] == key) return;
        if (c->set[s] == ORC_EMPTY_KEY) break;
        s = (s + 1) & c->mask;
    }
    c->set[s] = key;
    if (c->n >= c->cap) { c->overflow = 1; return; }
    c->out[c->n++] = key;
}

int32_t orc_touched_keys(const OrcVolume* v, const float* pos, int32_t n, float cx, float cy,
                         float cz, float radius_m, int64_t* out, int32_t cap) {
    if (!v || !pos || !out || n <= 0 || cap <= 0) return 0;
    const uint32_t size = next_pow2((uint32_t)cap * 2u);
    TouchCtx c = {v, out, cap, 0, (int64_t*)malloc((size_t)size * sizeof(int64_t)), size - 1, 0};
    if (!c.set) return -1;
    for (uint32_t i = 0; i < size; ++i) c.set[i] = ORC_EMPTY_KEY;
    const float cam[3] = {cx, cy, cz};
    walk_points(v, pos, n, cam, radius_m, touch_visit, &c);
    free(c.set);
    return c.overflow ? -1 : c.n;
}

/* ---- state access for parity checks -------------------------------------- */

/* Keys of allocated blocks in pool order. Returns the count written. */
int32_t orc_block_keys(const OrcVolume* v, int64_t* out, int32_t cap) {
    if (!v || !out) return -1;
    const int32_t n = v->block_count < cap ? v->block_count : cap;
    for (int32_t i = 0; i < n; ++i)
        out[i] = orc_pack_coord(v->coord[i * 3], v->coord[i * 3 + 1], v->coord[i * 3 + 2]);
    return n;
}

/* Pool index of a block key, or -1. */
int32_t orc_find_key(const OrcVolume* v, int64_t key) {
    int32_t c[3];
    orc_unpack_key(key, c);
    return find_block(v, c[0], c[1], c[2]);
}

/* Copy one pool block's 512 voxels of each array (any pointer may be NULL). */
int32_t orc_block_voxels(const OrcVolume* v, int32_t bi, float* tsdf, float* weight, float* r,
                         float* g, float* b) {
    if (!v || bi < 0 || bi >= v->block_count) return 1;
    const size_t o = (size_t)bi * ORC_BLOCK_VOXELS;
    const size_t nb = ORC_BLOCK_VOXELS * sizeof(float);
    if (tsdf) memcpy(tsdf, v->tsdf + o, nb);
    if (weight) memcpy(weight, v->weight + o, nb);
    if (r) memcpy(r, v->r + o, nb);
    if (g) memcpy(g, v->g + o, nb);
    if (b) memcpy(b, v->b + o, nb);
    return 0;
}

/* Load a block with given voxel data (used to run this extractor over a device
 * state dumped to the host). With mean_mode=1 `tsdf` holds means. rgb may be
 * NULL, which stores the neutral 128*weight (mean 128) the device ABI implies.
 * Returns the pool index, or -1 when the pool is full. */
int32_t orc_load_block(OrcVolume* v, int64_t key, const float* tsdf, const float* weight,
                       int32_t mean_mode) {
    if (!v) return -1;
    int32_t c[3];
    orc_unpack_key(key, c);
    const int32_t bi = find_or_insert(v, c[0], c[1], c[2]);
    if (bi < 0) return -1;
    v->mean_mode = mean_mode;
    const size_t o = (size_t)bi * ORC_BLOCK_VOXELS;
    for (int k = 0; k < ORC_BLOCK_VOXELS; ++k) {
        v->tsdf[o + k] = tsdf[k];
        v->weight[o + k] = weight[k];
        const float rgb = mean_mode ? 128.0f : 128.0f * weight[k];
        v->r[o + k] = rgb; v->g[o + k] = rgb; v->b[o + k] = rgb;
    }
    return bi;
}

/* ---- a8: read-side mean (volume.cpp:306-329) ------------------------------ */

static int sample(const OrcVolume* v, int32_t vx, int32_t vy, int32_t vz, float min_weight,
                  float* tsdf_out, float rgb_out[3]) {
    const int32_t bx = orc_floor_div(vx, ORC_BLOCK_DIM);
    const int32_t by = orc_floor_div(vy, ORC_BLOCK_DIM);
    const int32_t bz = orc_floor_div(vz, ORC_BLOCK_DIM);
    const int32_t bi = find_block(v, bx, by, bz);
    if (bi < 0) return 0;
    const int32_t lx = vx - bx * ORC_BLOCK_DIM;
    const int32_t ly = vy - by * ORC_BLOCK_DIM;
    const int32_t lz = vz - bz * ORC_BLOCK_DIM;
    const size_t idx = (size_t)bi * ORC_BLOCK_VOXELS + (size_t)((lz * 8 + ly) * 8 + lx);
    const float w = v->weight[idx];
    if (w < min_weight || w <= 0.0f) return 0;
    if (v->mean_mode) {
        *tsdf_out = v->tsdf[idx];
        rgb_out[0] = v->r[idx]; rgb_out[1] = v->g[idx]; rgb_out[2] = v->b[idx];
        return 1;
    }
    const float inv_w = 1.0f / w;
    *tsdf_out = v->tsdf[idx] * inv_w;
    rgb_out[0] = v->r[idx] * inv_w;
    rgb_out[1] = v->g[idx] * inv_w;
    rgb_out[2] = v->b[idx] * inv_w;
    return 1;
}

/* ---- a9: marching tetrahedra (marching_tets.cpp:37-243) ------------------- */

static const int kTets[6][4] = {
    {0, 5, 1, 7}, {0, 1, 3, 7}, {0, 3, 2, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 2, 6, 7},
};
static const int kCorner[8][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1},
};

typedef struct { float p[3]; float c[3]; } Vtx;

static float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* marching_tets.cpp:59-80: position-ordered so a shared edge is bit-identical. */
static Vtx interp(float iso, const float pa[3], const float pb[3], float va, float vb,
                  const float ca[3], const float cb[3]) {
    const int swap = (pb[0] < pa[0]) || (pb[0] == pa[0] && pb[1] < pa[1]) ||
                     (pb[0] == pa[0] && pb[1] == pa[1] && pb[2] < pa[2]);
    const float* p0 = swap ? pb : pa;
    const float* p1 = swap ? pa : pb;
    const float* c0 = swap ? cb : ca;
    const float* c1 = swap ? ca : cb;
    const float v0 = swap ? vb : va;
    const float v1 = swap ? va : vb;
    const float denom = v1 - v0;
    const float t = fabsf(denom) < 1e-12f ? 0.5f : (iso - v0) / denom;
    const float tc = clampf(t, 0.0f, 1.0f);
    Vtx o;
    for (int k = 0; k < 3; ++k) {
        o.p[k] = p0[k] + tc * (p1[k] - p0[k]);
        o.c[k] = c0[k] + tc * (c1[k] - c0[k]);
    }
    return o;
}

/* Triangles of one tetrahedron given its sign code (marching_tets.cpp:165-182).
 * Writes up to 6 vertices into tri and returns the triangle count. */
static int tet_triangles(int code, const int i[4], float iso, float pos[8][3],
                         const float val[8], float col[8][3], Vtx tri[6]) {
    int a, b, c, d, quad;
    switch (code) {
        case 0x01: case 0x0E: a = i[0]; b = i[1]; c = i[2]; d = i[3]; quad = 0; break;
        case 0x02: case 0x0D: a = i[1]; b = i[0]; c = i[2]; d = i[3]; quad = 0; break;
        case 0x04: case 0x0B: a = i[2]; b = i[0]; c = i[1]; d = i[3]; quad = 0; break;
        case 0x08: case 0x07: a = i[3]; b = i[0]; c = i[1]; d = i[2]; quad = 0; break;
        case 0x03: a = i[0]; b = i[1]; c = i[2]; d = i[3]; quad = 1; break;
        case 0x0C: a = i[2]; b = i[3]; c = i[0]; d = i[1]; quad = 1; break;
        case 0x05: a = i[0]; b = i[2]; c = i[1]; d = i[3]; quad = 1; break;
        case 0x0A: a = i[1]; b = i[3]; c = i[0]; d = i[2]; quad = 1; break;
        case 0x06: a = i[1]; b = i[2]; c = i[0]; d = i[3]; quad = 1; break;
        case 0x09: a = i[0]; b = i[3]; c = i[1]; d = i[2]; quad = 1; break;
        default: return 0;
    }
    if (!quad) {
        tri[0] = interp(iso, pos[a], pos[b], val[a], val[b], col[a], col[b]);
        tri[1] = interp(iso, pos[a], pos[c], val[a], val[c], col[a], col[c]);
        tri[2] = interp(iso, pos[a], pos[d], val[a], val[d], col[a], col[d]);
        return 1;
    }
    const Vtx ac = interp(iso, pos[a], pos[c], val[a], val[c], col[a], col[c]);
    const Vtx ad = interp(iso, pos[a], pos[d], val[a], val[d], col[a], col[d]);
    const Vtx bd = interp(iso, pos[b], pos[d], val[b], val[d], col[b], col[d]);
    const Vtx bc = interp(iso, pos[b], pos[c], val[b], val[c], col[b], col[c]);
    tri[0] = ac; tri[1] = ad; tri[2] = bd;
    tri[3] = ac; tri[4] = bd; tri[5] = bc;
    return 2;
}

static void write_triangle(float* dst_posnor, uint8_t* dst_rgb, const Vtx* t) {
    for (int k = 0; k < 3; ++k) {
        float* d = dst_posnor + k * 6;
        d[0] = t[k].p[0]; d[1] = t[k].p[1]; d[2] = t[k].p[2];
        if (dst_rgb) {
            uint8_t* c = dst_rgb + k * 3;
            c[0] = (uint8_t)clampf(t[k].c[0], 0.0f, 255.0f);
            c[1] = (uint8_t)clampf(t[k].c[1], 0.0f, 255.0f);
            c[2] = (uint8_t)clampf(t[k].c[2], 0.0f, 255.0f);
        }
    }
    /* face normal (marching_tets.cpp:205-218) */
    float* v0 = dst_posnor;
    float* v1 = v0 + 6;
    float* v2 = v0 + 12;
    const float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    float nx = e1[1] * e2[2] - e1[2] * e2[1];
    float ny = e1[2] * e2[0] - e1[0] * e2[2];
    float nz = e1[0] * e2[1] - e1[1] * e2[0];
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    if (len > 1e-20f) { nx /= len; ny /= len; nz /= len; }
    v0[3] = nx; v0[4] = ny; v0[5] = nz;
    v1[3] = nx; v1[4] = ny; v1[5] = nz;
    v2[3] = nx; v2[4] = ny; v2[5] = nz;
}

/* Returns the vertex count, or -1 on overflow (nothing is promised about the
 * buffer then) or bad arguments (c_api.cpp:335-347). When posnor is NULL and
 * cap is 0 it only counts. */
int32_t orc_extract_mesh(const OrcVolume* v, float* posnor, uint8_t* rgb, int32_t cap,
                         float min_weight, float iso) {
    if (!v) return -1;
    const int count_only = (posnor == NULL && cap == 0);
    if (!count_only && (!posnor || cap <= 0)) return -1;
    const float vs = v->voxel;
    int64_t written = 0;
    int overflow = 0;
    Vtx tri[6];
    for (int32_t bi = 0; bi < v->block_count; ++bi) {
        const int32_t bx = v->coord[bi * 3], by = v->coord[bi * 3 + 1], bz = v->coord[bi * 3 + 2];
        for (int32_t lz = 0; lz < 8; ++lz)
        for (int32_t ly = 0; ly < 8; ++ly)
        for (int32_t lx = 0; lx < 8; ++lx) {
            const int32_t vx = bx * 8 + lx, vy = by * 8 + ly, vz = bz * 8 + lz;
            float val[8], pos[8][3], col[8][3];
            int ok = 1;
            for (int c = 0; c < 8 && ok; ++c) {
                float rgb3[3];
                ok = sample(v, vx + kCorner[c][0], vy + kCorner[c][1], vz + kCorner[c][2],
                            min_weight, &val[c], rgb3);
                if (!ok) break;
                pos[c][0] = ((float)(vx + kCorner[c][0]) + 0.5f) * vs;
                pos[c][1] = ((float)(vy + kCorner[c][1]) + 0.5f) * vs;
                pos[c][2] = ((float)(vz + kCorner[c][2]) + 0.5f) * vs;
                col[c][0] = rgb3[0]; col[c][1] = rgb3[1]; col[c][2] = rgb3[2];
            }
            if (!ok) continue;
            for (int t = 0; t < 6; ++t) {
                const int* i = kTets[t];
                const int code = (val[i[0]] < iso ? 1 : 0) | (val[i[1]] < iso ? 2 : 0) |
                                 (val[i[2]] < iso ? 4 : 0) | (val[i[3]] < iso ? 8 : 0);
                const int n_tri = tet_triangles(code, i, iso, pos, val, col, tri);
                if (n_tri == 0) continue;
                const int64_t nv = (int64_t)n_tri * 3;
                if (count_only) { written += nv; continue; }
                if (written + nv > cap) { overflow = 1; written += nv; continue; }
                for (int k = 0; k < n_tri; ++k)
                    write_triangle(posnor + (size_t)(written + k * 3) * 6,
                                   rgb ? rgb + (size_t)(written + k * 3) * 3 : NULL, &tri[k * 3]);
                written += nv;
            }
        }
    }
    if (overflow) return -1;
    if (written > INT32_MAX) return -1;
    return (int32_t)written;
}

/* ---- analytic point scenes (cpp/tests/analytic_scenes.hpp:19-46) ---------- */

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* n_theta*n_phi points, xyz interleaved. Same float expressions as the header. */
void orc_scene_sphere(float radius, int32_t n_theta, int32_t n_phi, float* out) {
    size_t k = 0;
    for (int i = 0; i < n_theta; ++i) {
        const float theta = (float)M_PI * ((float)i + 0.5f) / (float)n_theta;
        for (int j = 0; j < n_phi; ++j) {
            const float phi = 2.0f * (float)M_PI * (float)j / (float)n_phi;
            out[k++] = radius * sinf(theta) * cosf(phi);
            out[k++] = radius * sinf(theta) * sinf(phi);
            out[k++] = radius * cosf(theta);
        }
    }
}

void orc_scene_plane(float z0, float half_extent, int32_t n, float* out) {
    size_t k = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            out[k++] = -half_extent + 2.0f * half_extent * (float)i / (float)(n - 1);
            out[k++] = -half_extent + 2.0f * half_extent * (float)j / (float)(n - 1);
            out[k++] = z0;
        }
}

/* ---- FP32 back-projection (the device kernel's definition) ----------------
 * Pixel (u, v) with planar depth z back-projects to camera coordinates
 * ((u - cx) * z / fx, (v - cy) * z / fy, z) -- pixel index, not index + 0.5,
 * as crates/tartanair/src/lib.rs:323-326 does -- and then to world as
 * R_wc * p + t_wc, evaluated left to right with no fused multiply-add. The
 * reference does this in FP64 on the host; the device does it in FP32, and
 * this function is that FP32 definition, bit for bit. Pixels with z <= 0,
 * non-finite z, or z > max_range (when max_range > 0) are skipped. Output is in
 * row-major pixel order. Returns the number of points written.
 */
int32_t orc_backproject_f32(const float* depth, int32_t W, int32_t H, float fx, float fy,
                            float cx, float cy, const float* R, const float* t,
                            float max_range, float* out) {
    int32_t n = 0;
    for (int32_t v = 0; v < H; ++v)
        for (int32_t u = 0; u < W; ++u) {
            const float z = depth[(size_t)v * W + u];
            if (!(z > 0.0f) || !isfinite(z) || (max_range > 0.0f && z > max_range)) continue;
            const float xc = ((float)u - cx) * z / fx;
            const float yc = ((float)v - cy) * z / fy;
            for (int i = 0; i < 3; ++i)
                out[(size_t)n * 3 + i] = R[i * 3 + 0] * xc + R[i * 3 + 1] * yc + R[i * 3 + 2] * z + t[i];
            ++n;
        }
    return n;
}

/* ---- projective update: FP64 oracle --------------------------------------
 * For every voxel of every block the frame touched (the distinct blocks its
 * band walk visits), with the camera transform R_cw (world -> camera) and the
 * camera centre c:
 *   p = R_cw * ((v + 0.5) * voxel - c)
 *   skip unless p.z > 0
 *   pixel = (floor(fx * p.x / p.z + cx + 0.5), floor(fy * p.y / p.z + cy + 0.5))
 *   skip if outside the image, or if its depth d is not > 0
 *   sdf = d - p.z; skip if sdf < -trunc; x = min(sdf * (1 / trunc), 1)
 *   skip if w >= w_max (the voxel is frozen at the cap, as the reference's
 *   weight_cap freezes it: types.hpp:99-101, volume.cpp:283-289)
 *   t <- (t * w + x) / (w + 1);  w <- w + 1
 * so after n observations t is the mean of the first min(n, w_max) of them
 * and w = min(n, w_max): the reference's read-side mean sum / w
 * (volume.cpp:322-324) of the same observations with unit weights.
 * The discrete decisions (pixel, the z and sdf gates) are evaluated in FP32
 * exactly as the device evaluates them, so that this oracle and the device
 * make the same choices; the continuous part (p, sdf, the running average) is
 * FP64 from the FP64 pose. Without that split the FP32-vs-FP64 error at a
 * depth discontinuity measures which side of a pixel edge a voxel projected
 * to, not storage or accumulation precision.
 *
 * The FP32 part is the device's exact expression sequence, in which the
 * pixel coordinates are ratios of two linear forms of the voxel index:
 *   A = fx * p.x + (cx + 0.5) * p.z,  B = fy * p.y + (cy + 0.5) * p.z,  Z = p.z
 *   per frame:  e_j = voxel * R_cw[:, j]        (j = x, y, z voxel axes)
 *               (A_j, B_j, Z_j) = (fx * e_j.x + cxh * e_j.z, fy * e_j.y + cyh * e_j.z, e_j.z)
 *               cxh = cx + 0.5, cyh = cy + 0.5
 *   per block:  o  = (float)(b * 8) * voxel + 0.5f * voxel - c   (block origin centre)
 *               d0 = R_cw * o                   (row dot, left to right)
 *               (A0, B0, Z0) = (fx * d0.x + cxh * d0.z, fy * d0.y + cyh * d0.z, d0.z)
 *   per run of 4 voxels along x (lx0 = 0 or 4):
 *               off = (ly * F_y + lz * F_z) + lx0 * F_x      (F = A, B, Z)
 *               F(lx0) = F0 + off,  F(lx + 1) = F(lx) + F_x
 *   per run:    iz(j) for its four Z values by batch inversion (run_reciprocals)
 *   per voxel:  skip unless 2^-10 <= Z <= 2^14; uf = A * iz; vf = B * iz
 *               skip unless 0 <= uf < W and 0 <= vf < H; pixel = ((int)uf, (int)vf)
 * with each product rounded separately (no FMA).
 */
typedef struct {
    float Rcw[9];   /* world -> camera, row-major */
    float c[3];     /* camera centre in world */
    double Rcw64[9];
    double c64[3];
    float fx, fy, cx, cy;
    int32_t W, H;
} OrcProjCam;

typedef struct {
    int32_t valid;     /* 1 when the voxel gets an observation */
    int32_t u, v;      /* chosen pixel */
    float z32;         /* FP32 camera z */
    float sdf32;       /* FP32 sdf (after the gates) */
} ProjDecision;

/* Observable camera depth of a voxel centre, metres, and the reciprocals of a
 * run of four depths by batch inversion (clamp to the range, running
 * products, one correctly rounded division, peel off with partial products),
 * every operation rounded on its own. Restates proj.hip run_reciprocals. */
#define PROJ_ZMIN (1.0f / 1024.0f)
#define PROJ_ZMAX 16384.0f

static void run_reciprocals(const float Z[4], float iz[4]) {
    float z[4];
    for (int j = 0; j < 4; ++j) z[j] = fminf(fmaxf(Z[j], PROJ_ZMIN), PROJ_ZMAX);
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

static ProjDecision proj_decide(const OrcProjCam* k, const float* depth, float voxel, float trunc,
                                int32_t bx, int32_t by, int32_t bz, int lx, int ly, int lz) {
    ProjDecision r = {0, 0, 0, 0.0f, 0.0f};
    const float cxh = k->cx + 0.5f, cyh = k->cy + 0.5f;
    float g[3][3];  /* g[F][j]: F = A, B, Z; j = voxel axis */
    for (int j = 0; j < 3; ++j) {
        const float ex = voxel * k->Rcw[0 * 3 + j];
        const float ey = voxel * k->Rcw[1 * 3 + j];
        const float ez = voxel * k->Rcw[2 * 3 + j];
        g[0][j] = k->fx * ex + cxh * ez;
        g[1][j] = k->fy * ey + cyh * ez;
        g[2][j] = ez;
    }
    const float h = 0.5f * voxel;
    const float ox = (float)(bx * 8) * voxel + h - k->c[0];
    const float oy = (float)(by * 8) * voxel + h - k->c[1];
    const float oz = (float)(bz * 8) * voxel + h - k->c[2];
    float d0[3];
    for (int i = 0; i < 3; ++i)
        d0[i] = k->Rcw[i * 3 + 0] * ox + k->Rcw[i * 3 + 1] * oy + k->Rcw[i * 3 + 2] * oz;
    const float f0[3] = {k->fx * d0[0] + cxh * d0[2], k->fy * d0[1] + cyh * d0[2], d0[2]};
    /* The voxel's run: the four voxels lx0..lx0+3 along x, whose forms are
     * the run's first value plus one step per voxel. */
    const int lx0 = lx & 4;
    float run[3][4];
    for (int q = 0; q < 3; ++q) {
        const float off = ((float)ly * g[q][1] + (float)lz * g[q][2]) + (float)lx0 * g[q][0];
        run[q][0] = f0[q] + off;
        for (int j = 1; j < 4; ++j) run[q][j] = run[q][j - 1] + g[q][0];
    }
    const int j = lx - lx0;
    const float Z = run[2][j];
    if (!(Z >= PROJ_ZMIN) || !(Z <= PROJ_ZMAX)) return r;
    float izr[4];
    run_reciprocals(run[2], izr);
    const float iz = izr[j];
    const float uf = run[0][j] * iz;
    const float vf = run[1][j] * iz;
    if (!(uf >= 0.0f) || !(vf >= 0.0f) || !(uf < (float)k->W) || !(vf < (float)k->H)) return r;
    const int32_t u = (int32_t)uf, v = (int32_t)vf;
    const float d = depth[(size_t)v * k->W + u];
    if (!(d > 0.0f)) return r;
    const float sdf = d - Z;
    if (sdf < -trunc) return r;
    r.valid = 1; r.u = u; r.v = v; r.z32 = Z; r.sdf32 = sdf;
    return r;
}

/* The FP32 device expressions for one voxel, returned for the device tests:
 * fills x (the clamped normalised sdf) and returns 1 when observed. */
int32_t orc_proj_observe_f32(const float* Rcw, const float* c, float fx, float fy, float cx,
                             float cy, int32_t W, int32_t H, const float* depth, float voxel,
                             float trunc, int32_t bx, int32_t by, int32_t bz, int32_t lx,
                             int32_t ly, int32_t lz, float* x_out) {
    OrcProjCam k;
    memcpy(k.Rcw, Rcw, sizeof(k.Rcw));
    memcpy(k.c, c, sizeof(k.c));
    k.fx = fx; k.fy = fy; k.cx = cx; k.cy = cy; k.W = W; k.H = H;
    const ProjDecision d = proj_decide(&k, depth, voxel, trunc, bx, by, bz, lx, ly, lz);
    if (!d.valid) return 0;
    const float x = d.sdf32 * (1.0f / trunc);
    *x_out = x < 1.0f ? x : 1.0f;
    return 1;
}

/* FP64 projective state over a fixed, caller-owned block list. */
typedef struct OrcProj {
    float voxel, trunc;
    int32_t w_max;
    int32_t n_blocks;
    int32_t cap_blocks;
    int64_t* keys;     /* per block */
    double* tsdf;      /* n*512 */
    int32_t* weight;   /* n*512 */
    /* key -> index map: open addressing over a power of two */
    int64_t* mkeys;
    int32_t* mval;
    uint32_t mmask;
} OrcProj;

void orc_proj_destroy(OrcProj* p) {
    if (!p) return;
    free(p->keys); free(p->tsdf); free(p->weight); free(p->mkeys); free(p->mval);
    free(p);
}

/* w_max: the weight at which a voxel freezes (ceil(weight_cap) when the cap is
 * > 0, else 65535, the u16 storage's largest weight). */
OrcProj* orc_proj_create(float voxel, float trunc, int32_t w_max, int32_t cap_blocks) {
    if (!(voxel > 0.0f) || cap_blocks <= 0) return NULL;
    OrcProj* p = (OrcProj*)calloc(1, sizeof(OrcProj));
    if (!p) return NULL;
    p->voxel = voxel;
    p->trunc = trunc > 0.0f ? trunc : 4.0f * voxel;
    p->w_max = w_max;
    p->cap_blocks = cap_blocks;
    p->keys = (int64_t*)malloc((size_t)cap_blocks * sizeof(int64_t));
    p->tsdf = (double*)calloc((size_t)cap_blocks * ORC_BLOCK_VOXELS, sizeof(double));
    p->weight = (int32_t*)calloc((size_t)cap_blocks * ORC_BLOCK_VOXELS, sizeof(int32_t));
    const uint32_t msize = next_pow2((uint32_t)cap_blocks * 2u);
    p->mmask = msize - 1;
    p->mkeys = (int64_t*)malloc((size_t)msize * sizeof(int64_t));
    p->mval = (int32_t*)malloc((size_t)msize * sizeof(int32_t));
    if (!p->keys || !p->tsdf || !p->weight || !p->mkeys || !p->mval) {
        orc_proj_destroy(p);
        return NULL;
    }
    for (uint32_t i = 0; i < msize; ++i) p->mkeys[i] = ORC_EMPTY_KEY;
    return p;
}

static int32_t proj_index(OrcProj* p, int64_t key, int insert) {
    uint64_t h = (uint64_t)key * 0x9E3779B97F4A7C15ull;
    uint32_t s = (uint32_t)(h >> 32) & p->mmask;
    for (;;) {
        if (p->mkeys[s] == key) return p->mval[s];
        if (p->mkeys[s] == ORC_EMPTY_KEY) {
            if (!insert || p->n_blocks >= p->cap_blocks) return -1;
            p->mkeys[s] = key;
            p->mval[s] = p->n_blocks;
            p->keys[p->n_blocks] = key;
            return p->n_blocks++;
        }
        s = (s + 1) & p->mmask;
    }
}

/* Integrate one frame over the listed touched blocks (keys). Rcw/c in FP64 and
 * in the FP32 values the device receives. Returns observations applied. */
int64_t orc_proj_integrate(OrcProj* p, const int64_t* touched, int32_t n_touched,
                           const float* depth, int32_t W, int32_t H, float fx, float fy,
                           float cx, float cy, const double* Rcw64, const double* c64,
                           const float* Rcw32, const float* c32) {
    if (!p) return -1;
    OrcProjCam k;
    memcpy(k.Rcw, Rcw32, sizeof(k.Rcw));
    memcpy(k.c, c32, sizeof(k.c));
    memcpy(k.Rcw64, Rcw64, sizeof(k.Rcw64));
    memcpy(k.c64, c64, sizeof(k.c64));
    k.fx = fx; k.fy = fy; k.cx = cx; k.cy = cy; k.W = W; k.H = H;
    int64_t n_obs = 0;
    const double vs = (double)p->voxel;
    for (int32_t i = 0; i < n_touched; ++i) {
        const int32_t bi = proj_index(p, touched[i], 1);
        if (bi < 0) continue;
        int32_t b[3];
        orc_unpack_key(touched[i], b);
        for (int lz = 0; lz < 8; ++lz)
        for (int ly = 0; ly < 8; ++ly)
        for (int lx = 0; lx < 8; ++lx) {
            const ProjDecision d = proj_decide(&k, depth, p->voxel, p->trunc, b[0], b[1], b[2], lx, ly, lz);
            if (!d.valid) continue;
            const double wx = ((double)(b[0] * 8 + lx) + 0.5) * vs - k.c64[0];
            const double wy = ((double)(b[1] * 8 + ly) + 0.5) * vs - k.c64[1];
            const double wz = ((double)(b[2] * 8 + lz) + 0.5) * vs - k.c64[2];
            const double z = k.Rcw64[6] * wx + k.Rcw64[7] * wy + k.Rcw64[8] * wz;
            const double dd = (double)depth[(size_t)d.v * W + d.u];
            double x = (dd - z) / (double)p->trunc;
            if (x > 1.0) x = 1.0;
            if (x < -1.0) x = -1.0;
            const size_t o = (size_t)bi * ORC_BLOCK_VOXELS + (size_t)((lz * 8 + ly) * 8 + lx);
            const int32_t w = p->weight[o];
#if ORC_PLANT == 1  /* planted: no freeze, the cap only bounds the divisor */
            p->tsdf[o] = (p->tsdf[o] * (double)w + x) / (double)(w + 1);
            p->weight[o] = w + 1 < p->w_max ? w + 1 : p->w_max;
#elif ORC_PLANT == 2  /* planted: cap applied before averaging */
            const int32_t wc = w + 1 < p->w_max ? w + 1 : p->w_max;
            p->tsdf[o] = (p->tsdf[o] * (double)w + x) / (double)wc;
            p->weight[o] = wc;
#elif ORC_PLANT == 3  /* planted: last observation only */
            if (w >= p->w_max) continue;
            p->tsdf[o] = x;
            p->weight[o] = w + 1;
#elif ORC_PLANT == 4  /* planted: frozen one observation early */
            if (w + 1 >= p->w_max) continue;
            p->tsdf[o] = (p->tsdf[o] * (double)w + x) / (double)(w + 1);
            p->weight[o] = w + 1;
#else
            /* a voxel whose weight has reached the cap is frozen: later
             * observations do not move it (types.hpp:99-101, volume.cpp:289) */
            if (w >= p->w_max) continue;
            p->tsdf[o] = (p->tsdf[o] * (double)w + x) / (double)(w + 1);
            p->weight[o] = w + 1;
#endif
            ++n_obs;
        }
    }
    return n_obs;
}

int32_t orc_proj_block_count(const OrcProj* p) { return p ? p->n_blocks : 0; }

/* Copy a block's state by key: tsdf (FP64 means) and integer weights. */
int32_t orc_proj_block(OrcProj* p, int64_t key, double* tsdf, int32_t* weight) {
    if (!p) return 1;
    const int32_t bi = proj_index(p, key, 0);
    if (bi < 0) return 1;
    const size_t o = (size_t)bi * ORC_BLOCK_VOXELS;
    if (tsdf) memcpy(tsdf, p->tsdf + o, ORC_BLOCK_VOXELS * sizeof(double));
    if (weight) memcpy(weight, p->weight + o, ORC_BLOCK_VOXELS * sizeof(int32_t));
    return 0;
}
