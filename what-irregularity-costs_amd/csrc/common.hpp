// Shared layout, constants and device helpers for the gfx950 TSDF library.
//
// Layout decisions that must match the reference exactly (cited):
//   * 64-bit packed block key, 21 bits per axis biased by 2^20
//     (cpp/include/osn_tsdf/types.hpp:52-69); -1 marks an empty slot (:50).
//   * 16-byte hash entry {int64 key, int32 block_idx, int32 pad}
//     (types.hpp:87-91). This library uses the pad word as the "last call that
//     touched this block" stamp; the reference never reads it.
//   * 8^3 blocks, voxel index (lz*8 + ly)*8 + lx (cpp/src/volume.cpp:271-276).
//   * Teschner multiply-xor hash masked to a power of two (types.hpp:150-155)
//     for tables created through the reference-shaped entry points, so the
//     device view handed to an external arm describes a table it can probe.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace osn_hip {

constexpr int kBlockDim = 8;
constexpr int kBlockVoxels = 512;
constexpr int64_t kEmptyKey = -1;
constexpr int64_t kCoordBias = int64_t(1) << 20;
constexpr int kCoordBits = 21;
// block_idx values other than a pool index
constexpr int32_t kIdxPending = -1;  // key inserted, index not yet assigned
constexpr int32_t kIdxDropped = -2;  // pool was full when the key was assigned

constexpr int kProbeHistBins = 64;  // depth 1..63, and 64 = "64 or more"

struct HashEntry {
    int64_t key;
    int32_t block_idx;
    int32_t stamp;
};
static_assert(sizeof(HashEntry) == 16, "hash entry must stay 16 bytes");

enum HashKind : int32_t {
    kHashTeschnerPow2 = 0,  // reference hash, capacity a power of two, mask
    kHashMixRange = 1,      // 64-bit mix of the key, any capacity, multiply-shift range
};

enum Storage : int32_t {
    kStoreSums = 0,  // reference: FP32 weighted sums tsdf, weight, r, g, b (SoA)
    kStoreF32 = 1,   // projective: {float tsdf_mean, float weight} per voxel
    kStoreF16 = 2,   // projective: {half tsdf_mean, uint16 weight} per voxel
    kStoreF64 = 3,   // projective FP64 oracle: {double tsdf_mean, double weight}
};

// Per-call list lengths. Allocate calls alternate between two sets (call
// parity): a call appends to lists[p] and its first workgroup clears
// lists[p ^ 1] for the next call, whose readers of lists[p ^ 1] (the previous
// call's update) have finished by then because they ran earlier on the same
// stream. Each set sits on its own 128-byte line so appends to one list do
// not serialise against another's.
struct alignas(128) ListCounters {
    int32_t new_count;       // keys inserted by the call
    int32_t pad0[31];
    int32_t touched_count;   // distinct slots the call visited
    int32_t pad1[31];
    int32_t assign_ticket;   // next work item of the call's index assignment
    int32_t pad2[31];
};

// Per-volume device counters, one allocation. block_count and drop_count are
// what the reference device view exposes as int32* and uint64*.
struct Counters {
    int32_t block_count;
    uint32_t fault;                  // nonzero: a bounded wait of the assignment gave up
    unsigned long long drop_count;
    unsigned long long cas_attempts;
    unsigned long long lane_steps;   // sum over lanes of probe steps
    unsigned long long wave_steps;   // sum over wave batches of 64 * deepest lane
    unsigned long long lookups;      // distinct-key probes issued to the table
    int32_t mesh_vertices;           // last extract's vertex count, -1 on overflow
    int32_t nb_blocks;               // blocks 0..nb_blocks-1 have their neighbour indices
                                     // in the extraction scratch (-1 entries are looked up again)
    unsigned long long probe_hist[kProbeHistBins];
    ListCounters lists[2];
};

// One block a call visited: its key, its table slot and its pool index when
// the block already existed (kIdxPending for blocks the call inserted, whose
// index the assignment pass writes to the table), so the update reaches the
// voxel record without walking slot -> index -> coordinate.
struct TouchEntry {
    int64_t key;
    uint32_t slot;
    int32_t bi;
};
static_assert(sizeof(TouchEntry) == 16, "touch entry must stay 16 bytes");

// Everything a kernel needs to address one volume. Passed by value.
struct VolDev {
    HashEntry* table;
    uint32_t capacity;     // table slots
    uint32_t mask;         // capacity - 1 when a power of two
    int32_t hash_kind;
    uint32_t probe_limit;  // max slots a lane inspects before it gives up
    Counters* ctr;
    int32_t* block_coord;  // 3 per pool slot
    int32_t pool_capacity;
    int32_t storage;
    // kStoreSums
    float* tsdf;
    float* weight;
    float* r;
    float* g;
    float* b;
    // kStoreF32 / kStoreF16 / kStoreF64
    void* vox;
    float voxel;
    float inv_voxel;
    float trunc;
    float weight_cap;
    int32_t steps;         // band half-width in samples
    int32_t w_max;         // projective: weight at which a voxel freezes, ceil(cap) or 65535
    // per-call work lists, capacity entries each; new_keys entries carry the
    // slot the key was inserted at (bi unused), sort_a/sort_b are the merge
    // buffers of the index assignment
    TouchEntry* new_keys;
    TouchEntry* touched;
    TouchEntry* sort_a;
    TouchEntry* sort_b;
    int32_t* assign_done;  // per parity: finished work items per run and level
};

__host__ __device__ inline int64_t pack_coord(int32_t x, int32_t y, int32_t z) {
    return ((int64_t(x) + kCoordBias) << (2 * kCoordBits)) |
           ((int64_t(y) + kCoordBias) << kCoordBits) | (int64_t(z) + kCoordBias);
}

__host__ __device__ inline void unpack_key(int64_t key, int32_t& x, int32_t& y, int32_t& z) {
    const int64_t m = (int64_t(1) << kCoordBits) - 1;
    x = int32_t(((key >> (2 * kCoordBits)) & m) - kCoordBias);
    y = int32_t(((key >> kCoordBits) & m) - kCoordBias);
    z = int32_t((key & m) - kCoordBias);
}

// floor(a / 8) for any sign: an arithmetic shift is exactly floor division by
// a power of two, and is what floor_div(a, 8) (types.hpp:162-165) computes.
__host__ __device__ inline int32_t floor_div8(int32_t a) { return a >> 3; }

__host__ __device__ inline uint32_t teschner(int32_t x, int32_t y, int32_t z) {
    return uint32_t(x) * 73856093u ^ uint32_t(y) * 19349663u ^ uint32_t(z) * 83492791u;
}

// Home slot of a key.
__device__ inline uint32_t home_slot(const VolDev& v, int64_t key) {
    if (v.hash_kind == kHashTeschnerPow2) {
        int32_t x, y, z;
        unpack_key(key, x, y, z);
        return teschner(x, y, z) & v.mask;
    }
    // murmur3 fmix64, then Lemire's multiply-shift range reduction
    uint64_t h = uint64_t(key);
    h ^= h >> 33; h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return uint32_t((uint64_t(uint32_t(h)) * uint64_t(v.capacity)) >> 32);
}

__device__ inline uint32_t next_slot(const VolDev& v, uint32_t s) {
    if (v.hash_kind == kHashTeschnerPow2) return (s + 1) & v.mask;
    return (s + 1 == v.capacity) ? 0u : s + 1;
}

// Lookup: pool index of a key, kIdxPending/kIdxDropped, or -3 when absent.
// Only called once every insert of the key has been assigned (a later kernel),
// so there is nothing to wait for.
constexpr int32_t kIdxAbsent = -3;
__device__ inline int32_t find_index(const VolDev& v, int64_t key) {
    uint32_t s = home_slot(v, key);
    for (uint32_t p = 0; p < v.probe_limit; ++p) {
        // key and index in one 16-byte load: a hit costs one memory round trip
        const int4 q = *reinterpret_cast<const int4*>(&v.table[s]);
        asm volatile("" ::"v"(q.z));  // keeps the index load beside the key's (not sunk past the compare)
        const int64_t k = (int64_t)(((uint64_t)(uint32_t)q.y << 32) | (uint32_t)q.x);
        if (k == key) return q.z;
        if (k == kEmptyKey) return kIdxAbsent;
        s = next_slot(v, s);
    }
    return kIdxAbsent;
}

}  // namespace osn_hip
