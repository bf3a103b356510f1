// Allocation for gfx950: band walk -> wave64 dedup -> workgroup LDS pre-hash ->
// table insert (phase A), then a deterministic sorted index assignment
// (phase B). Also the reference-semantics scatter update, which shares the
// band walk.
//
// Reference behaviour this reproduces: cpp/src/volume.cpp:147-198 (which blocks
// a batch allocates) and :200-299 (the scatter update). Insertion protocol:
// cpp/src/volume_cuda.cu:59-107 publishes a key by CAS, then takes a pool index
// from a counter, then publishes the index while readers spin on it. On
// wave64 hardware a lane spinning on a value another lane of its own wave has
// yet to publish can stall the wave; here nothing in phase A ever reads an
// index, so there is no spin, and phase B assigns indices in key order, which
// makes the pool order a function of the key sets alone.
//
// Compiled with -ffp-contract=off: the band walk must round exactly as the
// reference's x86-64 g++ build (SSE, no FMA) so that the allocated key set is
// bit-identical.

#include "internal.hpp"

namespace osn_hip {
namespace {

constexpr int kAllocThreads = 256;  // 4 waves; 16x16 pixels in depth mode
constexpr int kKeysPerLane = 8;     // distinct keys a lane buffers before a flush

__device__ inline int64_t readlane64(int64_t x, int lane) {
    const int lo = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(x))), lane);
    const int hi = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(x) >> 32)), lane);
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint64_t(uint32_t(lo)));
}

__device__ inline unsigned long long lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Append `mine` (when `want`) to a device list behind a counter with one
// atomic per wave. Must be called by the whole wave.
template <typename T>
__device__ inline void wave_append(bool want, T mine, int32_t* counter, T* list) {
    const unsigned long long m = __ballot(want);
    if (m == 0) return;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, __popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);
    if (want) list[base + __popcll(m & lanemask_lt())] = mine;
}

__device__ inline void wave_add_u64(unsigned long long x, unsigned long long* dst) {
    // reduce across the wave, one atomic
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (__lane_id() == 0 && x) atomicAdd(dst, x);
}

struct LaneKeys {
    int64_t k[kKeysPerLane];
    uint32_t valid = 0;
    int64_t last = kEmptyKey;

    __device__ inline void push(int64_t key) {
#pragma unroll
        for (int j = kKeysPerLane - 1; j > 0; --j) k[j] = k[j - 1];
        k[0] = key;
        valid = ((valid << 1) | 1u) & ((1u << kKeysPerLane) - 1u);
        last = key;
    }
    __device__ inline int64_t first() const {
        int64_t f = kEmptyKey;
#pragma unroll
        for (int j = kKeysPerLane - 1; j >= 0; --j)
            if ((valid >> j) & 1u) f = k[j];
        return f;
    }
    __device__ inline void remove(int64_t key) {
#pragma unroll
        for (int j = 0; j < kKeysPerLane; ++j)
            if (((valid >> j) & 1u) && k[j] == key) valid &= ~(1u << j);
    }
};

struct WgState {
    int64_t* set;     // LDS pre-hash, `set_size` entries
    int set_size;
    unsigned long long* hist;  // LDS histogram (probe mode)
    int32_t stamp;
};

// One batch of up to 64 distinct keys, one per lane (has = lane holds one).
template <int MODE, bool TOUCH>
__device__ void process_keys(const VolDev& v, const WgState& wg, bool has, int64_t key,
                             unsigned long long& cas_acc, unsigned long long& drop_acc,
                             unsigned long long& lane_acc, unsigned long long& lookup_acc,
                             unsigned long long& wave_acc) {
    // Workgroup pre-hash: only the lane that claims a key in LDS goes on to the
    // global table; the other waves of the workgroup drop their copy.
    bool go = has;
    if (has && wg.set_size > 0) {
        uint32_t h = uint32_t((uint64_t(key) * 0x9E3779B97F4A7C15ull) >> 40) & (wg.set_size - 1);
        for (int p = 0; p < 8; ++p) {
            const unsigned long long prev = atomicCAS(
                reinterpret_cast<unsigned long long*>(&wg.set[h]),
                (unsigned long long)kEmptyKey, (unsigned long long)key);
            if (prev == (unsigned long long)kEmptyKey) break;  // ours
            if (prev == (unsigned long long)key) { go = false; break; }
            h = (h + 1) & (wg.set_size - 1);
        }
    }

    bool inserted = false, found = false;
    uint32_t slot = 0;
    uint32_t steps = 0;
    if (go) {
        lookup_acc += 1;
        slot = home_slot(v, key);
        for (uint32_t p = 0; p < v.probe_limit; ++p) {
            ++steps;
            const int64_t cur =
                __hip_atomic_load(&v.table[slot].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == key) { found = true; break; }
            if (cur == kEmptyKey) {
                if (MODE == kAllocNoCas) break;
                if (MODE == kAllocCountCas) cas_acc += 1;
                const unsigned long long prev =
                    atomicCAS(reinterpret_cast<unsigned long long*>(&v.table[slot].key),
                              (unsigned long long)kEmptyKey, (unsigned long long)key);
                if (prev == (unsigned long long)kEmptyKey) { inserted = true; break; }
                if (prev == (unsigned long long)key) { found = true; break; }
            }
            slot = next_slot(v, slot);
        }
        if (MODE != kAllocNoCas && !found && !inserted) drop_acc += 1;  // probe bound hit
        if (found) {
            const int32_t idx =
                __hip_atomic_load(&v.table[slot].block_idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (idx == kIdxDropped) drop_acc += 1;
        }
    }

    if (MODE == kAllocProbe) {
        lane_acc += steps;
        if (go) {
            const int bin = steps < (uint32_t)kProbeHistBins ? int(steps) - 1 : kProbeHistBins - 1;
            atomicAdd(&wg.hist[bin], 1ull);
        }
        // divergence cost: the wave runs the probe loop until its deepest lane
        // finishes, for every lane
        uint32_t mx = steps;
        for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
        if (__lane_id() == 0) wave_acc += 64ull * mx;
    }

    bool touch = false;
    if (TOUCH && (found || inserted)) {
        const int32_t old = atomicExch(&v.table[slot].stamp, wg.stamp);
        touch = (old != wg.stamp);
    }
    if (TOUCH) wave_append<uint32_t>(touch, slot, &v.ctr->touched_count, v.touched);
    if (MODE != kAllocNoCas) wave_append<int64_t>(inserted, key, &v.ctr->new_count, v.new_keys);
}

// Drain every lane's buffered keys: repeatedly take the first live key of the
// first live lane, strike it from every lane, and hand it to lane `nd`. Each
// distinct key of the wave is therefore processed once, 64 at a time.
template <int MODE, bool TOUCH>
__device__ void flush_wave(const VolDev& v, const WgState& wg, LaneKeys& lk,
                           unsigned long long& cas_acc, unsigned long long& drop_acc,
                           unsigned long long& lane_acc, unsigned long long& lookup_acc,
                           unsigned long long& wave_acc) {
    int nd = 0;
    int64_t mine = kEmptyKey;
    for (;;) {
        const unsigned long long m = __ballot(lk.valid != 0);
        if (m == 0) break;
        const int leader = __ffsll((long long)m) - 1;
        const int64_t k = readlane64(lk.first(), leader);
        lk.remove(k);
        if (__lane_id() == nd) mine = k;
        if (++nd == 64) {
            process_keys<MODE, TOUCH>(v, wg, true, mine, cas_acc, drop_acc, lane_acc, lookup_acc,
                                      wave_acc);
            nd = 0;
            mine = kEmptyKey;
        }
    }
    if (nd > 0)
        process_keys<MODE, TOUCH>(v, wg, __lane_id() < nd, mine, cas_acc, drop_acc, lane_acc,
                                  lookup_acc, wave_acc);
}

// One point per thread. FROM_DEPTH: a 16x16 pixel tile per workgroup, 8x8 per
// wave; the point is the pixel back-projected in FP32 as
//   xc = (u - cx) * z / fx, yc = (v - cy) * z / fy, p = R * (xc, yc, z) + cam
// evaluated left to right without fused multiply-add.
enum SrcKind : int { kSrcPoints = 0, kSrcDepth = 1, kSrcCoords = 2 };

template <int MODE, bool TOUCH, int SRC>
__global__ __launch_bounds__(kAllocThreads) void k_alloc(VolDev v, PointSrc src, const int32_t* coords,
                                                         int32_t set_size) {
    __shared__ int64_t s_set[2048];
    __shared__ unsigned long long s_hist[kProbeHistBins];
    for (int i = threadIdx.x; i < set_size; i += kAllocThreads) s_set[i] = kEmptyKey;
    if (MODE == kAllocProbe)
        for (int i = threadIdx.x; i < kProbeHistBins; i += kAllocThreads) s_hist[i] = 0;
    __syncthreads();

    WgState wg{s_set, set_size, s_hist, __hip_atomic_load(&v.ctr->stamp, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT)};

    unsigned long long cas_acc = 0, drop_acc = 0, lane_acc = 0, lookup_acc = 0, wave_acc = 0;
    LaneKeys lk;
    if (SRC == kSrcCoords) {
        // allocator stress path: the keys are given, no band walk
        const int i = blockIdx.x * kAllocThreads + threadIdx.x;
        if (i < src.n) lk.push(pack_coord(coords[i * 3 + 0], coords[i * 3 + 1], coords[i * 3 + 2]));
        flush_wave<MODE, TOUCH>(v, wg, lk, cas_acc, drop_acc, lane_acc, lookup_acc, wave_acc);
    } else {
    bool ok = false;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (SRC == kSrcDepth) {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int u = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
        const int w = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
        if (u < src.W && w < src.H) {
            const float z = src.depth[(size_t)w * src.W + u];
            ok = (z > 0.0f) && __builtin_isfinite(z) && !(src.max_range > 0.0f && z > src.max_range);
            if (ok) {
                const float xc = ((float)u - src.cx) * z / src.fx;
                const float yc = ((float)w - src.cy) * z / src.fy;
                px = src.R[0] * xc + src.R[1] * yc + src.R[2] * z + src.cam[0];
                py = src.R[3] * xc + src.R[4] * yc + src.R[5] * z + src.cam[1];
                pz = src.R[6] * xc + src.R[7] * yc + src.R[8] * z + src.cam[2];
            }
        }
    } else {
        const int i = blockIdx.x * kAllocThreads + threadIdx.x;
        if (i < src.n) {
            px = src.pos[i * 3 + 0];
            py = src.pos[i * 3 + 1];
            pz = src.pos[i * 3 + 2];
            ok = true;
        }
    }

    // Band walk (volume.cpp:158-191).
    float dist = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
    if (ok) {
        const float dx = px - src.cam[0], dy = py - src.cam[1], dz = pz - src.cam[2];
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (src.radius_sq > 0.0f && d2 > src.radius_sq) ok = false;
        dist = sqrtf(d2);
        if (!(dist > 1e-6f)) ok = false;
        ux = dx / dist; uy = dy / dist; uz = dz / dist;
    }

    for (int32_t s = -v.steps; s <= v.steps; ++s) {
        if (ok) {
            const float t = (float)s * v.voxel;
            const int32_t vx = (int32_t)floorf((px + ux * t) * v.inv_voxel);
            const int32_t vy = (int32_t)floorf((py + uy * t) * v.inv_voxel);
            const int32_t vz = (int32_t)floorf((pz + uz * t) * v.inv_voxel);
            const float cx = ((float)vx + 0.5f) * v.voxel;
            const float cy = ((float)vy + 0.5f) * v.voxel;
            const float cz = ((float)vz + 0.5f) * v.voxel;
            const float ex = cx - src.cam[0], ey = cy - src.cam[1], ez = cz - src.cam[2];
            const float sdf = dist - sqrtf(ex * ex + ey * ey + ez * ez);
            if (!(sdf < -v.trunc)) {
                const int64_t key = pack_coord(floor_div8(vx), floor_div8(vy), floor_div8(vz));
                if (key != lk.last) lk.push(key);
            }
        }
        if (__ballot(lk.valid >> (kKeysPerLane - 1)))
            flush_wave<MODE, TOUCH>(v, wg, lk, cas_acc, drop_acc, lane_acc, lookup_acc, wave_acc);
    }
    flush_wave<MODE, TOUCH>(v, wg, lk, cas_acc, drop_acc, lane_acc, lookup_acc, wave_acc);
    }

    // the reference's counting variant tallies attempts into drop_count
    // (cpp/src/volume_cuda.cu:59-61,81); keep a separate total as well
    if (MODE == kAllocCountCas) {
        wave_add_u64(cas_acc, &v.ctr->cas_attempts);
        wave_add_u64(cas_acc, &v.ctr->drop_count);
    }
    wave_add_u64(drop_acc, &v.ctr->drop_count);
    wave_add_u64(lookup_acc, &v.ctr->lookups);
    if (MODE == kAllocProbe) {
        wave_add_u64(lane_acc, &v.ctr->lane_steps);
        if (__lane_id() == 0 && wave_acc) atomicAdd(&v.ctr->wave_steps, wave_acc);
        __syncthreads();
        for (int i = threadIdx.x; i < kProbeHistBins; i += kAllocThreads)
            if (s_hist[i]) atomicAdd(&v.ctr->probe_hist[i], s_hist[i]);
    }
}

// Start of an allocate call: new stamp, empty per-call lists. A kernel rather
// than a memset so that a captured graph replays it correctly.
__global__ void k_begin_call(Counters* c) {
    c->stamp = c->stamp + 1;
    c->new_count = 0;
    c->touched_count = 0;
}

// ---- phase B: sort the call's new keys, assign pool indices in key order ----

__device__ inline uint32_t find_slot(const VolDev& v, int64_t key) {
    uint32_t s = home_slot(v, key);
    for (uint32_t p = 0; p < v.capacity; ++p) {
        if (v.table[s].key == key) return s;
        s = next_slot(v, s);
    }
    return 0xffffffffu;
}

// Merge two sorted runs of distinct keys into dst with all threads of the
// workgroup: each thread owns an equal slice of the output and finds where it
// starts by binary search on the merge path.
__device__ void merge_runs(const int64_t* a, int la, const int64_t* b, int lb, int64_t* dst) {
    const int m = la + lb;
    const int per = (m + blockDim.x - 1) / blockDim.x;
    const int k0 = min(m, (int)threadIdx.x * per);
    const int k1 = min(m, k0 + per);
    if (k0 >= k1) return;
    int lo = max(0, k0 - lb), hi = min(k0, la);
    while (lo < hi) {  // i = elements taken from a among the first k0 outputs
        const int i = (lo + hi) >> 1;
        if (a[i] < b[k0 - i - 1]) lo = i + 1; else hi = i;
    }
    int i = lo, j = k0 - lo;
    for (int k = k0; k < k1; ++k) {
        const bool take_a = (j >= lb) || (i < la && a[i] < b[j]);
        dst[k] = take_a ? a[i++] : b[j++];
    }
}

__global__ __launch_bounds__(1024) void k_assign(VolDev v, int32_t flag_stride) {
    __shared__ int64_t s_keys[kSortTile];
    __shared__ int s_arrival;
    const int n = v.ctr->new_count;
    const int tile = blockIdx.x;
    if (n == 0 || tile * kSortTile >= n) return;

    // 1. bitonic sort of this tile in LDS, padded to a power of two
    const int lo = tile * kSortTile;
    const int cnt = min(kSortTile, n - lo);
    int P = 2;
    while (P < cnt) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x)
        s_keys[i] = i < cnt ? v.new_keys[lo + i] : INT64_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const int64_t x = s_keys[i], y = s_keys[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { s_keys[i] = y; s_keys[ixj] = x; }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) v.sort_a[lo + i] = s_keys[i];

    // 2. merge tree: the second of two sibling runs to finish merges them, so
    //    no workgroup ever waits on another
    int64_t* src = v.sort_a;
    int64_t* dst = v.sort_b;
    int run = tile;
    for (int level = 0; (kSortTile << level) < n; ++level) {
        const int width = kSortTile << level;
        const int pair = run >> 1;
        const int a_lo = pair * 2 * width;
        const int a_hi = min(a_lo + width, n);
        const int b_hi = min(a_lo + 2 * width, n);
        if (a_hi >= b_hi) {
            for (int i = a_lo + (int)threadIdx.x; i < a_hi; i += blockDim.x) dst[i] = src[i];
        } else {
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0) {
                int* flag = &v.sort_flags[level * flag_stride + pair];
                const int old = atomicAdd(flag, 1);
                if (old == 1) *flag = 0;  // reset for the next call
                s_arrival = old;
            }
            __syncthreads();
            if (s_arrival == 0) return;  // sibling not done; it will merge
            __threadfence();
            merge_runs(src + a_lo, a_hi - a_lo, src + a_hi, b_hi - a_hi, dst + a_lo);
        }
        __syncthreads();
        int64_t* t = src; src = dst; dst = t;
        run = pair;
    }
    __threadfence();
    __syncthreads();

    // 3. this workgroup holds all n keys in order: assign pool indices
    const int bc0 = v.ctr->block_count;
    unsigned long long drops = 0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const int64_t key = src[r];
        const uint32_t slot = find_slot(v, key);
        if (slot == 0xffffffffu) continue;
        const int64_t idx = (int64_t)bc0 + r;
        if (idx < v.pool_capacity) {
            int32_t x, y, z;
            unpack_key(key, x, y, z);
            v.block_coord[idx * 3 + 0] = x;
            v.block_coord[idx * 3 + 1] = y;
            v.block_coord[idx * 3 + 2] = z;
            v.table[slot].block_idx = (int32_t)idx;
        } else {
            v.table[slot].block_idx = kIdxDropped;
            drops += 1;
        }
    }
    for (int o = 32; o > 0; o >>= 1) drops += __shfl_xor(drops, o);
    if ((threadIdx.x & 63) == 0 && drops) atomicAdd(&v.ctr->drop_count, drops);
    if (threadIdx.x == 0) v.ctr->block_count = (int32_t)min((int64_t)bc0 + n, (int64_t)v.pool_capacity);
}

// ---- reference-semantics scatter update (volume.cpp:200-299) ---------------

__global__ __launch_bounds__(256) void k_scatter_update(VolDev v, PointSrc src) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= src.n) return;
    const float px = src.pos[i * 3 + 0], py = src.pos[i * 3 + 1], pz = src.pos[i * 3 + 2];
    const float dx = px - src.cam[0], dy = py - src.cam[1], dz = pz - src.cam[2];
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (src.radius_sq > 0.0f && d2 > src.radius_sq) return;
    const float dist = sqrtf(d2);
    if (!(dist > 1e-6f)) return;
    const float ux = dx / dist, uy = dy / dist, uz = dz / dist;
    const float inv_trunc = 1.0f / v.trunc;
    int64_t last_key = kEmptyKey;
    int32_t last_bi = kIdxAbsent;
    for (int32_t s = -v.steps; s <= v.steps; ++s) {
        const float t = (float)s * v.voxel;
        const int32_t vx = (int32_t)floorf((px + ux * t) * v.inv_voxel);
        const int32_t vy = (int32_t)floorf((py + uy * t) * v.inv_voxel);
        const int32_t vz = (int32_t)floorf((pz + uz * t) * v.inv_voxel);
        const float cx = ((float)vx + 0.5f) * v.voxel;
        const float cy = ((float)vy + 0.5f) * v.voxel;
        const float cz = ((float)vz + 0.5f) * v.voxel;
        const float ex = cx - src.cam[0], ey = cy - src.cam[1], ez = cz - src.cam[2];
        const float sdf = dist - sqrtf(ex * ex + ey * ey + ez * ez);
        if (sdf < -v.trunc) continue;
        const int32_t bx = floor_div8(vx), by = floor_div8(vy), bz = floor_div8(vz);
        const int64_t key = pack_coord(bx, by, bz);
        if (key != last_key) { last_key = key; last_bi = find_index(v, key); }
        if (last_bi < 0) continue;
        const float sdf_n = fminf(fmaxf(sdf * inv_trunc, -1.0f), 1.0f);
        const size_t idx = (size_t)last_bi * kBlockVoxels +
                           (size_t)(((vz & 7) * 8 + (vy & 7)) * 8 + (vx & 7));
        if (v.weight_cap > 0.0f && v.weight[idx] >= v.weight_cap) continue;
        unsafeAtomicAdd(&v.weight[idx], 1.0f);
        unsafeAtomicAdd(&v.tsdf[idx], sdf_n);
        unsafeAtomicAdd(&v.r[idx], 128.0f);
        unsafeAtomicAdd(&v.g[idx], 128.0f);
        unsafeAtomicAdd(&v.b[idx], 128.0f);
    }
}

__global__ void k_fill_table(HashEntry* t, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) t[i] = HashEntry{kEmptyKey, kIdxPending, 0};
}

__global__ void k_zero_words(uint32_t* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t k = i; k < n; k += stride) p[k] = 0u;
}

}  // namespace

int32_t sort_flag_stride(uint32_t capacity) {
    return (int32_t)((capacity + kSortTile - 1) / kSortTile / 2 + 1);
}

hipError_t launch_begin_call(const VolDev& v, hipStream_t s) {
    k_begin_call<<<1, 1, 0, s>>>(v.ctr);
    return hipGetLastError();
}

template <int MODE, bool TOUCH>
static void launch_alloc_t(const VolDev& v, const PointSrc& src, const int32_t* coords, int set,
                           hipStream_t s) {
    if (coords) {
        dim3 grid((src.n + kAllocThreads - 1) / kAllocThreads);
        k_alloc<MODE, TOUCH, kSrcCoords><<<grid, kAllocThreads, 0, s>>>(v, src, coords, set);
    } else if (src.depth) {
        dim3 grid((src.W + 15) / 16, (src.H + 15) / 16);
        k_alloc<MODE, TOUCH, kSrcDepth><<<grid, kAllocThreads, 0, s>>>(v, src, nullptr, set);
    } else {
        dim3 grid((src.n + kAllocThreads - 1) / kAllocThreads);
        k_alloc<MODE, TOUCH, kSrcPoints><<<grid, kAllocThreads, 0, s>>>(v, src, nullptr, set);
    }
}

hipError_t launch_alloc(const VolDev& v, const PointSrc& src, const AllocOpts& o,
                        const int32_t* coords, hipStream_t s) {
    int set = o.lds_set;
    if (set < 0) set = 0;
    if (set > 2048) set = 2048;
    if (set & (set - 1)) return hipErrorInvalidValue;
    switch (o.mode) {
        case kAllocNormal:
            if (o.track_touched) launch_alloc_t<kAllocNormal, true>(v, src, coords, set, s);
            else launch_alloc_t<kAllocNormal, false>(v, src, coords, set, s);
            break;
        case kAllocCountCas: launch_alloc_t<kAllocCountCas, false>(v, src, coords, set, s); break;
        case kAllocProbe: launch_alloc_t<kAllocProbe, false>(v, src, coords, set, s); break;
        case kAllocNoCas: launch_alloc_t<kAllocNoCas, false>(v, src, coords, set, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Back-project every pixel (the same FP32 expressions as the depth allocate
// path) into 3 floats per pixel, NaN for pixels that are skipped.
__global__ void k_backproject(PointSrc src, float* out) {
    const int u = blockIdx.x * 16 + (threadIdx.x & 15);
    const int w = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (u >= src.W || w >= src.H) return;
    const float z = src.depth[(size_t)w * src.W + u];
    float* o = out + ((size_t)w * src.W + u) * 3;
    const bool ok = (z > 0.0f) && __builtin_isfinite(z) && !(src.max_range > 0.0f && z > src.max_range);
    if (!ok) {
        o[0] = o[1] = o[2] = __builtin_nanf("");
        return;
    }
    const float xc = ((float)u - src.cx) * z / src.fx;
    const float yc = ((float)w - src.cy) * z / src.fy;
    o[0] = src.R[0] * xc + src.R[1] * yc + src.R[2] * z + src.cam[0];
    o[1] = src.R[3] * xc + src.R[4] * yc + src.R[5] * z + src.cam[1];
    o[2] = src.R[6] * xc + src.R[7] * yc + src.R[8] * z + src.cam[2];
}

hipError_t launch_backproject(const PointSrc& src, float* out, hipStream_t s) {
    dim3 grid((src.W + 15) / 16, (src.H + 15) / 16);
    k_backproject<<<grid, 256, 0, s>>>(src, out);
    return hipGetLastError();
}

hipError_t launch_assign(const VolDev& v, hipStream_t s) {
    const uint32_t tiles = (v.capacity + kSortTile - 1) / kSortTile;
    k_assign<<<tiles, 1024, 0, s>>>(v, sort_flag_stride(v.capacity));
    return hipGetLastError();
}

hipError_t launch_scatter_update(const VolDev& v, const PointSrc& src, hipStream_t s) {
    k_scatter_update<<<(src.n + 255) / 256, 256, 0, s>>>(v, src);
    return hipGetLastError();
}

hipError_t launch_fill_table(HashEntry* table, uint32_t n, hipStream_t s) {
    k_fill_table<<<(n + 255) / 256, 256, 0, s>>>(table, n);
    return hipGetLastError();
}

hipError_t launch_zero_words(uint32_t* p, size_t n_words, hipStream_t s) {
    size_t blocks = (n_words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    k_zero_words<<<(unsigned)blocks, 256, 0, s>>>(p, n_words);
    return hipGetLastError();
}

}  // namespace osn_hip
