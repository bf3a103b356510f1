// Allocation for gfx950: band walk -> wave64 dedup -> workgroup LDS key set
// (one per work unit) -> table insert (phase A), then a deterministic sorted index assignment
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
// Appends to the call's touched/new lists are buffered in LDS and published
// with one global atomic per workgroup: per-wave atomics on one counter
// serialise in L2 and were measured to dominate the kernel.
//
// Compiled with -ffp-contract=off: the band walk must round exactly as the
// reference's x86-64 g++ build (SSE, no FMA) so that the allocated key set is
// bit-identical.

#include "internal.hpp"

namespace osn_hip {
namespace {

constexpr int kAllocThreads = 256;  // 4 waves; an 8x8-pixel quarter tile each in depth mode
constexpr int kKeysPerLane = 4;     // distinct keys a lane buffers before a flush
// The walk's registers bound its occupancy (the probes and atomics want many
// waves in flight): at 6 waves per SIMD it fits 80 VGPRs with a few spilled
// bytes; 7 spills more than it gains, 5 leaves latency exposed.
constexpr int kAllocWavesPerSimd = 6;
constexpr int kWgList = 512;        // LDS touched-list capacity per workgroup
constexpr int kWgNew = 256;         // LDS new-key-list capacity per workgroup
constexpr int kMaxSet = 1024;       // LDS key-set capacity
// Slots a lane probes on its own before an unresolved lookup is handed to the
// whole wave (64 slots per round).
constexpr uint32_t kLaneProbes = 4;

enum SrcKind : int { kSrcPoints = 0, kSrcDepth = 1, kSrcCoords = 2 };

__device__ inline int64_t readlane64(int64_t x, int lane) {
    const int lo = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(x))), lane);
    const int hi = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(x) >> 32)), lane);
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint64_t(uint32_t(lo)));
}

__device__ inline unsigned long long lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace
}  // namespace osn_hip

// Wave-wide reductions over the active lanes (DPP), from the device library
// every HIP program links; the HIP headers declare the 64-bit forms only on
// request.
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_add_u64(
    unsigned long long);
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_or_u64(
    unsigned long long);

namespace osn_hip {
namespace {

// Table and list accesses go through the global address space (global, not
// flat, instructions: a flat access also counts against the LDS wait counter,
// so every LDS wait of the key set would wait for outstanding table probes).
#define OSN_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline T OSN_GLOBAL* glob(T* p) {
    return (T OSN_GLOBAL*)p;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

#define OSN_CONST __attribute__((address_space(4)))

// Copy of a descriptor that stays constant during the kernel, read through the
// constant address space (scalar loads, uniform over the wave).
template <typename T>
__device__ inline T load_const(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "descriptor of whole dwords");
    T out;
    const uint32_t OSN_CONST* src = (const uint32_t OSN_CONST*)p;
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
    return out;
}

// Store a 16-byte list entry with one global store.
__device__ inline void put(TouchEntry* p, const TouchEntry& e) {
    u32x4 w;
    __builtin_memcpy(&w, &e, sizeof(w));
    *(u32x4 OSN_GLOBAL*)p = w;
}

__device__ inline int32_t add_global(int32_t* p, int32_t x) {
    return __hip_atomic_fetch_add(glob(p), x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-wide sum (every lane active).
__device__ inline unsigned long long wave_sum_u64(unsigned long long x) {
    return __ockl_wfred_add_u64(x);
}

struct WgShared {
    int64_t set[kMaxSet];
    int64_t dense[kAllocThreads];  // a unit's keys, compacted for the table step
    int32_t n_dense;
    TouchEntry new_keys[kWgNew];
    TouchEntry touched[kWgList];
    int32_t n_new, n_touched;
    int32_t base_new, base_touched;
    unsigned long long drops, cas, lookups, lane_steps, wave_steps;
    unsigned long long hist[kProbeHistBins];
};

// Append `mine` (when `want`) to an LDS list, one LDS atomic per wave; past
// the LDS capacity, straight to the global list.
template <int CAP>
__device__ inline void wave_append(bool want, TouchEntry mine, int32_t* lds_count,
                                   TouchEntry* lds_list, int32_t* g_count, TouchEntry* g_list) {
    const unsigned long long m = __ballot(want);
    if (m == 0) return;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (__lane_id() == leader) base = atomicAdd(lds_count, __popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);
    if (!want) return;
    const int i = base + __popcll(m & lanemask_lt());
    if (i < CAP) lds_list[i] = mine;
    else put(g_list + add_global(g_count, 1), mine);
}

// Publish the workgroup's buffered lists: one global atomic per list.
__device__ void flush_lists(const VolDev& v, ListCounters* lc, WgShared& s) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nt = min(s.n_touched, kWgList), nn = min(s.n_new, kWgNew);
        s.base_touched = nt ? add_global(&lc->touched_count, nt) : 0;
        s.base_new = nn ? add_global(&lc->new_count, nn) : 0;
    }
    __syncthreads();
    const int nt = min(s.n_touched, kWgList), nn = min(s.n_new, kWgNew);
    for (int i = threadIdx.x; i < nt; i += blockDim.x) put(v.touched + s.base_touched + i, s.touched[i]);
    for (int i = threadIdx.x; i < nn; i += blockDim.x) put(v.new_keys + s.base_new + i, s.new_keys[i]);
    __syncthreads();
    if (threadIdx.x == 0) { s.n_touched = 0; s.n_new = 0; }
    __syncthreads();
}

struct LaneKeys {
    int64_t k[kKeysPerLane];
    uint32_t valid = 0;
    int64_t last = kEmptyKey;

    // Buffer a key unless it repeats the previous one. A lane walks one ray,
    // and a ray meets each (convex) block in one contiguous run of samples, so
    // comparing with the previous key removes every in-lane repeat.
    __device__ inline void push(int64_t key) {
        if (key == last) return;
        last = key;
#pragma unroll
        for (int j = kKeysPerLane - 1; j > 0; --j) k[j] = k[j - 1];
        k[0] = key;
        valid = ((valid << 1) | 1u) & ((1u << kKeysPerLane) - 1u);
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

struct TileCtx {
    const VolDev* v;
    ListCounters* lc;
    WgShared* s;
    int32_t set_size;
    int32_t stamp;
};

// Slot of a key in the workgroup set: the low bits of the three block
// coordinates interleaved (z in bits 0-2, y in 3-5, x from 6 up), so the
// blocks of a unit's few-block neighbourhood land in distinct slots; shifts
// and exclusive-ors only.
__device__ inline uint32_t set_slot(int64_t key, int set_size) {
    const uint32_t lo = uint32_t(uint64_t(key)), hi = uint32_t(uint64_t(key) >> 32);
    return (lo ^ (lo >> 18) ^ (hi >> 4)) & uint32_t(set_size - 1);
}

// Lookups still unresolved after their lanes' own probes (the lanes of
// `pend`, each with its next slot and the slots inspected so far), resolved
// one at a time by the whole wave: a round reads the next 64 slots of the
// key's probe sequence, one per lane, and takes the first slot holding the
// key or, before it, the first empty one (a compare-and-swap by the lane that
// read it; a slot another key took meanwhile continues the scan after it).
// Compiled only into the kernel that inserts given block coordinates (the
// allocate-only stress path, whose tables run at high load factors); the
// walks' tables stay sparse and keep per-lane probing to the probe bound.
// Every lane of the wave calls it; lane l's result is for its own lookup when
// l is in `pend`.
struct CoopProbe {
    bool found, inserted;
    uint32_t slot;
    int32_t idx, stamp;
    uint32_t steps, n_cas, rounds;
};
template <int MODE>
__device__ inline CoopProbe coop_probe(HashEntry OSN_GLOBAL* table,
                                                          uint32_t capacity, uint32_t mask,
                                                          int32_t hash_kind, uint32_t probe_limit,
                                                          unsigned long long pend, int64_t key,
                                                          uint32_t slot, uint32_t steps) {
    const int lane = __lane_id();
    CoopProbe out{false, false, slot, kIdxPending, 0, steps, 0u, 0u};
    auto wrap = [&](uint32_t s) {
        return hash_kind == kHashTeschnerPow2 ? (s & mask) : (s >= capacity ? s - capacity : s);
    };
    uint32_t rounds = 0;
    while (pend) {
        const int l = __ffsll((long long)pend) - 1;
        pend &= pend - 1;
        const int64_t kk = readlane64(key, l);
        uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)slot, l);   // next slot to inspect
        uint32_t depth = (uint32_t)__builtin_amdgcn_readlane((int)steps, l);  // slots inspected
        bool f = false, ins = false;
        int32_t fidx = kIdxPending, fst = 0;
        uint32_t fslot = base, n_cas = 0;
        while (depth < probe_limit) {
            ++rounds;
            const uint32_t room = probe_limit - depth;  // <= capacity: a round never wraps twice
            const bool in = (uint32_t)lane < room;
            const uint32_t s = wrap(base + (uint32_t)lane);
            int64_t cur = kEmptyKey;
            int32_t ci = kIdxPending, cs = 0;
            if (in) {
                HashEntry OSN_GLOBAL* ent = table + s;
                cur = __hip_atomic_load(&ent->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ci = __hip_atomic_load(&ent->block_idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                cs = __hip_atomic_load(&ent->stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const unsigned long long mk = __ballot(in && cur == kk);
            const unsigned long long me = __ballot(in && cur == kEmptyKey);
            const int at_k = mk ? __ffsll((long long)mk) - 1 : 64;
            const int at_e = me ? __ffsll((long long)me) - 1 : 64;
            if (at_k < at_e) {  // the key, before any empty slot
                f = true;
                fslot = (uint32_t)__builtin_amdgcn_readlane((int)s, at_k);
                fidx = __builtin_amdgcn_readlane(ci, at_k);
                fst = __builtin_amdgcn_readlane(cs, at_k);
                depth += (uint32_t)at_k + 1;
                break;
            }
            if (at_e < 64) {  // the first empty slot of the sequence
                depth += (uint32_t)at_e + 1;
                fslot = (uint32_t)__builtin_amdgcn_readlane((int)s, at_e);
                if (MODE == kAllocNoCas) break;
                ++n_cas;
                int64_t prev = kEmptyKey;
                if (lane == at_e)
                    __hip_atomic_compare_exchange_strong(&table[fslot].key, &prev, kk,
                                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                prev = readlane64(prev, at_e);
                if (prev == kEmptyKey) { ins = true; break; }
                if (prev == kk) { f = true; break; }  // index pending, stamp 0, as a lane's own CAS
                base = wrap(fslot + 1);  // another key took the slot: go on after it
                continue;
            }
            const uint32_t adv = room < 64u ? room : 64u;
            depth += adv;
            base = wrap(base + adv);
        }
        if (lane == l) {
            out.found = f;
            out.inserted = ins;
            out.slot = fslot;
            out.idx = fidx;
            out.stamp = fst;
            out.steps = depth;
            out.n_cas = n_cas;
        }
    }
    out.rounds = rounds;
    return out;
}

// One batch of up to 64 distinct keys, one per lane (go = lane holds one),
// against the global table: find or insert, the touched and new lists, and
// the per-workgroup statistics.
template <int MODE, bool TOUCH, bool COOP = false>
__device__ __attribute__((always_inline)) void table_keys(const TileCtx& c, bool go, int64_t key) {
    const VolDev& v = *c.v;
    WgShared& s = *c.s;
    bool inserted = false, found = false;
    uint32_t slot = 0;
    uint32_t steps = 0;
    int32_t idx = kIdxPending;
    int32_t stamp = 0;  // the slot's stamp as read with the key
    unsigned long long cas = 0, drops = 0;
    // Each lane first walks its own key's probe sequence for a few slots; the
    // few lanes still unresolved after that (deep probes at high load
    // factors) are then resolved one at a time by the whole wave, 64
    // consecutive slots per round trip, so a wave no longer runs as many
    // rounds as its deepest lane has slots to inspect.
    bool stopped = false;  // a lane that reached an empty slot without inserting (no-CAS mode)
    // The first call after a reset finds no key of an earlier call in the
    // table, so a lane may try its compare-and-swap at once: the old key it
    // returns says inserted (empty), found (a key another workgroup of this
    // call inserted: its index is still pending) or occupied (probe on), one
    // round trip a slot instead of a read and then the exchange. Calls that
    // track touched blocks need the slot's index and stamp, and keep the read.
    const bool cas_first = MODE == kAllocNormal && !TOUCH && c.stamp == 1;
    if (go) {
        slot = home_slot(v, key);
        const uint32_t own =
            COOP && v.probe_limit > kLaneProbes ? kLaneProbes : v.probe_limit;
        for (uint32_t p = 0; p < own; ++p) {
            ++steps;
            if (cas_first) {
                int64_t prev = kEmptyKey;
                __hip_atomic_compare_exchange_strong(&glob(v.table)[slot].key, &prev, key,
                                                     __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                if (prev == kEmptyKey) { inserted = true; break; }
                if (prev == key) { found = true; break; }
                slot = next_slot(v, slot);
                continue;
            }
            // key, index and stamp of the slot in one round trip: no index
            // changes during this phase, and a key inserted during it still
            // holds kIdxPending and stamp 0 from the empty slot
            HashEntry OSN_GLOBAL* ent = glob(v.table) + slot;
            const int64_t cur =
                __hip_atomic_load(&ent->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int32_t ci =
                __hip_atomic_load(&ent->block_idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int32_t cs =
                __hip_atomic_load(&ent->stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == key) { found = true; idx = ci; stamp = cs; break; }
            if (cur == kEmptyKey) {
                if (MODE == kAllocNoCas) { stopped = true; break; }
                if (MODE == kAllocCountCas) cas += 1;
                int64_t prev = kEmptyKey;  // the slot's key after the exchange attempt
                __hip_atomic_compare_exchange_strong(&ent->key, &prev, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (prev == kEmptyKey) { inserted = true; break; }
                if (prev == key) { found = true; break; }
            }
            slot = next_slot(v, slot);
        }
    }
    uint32_t wave_rounds = 0;  // probe rounds of the whole wave (probe mode)
    if (MODE == kAllocProbe) {
        uint32_t mx = steps;
        for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
        wave_rounds = mx;
    }
    const unsigned long long pend =
        COOP ? __ballot(go && !found && !inserted && !stopped && steps < v.probe_limit) : 0ull;
    if (COOP && pend) {
        const CoopProbe r = coop_probe<MODE>(glob(v.table), v.capacity, v.mask, v.hash_kind,
                                             v.probe_limit, pend, key, slot, steps);
        if ((pend >> __lane_id()) & 1ull) {
            found = r.found;
            inserted = r.inserted;
            slot = r.slot;
            if (r.found) { idx = r.idx; stamp = r.stamp; }
            steps = r.steps;
            if (MODE == kAllocCountCas) cas += r.n_cas;
        }
        wave_rounds += r.rounds;
    }
    if (go) {
        if (MODE != kAllocNoCas && !found && !inserted) drops += 1;  // probe bound hit
        if (found && idx == kIdxDropped) drops += 1;
    }

    // per-workgroup statistics (rare or measurement-only), one wave-level add
    if (__ballot(drops != 0)) {
        const unsigned long long wd = wave_sum_u64(drops);
        if (__lane_id() == 0) atomicAdd(&s.drops, wd);
    }
    if (MODE == kAllocCountCas) {
        const unsigned long long wc = wave_sum_u64(cas);
        if (__lane_id() == 0 && wc) atomicAdd(&s.cas, wc);
    }
    if (MODE == kAllocProbe) {
        const unsigned long long ws = wave_sum_u64(steps);
        const unsigned long long wl = wave_sum_u64(go ? 1ull : 0ull);
        if (__lane_id() == 0) {
            atomicAdd(&s.lane_steps, ws);
            atomicAdd(&s.lookups, wl);
            // the wave's probe rounds: its own-lane rounds, then one per
            // cooperative scan
            atomicAdd(&s.wave_steps, 64ull * wave_rounds);
        }
        if (go) {
            const int bin = steps < (uint32_t)kProbeHistBins ? int(steps) - 1 : kProbeHistBins - 1;
            atomicAdd(&s.hist[bin], 1ull);
        }
    }

    // The first lane of the call to stamp the slot records the block as
    // touched; a stamp already read as this call's means another workgroup
    // did, and needs no exchange.
    bool touch = false;
    if (TOUCH && (found || inserted) && stamp != c.stamp) {
        const int32_t old = __hip_atomic_exchange(&glob(v.table)[slot].stamp, c.stamp,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        touch = (old != c.stamp);
    }
    if (TOUCH)
        wave_append<kWgList>(touch, TouchEntry{key, slot, idx}, &s.n_touched, s.touched,
                             &c.lc->touched_count, v.touched);
    if (MODE != kAllocNoCas)
        wave_append<kWgNew>(inserted, TouchEntry{key, slot, kIdxPending}, &s.n_new, s.new_keys,
                            &c.lc->new_count, v.new_keys);
}

// One batch of up to 64 distinct keys of a wave (has = lane holds one). With
// a workgroup set, a key is only placed in it here (one LDS compare-and-swap,
// a key the set already holds stops there); the set's keys reach the table
// together in flush_set at the end of the work unit, so each distinct key of
// the unit costs one table lookup and the waves' batches of a few keys cost
// no table round trip. A key the crowded set cannot take after 8 probes goes
// to the table at once. Without a set every batch goes to the table.
template <int MODE, bool TOUCH, bool COOP = false>
__device__ __attribute__((always_inline)) void process_keys(const TileCtx& c, bool has, int64_t key) {
    if (c.set_size == 0) {
        table_keys<MODE, TOUCH, COOP>(c, has, key);
        return;
    }
    WgShared& s = *c.s;
    bool spill = false;
    if (has) {
        uint32_t h = set_slot(key, c.set_size);
        spill = true;
        for (int p = 0; p < 8; ++p) {
            const unsigned long long prev =
                atomicCAS(reinterpret_cast<unsigned long long*>(&s.set[h]),
                          (unsigned long long)kEmptyKey, (unsigned long long)key);
            if (prev == (unsigned long long)kEmptyKey || prev == (unsigned long long)key) {
                spill = false;
                break;
            }
            h = (h + 1) & (c.set_size - 1);
        }
    }
    if (__ballot(spill)) table_keys<MODE, TOUCH, COOP>(c, spill, key);
}

// End of a work unit: every key the set holds goes to the table, and the set
// is left empty for the next unit. The set's keys are first compacted into a
// dense list (one LDS atomic per wave), so a unit's few dozen keys fill one
// wave's lanes at the table instead of a few lanes of every wave; a key past
// the list's capacity goes to the table at once. Every thread of the
// workgroup calls it.
template <int MODE, bool TOUCH, bool COOP = false>
__device__ void flush_set(const TileCtx& c) {
    if (c.set_size == 0) return;
    WgShared& s = *c.s;
    __syncthreads();
    for (int base = 0; base < c.set_size; base += blockDim.x) {
        const int i = base + (int)threadIdx.x;
        const int64_t key = i < c.set_size ? s.set[i] : kEmptyKey;
        const bool has = key != kEmptyKey;
        const unsigned long long m = __ballot(has);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            int at = 0;
            if (__lane_id() == leader) at = atomicAdd(&s.n_dense, __popcll(m));
            at = __builtin_amdgcn_readlane(at, leader) + __popcll(m & lanemask_lt());
            const bool spill = has && at >= kAllocThreads;
            if (has && !spill) s.dense[at] = key;
            if (__ballot(spill)) table_keys<MODE, TOUCH, COOP>(c, spill, key);
        }
        if (i < c.set_size) s.set[i] = kEmptyKey;
    }
    __syncthreads();
    const int n = min(s.n_dense, kAllocThreads);
    for (int b0 = (int)(threadIdx.x & ~63); b0 < n; b0 += blockDim.x) {
        const int i = b0 + __lane_id();
        const bool has = i < n;
        table_keys<MODE, TOUCH, COOP>(c, has, has ? s.dense[i] : kEmptyKey);
    }
    __syncthreads();
    if (threadIdx.x == 0) s.n_dense = 0;
}

// Drain every lane's buffered keys: repeatedly take the first live key of the
// first live lane, strike it from every lane, and hand it to lane `nd`. Each
// distinct key of the wave is therefore processed once, 64 at a time.
//
// Most keys of a wave lie in a few neighbouring blocks, so a window pass goes
// first: every lane marks its keys that fall in the 4x4x4 blocks starting one
// block below the first live key, the 64-bit marks are OR-ed across the wave,
// and lane i takes window cell i. Only keys outside the window reach the
// one-key-per-iteration loop.

template <int MODE, bool TOUCH>
__device__ void flush_wave(const TileCtx& c, LaneKeys& lk) {
    const unsigned long long live = __ballot(lk.valid != 0);
    if (live == 0) return;
    {
        const int leader = __ffsll((long long)live) - 1;
        const uint64_t k0 = (uint64_t)readlane64(lk.first(), leader);
        constexpr uint64_t kM = (uint64_t(1) << kCoordBits) - 1;
        // biased coordinates of the window origin (wraps harmlessly at 0)
        const uint32_t ox = uint32_t(k0 >> (2 * kCoordBits)) - 1u;
        const uint32_t oy = uint32_t((k0 >> kCoordBits) & kM) - 1u;
        const uint32_t oz = uint32_t(k0 & kM) - 1u;
        unsigned long long cells = 0;
#pragma unroll
        for (int j = 0; j < kKeysPerLane; ++j) {
            if (!((lk.valid >> j) & 1u)) continue;
            const uint64_t key = (uint64_t)lk.k[j];
            const uint32_t dx = uint32_t(key >> (2 * kCoordBits)) - ox;
            const uint32_t dy = uint32_t((key >> kCoordBits) & kM) - oy;
            const uint32_t dz = uint32_t(key & kM) - oz;
            if ((dx | dy | dz) < 4u) {
                cells |= 1ull << (dx | (dy << 2) | (dz << 4));
                lk.valid &= ~(1u << j);
            }
        }
        cells = __ockl_wfred_or_u64(cells);
        const int lane = __lane_id();
        const uint64_t mine = (uint64_t(ox + (lane & 3)) << (2 * kCoordBits)) |
                              (uint64_t(oy + ((lane >> 2) & 3)) << kCoordBits) |
                              uint64_t(oz + (lane >> 4));
        process_keys<MODE, TOUCH>(c, (cells >> lane) & 1ull, (int64_t)mine);
    }
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
            process_keys<MODE, TOUCH>(c, true, mine);
            nd = 0;
            mine = kEmptyKey;
        }
    }
    if (nd > 0) process_keys<MODE, TOUCH>(c, __lane_id() < nd, mine);
}

// The wave's window cells and its lanes' buffered keys in as few
// process_keys calls as possible (usually one): lane i takes window cell i
// when it is marked, and each distinct buffered key (a key outside the window)
// goes to a lane whose cell is empty. A window holds a few blocks of a wave's
// 64 neighbouring pixels, so empty lanes abound; a second call follows only
// when they run out. Buffered keys are deduplicated across lanes first-live-
// key first, as in flush_wave, and lie outside the window, so no key reaches
// two lanes of one call.
template <int MODE, bool TOUCH>
__device__ void process_window_and_buffers(const TileCtx& c, unsigned long long cells, int32_t ox,
                                           int32_t oy, int32_t oz, LaneKeys& lk) {
    const int lane = __lane_id();
    bool has = (cells >> lane) & 1ull;
    int64_t key = pack_coord(ox + (lane & 3), oy + ((lane >> 2) & 3), oz + (lane >> 4));
    unsigned long long free_lanes = ~cells;  // wave-uniform
    for (;;) {
        const unsigned long long m = __ballot(lk.valid != 0);
        if (m == 0) break;
        if (free_lanes == 0) {
            process_keys<MODE, TOUCH>(c, has, key);
            has = false;
            free_lanes = ~0ull;
        }
        const int leader = __ffsll((long long)m) - 1;
        const int64_t k = readlane64(lk.first(), leader);
        lk.remove(k);
        const int dst = __ffsll((long long)free_lanes) - 1;
        free_lanes &= free_lanes - 1;
        if (lane == dst) {
            has = true;
            key = k;
        }
    }
    process_keys<MODE, TOUCH>(c, has, key);
}

// Band walk (volume.cpp:158-191) state of one point.
struct Ray {
    float px, py, pz;  // the point
    float ux, uy, uz;  // unit direction from the camera
    float dist;        // camera-to-point distance
    float thr;         // samples with B < thr cannot be culled (see make_ray)
};

// Sets the ray of point p up; false when the reference skips the point
// (radius cull, dist <= 1e-6).
__device__ inline bool make_ray(const VolDev& v, const PointSrc& src, float px, float py, float pz,
                                Ray& r) {
    const float dx = px - src.cam[0], dy = py - src.cam[1], dz = pz - src.cam[2];
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float dist = sqrtf(d2);
    r.px = px; r.py = py; r.pz = pz;
    r.dist = dist;
    r.ux = dx / dist; r.uy = dy / dist; r.uz = dz / dist;
    // Cull shortcut. The voxel centre chosen for a sample at signed offset t
    // along the ray lies within half a voxel diagonal of the sample, so in
    // exact arithmetic sdf >= -(B + 0.8661 voxel) with B = t when the sample
    // is on the camera's side of the point's far end (dist + t >= 0) and |t|
    // otherwise. Every FP32 rounding on the way is below 2^-23 of the
    // magnitudes involved, a few dozen times over; `slack` bounds them all with
    // a 64 * 2^-23 relative margin. A sample with B < thr cannot be culled, so
    // its square root is skipped; the others evaluate the reference's
    // expression as is.
    const float slack = 7.63e-6f * (fabsf(px) + fabsf(py) + fabsf(pz) + fabsf(src.cam[0]) +
                                    fabsf(src.cam[1]) + fabsf(src.cam[2]) + dist +
                                    (float)v.steps * v.voxel);
    r.thr = (v.trunc - 0.8661f * v.voxel) - slack;
    return !(src.radius_sq > 0.0f && d2 > src.radius_sq) && (dist > 1e-6f);
}

// Voxel coordinate of sample s along one axis: floor((p + u * t) * (1 / voxel))
// with t = s * voxel, every operation rounded on its own.
__device__ inline int32_t sample_voxel(float p, float u, float t, float inv_voxel) {
    return (int32_t)floorf((p + u * t) * inv_voxel);
}

// B of sample s (see make_ray): t = s * voxel with voxel > 0, so B = t for s >= 0.
__device__ inline float sample_b(const Ray& r, int s, float t) {
    return s >= 0 ? t : (r.dist + t >= 0.0f ? t : -t);
}

// The reference's cull of a sample in voxel (vx, vy, vz) whose B >= thr.
__device__ inline bool sample_culled(const VolDev& v, const PointSrc& src, const Ray& r, int32_t vx,
                                     int32_t vy, int32_t vz) {
    const float cx = ((float)vx + 0.5f) * v.voxel;
    const float cy = ((float)vy + 0.5f) * v.voxel;
    const float cz = ((float)vz + 0.5f) * v.voxel;
    const float ex = cx - src.cam[0], ey = cy - src.cam[1], ez = cz - src.cam[2];
    const float sdf = r.dist - sqrtf(ex * ex + ey * ey + ez * ez);
    return sdf < -v.trunc;
}

// Block (bx, by, bz) goes to the wave's 4x4x4-block window (cells) or, when
// outside it, to the lane buffer.
__device__ inline void mark_block(int32_t bx, int32_t by, int32_t bz, int32_t ox, int32_t oy,
                                  int32_t oz, unsigned long long& cells, LaneKeys& lk) {
    const uint32_t dx = uint32_t(bx - ox), dy = uint32_t(by - oy), dz = uint32_t(bz - oz);
    if ((dx | dy | dz) < 4u) cells |= 1ull << (dx | (dy << 2) | (dz << 4));
    else lk.push(pack_coord(bx, by, bz));
}

// Sample s of a live lane's band walk: the voxel (p + u * t) / voxel, unless
// culled, marks its block.
__device__ inline void mark_sample(const VolDev& v, const PointSrc& src, const Ray& r, int s,
                                   float t, int32_t ox, int32_t oy, int32_t oz,
                                   unsigned long long& cells, LaneKeys& lk) {
    const int32_t vx = sample_voxel(r.px, r.ux, t, v.inv_voxel);
    const int32_t vy = sample_voxel(r.py, r.uy, t, v.inv_voxel);
    const int32_t vz = sample_voxel(r.pz, r.uz, t, v.inv_voxel);
    const bool culled = sample_b(r, s, t) >= r.thr && sample_culled(v, src, r, vx, vy, vz);
    if (!culled) mark_block(floor_div8(vx), floor_div8(vy), floor_div8(vz), ox, oy, oz, cells, lk);
}

// Sample S's cull, decided without the square root where that is safe. The
// reference culls when RN(dist - RN(sqrt(e2))) < -trunc, e2 the FP32 squared
// distance of the voxel centre. With T = dist + trunc, e2 below T^2 (1 -
// 2^-19) means the square root is at least 2^-21 T short of T, far more than
// the few roundings of the sequence can move, so the sample is kept; above
// T^2 (1 + 2^-19) it is culled. Only lanes between the two (about 0.2 % of
// rays, checked on CPU in exact FP32 emulation over six scales) evaluate the
// reference expression.
__device__ inline bool last_sample_culled(const VolDev& v, const PointSrc& src, const Ray& r,
                                          int32_t vx, int32_t vy, int32_t vz) {
    const float cx = ((float)vx + 0.5f) * v.voxel;
    const float cy = ((float)vy + 0.5f) * v.voxel;
    const float cz = ((float)vz + 0.5f) * v.voxel;
    const float ex = cx - src.cam[0], ey = cy - src.cam[1], ez = cz - src.cam[2];
    const float e2 = ex * ex + ey * ey + ez * ez;
    const float T = r.dist + v.trunc;
    const float TT = T * T;
    const bool keep = e2 < TT * (1.0f - 0x1p-19f);
    const bool cull = e2 > TT * (1.0f + 0x1p-19f);
    if (__ballot(!keep && !cull)) {
        if (!keep && !cull) return r.dist - sqrtf(e2) < -v.trunc;
    }
    return cull;
}

// The blocks of samples -S..S of a live lane, exactly as marking every sample
// would find them, from a few samples. Each voxel coordinate of sample s is a
// monotone function of s (t = s * voxel grows with s, and every FP32
// operation of sample_voxel is monotone), so each block coordinate is too.
// In the usual case only sample S can be culled (no other sample reaches
// B >= thr) and the blocks of samples -S and S differ by at most one along
// every axis. Then each axis changes block once, at a sample c found by
// bisection over -S..S, and the blocks of the kept range -S..b (b = S, or
// S - 1 when sample S is culled) are those at -S and at each c <= b: at most
// 4, so the lane buffer cannot overflow. Returns false, marking nothing, in
// any other case; the caller then marks every sample.
template <int S>
__device__ inline bool mark_band(const VolDev& v, const PointSrc& src, const Ray& r, int32_t ox,
                                 int32_t oy, int32_t oz, unsigned long long& cells,
                                 LaneKeys& lk) {
    // a sample other than S with B >= thr: samples s >= 0 have B = t_s, at
    // most t_{S-1}; samples s < 0 on the camera's side (dist + t_s >= 0, true
    // for all of them when it holds for t_{-S}) have B = t_s < 0, below a
    // positive thr; otherwise every sample is checked
    const float ta = (float)(-S) * v.voxel, te = (float)S * v.voxel;
    bool other_b;
    if (r.dist + ta >= 0.0f && r.thr > 0.0f) {
        other_b = (float)(S - 1) * v.voxel >= r.thr;
    } else {
        other_b = false;
#pragma unroll
        for (int s = -S; s < S; ++s) other_b |= sample_b(r, s, (float)s * v.voxel) >= r.thr;
    }
    const int32_t a[3] = {sample_voxel(r.px, r.ux, ta, v.inv_voxel),
                          sample_voxel(r.py, r.uy, ta, v.inv_voxel),
                          sample_voxel(r.pz, r.uz, ta, v.inv_voxel)};
    const int32_t e[3] = {sample_voxel(r.px, r.ux, te, v.inv_voxel),
                          sample_voxel(r.py, r.uy, te, v.inv_voxel),
                          sample_voxel(r.pz, r.uz, te, v.inv_voxel)};
    int32_t ka[3], kb[3];
    bool simple = !other_b;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ka[q] = floor_div8(a[q]);
        kb[q] = floor_div8(e[q]);
        simple &= (uint32_t)(kb[q] - ka[q] + 1) <= 2u;
    }
    if (!simple) return false;
    const float p[3] = {r.px, r.py, r.pz}, u[3] = {r.ux, r.uy, r.uz};
    constexpr int kHalvings = ceil_log2(2 * S);  // the range -S..S has 2S steps
    int c[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        c[q] = S + 1;  // no change along q
        if (kb[q] != ka[q]) {
            int lo = -S, hi = S;  // block ka at lo, kb at hi
#pragma unroll
            for (int i = 0; i < kHalvings; ++i) {
                if (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    const int32_t km = floor_div8(sample_voxel(p[q], u[q], (float)mid * v.voxel,
                                                               v.inv_voxel));
                    if (km == kb[q]) hi = mid;
                    else lo = mid;
                }
            }
            c[q] = hi;
        }
    }
    const bool cull_last = sample_b(r, S, te) >= r.thr && last_sample_culled(v, src, r, e[0], e[1], e[2]);
    const int b = cull_last ? S - 1 : S;  // the kept range ends at b
    mark_block(ka[0], ka[1], ka[2], ox, oy, oz, cells, lk);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        bool fresh = c[q] <= b;
        for (int j = 0; j < q; ++j) fresh &= c[j] != c[q];
        if (fresh)
            mark_block(c[q] >= c[0] ? kb[0] : ka[0], c[q] >= c[1] ? kb[1] : ka[1],
                       c[q] >= c[2] ? kb[2] : ka[2], ox, oy, oz, cells, lk);
    }
    return true;
}

// Walks one point per lane (live lanes only) and hands every distinct block
// of the wave to process_keys once. Keys inside a 4x4x4-block window anchored
// one block below the first live lane's point are marked in a 64-bit mask
// (lane i later takes cell i), with no buffering; only keys outside it go
// through the lane buffer. STEPS is the band half-width in samples when known
// at compile time (0: read it from the volume).
template <int MODE, bool TOUCH, int STEPS>
__device__ void walk_points(const TileCtx& c, const PointSrc& src, bool ok, float px, float py,
                            float pz) {
    const VolDev& v = *c.v;
    Ray r{};
    if (ok) ok = make_ray(v, src, px, py, pz, r);
    const unsigned long long live = __ballot(ok);
    if (live == 0) return;
    const int leader = __ffsll((long long)live) - 1;
    const int32_t ox = __builtin_amdgcn_readlane(floor_div8((int32_t)floorf(px * v.inv_voxel)), leader) - 1;
    const int32_t oy = __builtin_amdgcn_readlane(floor_div8((int32_t)floorf(py * v.inv_voxel)), leader) - 1;
    const int32_t oz = __builtin_amdgcn_readlane(floor_div8((int32_t)floorf(pz * v.inv_voxel)), leader) - 1;
    unsigned long long cells = 0;
    LaneKeys lk;
    // The shortcut marks at most 4 blocks per lane (kKeysPerLane); when a
    // lane cannot take it, the whole wave walks sample by sample, draining
    // the lane buffers whenever one fills.
    static_assert(kKeysPerLane >= 4, "lane buffer bound");
    bool walked = false;
    if (STEPS > 0 && ok) walked = mark_band<STEPS>(v, src, r, ox, oy, oz, cells, lk);
    if (STEPS == 0 || __ballot(ok && !walked)) {
        for (int32_t s = -v.steps; s <= v.steps; ++s) {
            if (ok && !walked) mark_sample(v, src, r, s, (float)s * v.voxel, ox, oy, oz, cells, lk);
            if (__ballot(lk.valid >> (kKeysPerLane - 1))) flush_wave<MODE, TOUCH>(c, lk);
        }
    }
    cells = __ockl_wfred_or_u64(cells);
    process_window_and_buffers<MODE, TOUCH>(c, cells, ox, oy, oz, lk);
}

template <int MODE, bool TOUCH>
__device__ inline void walk_points_any(const TileCtx& c, const PointSrc& src, bool ok, float px,
                                       float py, float pz) {
    if (c.v->steps == 4) walk_points<MODE, TOUCH, 4>(c, src, ok, px, py, pz);
    else walk_points<MODE, TOUCH, 0>(c, src, ok, px, py, pz);
}

// Depth of pixel (u, w), 0 (invalid) outside the image.
__device__ inline float load_depth(const PointSrc& src, int u, int w) {
    return (u < src.W && w < src.H) ? glob(src.depth)[(size_t)w * src.W + u] : 0.0f;
}

// A depth pixel is back-projected in FP32 as
//   xc = (u - cx) * z / fx, yc = (v - cy) * z / fy, p = R * (xc, yc, z) + cam
// evaluated left to right without fused multiply-add; false for a pixel the
// walk skips (z <= 0, not finite, or beyond max_range when that is set).
__device__ inline bool backproject_pixel(const PointSrc& src, int u, int w, float z, float& px,
                                         float& py, float& pz) {
    const bool ok =
        (z > 0.0f) && __builtin_isfinite(z) && !(src.max_range > 0.0f && z > src.max_range);
    if (ok) {
        const float xc = ((float)u - src.cx) * z / src.fx;
        const float yc = ((float)w - src.cy) * z / src.fy;
        px = src.R[0] * xc + src.R[1] * yc + src.R[2] * z + src.cam[0];
        py = src.R[3] * xc + src.R[4] * yc + src.R[5] * z + src.cam[1];
        pz = src.R[6] * xc + src.R[7] * yc + src.R[8] * z + src.cam[2];
    }
    return ok;
}

// Depth images are walked in work units of sub x sub tiles of 16x16 pixels
// (sub = 1 for a single frame, whose 16x16 tiles already fill the chip; 2 in
// batches): the workgroup collects the unit's keys in its set, each of
// its four waves takes an 8x8 quarter of a tile, and the depth of the next
// tile is loaded before the current one is walked. With src.tiled set, each
// pixel is also stored to the update's 8x4-tiled copy: a wave's 8x8 quarter is
// two whole tiles, 256 contiguous bytes. The units cover the image rounded up
// to 16-pixel multiples, so every pixel of the copy (padding included, as 0)
// is written.
template <int MODE, bool TOUCH>
__device__ void alloc_depth_unit(const TileCtx& c, const PointSrc& src, int unit, int sub) {
    const int span = 16 * sub;
    const int per_row = (src.W + span - 1) / span;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int u0 = (unit % per_row) * span + (wave & 1) * 8 + (lane & 7);
    const int w0 = (unit / per_row) * span + (wave >> 1) * 8 + (lane >> 3);
    const int n = sub * sub;
    const int Wt = tiled_width(src.W), Ht = tiled_height(src.H);
    const uint32_t tk = (uint32_t)tiled_stride_k(src.W);
    float z = load_depth(src, u0, w0);
#pragma unroll 1
    for (int j = 0; j < n; ++j) {
        const int u = u0 + (j % sub) * 16, w = w0 + (j / sub) * 16;
        const int jn = j + 1;
        const float zn = jn < n ? load_depth(src, u0 + (jn % sub) * 16, w0 + (jn / sub) * 16) : 0.0f;
        if (src.tiled && u < Wt && w < Ht) glob(src.tiled)[tiled_index(u, w, tk)] = z;
        float px = 0.f, py = 0.f, pz = 0.f;
        const bool ok = backproject_pixel(src, u, w, z, px, py, pz);
        walk_points_any<MODE, TOUCH>(c, src, ok, px, py, pz);
        z = zn;
    }
}

// One tile of 256 points (points / coords), one per thread.
template <int MODE, bool TOUCH, int SRC>
__device__ void alloc_tile(const TileCtx& c, const PointSrc& src, const int32_t* coords, int tile) {
    const int i = tile * kAllocThreads + threadIdx.x;
    const bool has = i < src.n;
    if (SRC == kSrcCoords) {
        // Given keys have no spatial coherence to exploit within a wave, so
        // each lane goes straight to the workgroup key set with its key.
        const int64_t key =
            has ? pack_coord(glob(coords)[i * 3 + 0], glob(coords)[i * 3 + 1], glob(coords)[i * 3 + 2])
                : kEmptyKey;
        process_keys<MODE, TOUCH, true>(c, has, key);
        return;
    }
    float px = 0.f, py = 0.f, pz = 0.f;
    if (has) {
        px = glob(src.pos)[i * 3 + 0];
        py = glob(src.pos)[i * 3 + 1];
        pz = glob(src.pos)[i * 3 + 2];
    }
    walk_points_any<MODE, TOUCH>(c, src, has, px, py, pz);
}

__device__ void wg_init(WgShared& s, int set_size) {
    for (int i = threadIdx.x; i < set_size; i += blockDim.x) s.set[i] = kEmptyKey;
    for (int i = threadIdx.x; i < kProbeHistBins; i += blockDim.x) s.hist[i] = 0;
    if (threadIdx.x == 0) {
        s.n_new = s.n_touched = 0;
        s.n_dense = 0;
        s.drops = s.cas = s.lookups = s.lane_steps = s.wave_steps = 0;
    }
}

template <int MODE>
__device__ void flush_stats(const VolDev& v, WgShared& s) {
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s.drops) atomicAdd(&v.ctr->drop_count, s.drops);
        if (MODE == kAllocCountCas && s.cas) {
            // the reference's counting variant tallies attempts into
            // drop_count (cpp/src/volume_cuda.cu:59-61,81); keep a total too
            atomicAdd(&v.ctr->cas_attempts, s.cas);
            atomicAdd(&v.ctr->drop_count, s.cas);
        }
        if (MODE == kAllocProbe) {
            atomicAdd(&v.ctr->lane_steps, s.lane_steps);
            atomicAdd(&v.ctr->wave_steps, s.wave_steps);
            atomicAdd(&v.ctr->lookups, s.lookups);
        }
        s.drops = s.cas = s.lookups = s.lane_steps = s.wave_steps = 0;
    }
    if (MODE == kAllocProbe)
        for (int i = threadIdx.x; i < kProbeHistBins; i += blockDim.x) {
            if (s.hist[i]) atomicAdd(&v.ctr->probe_hist[i], s.hist[i]);
            s.hist[i] = 0;
        }
    __syncthreads();
}

__device__ inline void reset_other_parity(Counters* ctr, int parity) {
    ctr->lists[parity ^ 1].new_count = 0;
    ctr->lists[parity ^ 1].touched_count = 0;
    ctr->lists[parity ^ 1].assign_ticket = 0;
}

template <int MODE, bool TOUCH, int SRC>
__global__ __launch_bounds__(kAllocThreads, kAllocWavesPerSimd) void k_alloc(VolDev v, PointSrc src,
                                                         const int32_t* coords, int32_t set_size,
                                                         CallInfo call, int32_t n_tiles) {
    __shared__ WgShared s;
    wg_init(s, set_size);
    if (blockIdx.x == 0 && threadIdx.x == 0) reset_other_parity(v.ctr, call.parity);
    __syncthreads();
    const TileCtx c{&v, &v.ctr->lists[call.parity], &s, set_size, call.stamp};
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        if (SRC == kSrcDepth) alloc_depth_unit<MODE, TOUCH>(c, src, tile, 1);
        else alloc_tile<MODE, TOUCH, SRC>(c, src, coords, tile);
        flush_set<MODE, TOUCH, SRC == kSrcCoords>(c);
    }
    flush_lists(v, c.lc, s);
    flush_stats<MODE>(v, s);
}

// Many scenes in one launch: the depth work units of all active scenes form
// one work list (scene s owns [tile_prefix[s], tile_prefix[s + 1]), read once
// into LDS); each scene keeps its own table and lists, flushed when the
// workgroup moves on to another scene.
template <int MODE>
__global__ __launch_bounds__(kAllocThreads, kAllocWavesPerSimd) void k_alloc_batch(BatchDev b, int32_t set_size,
                                                               int32_t total_tiles) {
    __shared__ WgShared s;
    __shared__ int32_t s_prefix[kMaxBatchScenes + 1];
    wg_init(s, set_size);
    for (int i = threadIdx.x; i <= b.n_scenes; i += blockDim.x) s_prefix[i] = b.tile_prefix[i];
    __syncthreads();
    const int k = b.step_ctr ? *b.step_ctr : b.step;
    int cur = -1;
    int sc = 0;
    CallInfo call{0, 0};
    VolDev v{};
    PointSrc src{};
    for (int tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        while (sc + 1 < b.n_scenes && s_prefix[sc + 1] <= tile) ++sc;  // tiles ascend
        const SceneStep* st = b.steps + k * b.n_scenes + sc;
        const int local = tile - s_prefix[sc];
        if (local == 0 && threadIdx.x == 0)
            reset_other_parity(b.vols[sc].ctr, batch_call(b, sc, k).parity);
        if (!load_const(&st->active) || local >= load_const(&st->tiles)) continue;
        if (sc != cur) {
            if (cur >= 0) {
                flush_lists(v, &v.ctr->lists[call.parity], s);
                flush_stats<MODE>(v, s);
            }
            cur = sc;
            v = load_const(b.vols + sc);
            src = load_const(&st->src);
            call = batch_call(b, sc, k);
        }
        const TileCtx c{&v, &v.ctr->lists[call.parity], &s, set_size, call.stamp};
        alloc_depth_unit<MODE, true>(c, src, local, b.depth_sub);
        flush_set<MODE, true>(c);
    }
    if (cur >= 0) {
        flush_lists(v, &v.ctr->lists[call.parity], s);
        flush_stats<MODE>(v, s);
    }
}

// ---- phase B: sort the call's new keys, assign pool indices in key order ----
//
// The call's n new keys (each with the slot it was inserted at) form
// T = ceil(n / kSortTile) tiles. Work item (0, c) sorts tile c in LDS; item
// (l, c) with l >= 1 produces output chunk c (positions c * kSortTile ...) of
// level l by merging the two level-(l - 1) runs that cover it. Items of the
// top level L = ceil(log2 T) do not store their run: output position p
// becomes pool index bc0 + p, written to the key's table slot and to
// block_coord. Level l writes buffer l & 1 (sort_a, sort_b) and reads the
// other one.
//
// Workgroups take items in ticket order (level-major) from a per-call counter.
// An item waits only for items of the level below, whose tickets are smaller
// and so were claimed by workgroups already running; the lowest unfinished
// ticket can always proceed, hence every wait ends. Each wait is bounded all
// the same: past the bound the call records a fault, and every other waiter
// sees it and leaves, so the kernel always drains.

constexpr uint32_t kSpinLimit = 1u << 22;   // about 0.5 s of sleeps

struct AssignLds {
    int64_t key[kSortTile];
    uint32_t slot[kSortTile];
    int64_t key2[kSortTile];   // the sort's merge levels alternate between the two
    uint32_t slot2[kSortTile];
    int32_t ticket;
    int32_t flag;
    int32_t corank[2];
    unsigned long long drops;
    uint32_t box_min[3], box_max[3];  // the tile's key fields' ranges (bitmap ranking)
};

constexpr int kRankBitsMax = 16;  // bitmap ranking over boxes of at most 2^16 cells

// Rank a tile's keys through a bitmap when they span a small box of block
// coordinates: keys of one call are distinct, and a key's three 21-bit fields
// order it as the integer does, so with each field offset by the box's low
// corner and the three concatenated (x high, z low) into a cell index of at
// most kRankBitsMax bits, a key's rank is the number of occupied cells below
// its own: one bit set per key, a prefix count over the bitmap's words, no
// comparisons. The bitmap and its word prefixes reuse key2. False, with
// nothing moved, when the box is larger (the caller then sorts). Every
// thread of a 1024-thread workgroup calls it; cnt <= kSortTile.
__device__ bool bitmap_rank(AssignLds& s, int cnt) {
    static_assert(kSortTile <= 4 * 1024, "at most 4 keys per thread");
    static_assert((2 << kRankBitsMax) / 32 * 4 <= (int)sizeof(s.key2), "bitmap and prefixes fit");
    constexpr uint64_t kM = (uint64_t(1) << kCoordBits) - 1;
    auto field = [&](uint64_t k, int q) {
        return uint32_t((k >> (kCoordBits * (2 - q))) & kM);
    };
    if (threadIdx.x < 3) {
        s.box_min[threadIdx.x] = ~0u;
        s.box_max[threadIdx.x] = 0u;
    }
    __syncthreads();
    uint32_t mn[3] = {~0u, ~0u, ~0u}, mx[3] = {0u, 0u, 0u};
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const uint64_t k = (uint64_t)s.key[i];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            mn[q] = min(mn[q], field(k, q));
            mx[q] = max(mx[q], field(k, q));
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[q] = min(mn[q], (uint32_t)__shfl_xor((int)mn[q], o));
            mx[q] = max(mx[q], (uint32_t)__shfl_xor((int)mx[q], o));
        }
        if (__lane_id() == 0) {
            atomicMin(&s.box_min[q], mn[q]);
            atomicMax(&s.box_max[q], mx[q]);
        }
    }
    __syncthreads();
    uint32_t lo[3], bits[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        lo[q] = s.box_min[q];
        const uint32_t r = s.box_max[q] - lo[q];
        bits[q] = r ? 32u - (uint32_t)__clz(r) : 0u;
    }
    const uint32_t total = bits[0] + bits[1] + bits[2];
    if (total > (uint32_t)kRankBitsMax) return false;  // the same answer in every thread
    uint32_t* bm = reinterpret_cast<uint32_t*>(s.key2);
    uint32_t* pre = bm + (1 << kRankBitsMax) / 32;
    const int n_words = ((1 << total) + 31) / 32;
    for (int w = threadIdx.x; w < n_words; w += blockDim.x) bm[w] = 0u;
    __syncthreads();
    auto cell = [&](uint64_t k) {
        return ((field(k, 0) - lo[0]) << (bits[1] + bits[2])) |
               ((field(k, 1) - lo[1]) << bits[2]) | (field(k, 2) - lo[2]);
    };
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const uint32_t c = cell((uint64_t)s.key[i]);
        atomicOr(&bm[c >> 5], 1u << (c & 31u));
    }
    __syncthreads();
    // exclusive prefix of the words' popcounts: a thread takes words 2t, 2t + 1
    const int w0 = 2 * (int)threadIdx.x;
    const uint32_t p0 = w0 < n_words ? (uint32_t)__popc(bm[w0]) : 0u;
    const uint32_t p1 = w0 + 1 < n_words ? (uint32_t)__popc(bm[w0 + 1]) : 0u;
    uint32_t incl = p0 + p1;
    const int lane = __lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= o) incl += t;
    }
    uint32_t* wave_tot = s.slot2;
    if (lane == 63) wave_tot[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int wv = 0; wv < (int)(threadIdx.x >> 6); ++wv) before += wave_tot[wv];
    const uint32_t ex = before + incl - (p0 + p1);
    if (w0 < n_words) pre[w0] = ex;
    if (w0 + 1 < n_words) pre[w0 + 1] = ex + p0;
    __syncthreads();
    // one occupied cell per key, as distinct keys give; anything else is
    // left to the comparison sort (key2 is its scratch too)
    if (pre[n_words - 1] + (uint32_t)__popc(bm[n_words - 1]) != (uint32_t)cnt) return false;
    // ranks in registers, then every key moves to its place
    int64_t k[4];
    uint32_t sl[4], rank[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = (int)threadIdx.x + j * (int)blockDim.x;
        if (i < cnt) {
            k[j] = s.key[i];
            sl[j] = s.slot[i];
            const uint32_t c = cell((uint64_t)k[j]);
            rank[j] = pre[c >> 5] + (uint32_t)__popc(bm[c >> 5] & ((1u << (c & 31u)) - 1u));
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = (int)threadIdx.x + j * (int)blockDim.x;
        if (i < cnt) {
            s.key[rank[j]] = k[j];
            s.slot[rank[j]] = sl[j];
        }
    }
    __syncthreads();
    return true;
}

__device__ inline bool entry_less(int64_t ka, uint32_t sa, int64_t kb, uint32_t sb) {
    return ka < kb || (ka == kb && sa < sb);
}

__device__ inline int64_t shfl_xor64(int64_t x, int m) {
    const int lo = __shfl_xor((int)(uint32_t)(uint64_t)x, m);
    const int hi = __shfl_xor((int)(uint32_t)((uint64_t)x >> 32), m);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

// Sort key/slot[0, cnt) in place by (key, slot). Real keys are distinct; the
// padding (INT64_MAX, all-ones slot) compares above any real entry through
// the slot, since a real key may itself equal INT64_MAX. Up to 256 entries: a
// rank sort (one barrier). Above: the list padded to a power of two P is cut
// into runs of 64 that waves sort in registers with a bitonic network of
// lane exchanges (no barriers), then runs are merged pairwise in LDS, level
// by level (one barrier each): every thread finds where its stretch of the
// merged output starts by a binary search over the two runs (merge path)
// and merges that stretch sequentially, ties taken from the first run.
__device__ void lds_sort(AssignLds& s, int cnt) {
    if (cnt <= 256) {
        const bool has = (int)threadIdx.x < cnt;
        int64_t mine = 0;
        uint32_t mslot = 0;
        int rank = 0;
        if (has) {
            mine = s.key[threadIdx.x];
            mslot = s.slot[threadIdx.x];
            for (int j = 0; j < cnt; ++j) rank += s.key[j] < mine;
        }
        __syncthreads();
        if (has) { s.key[rank] = mine; s.slot[rank] = mslot; }
        __syncthreads();
        return;
    }
    if (bitmap_rank(s, cnt)) return;
    int P = 512;
    while (P < cnt) P <<= 1;
    for (int i = cnt + threadIdx.x; i < P; i += blockDim.x) {
        s.key[i] = INT64_MAX;
        s.slot[i] = 0xffffffffu;
    }
    __syncthreads();
    const int lane = __lane_id();
    const int waves = blockDim.x >> 6;
    for (int r = threadIdx.x >> 6; r < (P >> 6); r += waves) {
        int64_t k = s.key[r * 64 + lane];
        uint32_t sl = s.slot[r * 64 + lane];
        for (int kk = 2; kk <= 64; kk <<= 1) {
            for (int j = kk >> 1; j > 0; j >>= 1) {
                const int64_t ok = shfl_xor64(k, j);
                const uint32_t osl = (uint32_t)__shfl_xor((int)sl, j);
                // the pair's lower position keeps the smaller entry in an
                // ascending block (lane & kk clear) and the larger otherwise
                const bool keep_min = ((lane & j) != 0) == ((lane & kk) != 0);
                const bool other_less = entry_less(ok, osl, k, sl);
                if (keep_min == other_less) { k = ok; sl = osl; }
            }
        }
        s.key[r * 64 + lane] = k;
        s.slot[r * 64 + lane] = sl;
    }
    __syncthreads();
    int64_t* ka = s.key;
    uint32_t* sa = s.slot;
    int64_t* kb = s.key2;
    uint32_t* sb = s.slot2;
    const int per = P > (int)blockDim.x ? P / (int)blockDim.x : 1;  // outputs per thread
    for (int run = 64; run < P; run <<= 1) {
        for (int t = threadIdx.x; t * per < P; t += blockDim.x) {
            const int og = t * per;
            const int a0 = og / (2 * run) * (2 * run), b0 = a0 + run;
            const int o = og - a0;
            int lo = max(0, o - run), hi = min(o, run);
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                // A[mid] goes out before B[o - mid - 1]: it is among the first o
                if (!entry_less(ka[b0 + o - mid - 1], sa[b0 + o - mid - 1], ka[a0 + mid], sa[a0 + mid]))
                    lo = mid + 1;
                else
                    hi = mid;
            }
            int ia = lo, ib = o - lo;
            for (int q = 0; q < per; ++q) {
                const bool take_a = ib >= run ||
                                    (ia < run && !entry_less(ka[b0 + ib], sa[b0 + ib], ka[a0 + ia],
                                                             sa[a0 + ia]));
                const int from = take_a ? a0 + ia : b0 + ib;
                kb[og + q] = ka[from];
                sb[og + q] = sa[from];
                ia += take_a ? 1 : 0;
                ib += take_a ? 0 : 1;
            }
        }
        __syncthreads();
        int64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* ts = sa; sa = sb; sb = ts;
    }
    if (ka != s.key) {
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            s.key[i] = ka[i];
            s.slot[i] = sa[i];
        }
        __syncthreads();
    }
}

// Number of A entries among the first k outputs of merge(A, B), keys
// distinct, by one wave: each round tests 64 evenly spaced candidates of the
// monotone predicate A[i] < B[k - i - 1] and keeps the gap where it turns.
__device__ int corank_wave(const TouchEntry OSN_GLOBAL* A, int la, const TouchEntry OSN_GLOBAL* B,
                           int lb, int k) {
    int lo = max(0, k - lb), hi = min(k, la);
    const int lane = __lane_id();
    while (lo < hi) {
        const int stride = (hi - lo + 63) >> 6;
        const int i = lo + lane * stride;
        const bool p = i < hi && A[i].key < B[k - i - 1].key;
        const int t = __popcll(__ballot(p));
        if (t == 0) break;
        const int nhi = min(hi, lo + t * stride);
        lo = lo + (t - 1) * stride + 1;
        hi = nhi;
    }
    return lo;
}

// Same in LDS, one thread, binary search.
__device__ inline int corank_lds(const int64_t* a, int la, const int64_t* b, int lb, int k) {
    int lo = max(0, k - lb), hi = min(k, la);
    while (lo < hi) {
        const int i = (lo + hi) >> 1;
        if (a[i] < b[k - i - 1]) lo = i + 1; else hi = i;
    }
    return lo;
}

struct AssignCall {
    int32_t n, T, L, t_stride;
    int32_t* done;      // this parity's counters, level l at done[l * t_stride]
    int32_t bc0;
};

// Output position p of the call's sorted order: store it, or at the top level
// assign its pool index.
__device__ inline void emit_sorted(const VolDev& v, const AssignCall& a, bool top, TouchEntry* dst,
                                   int p, int64_t key, uint32_t slot, unsigned long long& drops) {
    if (!top) {
        TouchEntry OSN_GLOBAL* d = glob(dst) + p;
        d->key = key;
        d->slot = slot;
        d->bi = kIdxPending;
        return;
    }
    const int64_t idx = (int64_t)a.bc0 + p;
    if (idx < v.pool_capacity) {
        int32_t x, y, z;
        unpack_key(key, x, y, z);
        int32_t OSN_GLOBAL* bc = glob(v.block_coord) + idx * 3;
        bc[0] = x;
        bc[1] = y;
        bc[2] = z;
        glob(v.table)[slot].block_idx = (int32_t)idx;
    } else {
        glob(v.table)[slot].block_idx = kIdxDropped;
        drops += 1;
    }
}

// Thread 0 waits until *p >= target; false (for every thread) when the call
// has faulted. The spin reads with relaxed loads; one agent-scope acquire
// fence afterwards (an L1 + L2 invalidate on this XCD) makes the producers'
// runs visible to every wave of the workgroup, which read only after the
// barrier. An acquire load inside the spin would invalidate L2 on every
// round and evict what concurrent items on the XCD are merging.
__device__ bool wait_count(const VolDev& v, AssignLds& s, int32_t* p, int32_t target) {
    if (threadIdx.x == 0) {
        int ok = 1;
        for (uint32_t it = 0;; ++it) {
            if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
            if (__hip_atomic_load(&v.ctr->fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (it >= kSpinLimit) {
                atomicOr(&v.ctr->fault, 1u);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s.flag = ok;
    }
    __syncthreads();
    const bool ok = s.flag != 0;
    __syncthreads();
    return ok;
}

// Every wave's global stores reach L2 (the barrier alone waits only for LDS),
// then thread 0 writes this XCD's L2 back with one agent-scope release fence
// and bumps *ctr. Returns the old count in thread 0.
__device__ inline int publish_count(int32_t* ctr) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int old = 0;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        old = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return old;
}

__device__ inline int32_t chunks_in_run(const AssignCall& a, int level, int run) {
    return min((run + 1) << level, a.T) - (run << level);
}

// One work item; false when the call has faulted.
__device__ bool assign_item(const VolDev& v, const AssignCall& a, AssignLds& s, int level, int c) {
    const bool top = level == a.L;
    TouchEntry* dst = (level & 1) ? v.sort_b : v.sort_a;
    const TouchEntry* src = (level & 1) ? v.sort_a : v.sort_b;
    const int o0 = c * kSortTile;
    const int o1 = min(o0 + kSortTile, a.n);
    unsigned long long drops = 0;
    if (level == 0) {
        const int cnt = o1 - o0;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            const TouchEntry OSN_GLOBAL* e = glob(v.new_keys) + o0 + i;
            s.key[i] = e->key;
            s.slot[i] = e->slot;
        }
        __syncthreads();
        lds_sort(s, cnt);
        for (int i = threadIdx.x; i < cnt; i += blockDim.x)
            emit_sorted(v, a, top, dst, o0 + i, s.key[i], s.slot[i], drops);
    } else {
        const int w_in = kSortTile << (level - 1);
        const int q = c >> level;
        const int a_lo = q * 2 * w_in;
        const int a_hi = min(a_lo + w_in, a.n);
        const int b_hi = min(a_lo + 2 * w_in, a.n);
        int32_t* below = a.done + (level - 1) * a.t_stride;
        if (!wait_count(v, s, below + 2 * q, chunks_in_run(a, level - 1, 2 * q))) return false;
        if (a_hi < b_hi &&
            !wait_count(v, s, below + 2 * q + 1, chunks_in_run(a, level - 1, 2 * q + 1)))
            return false;
        const int m = o1 - o0;
        if (a_hi >= b_hi) {  // a lone run: copied (or assigned) as it is
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const TouchEntry OSN_GLOBAL* e = glob(src) + o0 + i;
                emit_sorted(v, a, top, dst, o0 + i, e->key, e->slot, drops);
            }
        } else {
            const TouchEntry OSN_GLOBAL* A = glob(src) + a_lo;
            const TouchEntry OSN_GLOBAL* B = glob(src) + a_hi;
            const int la = a_hi - a_lo, lb = b_hi - a_hi;
            const int wave = threadIdx.x >> 6;
            if (wave < 2) {
                const int r = corank_wave(A, la, B, lb, (wave ? o1 : o0) - a_lo);
                if (__lane_id() == 0) s.corank[wave] = r;
            }
            __syncthreads();
            const int i0 = s.corank[0], i1 = s.corank[1];
            const int na = i1 - i0, nb = m - na;
            const int j0 = (o0 - a_lo) - i0;
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const TouchEntry OSN_GLOBAL* e = i < na ? A + (i0 + i) : B + (j0 + i - na);
                s.key[i] = e->key;
                s.slot[i] = e->slot;
            }
            __syncthreads();
            const int per = (m + blockDim.x - 1) / blockDim.x;
            const int k0 = min(m, (int)threadIdx.x * per);
            const int k1 = min(m, k0 + per);
            if (k0 < k1) {
                const int64_t* ka = s.key;
                const int64_t* kb = s.key + na;
                int i = corank_lds(ka, na, kb, nb, k0);
                int j = k0 - i;
                for (int k = k0; k < k1; ++k) {
                    const bool take_a = j >= nb || (i < na && ka[i] < kb[j]);
                    const int from = take_a ? i++ : na + j++;
                    emit_sorted(v, a, top, dst, o0 + k, s.key[from], s.slot[from], drops);
                }
            }
        }
    }
    // publish: this item's stores, then its count
    if (top) {
        drops = wave_sum_u64(drops);
        if (__lane_id() == 0 && drops) atomicAdd(&s.drops, drops);
    }
    int32_t* mine = a.done + level * a.t_stride + (c >> level);
    int old = 0;
    if (!top) {
        old = publish_count(mine);
    } else {
        // nothing in this kernel reads a top item's stores: count it, no fence;
        // a one-tile call has nothing to count
        __syncthreads();
        if (threadIdx.x == 0 && a.T > 1)
            old = __hip_atomic_fetch_add(mine, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) {
        if (top && s.drops) atomicAdd(&v.ctr->drop_count, s.drops);
        s.drops = 0;
        s.flag = top && old == a.T - 1;  // the last top item finishes the call
    }
    __syncthreads();
    if (s.flag) {
        // every item has finished, so no one reads these counters any more
        for (int l = 0; l <= a.L; ++l) {
            const int runs = (a.T + (1 << l) - 1) >> l;
            for (int r = threadIdx.x; r < runs; r += blockDim.x) a.done[l * a.t_stride + r] = 0;
        }
        if (threadIdx.x == 0)
            v.ctr->block_count = (int32_t)min((int64_t)a.bc0 + a.n, (int64_t)v.pool_capacity);
    }
    __syncthreads();
    return true;
}

__device__ void assign_volume(const VolDev& v, int parity, int wg, AssignLds& s) {
    AssignCall a;
    a.n = v.ctr->lists[parity].new_count;
    if (a.n == 0) return;
    a.T = (a.n + kSortTile - 1) / kSortTile;
    a.L = ceil_log2(a.T);
    if (wg >= a.T * (a.L + 1)) return;
    a.t_stride = assign_tiles_max(v.capacity);
    a.done = v.assign_done + (size_t)parity * assign_done_per_parity(v.capacity);
    a.bc0 = v.ctr->block_count;
    if (threadIdx.x == 0) s.drops = 0;
    if (a.T == 1) {
        // one tile, one item: workgroup 0 (the only one past the check
        // above) takes it without a ticket
        __syncthreads();
        assign_item(v, a, s, 0, 0);
        return;
    }
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s.ticket = atomicAdd(&v.ctr->lists[parity].assign_ticket, 1);
        __syncthreads();
        const int t = s.ticket;
        if (t >= a.T * (a.L + 1)) break;
        const int level = t / a.T;
        if (!assign_item(v, a, s, level, t - level * a.T)) break;
    }
}

__global__ __launch_bounds__(1024) void k_assign(VolDev v, CallInfo call) {
    __shared__ AssignLds s;
    assign_volume(v, call.parity, blockIdx.x, s);
}

__global__ __launch_bounds__(1024) void k_assign_batch(BatchDev b) {
    __shared__ AssignLds s;
    const int k = b.step_ctr ? *b.step_ctr : b.step;
    const int sc = blockIdx.y;
    const SceneStep& st = b.steps[k * b.n_scenes + sc];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the scene's update parameters for this step, built once here (the
        // touched list is final) rather than by every update workgroup
        UpdParams p = make_params(b.vols[sc], st.cam, batch_call(b, sc, k).parity);
        if (!st.active) p.n = 0;
        b.upd[sc] = p;
    }
    if (!st.active) return;
    const VolDev v = b.vols[sc];
    assign_volume(v, batch_call(b, sc, k).parity, blockIdx.x, s);
}

// ---- reference-semantics scatter update (volume.cpp:200-299) ---------------

// Every point walks its 2 * steps + 1 samples along the camera ray; a sample
// inside the truncation band adds one visit (weight 1, the clamped
// normalised distance) to its voxel. Weights are exact integer counts in any
// order; a tsdf sum changes only in summation order, which is arbitrary in
// the reference too (its atomics race). The colour sums are 128 * count, as
// the reference's repeated 128s add exactly.
//
// Float atomics at device scope are performed at the memory side, serially
// per address and at a few tens of G/s for the whole chip: a single pass
// that flushes a workgroup's distinct voxels with five of them each spends
// most of its time there. Hence two passes over the scratch that
// ScatterScratch describes: the visit pass groups the workgroup's distinct
// voxels by block and writes them as entries, and the apply pass adds each
// touched block's totals with plain coalesced loads and stores, each block
// owned by one workgroup.

static_assert(kScatterBlocks == kScatterThreads, "one block slot per thread");
static_assert(kScatterBlocks == 256, "the block-slot hash takes the top 8 bits");
static_assert(kBlockVoxels == 512 && kScatterSet <= 4096 && kScatterSetDense <= kScatterSet,
              "a grouped entry packs voxel (9 bits), block slot (8) and rank (12)");

template <typename Key, int SET>
struct ScatterVisitLds {
    // voxel index in the pool, ~0 = empty; once grouped by block, the voxel
    // in its block | block slot << 9 | rank among the block's entries << 17
    Key key[SET];
    uint32_t w[SET];                 // visits
    float t[SET];                    // sum of the visits' normalised distances
    uint32_t bkey[kScatterBlocks];   // pool block, ~0 = empty
    uint32_t bcnt[kScatterBlocks];   // the block's entries, then its first entry
    uint32_t wsum[kScatterThreads / 64];
    uint8_t direct[kScatterBlocks];  // the block's segments are full: add directly
};

__device__ inline void scatter_global(const VolDev& v, size_t idx, float w, float t) {
    unsafeAtomicAdd(&v.weight[idx], w);
    unsafeAtomicAdd(&v.tsdf[idx], t);
    const float c = 128.0f * w;
    unsafeAtomicAdd(&v.r[idx], c);
    unsafeAtomicAdd(&v.g[idx], c);
    unsafeAtomicAdd(&v.b[idx], c);
}

// Weight caps: the reference skips a visit once the voxel's weight has
// reached the cap (volume.cpp:283-289), so its sequential CPU arm ends a
// voxel at min(cap, weight before the call + its visits), rounded up to
// whole visits. A group of w visits (sum t) takes take = min(w, ceil(cap -
// weight)) of them against the weight it finds; the same count results
// whatever the order in which groups arrive. When only part of the w visits
// fit, their sum is taken in proportion (t * take / w), a convex weighting of
// the same observations; which visits the reference keeps at the cap depends
// on its thread timing anyway.
__device__ inline float cap_take(float cur, float w, float cap) {
    return cur < cap ? fminf(w, ceilf(cap - cur)) : 0.0f;
}

// One group added with global atomics, concurrent with other groups of the
// same voxel: the weight is claimed with a compare-and-swap loop, so the
// claims of one voxel are serialised.
__device__ inline void scatter_global_capped(const VolDev& v, size_t idx, float w, float t) {
    unsigned int* wp = reinterpret_cast<unsigned int*>(&v.weight[idx]);
    unsigned int old = __atomic_load_n(wp, __ATOMIC_RELAXED);
    float take;
    for (;;) {
        const float cur = __uint_as_float(old);
        take = cap_take(cur, w, v.weight_cap);
        if (take <= 0.0f) return;
        const unsigned int seen = atomicCAS(wp, old, __float_as_uint(cur + take));
        if (seen == old) break;
        old = seen;
    }
    unsafeAtomicAdd(&v.tsdf[idx], take == w ? t : t * (take / w));
    const float c = 128.0f * take;
    unsafeAtomicAdd(&v.r[idx], c);
    unsafeAtomicAdd(&v.g[idx], c);
    unsafeAtomicAdd(&v.b[idx], c);
}

__device__ inline void scatter_direct(const VolDev& v, size_t idx, float w, float t) {
    if (v.weight_cap > 0.0f) scatter_global_capped(v, idx, w, t);
    else scatter_global(v, idx, w, t);
}

// Whether the lane to the left in the same 16-lane row holds the same key (a
// row's first lane: never). Callers give lanes without a key distinct values.
__device__ inline bool same_as_left(uint32_t me) {
    const uint32_t left =
        (uint32_t)__builtin_amdgcn_update_dpp((int)~me, (int)me, 0x111, 0xF, 0xF, false);  // row_shr:1
    return left == me;
}
__device__ inline bool same_as_left(unsigned long long me) {
    // both halves' moves run in every lane (no short-circuit)
    const bool lo = same_as_left((uint32_t)me), hi = same_as_left((uint32_t)(me >> 32));
    return lo && hi;
}

// One step of a suffix sum within runs: lane L adds the partial sums of lane
// L + K when that lane is in the same 16-lane row and the same run. After
// steps 1, 2, 4, 8 a run's first lane holds the whole run's visits and sum
// (runs are at most a row long). Every lane must be active.
template <int K>
__device__ inline void run_step(int run, int& w, float& ts) {
    constexpr int kRowShl = 0x100 + K;
    const int r_o = __builtin_amdgcn_update_dpp(-1, run, kRowShl, 0xF, 0xF, false);
    const int w_o = __builtin_amdgcn_update_dpp(0, w, kRowShl, 0xF, 0xF, false);
    const float t_o =
        __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(ts), kRowShl, 0xF, 0xF, false));
    if (r_o == run) {
        w += w_o;
        ts += t_o;
    }
}

// SPLIT (small calls): a lane takes one sample of one point instead of the
// point's whole band. Wave q of a point group's waves takes sample q of its
// 64 consecutive points, so the sample's chain (ray, voxel, lookup, set
// insert) is a lane's whole work and small calls fill more of the chip; the
// runs of neighbouring lanes are the same sample of neighbouring points, as
// in the band layout. A group's waves are dealt kScatterThreads / 64 to a
// workgroup (the last workgroup's spare waves have no sample).
template <typename Key, int SET, bool SPLIT>
__global__ __launch_bounds__(kScatterThreads) void k_scatter_visit(VolDev v, PointSrc src,
                                                                   ScatterScratch sc) {
    __shared__ ScatterVisitLds<Key, SET> s;
    constexpr Key kEmpty = ~Key(0);
    const int tid = threadIdx.x;
    for (int e = tid; e < SET; e += kScatterThreads) {
        s.key[e] = kEmpty;
        s.w[e] = 0u;
        s.t[e] = 0.0f;
    }
    s.bkey[tid] = ~0u;
    s.bcnt[tid] = 0u;
    // the next call's touched list starts empty (the apply pass of the call
    // before this one has read it)
    if (blockIdx.x == 0 && tid < kScatterLists)
        sc.touched_count[((sc.parity ^ 1) * kScatterLists + tid) * kScatterCtrStride] = 0;
    __syncthreads();

    int i = blockIdx.x * kScatterThreads + tid;
    int32_t st_lo = -v.steps, st_hi = v.steps;  // the lane's samples
    if constexpr (SPLIT) {
        constexpr int kWaves = kScatterThreads / 64;
        const int n_samples = 2 * v.steps + 1;
        const int per_group = (n_samples + kWaves - 1) / kWaves;  // workgroups per point group
        const int q = (int)(blockIdx.x % (uint32_t)per_group) * kWaves + (tid >> 6);
        i = (int)(blockIdx.x / (uint32_t)per_group) * 64 + (tid & 63);
        st_lo = q - v.steps;
        st_hi = q < n_samples ? st_lo : st_lo - 1;  // wave-uniform: spare waves skip
    }
    float px = 0.f, py = 0.f, pz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, d2 = 0.f, dist = 0.f;
    bool ok = i < src.n;
    if (ok) {
        px = glob(src.pos)[i * 3 + 0];
        py = glob(src.pos)[i * 3 + 1];
        pz = glob(src.pos)[i * 3 + 2];
        dx = px - src.cam[0]; dy = py - src.cam[1]; dz = pz - src.cam[2];
        d2 = dx * dx + dy * dy + dz * dz;
        ok = !(src.radius_sq > 0.0f && d2 > src.radius_sq);
        dist = sqrtf(d2);
        ok = ok && (dist > 1e-6f);
    }
    {
        // Every lane runs every sample (the runs below compare neighbouring
        // lanes); `val` says whether the lane's sample visits a voxel.
        const float ux = ok ? dx / dist : 0.0f, uy = ok ? dy / dist : 0.0f,
                    uz = ok ? dz / dist : 0.0f;
        const float inv_trunc = 1.0f / v.trunc;
        const int lane = tid & 63;
        int64_t last_key = kEmptyKey;
        int32_t last_bi = kIdxAbsent;
        for (int32_t st = st_lo; st <= st_hi; ++st) {
            const float t = (float)st * v.voxel;
            const int32_t vx = (int32_t)floorf((px + ux * t) * v.inv_voxel);
            const int32_t vy = (int32_t)floorf((py + uy * t) * v.inv_voxel);
            const int32_t vz = (int32_t)floorf((pz + uz * t) * v.inv_voxel);
            const float cx = ((float)vx + 0.5f) * v.voxel;
            const float cy = ((float)vy + 0.5f) * v.voxel;
            const float cz = ((float)vz + 0.5f) * v.voxel;
            const float ex = cx - src.cam[0], ey = cy - src.cam[1], ez = cz - src.cam[2];
            const float sdf = dist - sqrtf(ex * ex + ey * ey + ez * ez);
            bool val = ok && !(sdf < -v.trunc);
            const int32_t bx = floor_div8(vx), by = floor_div8(vy), bz = floor_div8(vz);
            const int64_t key = pack_coord(bx, by, bz);
            if (val && key != last_key) { last_key = key; last_bi = find_index(v, key); }
            val = val && last_bi >= 0 && last_bi < v.pool_capacity;
            const float sdf_n = fminf(fmaxf(sdf * inv_trunc, -1.0f), 1.0f);
            const Key idx = (Key)(val ? last_bi : 0) * kBlockVoxels +
                            (Key)(((vz & 7) * 8 + (vy & 7)) * 8 + (vx & 7));
            // Neighbouring points of a wave often visit the same voxel at the
            // same sample: runs of equal voxels in consecutive lanes (within
            // a 16-lane row) are summed across lanes first, and only a run's
            // first lane meets the LDS set.
            const bool head = !same_as_left(val ? idx : ~Key(0) - (Key)lane);
            const unsigned long long heads = __ballot(head);
            // run number: the heads at or before this lane
            const int run = __builtin_amdgcn_mbcnt_hi((uint32_t)(heads >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)heads, 0u)) +
                            (head ? 1 : 0);
            int w = val ? 1 : 0;
            float ts = val ? sdf_n : 0.0f;
            run_step<1>(run, w, ts);
            run_step<2>(run, w, ts);
            run_step<4>(run, w, ts);
            run_step<8>(run, w, ts);
            if (!(val && head)) continue;
            // multiply-shift onto [0, SET), which need not be a power of two
            uint32_t h = uint32_t((uint64_t(uint32_t((uint64_t(idx) * 0x9E3779B97F4A7C15ull) >> 32)) *
                                   SET) >> 32);
            bool placed = false;
            for (int p = 0; p < 16; ++p) {
                const Key prev = atomicCAS(&s.key[h], kEmpty, idx);
                if (prev == kEmpty || prev == idx) {
                    atomicAdd(&s.w[h], (uint32_t)w);
                    atomicAdd(&s.t[h], ts);
                    placed = true;
                    break;
                }
                h = h + 1 == SET ? 0u : h + 1;
            }
            if (!placed) scatter_direct(v, (size_t)idx, (float)w, ts);  // LDS set crowded
        }
    }
    __syncthreads();

    // the distinct voxels by block: a slot per block, a rank per entry
    for (int e = tid; e < SET; e += kScatterThreads) {
        const Key k = s.key[e];
        if (k == kEmpty) continue;
        const uint32_t bi = (uint32_t)(k / kBlockVoxels);
        uint32_t h = (bi * 0x9E3779B1u) >> 24;  // kScatterBlocks = 256 slots
        int found = -1;
        for (int p = 0; p < kScatterBlocks; ++p) {
            const uint32_t prev = atomicCAS(&s.bkey[h], ~0u, bi);
            if (prev == ~0u || prev == bi) {
                found = (int)h;
                break;
            }
            h = (h + 1) & (kScatterBlocks - 1);
        }
        if (found < 0) {  // more distinct blocks than slots
            scatter_direct(v, (size_t)k, (float)s.w[e], s.t[e]);
            s.key[e] = kEmpty;
            continue;
        }
        const uint32_t rank = atomicAdd(&s.bcnt[found], 1u);
        s.key[e] = (Key)((uint32_t)(k % kBlockVoxels) | ((uint32_t)found << 9) | (rank << 17));
    }
    __syncthreads();

    // thread j owns block slot j: its first entry in the region (exclusive
    // scan of the counts) and the block's segment for this workgroup
    const uint32_t c = s.bcnt[tid];
    const int lane = tid & 63;
    uint32_t incl = c;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s.wsum[tid >> 6] = incl;
    __syncthreads();
    uint32_t first = incl - c;
    for (int w = 0; w < (tid >> 6); ++w) first += s.wsum[w];
    const uint32_t region = blockIdx.x * (uint32_t)sc.region;
    uint8_t direct = 0;
    if (c > 0) {
        const uint32_t bi = s.bkey[tid];
        const int seg = atomicAdd(&sc.seg_count[bi], 1);
        if (seg == 0) {
            const uint32_t j = bi % kScatterLists;
            const int at = atomicAdd(
                &sc.touched_count[(sc.parity * kScatterLists + j) * kScatterCtrStride], 1);
            sc.touched[(size_t)j * sc.list_cap + at] = (int32_t)bi;
        }
        if (seg < kScatterSegs)
            sc.segs[(size_t)bi * kScatterSegs + seg] = ScatterSeg{region + first, c};
        else
            direct = 1;
    }
    s.bcnt[tid] = first;
    s.direct[tid] = direct;
    __syncthreads();

    for (int e = tid; e < SET; e += kScatterThreads) {
        const Key k = s.key[e];
        if (k == kEmpty) continue;
        const uint32_t f = (uint32_t)k;
        const uint32_t vox = f & (kBlockVoxels - 1), sl = (f >> 9) & 255u, rank = f >> 17;
        if (s.direct[sl])
            scatter_direct(v, (size_t)s.bkey[sl] * kBlockVoxels + vox, (float)s.w[e], s.t[e]);
        else
            sc.entries[region + s.bcnt[sl] + rank] = make_uint2(vox | (s.w[e] << 9), __float_as_uint(s.t[e]));
    }
}

// One touched block per workgroup at a time: its segments summed per voxel
// in LDS, then each visited voxel read, capped and written once.
__global__ __launch_bounds__(kScatterThreads) void k_scatter_apply(VolDev v, ScatterScratch sc) {
    static_assert(kScatterLists == 64, "one lane per list");
    static_assert(kScatterSegs == 2 * 64 && kBlockVoxels == 2 * kScatterThreads,
                  "two waves hold a block's segments, two voxels per thread");
    __shared__ uint32_t s_w[kBlockVoxels];
    __shared__ float s_t[kBlockVoxels];
    __shared__ int s_first[kScatterLists + 1];       // exclusive prefix of the list lengths
    __shared__ uint32_t s_soff[kScatterSegs];        // a segment's first entry
    __shared__ uint32_t s_spre[kScatterSegs + 1];    // exclusive prefix of the segment lengths
    __shared__ uint32_t s_wave0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (wave == 0) {
        const int c = sc.touched_count[(sc.parity * kScatterLists + lane) * kScatterCtrStride];
        int incl = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        s_first[lane + 1] = incl;
        if (lane == 0) s_first[0] = 0;
    }
    __syncthreads();
    const int n = s_first[kScatterLists];
    // the call's touched blocks, for the host's choice of set size next call
    if (blockIdx.x == 0 && tid == 0 && sc.touched_out) *sc.touched_out = n;
    const bool capped = v.weight_cap > 0.0f;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        int j = 0;  // the list holding item k: s_first[j] <= k < s_first[j + 1]
        for (int half = kScatterLists / 2; half > 0; half >>= 1)
            if (s_first[j + half] <= k) j += half;
        const int32_t bi = sc.touched[(size_t)j * sc.list_cap + (k - s_first[j])];
        const size_t base = (size_t)bi * kBlockVoxels;
        // the thread's two voxels' sums, in flight while the segments are summed
        float cw[2], ct[2], cr[2], cg[2], cb[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const size_t idx = base + tid + q * kScatterThreads;
            cw[q] = v.weight[idx];
            ct[q] = v.tsdf[idx];
            cr[q] = v.r[idx];
            cg[q] = v.g[idx];
            cb[q] = v.b[idx];
            s_w[tid + q * kScatterThreads] = 0u;
            s_t[tid + q * kScatterThreads] = 0.0f;
        }
        // segments: first entries and an exclusive prefix of their lengths
        // (every slot's descriptor read beside the count, those past it unused)
        ScatterSeg d{0u, 0u};
        if (tid < kScatterSegs) d = sc.segs[(size_t)bi * kScatterSegs + tid];
        const int nseg = min(sc.seg_count[bi], kScatterSegs);
        uint32_t c = 0;
        if (tid < nseg) {
            s_soff[tid] = d.off;
            c = d.cnt;
        }
        uint32_t incl = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (tid == 63) s_wave0 = incl;
        __syncthreads();
        if (tid < kScatterSegs) s_spre[tid + 1] = incl + (wave == 1 ? s_wave0 : 0u);
        if (tid == 0) s_spre[0] = 0u;
        __syncthreads();
        // every entry of every segment, all loads independent
        const uint32_t total = s_spre[nseg];
        for (uint32_t e = tid; e < total; e += kScatterThreads) {
            int g = 0;  // the segment holding entry e: s_spre[g] <= e < s_spre[g + 1]
            for (int half = kScatterSegs / 2; half > 0; half >>= 1)
                if (g + half < nseg && s_spre[g + half] <= e) g += half;
            const uint2 en = sc.entries[s_soff[g] + (e - s_spre[g])];
            atomicAdd(&s_w[en.x & (kBlockVoxels - 1)], en.x >> 9);
            atomicAdd(&s_t[en.x & (kBlockVoxels - 1)], __uint_as_float(en.y));
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int l = tid + q * kScatterThreads;
            const uint32_t W = s_w[l];
            if (W == 0u) continue;
            const size_t idx = base + l;
            const float w = (float)W, T = s_t[l];
            const float take = capped ? cap_take(cw[q], w, v.weight_cap) : w;
            if (take <= 0.0f) continue;
            v.weight[idx] = cw[q] + take;
            v.tsdf[idx] = ct[q] + (take == w ? T : T * (take / w));
            const float cc = 128.0f * take;
            v.r[idx] = cr[q] + cc;
            v.g[idx] = cg[q] + cc;
            v.b[idx] = cb[q] + cc;
        }
        if (tid == 0) sc.seg_count[bi] = 0;
        __syncthreads();
    }
}

__global__ void k_fill_table(HashEntry* t, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) t[i] = HashEntry{kEmptyKey, kIdxPending, 0};
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

}  // namespace

template <int MODE, bool TOUCH>
static void launch_alloc_t(const VolDev& v, const PointSrc& src, const int32_t* coords, int set,
                           CallInfo call, int grid_cap, hipStream_t s) {
    int tiles;
    if (src.depth && !coords) tiles = depth_units(src.W, src.H, 1);
    else tiles = (src.n + kAllocThreads - 1) / kAllocThreads;
    const int grid = grid_cap > 0 ? min(tiles, grid_cap) : tiles;
    if (coords)
        k_alloc<MODE, TOUCH, kSrcCoords><<<grid, kAllocThreads, 0, s>>>(v, src, coords, set, call, tiles);
    else if (src.depth)
        k_alloc<MODE, TOUCH, kSrcDepth><<<grid, kAllocThreads, 0, s>>>(v, src, nullptr, set, call, tiles);
    else
        k_alloc<MODE, TOUCH, kSrcPoints><<<grid, kAllocThreads, 0, s>>>(v, src, nullptr, set, call, tiles);
}

static int clamp_set(int set) {
    if (set < 0) set = 0;
    if (set > kMaxSet) set = kMaxSet;
    return set;
}

hipError_t launch_alloc(const VolDev& v, const PointSrc& src, const AllocOpts& o,
                        const int32_t* coords, CallInfo call, hipStream_t s) {
    const int set = clamp_set(o.lds_set);
    if (set & (set - 1)) return hipErrorInvalidValue;
    switch (o.mode) {
        case kAllocNormal:
            if (o.track_touched) launch_alloc_t<kAllocNormal, true>(v, src, coords, set, call, o.grid, s);
            else launch_alloc_t<kAllocNormal, false>(v, src, coords, set, call, o.grid, s);
            break;
        case kAllocCountCas: launch_alloc_t<kAllocCountCas, false>(v, src, coords, set, call, o.grid, s); break;
        case kAllocProbe: launch_alloc_t<kAllocProbe, false>(v, src, coords, set, call, o.grid, s); break;
        case kAllocNoCas: launch_alloc_t<kAllocNoCas, false>(v, src, coords, set, call, o.grid, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_alloc_batch(const BatchDev& b, int32_t total_tiles, const AllocOpts& o,
                              hipStream_t s, const LaunchEvents& ev) {
    const int set = clamp_set(o.lds_set);
    if (set & (set - 1)) return hipErrorInvalidValue;
    const int grid = o.grid > 0 ? min(total_tiles, o.grid) : total_tiles;
    switch (o.mode) {
        case kAllocNormal:
            launch_kernel(k_alloc_batch<kAllocNormal>, dim3(grid), dim3(kAllocThreads), s, ev, b, set,
                          total_tiles);
            break;
        case kAllocCountCas:
            launch_kernel(k_alloc_batch<kAllocCountCas>, dim3(grid), dim3(kAllocThreads), s, ev, b,
                          set, total_tiles);
            break;
        case kAllocProbe:
            launch_kernel(k_alloc_batch<kAllocProbe>, dim3(grid), dim3(kAllocThreads), s, ev, b, set,
                          total_tiles);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_assign(const VolDev& v, CallInfo call, hipStream_t s) {
    const uint32_t tiles = (uint32_t)assign_tiles_max(v.capacity);
    const uint32_t grid = tiles < (uint32_t)kAssignMaxGrid ? tiles : (uint32_t)kAssignMaxGrid;
    k_assign<<<grid, 1024, 0, s>>>(v, call);
    return hipGetLastError();
}

hipError_t launch_assign_batch(const BatchDev& b, uint32_t max_capacity, hipStream_t s,
                               const LaunchEvents& ev) {
    // Items are taken by ticket, so any width is correct. A frame after a
    // scene's first few adds a few hundred new keys (one tile), so a narrow
    // launch: dispatching idle 1024-thread workgroups was measured to cost
    // more than the first frames gain from a wide one.
    const uint32_t tiles = (uint32_t)assign_tiles_max(max_capacity);
    const uint32_t gx = tiles < 4u ? tiles : 4u;
    launch_kernel(k_assign_batch, dim3(gx, b.n_scenes), dim3(1024), s, ev, b);
    return hipGetLastError();
}

hipError_t launch_scatter_update(const VolDev& v, const PointSrc& src, const ScatterScratch& sc,
                                 bool dense, bool split, hipStream_t s) {
    // the region of workgroup g starts at entry g * sc.region (32-bit); it
    // holds the workgroup's distinct voxels
    const int64_t grid = scatter_visit_grid(src.n, v.steps, split);
    const int32_t need = split ? scatter_region_split() : scatter_region(v.steps);
    if (src.n <= 0 || sc.region < need || grid * sc.region > (int64_t)UINT32_MAX)
        return hipErrorInvalidValue;
    // voxel indices as 32-bit LDS keys while they fit (~0 stays free)
    const bool small_keys = (uint64_t)v.pool_capacity * kBlockVoxels < 0xFFFFFFFFull;
    if (split) {
        if (small_keys)
            k_scatter_visit<uint32_t, kScatterSetSplit, true><<<(int)grid, kScatterThreads, 0, s>>>(v, src, sc);
        else
            k_scatter_visit<unsigned long long, kScatterSetSplit, true>
                <<<(int)grid, kScatterThreads, 0, s>>>(v, src, sc);
    } else if (small_keys && dense) {
        k_scatter_visit<uint32_t, kScatterSetDense, false>
            <<<(int)grid, kScatterThreads, 0, s>>>(v, src, sc);
    } else if (small_keys) {
        k_scatter_visit<uint32_t, kScatterSet, false><<<(int)grid, kScatterThreads, 0, s>>>(v, src, sc);
    } else if (dense) {
        k_scatter_visit<unsigned long long, kScatterSetDense, false>
            <<<(int)grid, kScatterThreads, 0, s>>>(v, src, sc);
    } else {
        k_scatter_visit<unsigned long long, kScatterSet, false>
            <<<(int)grid, kScatterThreads, 0, s>>>(v, src, sc);
    }
    const int ag = v.pool_capacity < kScatterApplyGrid ? v.pool_capacity : kScatterApplyGrid;
    k_scatter_apply<<<ag, kScatterThreads, 0, s>>>(v, sc);
    return hipGetLastError();
}

hipError_t launch_fill_table(HashEntry* table, uint32_t n, hipStream_t s) {
    k_fill_table<<<(n + 255) / 256, 256, 0, s>>>(table, n);
    return hipGetLastError();
}

namespace {

// Copy a row-major depth image into the update's 8x4-tiled layout, one thread
// per element of the copy (coalesced stores, 32-byte row pieces loaded).
__global__ __launch_bounds__(256) void k_tile_depth(const float* depth, int32_t W, int32_t H,
                                                    float* tiled) {
    const uint32_t wt8 = (uint32_t)tiled_width(W) >> 3;
    const uint32_t total = wt8 * 8u * (uint32_t)tiled_height(H);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t t = i >> 5, within = i & 31u;
    const uint32_t u = (t % wt8) * 8u + (within & 7u), w = (t / wt8) * 4u + (within >> 3);
    glob(tiled)[i] = (u < (uint32_t)W && w < (uint32_t)H) ? glob(depth)[(size_t)w * W + u] : 0.0f;
}

}  // namespace

hipError_t launch_tile_depth(const float* depth, int32_t W, int32_t H, float* tiled, hipStream_t s) {
    if (W <= 0 || H <= 0 || W > kMaxFrameDim || H > kMaxFrameDim) return hipErrorInvalidValue;
    const uint32_t total = (uint32_t)tiled_width(W) * (uint32_t)tiled_height(H);
    k_tile_depth<<<(total + 255) / 256, 256, 0, s>>>(depth, W, H, tiled);
    return hipGetLastError();
}

hipError_t launch_backproject(const PointSrc& src, float* out, hipStream_t s) {
    dim3 grid((src.W + 15) / 16, (src.H + 15) / 16);
    k_backproject<<<grid, 256, 0, s>>>(src, out);
    return hipGetLastError();
}

// Uncontended 64-bit compare-exchange ceiling: op i swaps slot
// (i * 2654435761) mod n_slots (a bijection when n_slots is a power of two),
// so no two lanes in flight touch the same address and the addresses are
// scattered over the whole buffer, as table inserts are.
__global__ void k_cas_ceiling(unsigned long long* slots, uint32_t mask, int32_t n_ops) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_ops) return;
    const uint32_t s = (uint32_t)((uint64_t)(uint32_t)i * 2654435761ull) & mask;
    atomicCAS(&slots[s], ~0ull, (unsigned long long)i);
}

hipError_t launch_cas_ceiling(unsigned long long* slots, uint32_t mask, int32_t n_ops,
                              hipStream_t s) {
    k_cas_ceiling<<<(n_ops + 255) / 256, 256, 0, s>>>(slots, mask, n_ops);
    return hipGetLastError();
}

}  // namespace osn_hip
