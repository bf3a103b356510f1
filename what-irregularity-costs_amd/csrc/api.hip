// Host side of the gfx950 TSDF arm: volume lifetime, stream discipline and the
// extern "C" entry points declared in include/osn_tsdf_hip.h.
//
// Mirrors the reference device ABI's behaviour (cpp/src/c_api_cuda.cu:151-246,
// cpp/src/cuda_impl.cuh:50-111, cpp/src/volume_cuda.cu:343-478): one private
// stream per volume, the library owns every device allocation, getters
// synchronise, device failures are printed as "[osn_tsdf_hip] file:line" and
// mark the volume invalid. Differences, by design: allocation runs as
// begin -> insert -> assign-in-key-order (no index publication spin), and
// extraction scratch is allocated once and grown, never per call.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/osn_tsdf_hip.h"
#include "internal.hpp"

using namespace osn_hip;

static_assert(sizeof(OsnTsdfDeviceView) == 104, "OsnTsdfDeviceView must stay 104 bytes");

namespace {

constexpr uint32_t kScratchSlots = 1u << 20;  // reference types.hpp:30-33
constexpr int kStages = 4;
// scatter-update entries of one pass pair at most (8 B each): calls with more
// points run as several pass pairs
constexpr int64_t kScatterEntriesMax = int64_t(1) << 24;

#define HIP_OK(expr)                                                                        \
    do {                                                                                    \
        const hipError_t err_ = (expr);                                                     \
        if (err_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "[osn_tsdf_hip] %s:%d: %s -> %s\n", __FILE__, __LINE__,   \
                         #expr, hipGetErrorString(err_));                                   \
            return false;                                                                   \
        }                                                                                   \
    } while (0)

uint32_t next_pow2(uint32_t v) {
    if (v < 2) return 2;
    --v;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}

size_t voxel_bytes(int32_t storage) {
    switch (storage) {
        case kStoreSums: return 20;
        case kStoreF32: return 8;
        case kStoreF16: return 4;
        case kStoreF64: return 16;
        default: return 0;
    }
}

// Copies into a caller's pageable host memory, split over a few threads: the
// reference ABI's extract_mesh fills such a buffer, and one thread's memcpy
// (what a pageable device-to-host copy ends in) runs well below what the
// link and the host memory allow. One job at a time; the calling thread takes
// a share too.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool pool;
        return pool;
    }

    void copy(void* dst, const void* src, size_t n) {
        if (n < (size_t(1) << 20) || workers_.empty()) {
            std::memcpy(dst, src, n);
            return;
        }
        std::lock_guard<std::mutex> job(job_m_);
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            n_ = n;
            pending_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        const int n = (int)std::min(7u, hw > 1 ? hw - 1 : 0u);
        // a thread that cannot be started leaves the pool smaller (no
        // exception may leave a C entry point)
        for (int i = 0; i < n; ++i) {
            try {
                workers_.emplace_back([this, i] { work(i + 1); });
            } catch (...) {
                break;
            }
        }
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    // part p of parts() equal pieces, 64-byte aligned
    void part(int p) {
        const size_t parts = workers_.size() + 1;
        const size_t step = ((n_ + parts - 1) / parts + 63) & ~size_t(63);
        const size_t lo = std::min(n_, step * p), hi = std::min(n_, lo + step);
        if (hi > lo) std::memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void work(int id) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            part(id);
            std::lock_guard<std::mutex> lk(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex job_m_, m_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t n_ = 0;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

// Pinned host staging for device-to-pageable copies: two halves of this size.
constexpr size_t kPinnedChunk = size_t(4) << 20;

}  // namespace

struct OsnTsdfHipVolume {
    OsnTsdfHipConfig cfg{};
    bool reference_shape = false;  // created by osn_tsdf_hip_create
    mutable bool ok = false;       // cleared by a launch failure or a device-side fault
    hipStream_t stream = nullptr;
    VolDev d{};
    uint32_t table_alloc = 0;      // slots allocated (live table + scratch region)
    int32_t lds_set = 1024;
    int32_t update_threads = 256;
    int32_t update_grid = 0;
    int32_t alloc_grid = 0;
    int64_t calls = 0;  // allocate calls since create/reset: call c stamps c + 1, uses list set c & 1
    bool in_batch = false;
    hipStream_t batch_stream = nullptr;  // the batch's stream while in a batch
    int device = 0;                      // the device current at create
    // extraction scratch, grown on demand
    int32_t* ex_vcount = nullptr;
    int32_t* ex_voff = nullptr;
    int32_t* ex_chunks = nullptr;
    int32_t* ex_nb = nullptr;  // 8 neighbour indices per block
    int32_t* ex_flag_count = nullptr;  // [0] = arrival counter, [1] = vertex count
    int32_t ex_blocks_cap = 0;
    float* stage_posnor = nullptr;
    uint8_t* stage_rgb = nullptr;
    // pinned host staging (2 x kPinnedChunk) and its chunk events, at the first
    // extract_mesh into a host buffer
    char* pinned = nullptr;
    hipEvent_t pinned_ev[2] = {nullptr, nullptr};
    int64_t stage_cap = 0;
    // the depth image in the update's 8x4-tiled layout, grown on demand
    float* d_tiled = nullptr;
    size_t tiled_cap = 0;
    // scatter update scratch: per-block parts at the first call, entries
    // grown to the largest chunk of points
    ScatterScratch sc{};
    int64_t sc_entries_cap = 0;  // entries allocated
    // pinned, device-mapped: the touched-block count of the last update pass pair
    // (written by its apply pass, read here without waiting: a hint for the next
    // call's LDS set size, whatever call it is from)
    int32_t* sc_touched_host = nullptr;
    int32_t scatter_chunk() const {  // points per pass pair
        const int64_t groups = kScatterEntriesMax / scatter_region(d.steps);
        return (int32_t)(groups * kScatterThreads);
    }
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double ms[kStages] = {0, 0, 0, 0};
    int64_t stage_calls[kStages] = {0, 0, 0, 0};

    ~OsnTsdfHipVolume();
    bool alloc();
    bool reset();
    bool fail(const char* what, hipError_t e);
    hipEvent_t take_event();
    void stage_begin(int stage, hipEvent_t& a);
    void stage_end(int stage, hipEvent_t a);
    bool grow_extract(int32_t blocks);
    bool grow_stage(int64_t verts, bool rgb);
    bool grow_tiled(int32_t W, int32_t H);
    bool copy_to_host(void* dst, const void* src, size_t n);
    bool grow_scatter(int64_t entries);
    // Wait for every stream that writes this volume: its own and, while it
    // belongs to a batch, the batch's (a non-blocking stream of its own).
    hipError_t sync() const {
        hipError_t e = hipStreamSynchronize(stream);
        if (e == hipSuccess && batch_stream) e = hipStreamSynchronize(batch_stream);
        return e;
    }
};

bool OsnTsdfHipVolume::fail(const char* what, hipError_t e) {
    std::fprintf(stderr, "[osn_tsdf_hip] %s -> %s\n", what, hipGetErrorString(e));
    ok = false;
    return false;
}

bool OsnTsdfHipVolume::alloc() {
    const uint32_t cap = d.capacity;
    HIP_OK(hipGetDevice(&device));
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    table_alloc = cap + (reference_shape ? kScratchSlots : 0u);
    HIP_OK(hipMalloc(&d.table, (size_t)table_alloc * sizeof(HashEntry)));
    HIP_OK(hipMalloc(&d.ctr, sizeof(Counters)));
    HIP_OK(hipMalloc(&d.block_coord, (size_t)cfg.pool_capacity_blocks * 3 * sizeof(int32_t)));
    const size_t n_vox = (size_t)cfg.pool_capacity_blocks * kBlockVoxels;
    if (d.storage == kStoreSums) {
        HIP_OK(hipMalloc(&d.tsdf, n_vox * sizeof(float)));
        HIP_OK(hipMalloc(&d.weight, n_vox * sizeof(float)));
        HIP_OK(hipMalloc(&d.r, n_vox * sizeof(float)));
        HIP_OK(hipMalloc(&d.g, n_vox * sizeof(float)));
        HIP_OK(hipMalloc(&d.b, n_vox * sizeof(float)));
    } else {
        HIP_OK(hipMalloc(&d.vox, n_vox * voxel_bytes(d.storage)));
    }
    HIP_OK(hipMalloc(&d.new_keys, (size_t)cap * sizeof(TouchEntry)));
    HIP_OK(hipMalloc(&d.touched, (size_t)cap * sizeof(TouchEntry)));
    HIP_OK(hipMalloc(&d.sort_a, (size_t)cap * sizeof(TouchEntry)));
    HIP_OK(hipMalloc(&d.sort_b, (size_t)cap * sizeof(TouchEntry)));
    HIP_OK(hipMalloc(&d.assign_done, 2 * (size_t)assign_done_per_parity(cap) * sizeof(int32_t)));
    if (reference_shape) {
        // the sentinel region past the live table (cpp/src/cuda_impl.cuh:61-67,
        // 84-88): keys -2 can never satisfy a CAS against the empty key, so
        // the region is write-once and reset() leaves it alone
        std::vector<HashEntry> host(kScratchSlots, HashEntry{-2, -1, 0});
        HIP_OK(hipMemcpyAsync(d.table + cap, host.data(), kScratchSlots * sizeof(HashEntry),
                              hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    return reset();
}

bool OsnTsdfHipVolume::reset() {
    // The table is refilled with the empty key (a zeroed table would read as
    // occupied at block (0,0,0)); counters and voxel storage are zeroed.
    HIP_OK(launch_fill_table(d.table, d.capacity, stream));
    HIP_OK(hipMemsetAsync(d.ctr, 0, sizeof(Counters), stream));
    HIP_OK(hipMemsetAsync(d.assign_done, 0,
                          2 * (size_t)assign_done_per_parity(d.capacity) * sizeof(int32_t), stream));
    const size_t n_vox = (size_t)cfg.pool_capacity_blocks * kBlockVoxels;
    if (d.storage == kStoreSums) {
        HIP_OK(hipMemsetAsync(d.tsdf, 0, n_vox * sizeof(float), stream));
        HIP_OK(hipMemsetAsync(d.weight, 0, n_vox * sizeof(float), stream));
        HIP_OK(hipMemsetAsync(d.r, 0, n_vox * sizeof(float), stream));
        HIP_OK(hipMemsetAsync(d.g, 0, n_vox * sizeof(float), stream));
        HIP_OK(hipMemsetAsync(d.b, 0, n_vox * sizeof(float), stream));
    } else {
        HIP_OK(hipMemsetAsync(d.vox, 0, n_vox * voxel_bytes(d.storage), stream));
    }
    if (sc.seg_count) {  // the scatter scratch's zero state, should a call have failed
        HIP_OK(hipMemsetAsync(sc.seg_count, 0, (size_t)cfg.pool_capacity_blocks * sizeof(int32_t),
                              stream));
        HIP_OK(hipMemsetAsync(sc.touched_count, 0, kScatterCtrInts * sizeof(int32_t), stream));
        sc.parity = 0;
    }
    HIP_OK(hipStreamSynchronize(stream));
    calls = 0;
    ok = true;
    return true;
}

OsnTsdfHipVolume::~OsnTsdfHipVolume() {
    if (stream) hipStreamSynchronize(stream);
    for (auto& p : ev_pending) {
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    for (auto e : ev_pool) hipEventDestroy(e);
    for (auto e : pinned_ev)
        if (e) hipEventDestroy(e);
    if (pinned) hipHostFree(pinned);
    if (sc_touched_host) hipHostFree(sc_touched_host);
    void* ptrs[] = {d.table, d.ctr, d.block_coord, d.tsdf, d.weight, d.r, d.g, d.b, d.vox,
                    d.new_keys, d.touched, d.sort_a, d.sort_b, d.assign_done, ex_vcount, ex_voff,
                    ex_chunks, ex_nb, ex_flag_count, stage_posnor, stage_rgb, d_tiled,
                    sc.entries, sc.segs, sc.seg_count, sc.touched, sc.touched_count};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (stream) hipStreamDestroy(stream);
}

hipEvent_t OsnTsdfHipVolume::take_event() {
    if (!ev_pool.empty()) {
        hipEvent_t e = ev_pool.back();
        ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void OsnTsdfHipVolume::stage_begin(int stage, hipEvent_t& a) {
    (void)stage;
    a = nullptr;
    if (!timing) return;
    a = take_event();
    if (a) hipEventRecord(a, stream);
}

void OsnTsdfHipVolume::stage_end(int stage, hipEvent_t a) {
    if (!timing || !a) return;
    hipEvent_t b = take_event();
    if (!b) return;
    hipEventRecord(b, stream);
    ev_pending.push_back({stage, {a, b}});
}

bool OsnTsdfHipVolume::grow_extract(int32_t blocks) {
    if (blocks <= ex_blocks_cap && ex_vcount) return true;
    // geometric growth (to the pool at most): a growing volume regrows, and
    // loses its cached neighbour indices, a few times rather than every call
    int64_t want = blocks > 1024 ? blocks : 1024;
    if (want < 2 * (int64_t)ex_blocks_cap) want = 2 * (int64_t)ex_blocks_cap;
    if (want > cfg.pool_capacity_blocks && blocks <= cfg.pool_capacity_blocks)
        want = cfg.pool_capacity_blocks;
    if (ex_vcount) { hipFree(ex_vcount); hipFree(ex_voff); hipFree(ex_chunks); hipFree(ex_nb); }
    ex_vcount = ex_voff = ex_chunks = ex_nb = nullptr;
    HIP_OK(hipMalloc(&ex_nb, (size_t)want * 8 * sizeof(int32_t)));
    HIP_OK(hipMalloc(&ex_vcount, (size_t)want * sizeof(int32_t)));
    HIP_OK(hipMalloc(&ex_voff, (size_t)want * sizeof(int32_t)));
    HIP_OK(hipMalloc(&ex_chunks, ((size_t)want / 1024 + 2) * sizeof(int32_t)));
    if (!ex_flag_count) {
        HIP_OK(hipMalloc(&ex_flag_count, 2 * sizeof(int32_t)));
        HIP_OK(hipMemsetAsync(ex_flag_count, 0, 2 * sizeof(int32_t), stream));
    }
    ex_blocks_cap = (int32_t)want;
    // the neighbour indices cached in the old scratch are gone; complete
    // before returning, since a batch's launches run on another stream
    HIP_OK(hipMemsetAsync(&d.ctr->nb_blocks, 0, sizeof(int32_t), stream));
    HIP_OK(hipStreamSynchronize(stream));
    return true;
}

bool OsnTsdfHipVolume::grow_stage(int64_t verts, bool rgb) {
    if (verts <= stage_cap && (!rgb || stage_rgb)) return true;
    const int64_t want = verts > stage_cap ? verts : stage_cap;
    if (stage_posnor) hipFree(stage_posnor);
    if (stage_rgb) hipFree(stage_rgb);
    stage_posnor = nullptr;
    stage_rgb = nullptr;
    HIP_OK(hipMalloc(&stage_posnor, (size_t)want * 6 * sizeof(float)));
    if (rgb) HIP_OK(hipMalloc(&stage_rgb, (size_t)want * 3));
    stage_cap = want;
    return true;
}

bool OsnTsdfHipVolume::grow_tiled(int32_t W, int32_t H) {
    const size_t need = (size_t)tiled_width(W) * (size_t)tiled_height(H);
    if (need <= tiled_cap) return true;
    if (d_tiled) hipFree(d_tiled);
    d_tiled = nullptr;
    tiled_cap = 0;
    HIP_OK(hipMalloc(&d_tiled, need * sizeof(float)));
    tiled_cap = need;
    return true;
}

// n bytes from device memory to a pageable host buffer: chunks go by DMA into
// the two pinned halves in turn, and each chunk's copy into dst (on the copy
// pool) runs while the next one's DMA does. Ends with the stream drained.
bool OsnTsdfHipVolume::copy_to_host(void* dst, const void* src, size_t n) {
    if (n == 0) return true;
    if (!pinned) {
        if (hipHostMalloc((void**)&pinned, 2 * kPinnedChunk, hipHostMallocDefault) != hipSuccess) {
            pinned = nullptr;  // no pinned memory: one pageable copy
            HIP_OK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            return true;
        }
        for (auto& e : pinned_ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    CopyPool& pool = CopyPool::get();
    char* out = static_cast<char*>(dst);
    const char* in = static_cast<const char*>(src);
    size_t prev_off = 0, prev_len = 0;
    int k = 0;
    for (size_t off = 0; off < n; off += kPinnedChunk, ++k) {
        const size_t len = std::min(kPinnedChunk, n - off);
        HIP_OK(hipMemcpyAsync(pinned + (k & 1) * kPinnedChunk, in + off, len,
                              hipMemcpyDeviceToHost, stream));
        HIP_OK(hipEventRecord(pinned_ev[k & 1], stream));
        if (k > 0) {
            HIP_OK(hipEventSynchronize(pinned_ev[(k - 1) & 1]));
            pool.copy(out + prev_off, pinned + ((k - 1) & 1) * kPinnedChunk, prev_len);
        }
        prev_off = off;
        prev_len = len;
    }
    HIP_OK(hipEventSynchronize(pinned_ev[(k - 1) & 1]));
    pool.copy(out + prev_off, pinned + ((k - 1) & 1) * kPinnedChunk, prev_len);
    return true;
}

// The update's scratch: per-block parts at the first call, entries grown to
// `entries` (a pass pair's visit workgroups times their region).
bool OsnTsdfHipVolume::grow_scatter(int64_t entries) {
    if (!sc.seg_count) {
        const size_t pool = (size_t)cfg.pool_capacity_blocks;
        HIP_OK(hipMalloc(&sc.segs, pool * kScatterSegs * sizeof(ScatterSeg)));
        HIP_OK(hipMalloc(&sc.seg_count, pool * sizeof(int32_t)));
        sc.list_cap = (int32_t)((pool + kScatterLists - 1) / kScatterLists);
        HIP_OK(hipMalloc(&sc.touched, (size_t)sc.list_cap * kScatterLists * sizeof(int32_t)));
        HIP_OK(hipMalloc(&sc.touched_count, kScatterCtrInts * sizeof(int32_t)));
        HIP_OK(hipMemsetAsync(sc.seg_count, 0, pool * sizeof(int32_t), stream));
        HIP_OK(hipMemsetAsync(sc.touched_count, 0, kScatterCtrInts * sizeof(int32_t), stream));
        sc.parity = 0;
        // without the hint every call takes the larger set
        sc.touched_out = nullptr;
        if (hipHostMalloc((void**)&sc_touched_host, sizeof(int32_t), hipHostMallocMapped) ==
            hipSuccess) {
            *sc_touched_host = 0;
            if (hipHostGetDevicePointer((void**)&sc.touched_out, sc_touched_host, 0) != hipSuccess)
                sc.touched_out = nullptr;
        } else {
            sc_touched_host = nullptr;
        }
    }
    const int64_t want = entries;
    if (want <= sc_entries_cap) return true;
    // the stream may still read the old buffer
    HIP_OK(hipStreamSynchronize(stream));
    if (sc.entries) hipFree(sc.entries);
    sc.entries = nullptr;
    sc_entries_cap = 0;
    HIP_OK(hipMalloc(&sc.entries, (size_t)want * sizeof(uint2)));
    sc_entries_cap = want;
    return true;
}

namespace {

OsnTsdfHipVolume* make_volume(const OsnTsdfHipConfig& c, bool reference_shape) {
    if (!(c.voxel_size_m > 0.0f) || c.pool_capacity_blocks <= 0) return nullptr;
    if (c.storage < kStoreSums || c.storage > kStoreF64) return nullptr;
    uint32_t slots = c.table_slots;
    if (slots == 0) slots = next_pow2((uint32_t)c.pool_capacity_blocks * 2u);
    if (c.hash_kind == kHashTeschnerPow2 && (slots & (slots - 1))) return nullptr;
    if (c.hash_kind != kHashTeschnerPow2 && c.hash_kind != kHashMixRange) return nullptr;
    if (slots < 2) return nullptr;
    if (c.lds_set > 2048 || (c.lds_set > 0 && (c.lds_set & (c.lds_set - 1)))) return nullptr;

    auto* v = new (std::nothrow) OsnTsdfHipVolume();
    if (!v) return nullptr;
    v->cfg = c;
    v->cfg.table_slots = slots;
    v->reference_shape = reference_shape;
    const float trunc = c.trunc_m > 0.0f ? c.trunc_m : 4.0f * c.voxel_size_m;
    VolDev& d = v->d;
    d.capacity = slots;
    d.mask = slots - 1;
    d.hash_kind = c.hash_kind;
    d.probe_limit = c.probe_limit ? (c.probe_limit < slots ? c.probe_limit : slots) : slots;
    d.pool_capacity = c.pool_capacity_blocks;
    d.storage = c.storage;
    d.voxel = c.voxel_size_m;
    d.inv_voxel = 1.0f / c.voxel_size_m;     // volume.cpp:148
    d.trunc = trunc;
    d.weight_cap = c.weight_cap;
    const int32_t cs = (int32_t)std::ceil(trunc * d.inv_voxel);  // volume.cpp:171
    d.steps = cs > 1 ? cs : 1;
    // projective storage: a voxel freezes once its weight reaches the cap. The
    // reference skips an observation when weight >= cap (volume.cpp:289), so
    // unit weights stop at ceil(cap); no cap means the u16 weight's largest
    // value.
    d.w_max = c.weight_cap > 0.0f
                  ? (int32_t)std::ceil(c.weight_cap < 65535.0f ? c.weight_cap : 65535.0f)
                  : 65535;
    if (d.w_max < 1) d.w_max = 1;
    v->lds_set = c.lds_set < 0 ? 1024 : c.lds_set;
    v->update_threads = c.update_threads > 0 ? c.update_threads : 256;
    v->update_grid = c.update_grid;
    v->alloc_grid = c.alloc_grid;
    if (!v->alloc()) {
        delete v;
        return nullptr;
    }
    return v;
}

PointSrc points_src(const float* pos, int32_t n, float cx, float cy, float cz, float radius) {
    PointSrc s{};
    s.pos = pos;
    s.n = n;
    s.cam[0] = cx; s.cam[1] = cy; s.cam[2] = cz;
    s.radius_sq = radius > 0.0f ? radius * radius : 0.0f;
    return s;
}

PointSrc depth_src(const float* depth, const OsnTsdfHipFrame* f) {
    PointSrc s{};
    s.depth = depth;
    s.W = f->width; s.H = f->height;
    s.fx = f->fx; s.fy = f->fy; s.cx = f->cx; s.cy = f->cy;
    std::memcpy(s.R, f->R_wc, sizeof(s.R));
    s.cam[0] = f->t_wc[0]; s.cam[1] = f->t_wc[1]; s.cam[2] = f->t_wc[2];
    s.max_range = f->max_range_m;
    s.radius_sq = f->radius_m > 0.0f ? f->radius_m * f->radius_m : 0.0f;
    return s;
}

ProjCam proj_cam(const float* depth, const OsnTsdfHipFrame* f) {
    ProjCam k{};
    // R_cw = R_wc^T
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            k.Rcw[i * 3 + j] = f->R_wc[j * 3 + i];
            k.Rcw64[i * 3 + j] = f->R_wc64[j * 3 + i];
        }
    for (int i = 0; i < 3; ++i) { k.c[i] = f->t_wc[i]; k.c64[i] = f->t_wc64[i]; }
    k.fx = f->fx; k.fy = f->fy; k.cx = f->cx; k.cy = f->cy;
    k.W = f->width; k.H = f->height;
    k.depth = depth;
    return k;
}

int32_t run_allocate(OsnTsdfHipVolume* v, const PointSrc& src, int mode, bool touched,
                     const int32_t* coords = nullptr) {
    if (v->in_batch) return 3;
    hipEvent_t e0, e1;
    const CallInfo call{(int32_t)(v->calls + 1), (int32_t)(v->calls & 1)};
    v->calls += 1;
    v->stage_begin(0, e0);
    AllocOpts o;
    o.mode = mode;
    o.track_touched = touched;
    o.lds_set = v->lds_set;
    o.grid = v->alloc_grid;
    hipError_t err = launch_alloc(v->d, src, o, coords, call, v->stream);
    v->stage_end(0, e0);
    if (err != hipSuccess) { v->fail("allocate launch", err); return 2; }
    if (mode == kAllocNoCas) return 0;
    v->stage_begin(1, e1);
    err = launch_assign(v->d, call, v->stream);
    v->stage_end(1, e1);
    if (err != hipSuccess) { v->fail("assign launch", err); return 2; }
    return 0;
}

bool valid_frame(const float* d_depth, const OsnTsdfHipFrame* f) {
    return d_depth && f && f->width > 0 && f->height > 0 && f->width <= kMaxFrameDim &&
           f->height <= kMaxFrameDim && f->fx != 0.0f && f->fy != 0.0f;
}

// Synchronise, then read one counter. The head of the counters (block_count,
// fault) is read first: a fault means an index assignment abandoned a bounded
// wait, so the pool order is incomplete and the volume is marked invalid.
template <typename T>
int32_t read_counter(const OsnTsdfHipVolume* v, size_t offset, T* out) {
    if (v->sync() != hipSuccess) return 2;
    int32_t head[2];
    if (hipMemcpy(head, v->d.ctr, sizeof(head), hipMemcpyDeviceToHost) != hipSuccess) return 2;
    if (head[1] != 0) {
        std::fprintf(stderr, "[osn_tsdf_hip] index assignment gave up a wait; volume invalid\n");
        v->ok = false;
        return 2;
    }
    if (offset == offsetof(Counters, block_count) && sizeof(T) == sizeof(int32_t)) {
        std::memcpy(out, head, sizeof(T));
        return 0;
    }
    if (hipMemcpy(out, reinterpret_cast<const char*>(v->d.ctr) + offset, sizeof(T),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return 2;
    return 0;
}

}  // namespace

extern "C" {

OsnTsdfHipVolume* osn_tsdf_hip_create(float voxel_size_m, int32_t pool_capacity_blocks,
                                      float trunc_m, float weight_cap) {
    OsnTsdfHipConfig c{};
    c.voxel_size_m = voxel_size_m;
    c.pool_capacity_blocks = pool_capacity_blocks;
    c.trunc_m = trunc_m;
    c.weight_cap = weight_cap;
    c.storage = kStoreSums;
    c.hash_kind = kHashTeschnerPow2;
    c.lds_set = -1;
    return make_volume(c, true);
}

OsnTsdfHipVolume* osn_tsdf_hip_create_ex(const OsnTsdfHipConfig* cfg) {
    if (!cfg) return nullptr;
    return make_volume(*cfg, false);
}

void osn_tsdf_hip_destroy(OsnTsdfHipVolume* v) { delete v; }

int32_t osn_tsdf_hip_valid(const OsnTsdfHipVolume* v) { return v && v->ok ? 1 : 0; }

int32_t osn_tsdf_hip_device_view(const OsnTsdfHipVolume* v, OsnTsdfDeviceView* out) {
    if (!v || !out || !v->ok) return 1;
    if (v->d.storage != kStoreSums || v->d.hash_kind != kHashTeschnerPow2) return 3;
    out->table = v->d.table;
    out->block_count = &v->d.ctr->block_count;
    out->drop_count = &v->d.ctr->drop_count;
    out->block_coord = v->d.block_coord;
    out->tsdf = v->d.tsdf;
    out->weight = v->d.weight;
    out->r = v->d.r;
    out->g = v->d.g;
    out->b = v->d.b;
    out->hash_mask = v->d.mask;
    out->scratch_base = v->d.capacity;
    out->pool_capacity = v->d.pool_capacity;
    out->block_dim = kBlockDim;
    out->voxel_size_m = v->d.voxel;
    out->trunc_m = v->d.trunc;
    out->weight_cap = v->d.weight_cap;
    return 0;
}

int32_t osn_tsdf_hip_allocate_blocks(OsnTsdfHipVolume* v, const float* d_positions,
                                     int32_t n_points, float cam_x, float cam_y, float cam_z,
                                     float radius_m) {
    if (!v || !d_positions || n_points <= 0) return 1;
    if (!v->ok) return 2;
    return run_allocate(v, points_src(d_positions, n_points, cam_x, cam_y, cam_z, radius_m),
                        kAllocNormal, false);
}

int32_t osn_tsdf_hip_allocate_slot_index(OsnTsdfHipVolume* v, const float* d_positions,
                                         int32_t n_points, float cam_x, float cam_y,
                                         float cam_z, float radius_m) {
    return osn_tsdf_hip_allocate_blocks(v, d_positions, n_points, cam_x, cam_y, cam_z, radius_m);
}

int32_t osn_tsdf_hip_allocate_counting_cas(OsnTsdfHipVolume* v, const float* d_positions,
                                           int32_t n_points, float cam_x, float cam_y,
                                           float cam_z, float radius_m) {
    if (!v || !d_positions || n_points <= 0) return 1;
    if (!v->ok) return 2;
    return run_allocate(v, points_src(d_positions, n_points, cam_x, cam_y, cam_z, radius_m),
                        kAllocCountCas, false);
}

int32_t osn_tsdf_hip_allocate_no_cas(OsnTsdfHipVolume* v, const float* d_positions,
                                     int32_t n_points, float cam_x, float cam_y, float cam_z,
                                     float radius_m) {
    if (!v || !d_positions || n_points <= 0) return 1;
    if (!v->ok) return 2;
    return run_allocate(v, points_src(d_positions, n_points, cam_x, cam_y, cam_z, radius_m),
                        kAllocNoCas, false);
}

int32_t osn_tsdf_hip_allocate_probe_depth(OsnTsdfHipVolume* v, const float* d_positions,
                                          int32_t n_points, float cam_x, float cam_y,
                                          float cam_z, float radius_m, uint64_t* lane_steps,
                                          uint64_t* wave_steps, uint64_t* hist) {
    if (lane_steps) *lane_steps = 0;
    if (wave_steps) *wave_steps = 0;
    if (!v || !d_positions || n_points <= 0) return 1;
    if (!v->ok) return 2;
    const int32_t rc = run_allocate(
        v, points_src(d_positions, n_points, cam_x, cam_y, cam_z, radius_m), kAllocProbe, false);
    if (rc) return rc;
    return osn_tsdf_hip_probe_stats(v, hist, lane_steps, wave_steps, nullptr);
}

int32_t osn_tsdf_hip_update_voxels(OsnTsdfHipVolume* v, const float* d_positions,
                                   int32_t n_points, float cam_x, float cam_y, float cam_z,
                                   float radius_m) {
    if (!v || !d_positions || n_points <= 0) return 1;
    if (!v->ok) return 2;
    if (v->d.storage != kStoreSums) return 3;
    // Chunks of points, each a call of the two-pass update: visits add up
    // and the cap takes against the weight each chunk finds, so the sums and
    // counts are those of one call over all points.
    // Small calls take one sample per lane (kScatterSplitPoints < per).
    const int32_t per = v->scatter_chunk();
    const bool split = n_points <= kScatterSplitPoints;
    const int32_t region = split ? scatter_region_split() : scatter_region(v->d.steps);
    const int32_t first = n_points < per ? n_points : per;
    if (!v->grow_scatter(scatter_visit_grid(first, v->d.steps, split) * region)) {
        v->ok = false;
        return 2;
    }
    v->sc.region = region;
    hipEvent_t e;
    v->stage_begin(2, e);
    hipError_t err = hipSuccess;
    for (int64_t at = 0; at < n_points && err == hipSuccess; at += per) {
        const int32_t m = (int32_t)(n_points - at < per ? n_points - at : per);
        // dense when the last pass pair saw many points per touched block
        const int32_t last = v->sc.touched_out ? __atomic_load_n(v->sc_touched_host, __ATOMIC_RELAXED) : 0;
        const bool dense = last > 0 && (int64_t)m >= (int64_t)kScatterDensePoints * last;
        err = launch_scatter_update(
            v->d, points_src(d_positions + (size_t)at * 3, m, cam_x, cam_y, cam_z, radius_m),
            v->sc, dense, split, v->stream);
        v->sc.parity ^= 1;
    }
    v->stage_end(2, e);
    if (err != hipSuccess) { v->fail("update launch", err); return 2; }
    return 0;
}

namespace {

// The allocate walk also leaves the volume's tiled copy of the depth image,
// which an update of the same frame reads (integrate_depth); a separate
// update_depth call tiles its own image first.
int32_t allocate_depth_tiled(OsnTsdfHipVolume* v, const float* d_depth,
                             const OsnTsdfHipFrame* frame) {
    if (!v || !valid_frame(d_depth, frame)) return 1;
    if (!v->ok) return 2;
    if (!v->in_batch && !v->grow_tiled(frame->width, frame->height)) return 2;
    PointSrc src = depth_src(d_depth, frame);
    src.tiled = v->d_tiled;
    return run_allocate(v, src, kAllocNormal, true);
}

int32_t update_depth_tiled(OsnTsdfHipVolume* v, const float* d_depth,
                           const OsnTsdfHipFrame* frame, bool tile_first) {
    if (!v || !valid_frame(d_depth, frame)) return 1;
    if (!v->ok) return 2;
    if (v->d.storage == kStoreSums) return 3;
    if (v->calls == 0 || v->in_batch) return 3;
    if (!v->grow_tiled(frame->width, frame->height)) return 2;
    hipEvent_t e;
    v->stage_begin(2, e);
    hipError_t err = hipSuccess;
    if (tile_first) err = launch_tile_depth(d_depth, frame->width, frame->height, v->d_tiled, v->stream);
    const CallInfo last{(int32_t)v->calls, (int32_t)((v->calls - 1) & 1)};
    if (err == hipSuccess)
        err = launch_proj_update(v->d, proj_cam(v->d_tiled, frame), last, v->update_grid,
                                 v->update_threads, v->stream);
    v->stage_end(2, e);
    if (err != hipSuccess) { v->fail("projective update launch", err); return 2; }
    return 0;
}

}  // namespace

int32_t osn_tsdf_hip_allocate_depth(OsnTsdfHipVolume* v, const float* d_depth,
                                    const OsnTsdfHipFrame* frame) {
    return allocate_depth_tiled(v, d_depth, frame);
}

int32_t osn_tsdf_hip_update_depth(OsnTsdfHipVolume* v, const float* d_depth,
                                  const OsnTsdfHipFrame* frame) {
    return update_depth_tiled(v, d_depth, frame, true);
}

int32_t osn_tsdf_hip_integrate_depth(OsnTsdfHipVolume* v, const float* d_depth,
                                     const OsnTsdfHipFrame* frame) {
    const int32_t rc = allocate_depth_tiled(v, d_depth, frame);
    if (rc) return rc;
    return update_depth_tiled(v, d_depth, frame, false);
}

int32_t osn_tsdf_hip_backproject(const float* d_depth, const OsnTsdfHipFrame* frame,
                                 float* d_points_out) {
    if (!valid_frame(d_depth, frame) || !d_points_out) return 1;
    const hipError_t err = launch_backproject(depth_src(d_depth, frame), d_points_out, nullptr);
    if (err != hipSuccess) return 2;
    return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

int32_t osn_tsdf_hip_insert_blocks(OsnTsdfHipVolume* v, const int32_t* d_coords, int32_t n) {
    return osn_tsdf_hip_insert_blocks_mode(v, d_coords, n, 0);
}

int32_t osn_tsdf_hip_insert_blocks_mode(OsnTsdfHipVolume* v, const int32_t* d_coords, int32_t n,
                                        int32_t mode) {
    if (!v || !d_coords || n <= 0 || mode < 0 || mode > 2) return 1;
    if (!v->ok) return 2;
    PointSrc s{};
    s.n = n;
    const int m = mode == 1 ? kAllocCountCas : (mode == 2 ? kAllocProbe : kAllocNormal);
    return run_allocate(v, s, m, false, d_coords);
}

int32_t osn_tsdf_hip_cas_attempts(const OsnTsdfHipVolume* v, uint64_t* out) {
    if (!v || !out || !v->ok) return 1;
    unsigned long long n = 0;
    if (read_counter(v, offsetof(Counters, cas_attempts), &n)) return 2;
    *out = n;
    return 0;
}

int32_t osn_tsdf_hip_cas_ceiling(int32_t log2_slots, int32_t reps, double* ms_per_op_batch) {
    if (log2_slots < 10 || log2_slots > 30 || reps <= 0 || !ms_per_op_batch) return 1;
    const uint32_t n = 1u << log2_slots;
    unsigned long long* slots = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s = nullptr;
    int32_t rc = 0;
    if (hipMalloc(&slots, (size_t)n * 8) != hipSuccess) return 2;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        rc = 2;
    } else {
        float total = 0.0f;
        for (int r = 0; r <= reps && rc == 0; ++r) {  // r == 0 is a warmup
            if (hipMemsetAsync(slots, 0xff, (size_t)n * 8, s) != hipSuccess) { rc = 2; break; }
            hipEventRecord(a, s);
            if (launch_cas_ceiling(slots, n - 1, (int32_t)n, s) != hipSuccess) { rc = 2; break; }
            hipEventRecord(b, s);
            float ms = 0.0f;
            if (hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(&ms, a, b) != hipSuccess) { rc = 2; break; }
            if (r > 0) total += ms;
        }
        *ms_per_op_batch = total / reps;
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (s) hipStreamDestroy(s);
    hipFree(slots);
    return rc;
}

int32_t osn_tsdf_hip_extract_mesh_device(OsnTsdfHipVolume* v, float* d_posnor, uint8_t* d_rgb,
                                         int32_t vert_cap, float min_weight, float iso,
                                         int32_t* d_count) {
    if (!v || !d_posnor || vert_cap <= 0 || !d_count) return 1;
    if (!v->ok) return 2;
    // a batch writes this volume on its own stream: let that work finish first
    if (v->batch_stream && hipStreamSynchronize(v->batch_stream) != hipSuccess) return 2;
    if (!v->grow_extract(v->d.pool_capacity)) { v->ok = false; return 2; }
    ExtractArgs a{};
    a.min_weight = min_weight;
    a.iso = iso;
    a.n_blocks = -1;  // read from the device: no host synchronisation
    a.block_vcount = v->ex_vcount;
    a.block_voff = v->ex_voff;
    a.block_nb = v->ex_nb;
    a.nb_cache = 1;
    a.chunk_sums = v->ex_chunks;
    a.scan_flag = v->ex_flag_count;
    a.posnor = d_posnor;
    a.rgb = d_rgb;
    a.cap = vert_cap;
    a.count_out = d_count;
    hipEvent_t e;
    v->stage_begin(3, e);
    hipError_t err = launch_extract(v->d, a, false, v->stream);
    if (err == hipSuccess) err = launch_extract(v->d, a, true, v->stream);
    v->stage_end(3, e);
    if (err != hipSuccess) { v->fail("extract launch", err); return 2; }
    return 0;
}

}  // extern "C"

// ---- extraction that allocates nothing ----------------------------------
//
// The per-block scratch of a device-resident extraction (neighbour indices,
// vertex counts, offsets, chunk sums, the scan's arrival counter) for volumes
// of up to OSN_TSDF_HIP_NOALLOC_MAX_BLOCKS pool blocks, in one file-scope
// device array. Each device that loads the library holds its own copy until
// the process ends; every no-allocation call on a device shares it.
namespace {

constexpr size_t kNaBlocks = OSN_TSDF_HIP_NOALLOC_MAX_BLOCKS;
constexpr size_t kNaNbOff = 0;                                  // 8 x int32 per block
constexpr size_t kNaCountOff = kNaNbOff + kNaBlocks * 8 * 4;    // int32 per block
constexpr size_t kNaVoffOff = kNaCountOff + kNaBlocks * 4;      // int32 per block
constexpr size_t kNaChunkOff = kNaVoffOff + kNaBlocks * 4;      // int32 per 1024 blocks, + 2
constexpr size_t kNaFlagOff =
    (kNaChunkOff + (kNaBlocks / kScanChunk + 2) * 4 + 255) / 256 * 256;  // 2 x int32
constexpr size_t kNaBytes = kNaFlagOff + 256;
static_assert(kNaBytes == OSN_TSDF_HIP_NOALLOC_SCRATCH_BYTES, "header states the scratch size");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// The array's address on each device, looked up once with that device
// current: in host code `scratch` only names the symbol, it is not a device
// address.
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
unsigned char* g_na_ptr[kNaMaxDevices] = {};

unsigned char* noalloc_scratch(int device) {
    if (device < 0 || device >= kNaMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (g_na_ptr[device]) return g_na_ptr[device];
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (cur != device && hipSetDevice(device) != hipSuccess) return nullptr;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) hipSetDevice(cur);
    if (e != hipSuccess || !ptr) {
        std::fprintf(stderr, "[osn_tsdf_hip] scratch symbol lookup failed: %s\n",
                     hipGetErrorString(e));
        return nullptr;
    }
    g_na_ptr[device] = static_cast<unsigned char*>(ptr);
    return g_na_ptr[device];
}

}  // namespace

extern "C" {

int32_t osn_tsdf_hip_extract_mesh_device_noalloc(OsnTsdfHipVolume* v, float* d_posnor,
                                                 uint8_t* d_rgb, int32_t vert_cap,
                                                 float min_weight, float iso,
                                                 int32_t* d_count) {
    if (!v || !d_posnor || vert_cap <= 0 || !d_count) return 1;
    if (!v->ok) return 2;
    if ((size_t)v->d.pool_capacity > kNaBlocks) {
        std::fprintf(stderr,
                     "[osn_tsdf_hip] extract_mesh_device_noalloc: pool of %d blocks exceeds the "
                     "%zu-block scratch\n", v->d.pool_capacity, kNaBlocks);
        return 4;
    }
    unsigned char* base = noalloc_scratch(v->device);
    if (!base) return 2;
    // a batch writes this volume on its own stream: let that work finish first
    if (v->batch_stream && hipStreamSynchronize(v->batch_stream) != hipSuccess) return 2;
    ExtractArgs a{};
    a.min_weight = min_weight;
    a.iso = iso;
    a.n_blocks = -1;  // read from the device: no host synchronisation
    a.block_nb = reinterpret_cast<int32_t*>(base + kNaNbOff);
    a.block_vcount = reinterpret_cast<int32_t*>(base + kNaCountOff);
    a.block_voff = reinterpret_cast<int32_t*>(base + kNaVoffOff);
    a.chunk_sums = reinterpret_cast<int32_t*>(base + kNaChunkOff);
    a.scan_flag = reinterpret_cast<int32_t*>(base + kNaFlagOff);
    a.nb_cache = 0;    // the scratch holds another call's indices: look every neighbour up
    a.clear_flag = 1;  // and start the arrival counter at zero
    a.posnor = d_posnor;
    a.rgb = d_rgb;
    a.cap = vert_cap;
    a.count_out = d_count;
    hipError_t err = launch_extract(v->d, a, false, v->stream);
    if (err == hipSuccess) err = launch_extract(v->d, a, true, v->stream);
    if (err != hipSuccess) { v->fail("extract launch", err); return 2; }
    return 0;
}

int32_t osn_tsdf_hip_extract_mesh(OsnTsdfHipVolume* v, float* buffer_posnor, uint8_t* buffer_rgb,
                                  int32_t vert_cap, float min_weight, float iso) {
    if (!v || !buffer_posnor || vert_cap <= 0) return -1;
    if (!v->ok) return -1;
    const int32_t n_blocks = osn_tsdf_hip_block_count(v);  // synchronises
    if (n_blocks <= 0) return 0;
    if (!v->grow_extract(n_blocks)) { v->ok = false; return -1; }
    ExtractArgs a{};
    a.min_weight = min_weight;
    a.iso = iso;
    a.n_blocks = n_blocks;
    a.block_vcount = v->ex_vcount;
    a.block_voff = v->ex_voff;
    a.block_nb = v->ex_nb;
    a.nb_cache = 1;
    a.chunk_sums = v->ex_chunks;
    a.scan_flag = v->ex_flag_count;
    a.cap = vert_cap;
    a.count_out = v->ex_flag_count + 1;
    hipEvent_t e;
    v->stage_begin(3, e);
    hipError_t err = launch_extract(v->d, a, false, v->stream);
    if (err != hipSuccess) { v->fail("extract count launch", err); return -1; }
    int32_t total = 0;
    if (hipStreamSynchronize(v->stream) != hipSuccess ||
        hipMemcpy(&total, a.count_out, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) {
        v->ok = false;
        return -1;
    }
    if (total <= 0) { v->stage_end(3, e); return total; }  // 0, or -1 on overflow
    if (!v->grow_stage(total, buffer_rgb != nullptr)) { v->ok = false; return -1; }
    a.posnor = v->stage_posnor;
    a.rgb = buffer_rgb ? v->stage_rgb : nullptr;
    err = launch_extract(v->d, a, true, v->stream);
    v->stage_end(3, e);
    if (err != hipSuccess) { v->fail("extract emit launch", err); return -1; }
    if (!v->copy_to_host(buffer_posnor, v->stage_posnor, (size_t)total * 6 * sizeof(float)))
        return -1;
    if (buffer_rgb && !v->copy_to_host(buffer_rgb, v->stage_rgb, (size_t)total * 3)) return -1;
    return total;
}

int32_t osn_tsdf_hip_reset(OsnTsdfHipVolume* v) {
    if (!v) return 1;
    return v->reset() ? 0 : 2;
}

int32_t osn_tsdf_hip_block_count(const OsnTsdfHipVolume* v) {
    if (!v || !v->ok) return 0;
    int32_t n = 0;
    if (read_counter(v, offsetof(Counters, block_count), &n)) return 0;
    return n;
}

uint64_t osn_tsdf_hip_drop_count(const OsnTsdfHipVolume* v) {
    if (!v || !v->ok) return 0;
    unsigned long long n = 0;
    if (read_counter(v, offsetof(Counters, drop_count), &n)) return 0;
    return n;
}

int32_t osn_tsdf_hip_touched_count(const OsnTsdfHipVolume* v) {
    if (!v || !v->ok) return 0;
    int32_t n = 0;
    if (v->calls == 0) return 0;
    const size_t off = offsetof(Counters, lists) + sizeof(ListCounters) * ((v->calls - 1) & 1) +
                       offsetof(ListCounters, touched_count);
    if (read_counter(v, off, &n)) return 0;
    return n;
}

void osn_tsdf_hip_synchronize(const OsnTsdfHipVolume* v) {
    if (v && v->stream) v->sync();
}

int32_t osn_tsdf_hip_block_keys(const OsnTsdfHipVolume* v, int64_t* host_keys, int32_t cap) {
    if (!v || !host_keys || cap < 0 || !v->ok) return -1;
    const int32_t n_all = osn_tsdf_hip_block_count(v);
    const int32_t n = n_all < cap ? n_all : cap;
    if (n <= 0) return 0;
    std::vector<int32_t> c((size_t)n * 3);
    if (hipMemcpy(c.data(), v->d.block_coord, c.size() * sizeof(int32_t),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    for (int32_t i = 0; i < n; ++i) host_keys[i] = pack_coord(c[i * 3], c[i * 3 + 1], c[i * 3 + 2]);
    return n;
}

int32_t osn_tsdf_hip_read_voxels(const OsnTsdfHipVolume* v, int32_t first, int32_t n,
                                 float* host_tsdf, float* host_weight) {
    if (!v || !v->ok || first < 0 || n <= 0 || first + n > v->d.pool_capacity) return 1;
    if (v->sync() != hipSuccess) return 2;
    const size_t o = (size_t)first * kBlockVoxels;
    const size_t cnt = (size_t)n * kBlockVoxels;
    if (v->d.storage == kStoreSums) {
        if (host_tsdf && hipMemcpy(host_tsdf, v->d.tsdf + o, cnt * 4, hipMemcpyDeviceToHost)) return 2;
        if (host_weight && hipMemcpy(host_weight, v->d.weight + o, cnt * 4, hipMemcpyDeviceToHost)) return 2;
        return 0;
    }
    if (v->d.storage == kStoreF32) {
        std::vector<float> raw(cnt * 2);
        if (hipMemcpy(raw.data(), static_cast<float*>(v->d.vox) + o * 2, cnt * 8,
                      hipMemcpyDeviceToHost))
            return 2;
        for (size_t i = 0; i < cnt; ++i) {
            if (host_tsdf) host_tsdf[i] = raw[i * 2];
            if (host_weight) host_weight[i] = raw[i * 2 + 1];
        }
        return 0;
    }
    if (v->d.storage == kStoreF16) {
        std::vector<uint32_t> raw(cnt);
        if (hipMemcpy(raw.data(), static_cast<uint32_t*>(v->d.vox) + o, cnt * 4,
                      hipMemcpyDeviceToHost))
            return 2;
        for (size_t i = 0; i < cnt; ++i) {
            // IEEE binary16 -> float, exact
            const uint32_t h = raw[i] & 0xffffu;
            const uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 31u, m = h & 1023u;
            float f;
            if (e == 0) f = std::ldexp((float)m, -24);
            else if (e == 31) f = m ? NAN : INFINITY;
            else f = std::ldexp((float)(m | 1024u), (int)e - 25);
            if (host_tsdf) host_tsdf[i] = s ? -f : f;
            if (host_weight) host_weight[i] = (float)(raw[i] >> 16);
        }
        return 0;
    }
    std::vector<double> raw(cnt * 2);
    if (hipMemcpy(raw.data(), static_cast<double*>(v->d.vox) + o * 2, cnt * 16,
                  hipMemcpyDeviceToHost))
        return 2;
    for (size_t i = 0; i < cnt; ++i) {
        if (host_tsdf) host_tsdf[i] = (float)raw[i * 2];
        if (host_weight) host_weight[i] = (float)raw[i * 2 + 1];
    }
    return 0;
}

int32_t osn_tsdf_hip_read_voxels_f64(const OsnTsdfHipVolume* v, int32_t first, int32_t n,
                                     double* host_tsdf, double* host_weight) {
    if (!v || !v->ok || first < 0 || n <= 0 || first + n > v->d.pool_capacity) return 1;
    if (v->d.storage != kStoreF64) return 3;
    if (v->sync() != hipSuccess) return 2;
    const size_t cnt = (size_t)n * kBlockVoxels;
    std::vector<double> raw(cnt * 2);
    if (hipMemcpy(raw.data(), static_cast<double*>(v->d.vox) + (size_t)first * kBlockVoxels * 2,
                  cnt * 16, hipMemcpyDeviceToHost))
        return 2;
    for (size_t i = 0; i < cnt; ++i) {
        if (host_tsdf) host_tsdf[i] = raw[i * 2];
        if (host_weight) host_weight[i] = raw[i * 2 + 1];
    }
    return 0;
}

int32_t osn_tsdf_hip_probe_stats(const OsnTsdfHipVolume* v, uint64_t* hist64,
                                 uint64_t* lane_steps, uint64_t* wave_steps, uint64_t* lookups) {
    if (!v || !v->ok) return 1;
    Counters c;
    if (v->sync() != hipSuccess) return 2;
    if (hipMemcpy(&c, v->d.ctr, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) return 2;
    if (hist64) for (int i = 0; i < kProbeHistBins; ++i) hist64[i] = c.probe_hist[i];
    if (lane_steps) *lane_steps = c.lane_steps;
    if (wave_steps) *wave_steps = c.wave_steps;
    if (lookups) *lookups = c.lookups;
    return 0;
}

int32_t osn_tsdf_hip_set_timing(OsnTsdfHipVolume* v, int32_t enable) {
    if (!v) return 1;
    v->timing = enable != 0;
    return 0;
}

int32_t osn_tsdf_hip_read_timing(OsnTsdfHipVolume* v, double* ms4, int64_t* calls4) {
    if (!v) return 1;
    if (hipStreamSynchronize(v->stream) != hipSuccess) return 2;
    for (auto& p : v->ev_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            v->ms[p.first] += ms;
            v->stage_calls[p.first] += 1;
        }
        v->ev_pool.push_back(p.second.first);
        v->ev_pool.push_back(p.second.second);
    }
    v->ev_pending.clear();
    for (int i = 0; i < kStages; ++i) {
        if (ms4) ms4[i] = v->ms[i];
        if (calls4) calls4[i] = v->stage_calls[i];
        v->ms[i] = 0.0;
        v->stage_calls[i] = 0;
    }
    return 0;
}

void* osn_tsdf_hip_stream(const OsnTsdfHipVolume* v) { return v ? (void*)v->stream : nullptr; }

}  // extern "C"

// ---- batches ---------------------------------------------------------------

struct OsnTsdfHipBatch {
    std::vector<OsnTsdfHipVolume*> vols;
    int32_t storage = kStoreF16;
    hipStream_t stream = nullptr;
    VolDev* d_vols = nullptr;
    int32_t* d_call_base = nullptr;
    int32_t* d_tile_prefix = nullptr;
    int32_t* d_step = nullptr;
    SceneStep* d_steps = nullptr;
    UpdParams* d_upd = nullptr;
    int32_t n_steps = 0;
    int32_t cursor = 0;
    // the device step counter equals cursor (a graph run left it there); an
    // eager run or a new schedule leaves it stale
    bool step_on_device = false;
    int32_t total_tiles = 0;
    uint32_t max_capacity = 0;
    uint32_t max_pool_capacity = 0;
    int32_t max_w_max = 0;  // the largest projective weight cap of the volumes
    std::vector<int64_t> base_calls;
    int32_t alloc_grid = 0, update_grid = 0, update_threads = 256, lds_set = 1024;
    int32_t alloc_mode = kAllocNormal;  // instrumentation of the allocate kernel
    int32_t depth_sub = kBatchDepthSub;  // allocate work-unit size, set with the schedule
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // mesh extraction every `extract_every` steps (0: never), per scene into
    // caller-owned device buffers; steps that extract replay their own graph
    std::vector<ExtractArgs> ex_args;
    ExtractArgs* d_ex_args = nullptr;  // the same, on the device, for the batched launches
    int32_t extract_every = 0;
    int32_t extract_grid = kExtractGridPerScene;
    hipGraph_t graph_x = nullptr;
    hipGraphExec_t exec_x = nullptr;
    bool timing = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double ms[kStages] = {0, 0, 0, 0};
    int64_t stage_calls[kStages] = {0, 0, 0, 0};

    // Runs of kGraphSteps steps alike in extraction are replayed from one
    // graph of that many steps (index: with extraction), so the per-launch
    // cost of a graph is paid once per run rather than once per step.
    static constexpr int kGraphSteps = 8;
    hipGraph_t graph_m[2] = {nullptr, nullptr};
    hipGraphExec_t exec_m[2] = {nullptr, nullptr};

    void drop_graph() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (exec_x) hipGraphExecDestroy(exec_x);
        if (graph_x) hipGraphDestroy(graph_x);
        exec = nullptr;
        graph = nullptr;
        exec_x = nullptr;
        graph_x = nullptr;
        for (int i = 0; i < 2; ++i) {
            if (exec_m[i]) hipGraphExecDestroy(exec_m[i]);
            if (graph_m[i]) hipGraphDestroy(graph_m[i]);
            exec_m[i] = nullptr;
            graph_m[i] = nullptr;
        }
    }
    // Whether steps first .. first + kGraphSteps - 1 all extract (x) or none does.
    bool uniform_run(int32_t first, bool& x) const {
        x = extracts_after(first);
        for (int j = 1; j < kGraphSteps; ++j)
            if (extracts_after(first + j) != x) return false;
        return true;
    }
    bool extracts_after(int32_t step) const {
        return extract_every > 0 && (step + 1) % extract_every == 0;
    }
    // Count, scan and emit every scene's mesh (device-side block counts, no
    // host synchronisation): one launch per pass for the whole batch.
    hipError_t enqueue_extract() {
        hipEvent_t e[2] = {};
        if (timing && hipEventCreate(&e[0]) == hipSuccess) hipEventRecord(e[0], stream);
        const int32_t n = (int32_t)vols.size();
        hipError_t err = launch_extract_batch(d_vols, d_ex_args, n, storage, max_pool_capacity,
                                              extract_grid, false, stream);
        if (err == hipSuccess)
            err = launch_extract_batch(d_vols, d_ex_args, n, storage, max_pool_capacity,
                                       extract_grid, true, stream);
        if (timing && hipEventCreate(&e[1]) == hipSuccess) {
            hipEventRecord(e[1], stream);
            ev_pending.push_back({3, {e[0], e[1]}});
        }
        return err;
    }
    // Capture one step (and the extraction after it when `with_extract`) as
    // a graph whose kernels read the step number from d_step.
    bool capture(bool with_extract, hipGraph_t& g_out, hipGraphExec_t& e_out, int steps = 1) {
        const bool t = timing;
        timing = false;
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            timing = t;
            return false;
        }
        // every step reads its index from the device counter the update advances
        hipError_t err = hipSuccess;
        for (int i = 0; i < steps && err == hipSuccess; ++i) {
            err = enqueue_step(dev(0, true));
            if (err == hipSuccess && with_extract) err = enqueue_extract();
        }
        hipGraph_t g = nullptr;
        const hipError_t end = hipStreamEndCapture(stream, &g);
        timing = t;
        if (err != hipSuccess || end != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return false;
        }
        if (hipGraphInstantiate(&e_out, g, nullptr, nullptr, 0) != hipSuccess) {
            hipGraphDestroy(g);
            return false;
        }
        g_out = g;
        return true;
    }
    BatchDev dev(int step, bool from_device) const {
        BatchDev b{};
        b.vols = d_vols;
        b.steps = d_steps;
        b.tile_prefix = d_tile_prefix;
        b.step_ctr = from_device ? d_step : nullptr;
        b.step_advance = from_device ? d_step : nullptr;  // the update kernel advances it
        b.call_base = d_call_base;
        b.upd = d_upd;
        b.step = step;
        b.n_scenes = (int32_t)vols.size();
        b.depth_sub = depth_sub;
        return b;
    }
    hipError_t enqueue_step(const BatchDev& b) {
        AllocOpts o;
        o.lds_set = lds_set;
        o.grid = alloc_grid;
        o.mode = alloc_mode;
        // With timing on, each stage's kernel records its own start and end
        // (stage i: events 2i, 2i + 1).
        hipEvent_t e[6] = {};
        LaunchEvents ev[3];
        if (timing) {
            bool made = true;
            for (int i = 0; i < 6 && made; ++i) made = hipEventCreate(&e[i]) == hipSuccess;
            if (made)
                for (int i = 0; i < 3; ++i) ev[i] = LaunchEvents{e[2 * i], e[2 * i + 1]};
        }
        hipError_t err = launch_alloc_batch(b, total_tiles, o, stream, ev[0]);
        if (err == hipSuccess) err = launch_assign_batch(b, max_capacity, stream, ev[1]);
        if (err == hipSuccess)
            err = launch_proj_update_batch(b, storage, update_grid, update_threads, stream, ev[2],
                                           max_w_max);
        if (timing)
            for (int i = 0; i < 3; ++i) ev_pending.push_back({i, {e[2 * i], e[2 * i + 1]}});
        return err;
    }
    ~OsnTsdfHipBatch() {
        if (stream) hipStreamSynchronize(stream);
        drop_graph();
        for (auto& p : ev_pending) { hipEventDestroy(p.second.first); hipEventDestroy(p.second.second); }
        for (void* p : {(void*)d_vols, (void*)d_call_base, (void*)d_tile_prefix, (void*)d_step,
                        (void*)d_steps, (void*)d_upd, (void*)d_ex_args})
            if (p) hipFree(p);
        if (stream) hipStreamDestroy(stream);
        for (auto* v : vols) {
            v->in_batch = false;
            v->batch_stream = nullptr;
        }
    }
};

extern "C" {

OsnTsdfHipBatch* osn_tsdf_hip_batch_create(OsnTsdfHipVolume* const* volumes, int32_t n) {
    if (!volumes || n <= 0 || n > kMaxBatchScenes) return nullptr;
    for (int i = 0; i < n; ++i) {
        if (!volumes[i] || !volumes[i]->ok || volumes[i]->in_batch) return nullptr;
        if (volumes[i]->d.storage == kStoreSums) return nullptr;
        if (volumes[i]->d.storage != volumes[0]->d.storage) return nullptr;
        for (int j = 0; j < i; ++j)
            if (volumes[j] == volumes[i]) return nullptr;
    }
    auto* b = new (std::nothrow) OsnTsdfHipBatch();
    if (!b) return nullptr;
    b->storage = volumes[0]->d.storage;
    std::vector<VolDev> hv;
    for (int i = 0; i < n; ++i) {
        OsnTsdfHipVolume* v = volumes[i];
        hipStreamSynchronize(v->stream);
        b->vols.push_back(v);
        hv.push_back(v->d);
        if (v->d.capacity > b->max_capacity) b->max_capacity = v->d.capacity;
        if ((uint32_t)v->d.pool_capacity > b->max_pool_capacity)
            b->max_pool_capacity = (uint32_t)v->d.pool_capacity;
        if (v->d.w_max > b->max_w_max) b->max_w_max = v->d.w_max;
    }
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&b->d_vols, sizeof(VolDev) * n) == hipSuccess &&
              hipMalloc(&b->d_call_base, sizeof(int32_t) * n) == hipSuccess &&
              hipMalloc(&b->d_tile_prefix, sizeof(int32_t) * (n + 1)) == hipSuccess &&
              hipMalloc(&b->d_step, sizeof(int32_t)) == hipSuccess &&
              hipMalloc(&b->d_upd, sizeof(UpdParams) * n) == hipSuccess &&
              hipMemcpy(b->d_vols, hv.data(), sizeof(VolDev) * n, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) { delete b; return nullptr; }
    for (auto* v : b->vols) {
        v->in_batch = true;
        v->batch_stream = b->stream;
    }
    return b;
}

void osn_tsdf_hip_batch_destroy(OsnTsdfHipBatch* b) { delete b; }

int32_t osn_tsdf_hip_batch_set_schedule(OsnTsdfHipBatch* b, int32_t n_steps,
                                        const float* const* d_depths,
                                        const OsnTsdfHipFrame* frames) {
    if (!b || n_steps <= 0 || !d_depths || !frames) return 1;
    const int n = (int)b->vols.size();
    if (hipStreamSynchronize(b->stream) != hipSuccess) return 2;
    b->drop_graph();
    std::vector<SceneStep> hs((size_t)n_steps * n);
    std::vector<int32_t> tiles(n, 0);
    // the work-unit size: 2x2 tiles unless that leaves the launch short of units
    std::vector<int32_t> widest(n, 0);
    for (int k = 0; k < n_steps; ++k)
        for (int s = 0; s < n; ++s) {
            const size_t i = (size_t)k * n + s;
            if (!d_depths[i]) continue;
            if (!valid_frame(d_depths[i], &frames[i])) return 1;
            widest[s] = std::max(widest[s], depth_units(frames[i].width, frames[i].height,
                                                        kBatchDepthSub));
        }
    int64_t units = 0;
    for (int s = 0; s < n; ++s) units += widest[s];
    const int32_t sub = units < kBatchMinUnits ? 1 : kBatchDepthSub;
    for (int k = 0; k < n_steps; ++k)
        for (int s = 0; s < n; ++s) {
            const size_t i = (size_t)k * n + s;
            SceneStep& st = hs[i];
            st.active = 0;
            if (!d_depths[i]) continue;
            if (!valid_frame(d_depths[i], &frames[i])) return 1;
            st.active = 1;
            st.src = depth_src(d_depths[i], &frames[i]);
            st.cam = proj_cam(nullptr, &frames[i]);  // tiled copy, set below
            st.tiles = depth_units(frames[i].width, frames[i].height, sub);
            if (st.tiles > tiles[s]) tiles[s] = st.tiles;
        }
    // Each volume's tiled depth copy is sized for its largest frame before any
    // pointer to it is recorded; allocate writes it and update reads it in the
    // same step.
    for (size_t i = 0; i < hs.size(); ++i)
        if (hs[i].active && !b->vols[i % n]->grow_tiled(hs[i].src.W, hs[i].src.H)) return 2;
    for (size_t i = 0; i < hs.size(); ++i) {
        hs[i].src.tiled = b->vols[i % n]->d_tiled;
        hs[i].cam.depth = b->vols[i % n]->d_tiled;
    }
    std::vector<int32_t> prefix(n + 1, 0);
    for (int s = 0; s < n; ++s) prefix[s + 1] = prefix[s] + tiles[s];
    b->base_calls.assign(n, 0);
    std::vector<int32_t> base(n);
    for (int s = 0; s < n; ++s) {
        b->base_calls[s] = b->vols[s]->calls;
        base[s] = (int32_t)b->vols[s]->calls;
    }
    if (b->d_steps) hipFree(b->d_steps);
    b->d_steps = nullptr;
    if (hipMalloc(&b->d_steps, sizeof(SceneStep) * hs.size()) != hipSuccess) return 2;
    if (hipMemcpy(b->d_steps, hs.data(), sizeof(SceneStep) * hs.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(b->d_tile_prefix, prefix.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) ||
        hipMemcpy(b->d_call_base, base.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice))
        return 2;
    b->n_steps = n_steps;
    b->cursor = 0;
    b->step_on_device = false;
    b->total_tiles = prefix[n];
    b->depth_sub = sub;
    return 0;
}

int32_t osn_tsdf_hip_batch_run(OsnTsdfHipBatch* b, int32_t n_steps, int32_t use_graph) {
    if (!b || n_steps <= 0 || !b->d_steps) return 1;
    if (b->cursor + n_steps > b->n_steps) return 1;
    if (!use_graph) {
        b->step_on_device = false;
        for (int i = 0; i < n_steps; ++i) {
            if (b->enqueue_step(b->dev(b->cursor + i, false)) != hipSuccess) return 2;
            if (b->extracts_after(b->cursor + i) && b->enqueue_extract() != hipSuccess) return 2;
        }
    } else {
        const int32_t first = b->cursor;
        // the graphs' steps read their index from the device counter, which
        // their update kernels advance: set it unless a graph run left it here
        if (!b->step_on_device &&
            (hipMemcpyAsync(b->d_step, &first, sizeof(int32_t), hipMemcpyHostToDevice, b->stream) ||
             hipStreamSynchronize(b->stream)))
            return 2;
        b->step_on_device = false;
        if (!b->exec && !b->capture(false, b->graph, b->exec)) { b->drop_graph(); return 2; }
        if (b->extract_every > 0 && !b->exec_x &&
            !b->capture(true, b->graph_x, b->exec_x)) {
            b->drop_graph();
            return 2;
        }
        // the runs of kGraphSteps alike steps this call replays: one graph each
        for (int i = 0; i + OsnTsdfHipBatch::kGraphSteps <= n_steps;) {
            bool x = false;
            if (!b->uniform_run(first + i, x)) { ++i; continue; }
            if (!b->exec_m[x] &&
                !b->capture(x, b->graph_m[x], b->exec_m[x], OsnTsdfHipBatch::kGraphSteps)) {
                b->drop_graph();
                return 2;
            }
            i += OsnTsdfHipBatch::kGraphSteps;
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (b->timing) { hipEventCreate(&e0); hipEventRecord(e0, b->stream); }
        for (int i = 0; i < n_steps;) {
            bool x = false;
            if (i + OsnTsdfHipBatch::kGraphSteps <= n_steps && b->uniform_run(first + i, x)) {
                if (hipGraphLaunch(b->exec_m[x], b->stream) != hipSuccess) return 2;
                i += OsnTsdfHipBatch::kGraphSteps;
                continue;
            }
            hipGraphExec_t g = b->extracts_after(first + i) ? b->exec_x : b->exec;
            if (hipGraphLaunch(g, b->stream) != hipSuccess) return 2;
            ++i;
        }
        if (b->timing) {
            hipEventCreate(&e1);
            hipEventRecord(e1, b->stream);
            b->ev_pending.push_back({0, {e0, e1}});
        }
        b->step_on_device = true;
    }
    b->cursor += n_steps;
    for (size_t s = 0; s < b->vols.size(); ++s) b->vols[s]->calls = b->base_calls[s] + b->cursor;
    return 0;
}

void osn_tsdf_hip_batch_synchronize(const OsnTsdfHipBatch* b) {
    if (b && b->stream) hipStreamSynchronize(b->stream);
}

int32_t osn_tsdf_hip_batch_tune(OsnTsdfHipBatch* b, int32_t alloc_grid, int32_t update_grid,
                                int32_t update_threads, int32_t lds_set) {
    if (!b) return 1;
    if (update_threads && update_threads != 128 && update_threads != 256 && update_threads != 512)
        return 1;
    if (lds_set > 1024 || (lds_set > 0 && (lds_set & (lds_set - 1)))) return 1;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return 2;
    b->drop_graph();
    if (alloc_grid >= 0) b->alloc_grid = alloc_grid;
    if (update_grid >= 0) b->update_grid = update_grid;
    if (update_threads) b->update_threads = update_threads;
    if (lds_set >= 0) b->lds_set = lds_set;
    return 0;
}

int32_t osn_tsdf_hip_batch_set_alloc_mode(OsnTsdfHipBatch* b, int32_t mode) {
    if (!b || mode < 0 || mode > 2) return 1;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return 2;
    b->drop_graph();
    b->alloc_mode = mode == 1 ? kAllocCountCas : (mode == 2 ? kAllocProbe : kAllocNormal);
    return 0;
}

int32_t osn_tsdf_hip_batch_set_extract(OsnTsdfHipBatch* b, float* const* d_posnor,
                                       const int32_t* vert_caps, int32_t* d_counts,
                                       float min_weight, float iso, int32_t every) {
    if (!b || every < 0) return 1;
    const size_t n = b->vols.size();
    if (every > 0 && (!d_posnor || !vert_caps || !d_counts)) return 1;
    if (every > 0)
        for (size_t s = 0; s < n; ++s)
            if (!d_posnor[s] || vert_caps[s] <= 0) return 1;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return 2;
    b->drop_graph();
    b->extract_every = every;
    b->ex_args.assign(n, ExtractArgs{});
    if (every == 0) return 0;
    for (size_t s = 0; s < n; ++s) {
        OsnTsdfHipVolume* v = b->vols[s];
        // scratch for a full pool, allocated now: nothing allocates while a
        // graph is captured or replayed
        if (!v->grow_extract(v->d.pool_capacity)) {
            v->ok = false;
            b->extract_every = 0;
            return 2;
        }
        ExtractArgs& a = b->ex_args[s];
        a.min_weight = min_weight;
        a.iso = iso;
        a.n_blocks = -1;
        a.block_vcount = v->ex_vcount;
        a.block_voff = v->ex_voff;
        a.block_nb = v->ex_nb;
        a.nb_cache = 1;
        a.chunk_sums = v->ex_chunks;
        a.scan_flag = v->ex_flag_count;
        a.posnor = d_posnor[s];
        a.rgb = nullptr;
        a.cap = vert_caps[s];
        a.count_out = d_counts + s;
    }
    if (!b->d_ex_args && hipMalloc(&b->d_ex_args, sizeof(ExtractArgs) * n) != hipSuccess) {
        b->d_ex_args = nullptr;
        b->extract_every = 0;
        return 2;
    }
    if (hipMemcpy(b->d_ex_args, b->ex_args.data(), sizeof(ExtractArgs) * n,
                  hipMemcpyHostToDevice) != hipSuccess) {
        b->extract_every = 0;
        return 2;
    }
    return 0;
}

int32_t osn_tsdf_hip_batch_set_timing(OsnTsdfHipBatch* b, int32_t enable) {
    if (!b) return 1;
    b->timing = enable != 0;
    return 0;
}

int32_t osn_tsdf_hip_batch_read_timing(OsnTsdfHipBatch* b, double* ms4, int64_t* calls4) {
    if (!b) return 1;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return 2;
    for (auto& p : b->ev_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            b->ms[p.first] += ms;
            b->stage_calls[p.first] += 1;
        }
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    b->ev_pending.clear();
    for (int i = 0; i < kStages; ++i) {
        if (ms4) ms4[i] = b->ms[i];
        if (calls4) calls4[i] = b->stage_calls[i];
        b->ms[i] = 0.0;
        b->stage_calls[i] = 0;
    }
    return 0;
}

}  // extern "C"
