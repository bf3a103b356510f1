// Memory ceiling of the projective update's access pattern, without its
// arithmetic: one wave per 2 KiB block (16 bytes per lane at lane * 16 and
// 1 KiB later, as a block of FP16 + u16 voxel records is laid out), blocks
// taken from an ascending list, the same grid-stride order the update uses.
//
//   read      load the block
//   write     store the block
//   rmw       load, modify, store the block
//   rmw-next  the same, with the next block's load issued before this block's
//             store (the update's pipeline)
//
// For each: contiguous blocks 0..n-1, n blocks drawn at random from a pool
// in ascending order (a list in pool order), and the same blocks in random
// order. Times are the average of 20 launches from HIP events; GB/s counts
// the bytes the pattern must move (n * 2 KiB per direction).
//
//   rmw_ceiling [n_blocks=100000] [pool_blocks=1048576] [grid=16384]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kThreads = 128;  // two waves per workgroup
constexpr int kRecs = 128;     // 16-byte records per block

enum Mode { kRead = 0, kWrite = 1, kRmw = 2 };

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_blocks(u32x4* pool, const int* list, int n,
                                                     unsigned* sink) {
    const int lane = threadIdx.x & 63;
    const int stride = gridDim.x * (kThreads / 64);
    unsigned acc = 0;
    for (int item = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); item < n; item += stride) {
        u32x4* rec = pool + (size_t)list[item] * kRecs + lane;
        if (MODE == kWrite) {
            rec[0] = u32x4{1u, 2u, 3u, (unsigned)item};
            rec[64] = u32x4{5u, 6u, 7u, 8u};
            continue;
        }
        u32x4 a = rec[0], b = rec[64];
        if (MODE == kRead) {
            acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w;
            continue;
        }
        a.x += 1u;
        b.x += 1u;
        rec[0] = a;
        rec[64] = b;
    }
    if (MODE == kRead && acc == 0x9e3779b9u) sink[0] = acc;  // keeps the loads
}

__global__ __launch_bounds__(kThreads) void k_rmw_next(u32x4* pool, const int* list, int n) {
    const int lane = threadIdx.x & 63;
    const int stride = gridDim.x * (kThreads / 64);
    int item = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (item >= n) return;
    u32x4* rec = pool + (size_t)list[item] * kRecs + lane;
    u32x4 a = rec[0], b = rec[64];
    while (true) {
        const int next = item + stride;
        u32x4* recn = rec;
        u32x4 an{}, bn{};
        if (next < n) {
            recn = pool + (size_t)list[next] * kRecs + lane;
            an = recn[0];
            bn = recn[64];
        }
        a.x += 1u;
        b.x += 1u;
        rec[0] = a;
        rec[64] = b;
        if (next >= n) break;
        item = next;
        rec = recn;
        a = an;
        b = bn;
    }
}

bool ok(hipError_t e, const char* what) {
    if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return e == hipSuccess;
}

}  // namespace

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 100000;
    const int pool_blocks = argc > 2 ? std::atoi(argv[2]) : (1 << 20);
    const int grid = argc > 3 ? std::atoi(argv[3]) : 16384;
    if (n <= 0 || pool_blocks < n || grid <= 0 || (size_t)pool_blocks * 2048 > ((size_t)16 << 30)) {
        std::fprintf(stderr, "need 0 < n_blocks <= pool_blocks <= 8M, grid > 0\n");
        return 1;
    }
    u32x4* pool = nullptr;
    int* list = nullptr;
    unsigned* sink = nullptr;
    if (!ok(hipMalloc(&pool, (size_t)pool_blocks * 2048), "hipMalloc pool") ||
        !ok(hipMalloc(&list, sizeof(int) * n), "hipMalloc list") ||
        !ok(hipMalloc(&sink, sizeof(unsigned)), "hipMalloc sink") ||
        !ok(hipMemset(pool, 0, (size_t)pool_blocks * 2048), "hipMemset"))
        return 2;
    hipEvent_t e0, e1;
    if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return 2;
    auto time_us = [&](auto&& launch) -> double {
        launch();
        launch();
        if (!ok(hipDeviceSynchronize(), "warmup")) return -1.0;
        const int reps = 20;
        hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r) launch();
        hipEventRecord(e1, nullptr);
        if (!ok(hipEventSynchronize(e1), "timed launches")) return -1.0;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        return 1e3 * ms / reps;
    };
    std::vector<int> ids(pool_blocks);
    std::iota(ids.begin(), ids.end(), 0);
    std::mt19937 rng(1);
    std::shuffle(ids.begin(), ids.end(), rng);
    const double mb = n * 2048.0 / 1e6;
    std::printf("n_blocks %d, pool %d blocks, grid %d x %d threads\n", n, pool_blocks, grid,
                kThreads);
    const char* names[3] = {"contiguous", "random", "unordered"};
    for (int layout = 0; layout < 3; ++layout) {
        std::vector<int> h(n);
        if (layout == 0)
            std::iota(h.begin(), h.end(), 0);
        else {
            std::copy(ids.begin(), ids.begin() + n, h.begin());
            if (layout == 1) std::sort(h.begin(), h.end());
        }
        if (!ok(hipMemcpy(list, h.data(), sizeof(int) * n, hipMemcpyHostToDevice), "list copy"))
            return 2;
        const double r = time_us([&] { k_blocks<kRead><<<grid, kThreads>>>(pool, list, n, sink); });
        const double w = time_us([&] { k_blocks<kWrite><<<grid, kThreads>>>(pool, list, n, sink); });
        const double m = time_us([&] { k_blocks<kRmw><<<grid, kThreads>>>(pool, list, n, sink); });
        const double p = time_us([&] { k_rmw_next<<<grid, kThreads>>>(pool, list, n); });
        if (r < 0 || w < 0 || m < 0 || p < 0) return 3;
        std::printf("%-10s read %7.1f us %5.0f GB/s | write %7.1f us %5.0f GB/s | "
                    "rmw %7.1f us %5.0f GB/s | rmw-next %7.1f us %5.0f GB/s\n",
                    names[layout], r, mb / r * 1e3, w, mb / w * 1e3, m,
                    2 * mb / m * 1e3, p, 2 * mb / p * 1e3);
    }
    hipFree(pool);
    hipFree(list);
    hipFree(sink);
    return 0;
}
