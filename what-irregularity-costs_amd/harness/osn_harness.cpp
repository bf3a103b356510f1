// Timing harness for the gfx950 TSDF arms over the reference's workload matrix.
//
// What it guarantees, following the reference harness (cpp/bench/bench_arms.cu:
// 1-93, 525-863, 1095-1125), on this library's C ABI:
//   * correctness gates timing: a cell whose arms disagree on the block count,
//     that dropped blocks, or that misses its published block count is reported
//     INVALID and not recorded;
//   * distributions: nearest-rank p50 / p95 / p99 / min / mean per stage, warmup
//     repetitions discarded;
//   * BATCH-amortised windows: a timed window holds BATCH launches and is
//     divided by BATCH; allocate goes into BATCH distinct volumes, reset outside
//     the window, so no launch finds the table already filled;
//   * interleaved arms: arms take turns within a repetition and the first arm
//     rotates; a first-half / second-half p50 drift self-check is printed;
//   * exclusive device: refuses to run while other work holds more than
//     --max-foreign-gib of device memory (exit 3);
//   * CSV with the reference's columns, appended across processes when
//     OSN_BENCH_CSV_APPEND is set;
//   * OSN1 export of each arm's mesh, untimed, from a volume rebuilt after the
//     timed window.
//
// Arms: "hip" (this library's reference-shaped point path, whose allocator
// collects a work unit's keys in a workgroup LDS set before the table) and
// "hip-noset" (the same path with the set off, every wave's keys straight to
// the table: its ablation).
// Windows are timed on the host clock around launch + synchronisation of all
// BATCH volumes (each volume has its own stream, so their launches overlap).
//
// Usage: osn_harness [--reps N] [--warmup N] [--batch N] [--voxel M]
//                    [--cells a,b,...] [--axis NAME] [--osnp PATH [--warm PATH]]
//                    [--csv PATH] [--export DIR] [--no-extract] [--list]
//                    [--max-foreign-gib G]

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/osn_io.h"
#include "../../include/osn_tsdf_hip.h"

namespace {

enum class Shape { Sphere, Plane, File };

struct Cell {
    const char* name;
    const char* axis;
    Shape shape;
    float extent;          // sphere radius or plane half-extent (m)
    int a, b;              // sphere: n_theta, n_phi; plane: n
    int32_t pool_blocks;   // 0: from the measured block count x pool_factor
    float pool_factor;
    int32_t known_blocks;  // published block count at voxel 1 cm, 0 = none
};

// The reference's synthetic matrix (bench_arms.cu:209-244). Known block counts
// are the published ones (results/sweep_*/sweep.csv) at 1 cm voxels.
const std::vector<Cell> kCells = {
    {"base", "baseline", Shape::Sphere, 0.5f, 400, 800, 1 << 14, 0, 1160},
    {"pts-20k", "points", Shape::Sphere, 0.5f, 100, 200, 1 << 14, 0, 1104},
    {"pts-80k", "points", Shape::Sphere, 0.5f, 200, 400, 1 << 14, 0, 1160},
    {"pts-720k", "points", Shape::Sphere, 0.5f, 600, 1200, 1 << 14, 0, 1160},
    {"pts-1280k", "points", Shape::Sphere, 0.5f, 800, 1600, 1 << 14, 0, 1160},
    {"r-0.25", "extent", Shape::Sphere, 0.25f, 400, 800, 1 << 14, 0, 280},
    {"r-1.0", "extent", Shape::Sphere, 1.0f, 400, 800, 1 << 16, 0, 4752},
    {"r-2.0", "extent", Shape::Sphere, 2.0f, 400, 800, 1 << 17, 0, 18160},
    {"lf-hi", "loadfactor", Shape::Sphere, 0.5f, 400, 800, 0, 1.05f, 1160},
    {"lf-mid", "loadfactor", Shape::Sphere, 0.5f, 400, 800, 0, 2.0f, 1160},
    {"lf-lo", "loadfactor", Shape::Sphere, 0.5f, 400, 800, 0, 4.0f, 1160},
    {"lf-sparse", "loadfactor", Shape::Sphere, 0.5f, 400, 800, 0, 16.0f, 1160},
    {"plane-320k", "shape", Shape::Plane, 1.0f, 566, 0, 1 << 14, 0, 1060},
};

struct Stats {
    double p50 = 0, p95 = 0, p99 = 0, min = 0, mean = 0;
    int n = 0;
};

// Nearest-rank percentiles (no interpolation: the sample counts do not
// support more precision).
Stats summarize(std::vector<double> v) {
    Stats s;
    if (v.empty()) return s;
    std::sort(v.begin(), v.end());
    s.n = (int)v.size();
    s.min = v.front();
    for (double x : v) s.mean += x;
    s.mean /= (double)v.size();
    auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)((double)v.size() * p))]; };
    s.p50 = pct(0.50);
    s.p95 = pct(0.95);
    s.p99 = pct(0.99);
    return s;
}

struct Row {
    std::string cell, axis, arm, stage, gpu;
    int32_t points = 0, pool = 0, blocks = 0, batch = 0;
    double load_factor = 0;
    Stats s;
};

struct Scene {
    std::vector<float> pts;
    float cam[3] = {0, 0, 0};
};

bool load_osnp(const char* path, Scene& out) {
    int32_t n = 0;
    if (osn_osnp_info(path, &n, out.cam)) {
        std::printf("  %s: cannot read an OSNP header\n", path);
        return false;
    }
    out.pts.resize((size_t)n * 3);
    int32_t got = 0;
    if (osn_osnp_read(path, out.pts.data(), n, &got, out.cam)) {
        std::printf("  %s: truncated or unreadable\n", path);
        return false;
    }
    return true;
}

bool make_scene(const Cell& c, const char* osnp, Scene& out) {
    if (c.shape == Shape::File) return load_osnp(osnp, out);
    if (c.shape == Shape::Sphere) {
        out.pts.resize((size_t)c.a * c.b * 3);
        osn_points_sphere(c.extent, c.a, c.b, out.pts.data());
    } else {
        out.pts.resize((size_t)c.a * c.a * 3);
        osn_points_plane(0.5f, c.extent, c.a, out.pts.data());
    }
    return true;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Clock and temperature as rocm-smi reports them (recorded, not pinned).
std::string gpu_state() {
    std::FILE* p = popen("rocm-smi --showtemp --showclocks 2>/dev/null", "r");
    if (!p) return "n/a";
    std::string out, line;
    char buf[512];
    while (std::fgets(buf, sizeof(buf), p)) {
        line = buf;
        if (line.find("sclk") != std::string::npos || line.find("(Edge)") != std::string::npos ||
            line.find("(Junction)") != std::string::npos) {
            while (!line.empty() && (line.back() == '\n' || line.back() == ' ')) line.pop_back();
            const size_t c = line.find("GPU[");
            out += (out.empty() ? "" : "; ") + line.substr(c == std::string::npos ? 0 : c);
        }
    }
    pclose(p);
    if (out.empty()) return "n/a";
    for (char& ch : out)
        if (ch == '"') ch = '\'';
    return out;
}

struct Arm {
    const char* name;
    int32_t lds_set;  // workgroup key-set entries, -1 = library default, 0 = off
};
const Arm kArms[] = {{"hip", -1}, {"hip-noset", 0}};
constexpr int kNumArms = 2;

OsnTsdfHipVolume* make_volume(float voxel, int32_t pool, const Arm& arm) {
    if (arm.lds_set < 0) return osn_tsdf_hip_create(voxel, pool, -1.0f, 32.0f);
    OsnTsdfHipConfig c{};
    c.voxel_size_m = voxel;
    c.pool_capacity_blocks = pool;
    c.trunc_m = -1.0f;
    c.weight_cap = 32.0f;
    c.storage = OSN_TSDF_HIP_STORE_SUMS;
    c.hash_kind = OSN_TSDF_HIP_HASH_REFERENCE;
    c.lds_set = arm.lds_set;
    return osn_tsdf_hip_create_ex(&c);
}

struct Options {
    int reps = 30, warmup = 5, batch = 8;
    float voxel = 0.01f;
    std::string cells, axis, csv, export_dir;
    const char* osnp = nullptr;
    const char* warm = nullptr;
    bool extract = true;
    double max_foreign_gib = 2.0;
};

bool run_cell(const Cell& cell, const Options& o, std::vector<Row>& rows) {
    Scene scene;
    if (!make_scene(cell, o.osnp, scene)) {
        std::printf("=== cell %s (axis %s) ===\n  SKIPPED: scene unavailable\n\n", cell.name, cell.axis);
        return false;
    }
    Scene warm;
    const bool has_warm = o.warm && cell.shape == Shape::File;
    if (has_warm && !load_osnp(o.warm, warm)) return false;
    const int32_t n = (int32_t)(scene.pts.size() / 3);
    const float cx = scene.cam[0], cy = scene.cam[1], cz = scene.cam[2];

    float *d_pts = nullptr, *d_warm = nullptr;
    hipMalloc(&d_pts, scene.pts.size() * sizeof(float));
    hipMemcpy(d_pts, scene.pts.data(), scene.pts.size() * sizeof(float), hipMemcpyHostToDevice);
    const int32_t n_warm = has_warm ? (int32_t)(warm.pts.size() / 3) : 0;
    if (n_warm) {
        hipMalloc(&d_warm, warm.pts.size() * sizeof(float));
        hipMemcpy(d_warm, warm.pts.data(), warm.pts.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    auto cleanup = [&] {
        hipFree(d_pts);
        if (d_warm) hipFree(d_warm);
    };

    // pool: stated, or measured with a generous probe volume and scaled
    int32_t pool = cell.pool_blocks;
    if (pool == 0) {
        OsnTsdfHipVolume* probe = osn_tsdf_hip_create(o.voxel, 1 << 18, -1.0f, 32.0f);
        if (!probe) { cleanup(); return false; }
        if (n_warm) osn_tsdf_hip_allocate_blocks(probe, d_warm, n_warm, cx, cy, cz, 0.0f);
        osn_tsdf_hip_allocate_blocks(probe, d_pts, n, cx, cy, cz, 0.0f);
        const int32_t nb = osn_tsdf_hip_block_count(probe);
        osn_tsdf_hip_destroy(probe);
        if (nb <= 0) { cleanup(); return false; }
        pool = (int32_t)std::max(64.0f, std::ceil((float)nb * cell.pool_factor));
    }

    // BATCH volumes per arm must fit in free memory; a reduced batch is reported
    size_t free_b = 0, total_b = 0;
    hipMemGetInfo(&free_b, &total_b);
    uint32_t table = 1;
    while (table < (uint32_t)pool * 2u) table <<= 1;
    const size_t per_vol = (size_t)(table + (1u << 20)) * 16 + (size_t)pool * 512 * 20 +
                           (size_t)pool * 12 + (size_t)table * 64;
    int batch = o.batch;
    while (batch > 1 && (size_t)batch * kNumArms * per_vol > free_b * 3 / 4) --batch;

    std::printf("=== cell %s (axis %s) ===\n", cell.name, cell.axis);
    std::printf("  %d points, voxel %.3f m, pool %d blocks, batch %d%s%s\n", n, o.voxel, pool,
                batch, batch != o.batch ? "  <-- REDUCED TO FIT" : "",
                has_warm ? ", pre-warmed outside the timed window" : "");

    // Volumes are created interleaved across arms: the first volumes a process
    // creates were measured to allocate ~20 % slower than later ones whatever
    // entry point created them, so creating one arm's batch first biases it.
    std::vector<OsnTsdfHipVolume*> vols[kNumArms];
    for (auto& vs : vols) vs.assign((size_t)batch, nullptr);
    bool ok = true;
    for (int j = 0; j < batch && ok; ++j)
        for (int a = 0; a < kNumArms && ok; ++a) {
            vols[a][j] = make_volume(o.voxel, pool, kArms[a]);
            ok = vols[a][j] != nullptr;
        }
    auto destroy = [&] {
        for (auto& vs : vols)
            for (auto* v : vs)
                if (v) osn_tsdf_hip_destroy(v);
    };
    if (!ok) {
        std::printf("  SKIPPED: volume allocation failed\n\n");
        destroy();
        cleanup();
        return false;
    }

    auto reset_all = [&](int a) {
        for (auto* v : vols[a]) {
            osn_tsdf_hip_reset(v);
            if (n_warm) {  // the same warm state for every arm, outside the window
                osn_tsdf_hip_allocate_blocks(v, d_warm, n_warm, cx, cy, cz, 0.0f);
                osn_tsdf_hip_update_voxels(v, d_warm, n_warm, cx, cy, cz, 0.0f);
            }
        }
        for (auto* v : vols[a]) osn_tsdf_hip_synchronize(v);
    };
    auto sync_all = [&](int a) {
        for (auto* v : vols[a]) osn_tsdf_hip_synchronize(v);
    };

    const int32_t mesh_cap = 4 << 20;
    std::vector<float> mesh((size_t)mesh_cap * 6);
    std::vector<double> smp[kNumArms][3];
    int32_t blocks_seen[kNumArms] = {-1, -1};
    uint64_t drops_seen[kNumArms] = {0, 0};
    bool launch_failed = false;
    for (int i = 0; i < o.warmup + o.reps; ++i)
        for (int k = 0; k < kNumArms; ++k) {
            const int a = (i + k) % kNumArms;  // the first arm rotates
            reset_all(a);
            double t0 = now_ms();
            for (auto* v : vols[a])
                launch_failed |= osn_tsdf_hip_allocate_blocks(v, d_pts, n, cx, cy, cz, 0.0f) != 0;
            sync_all(a);
            const double ta = (now_ms() - t0) / batch;
            t0 = now_ms();
            for (auto* v : vols[a])
                launch_failed |= osn_tsdf_hip_update_voxels(v, d_pts, n, cx, cy, cz, 0.0f) != 0;
            sync_all(a);
            const double tu = (now_ms() - t0) / batch;
            blocks_seen[a] = osn_tsdf_hip_block_count(vols[a][0]);
            drops_seen[a] = osn_tsdf_hip_drop_count(vols[a][0]);
            double te = -1.0;
            if (o.extract && a == 0) {  // extraction does not depend on the arm
                t0 = now_ms();
                for (int j = 0; j < batch; ++j)
                    osn_tsdf_hip_extract_mesh(vols[a][0], mesh.data(), nullptr, mesh_cap, 0.5f, 0.0f);
                te = (now_ms() - t0) / batch;
            }
            if (i >= o.warmup) {
                smp[a][0].push_back(ta);
                smp[a][1].push_back(tu);
                if (te >= 0) smp[a][2].push_back(te);
            }
        }

    // untimed OSN1 export of a volume rebuilt per arm
    if (!o.export_dir.empty())
        for (int a = 0; a < kNumArms; ++a) {
            reset_all(a);
            osn_tsdf_hip_allocate_blocks(vols[a][0], d_pts, n, cx, cy, cz, 0.0f);
            osn_tsdf_hip_update_voxels(vols[a][0], d_pts, n, cx, cy, cz, 0.0f);
            int32_t nv = -1;
            std::vector<float> buf;
            for (int32_t cap = 4 << 20; cap <= (64 << 20); cap *= 2) {
                buf.assign((size_t)cap * 6, 0.0f);
                nv = osn_tsdf_hip_extract_mesh(vols[a][0], buf.data(), nullptr, cap, 0.5f, 0.0f);
                if (nv >= 0) break;
            }
            const std::string path = o.export_dir + "/" + cell.name + "." + kArms[a].name + ".raw";
            if (nv >= 0 && osn_osn1_write(path.c_str(), buf.data(), nv) == 0)
                std::printf("  export: %s  %d verts\n", path.c_str(), nv);
            else
                std::printf("  export: %s failed (%d)\n", path.c_str(), nv);
        }

    // validity gates
    bool valid = !launch_failed;
    if (launch_failed) std::printf("  INVALID: a launch returned an error\n");
    for (int a = 0; a < kNumArms; ++a) {
        if (drops_seen[a] > 0) {
            std::printf("  INVALID: %s dropped %llu contributions (pool %d full)\n", kArms[a].name,
                        (unsigned long long)drops_seen[a], pool);
            valid = false;
        }
        if (blocks_seen[a] != blocks_seen[0]) {
            std::printf("  INVALID: %s built %d blocks, %s built %d\n", kArms[a].name,
                        blocks_seen[a], kArms[0].name, blocks_seen[0]);
            valid = false;
        }
    }
    if (cell.known_blocks > 0 && std::fabs(o.voxel - 0.01f) < 1e-7f && !has_warm &&
        blocks_seen[0] != cell.known_blocks) {
        std::printf("  INVALID: %d blocks, the published count is %d\n", blocks_seen[0],
                    cell.known_blocks);
        valid = false;
    }
    const double lf = (double)blocks_seen[0] / (double)table;
    std::printf("  blocks %d / %u slots -> load factor %.3f%s\n", blocks_seen[0], table, lf,
                valid ? "" : "   [CELL INVALID, NOT RECORDED]");
    if (valid) {
        Row proto;
        proto.cell = cell.name;
        proto.axis = cell.axis;
        proto.points = n;
        proto.pool = pool;
        proto.blocks = blocks_seen[0];
        proto.load_factor = lf;
        proto.batch = batch;
        proto.gpu = gpu_state();
        const char* stage[3] = {"allocate", "update", "extract-hip-only"};
        for (int a = 0; a < kNumArms; ++a)
            for (int st = 0; st < 3; ++st) {
                if (smp[a][st].empty()) continue;
                Row r = proto;
                r.arm = kArms[a].name;
                r.stage = stage[st];
                r.s = summarize(smp[a][st]);
                rows.push_back(r);
                std::printf("  %-10s %-16s p50 %8.4f  p95 %8.4f  p99 %8.4f  min %8.4f  (n=%d)\n",
                            r.arm.c_str(), r.stage.c_str(), r.s.p50, r.s.p95, r.s.p99, r.s.min,
                            r.s.n);
            }
        std::printf("  drift (first vs second half p50, update):");
        for (int a = 0; a < kNumArms; ++a) {
            const auto& v = smp[a][1];
            if (v.size() < 4) continue;
            const double x = summarize(std::vector<double>(v.begin(), v.begin() + v.size() / 2)).p50;
            const double y = summarize(std::vector<double>(v.begin() + v.size() / 2, v.end())).p50;
            const double d = x > 0 ? 100.0 * (y - x) / x : 0.0;
            std::printf("  %s %+.1f%%%s", kArms[a].name, d, std::fabs(d) > 5.0 ? "(!)" : "");
        }
        std::printf("\n");
    }
    std::printf("\n");
    destroy();
    cleanup();
    return valid;
}

bool listed(const std::string& list, const char* name) {
    if (list.empty()) return true;
    size_t p = 0;
    while (p <= list.size()) {
        const size_t q = list.find(',', p);
        const std::string item = list.substr(p, q == std::string::npos ? std::string::npos : q - p);
        if (item == name) return true;
        if (q == std::string::npos) break;
        p = q + 1;
    }
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    bool list = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) { std::fprintf(stderr, "%s needs a value\n", a.c_str()); std::exit(1); }
            return argv[++i];
        };
        if (a == "--reps") o.reps = std::atoi(next());
        else if (a == "--warmup") o.warmup = std::atoi(next());
        else if (a == "--batch") o.batch = std::max(1, std::atoi(next()));
        else if (a == "--voxel") o.voxel = (float)std::atof(next());
        else if (a == "--cells") o.cells = next();
        else if (a == "--axis") o.axis = next();
        else if (a == "--osnp") o.osnp = next();
        else if (a == "--warm") o.warm = next();
        else if (a == "--csv") o.csv = next();
        else if (a == "--export") o.export_dir = next();
        else if (a == "--no-extract") o.extract = false;
        else if (a == "--max-foreign-gib") o.max_foreign_gib = std::atof(next());
        else if (a == "--list") list = true;
        else {
            std::fprintf(stderr,
                         "usage: osn_harness [--reps N] [--warmup N] [--batch N] [--voxel M] "
                         "[--cells a,b] [--axis NAME] [--osnp PATH [--warm PATH]] [--csv PATH] "
                         "[--export DIR] [--no-extract] [--max-foreign-gib G] [--list]\n");
            return 1;
        }
    }
    if (list) {
        for (const auto& c : kCells) std::printf("%-12s %s\n", c.name, c.axis);
        return 0;
    }
    if (o.reps < 1 || o.warmup < 0) return 1;

    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        std::fprintf(stderr, "no HIP device\n");
        return 2;
    }
    size_t free_b = 0, total_b = 0;
    hipMemGetInfo(&free_b, &total_b);
    const double foreign = (double)(total_b - free_b) / (1u << 30);
    std::printf("device %s (%s), %d CUs, %.1f GiB, %.2f GiB held by other work\n", prop.name,
                prop.gcnArchName, prop.multiProcessorCount, total_b / double(1u << 30), foreign);
    if (foreign > o.max_foreign_gib) {
        std::printf("REFUSING TO RUN: %.2f GiB of device memory is held by other work "
                    "(limit %.2f); timings would measure the sharing\n", foreign, o.max_foreign_gib);
        return 3;
    }
    std::printf("state at start: %s\n\n", gpu_state().c_str());

    std::vector<Cell> cells;
    for (const auto& c : kCells)
        if (listed(o.cells, c.name) && (o.axis.empty() || o.axis == c.axis)) cells.push_back(c);
    if (o.osnp) cells.push_back({"osnp", "real", Shape::File, 0, 0, 0, 1 << 16, 0, 0});

    std::vector<Row> rows;
    int good = 0, bad = 0;
    for (const auto& c : cells) (run_cell(c, o, rows) ? good : bad)++;
    std::printf("cells: %d recorded, %d skipped or invalid\n", good, bad);
    std::printf("state at end: %s\n", gpu_state().c_str());

    std::string csv = o.csv;
    if (csv.empty() && std::getenv("OSN_BENCH_CSV")) csv = std::getenv("OSN_BENCH_CSV");
    if (!csv.empty()) {
        const bool append = std::getenv("OSN_BENCH_CSV_APPEND") != nullptr;
        std::FILE* f = std::fopen(csv.c_str(), append ? "a" : "w");
        if (f) {
            if (!append || std::ftell(f) == 0)
                std::fprintf(f, "run,device,device_name,sm,sm_count,cell,axis,points,voxel_m,"
                                "pool_blocks,blocks,load_factor,batch,scratch_slots,arm,stage,"
                                "p50_ms,p95_ms,p99_ms,min_ms,mean_ms,n,gpu_state\n");
            const long run = (long)std::time(nullptr);
            for (const auto& r : rows)
                std::fprintf(f,
                             "%ld,0,\"%s\",%s,%d,%s,%s,%d,%.6f,%d,%d,%.6f,%d,%d,%s,%s,"
                             "%.6f,%.6f,%.6f,%.6f,%.6f,%d,\"%s\"\n",
                             run, prop.name, prop.gcnArchName, prop.multiProcessorCount,
                             r.cell.c_str(), r.axis.c_str(), r.points, o.voxel, r.pool, r.blocks,
                             r.load_factor, r.batch, 1 << 20, r.arm.c_str(), r.stage.c_str(),
                             r.s.p50, r.s.p95, r.s.p99, r.s.min, r.s.mean, r.s.n, r.gpu.c_str());
            std::fclose(f);
            std::printf("wrote %s (%zu rows%s)\n", csv.c_str(), rows.size(), append ? ", appended" : "");
        }
    }
    return bad > 0 ? 4 : 0;
}
