// Mesh-comparison gate (host code). Restates crates/mesh-metrics/src/lib.rs:
//   sample_surface        lib.rs:102-150   seeded xorshift, cumulative area,
//                                          reflected barycentric sample
//   point_triangle_dist2  lib.rs:161-222   closest point on a triangle
//   TriGrid               lib.rs:230-300   triangles binned by AABB cells,
//                                          shell search until provably nearest
//   stats / compare /     lib.rs:315-463
//   gate
// The spatial hash here is an open-addressing table over the occupied cells
// with the triangles of each cell stored contiguously (two passes, no sort),
// and queries run on all host threads; neither changes any result: every
// distance is the exact minimum over the same candidate triangles.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/osn_mesh_metrics.h"

namespace {

struct V3 {
    double x, y, z;
};

inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Mesh {
    const float* xyz;
    int64_t nv;
    int32_t stride;
    int64_t n_tri() const { return nv / 3; }
    V3 vertex(int64_t i) const {
        const float* p = xyz + i * stride;
        return {(double)p[0], (double)p[1], (double)p[2]};
    }
};

// xorshift64 with the seed forced odd (lib.rs:79-100).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed | 1ull) {}
    uint64_t next_u64() {
        uint64_t x = s;
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        s = x;
        return x;
    }
    double next_f64() { return (double)(next_u64() >> 11) / (double)(1ull << 53); }
};

std::vector<V3> sample_surface(const Mesh& m, int64_t n_samples, uint64_t seed) {
    std::vector<V3> out;
    const int64_t nt = m.n_tri();
    if (nt == 0 || n_samples <= 0) return out;
    std::vector<double> cum((size_t)nt);
    double total = 0.0;
    for (int64_t t = 0; t < nt; ++t) {
        const V3 a = m.vertex(t * 3), b = m.vertex(t * 3 + 1), c = m.vertex(t * 3 + 2);
        const V3 ab = sub(b, a), ac = sub(c, a);
        const V3 cr = {ab.y * ac.z - ab.z * ac.y, ab.z * ac.x - ab.x * ac.z,
                       ab.x * ac.y - ab.y * ac.x};
        total += 0.5 * std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
        cum[(size_t)t] = total;
    }
    if (total <= 0.0) {  // zero-area soup: its vertices stand in for samples
        for (int64_t i = 0; i < m.nv; ++i) out.push_back(m.vertex(i));
        return out;
    }
    Rng rng(seed);
    out.reserve((size_t)n_samples);
    for (int64_t i = 0; i < n_samples; ++i) {
        const double r = rng.next_f64() * total;
        // first triangle whose cumulative area reaches r
        int64_t t = std::lower_bound(cum.begin(), cum.end(), r) - cum.begin();
        if (t > nt - 1) t = nt - 1;
        const V3 a = m.vertex(t * 3), b = m.vertex(t * 3 + 1), c = m.vertex(t * 3 + 2);
        double u = rng.next_f64(), v = rng.next_f64();
        if (u + v > 1.0) {
            u = 1.0 - u;
            v = 1.0 - v;
        }
        const double w = 1.0 - u - v;
        out.push_back({a.x * w + b.x * u + c.x * v, a.y * w + b.y * u + c.y * v,
                       a.z * w + b.z * u + c.z * v});
    }
    return out;
}

// Squared distance from p to triangle (a, b, c): the closest point lies in
// one of the seven Voronoi regions (three vertices, three edges, interior).
double point_triangle_dist2(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return dot(ap, ap);
    const V3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return dot(bp, bp);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double v = d1 / (d1 - d3);
        const V3 q = {a.x + ab.x * v, a.y + ab.y * v, a.z + ab.z * v};
        const V3 d = sub(p, q);
        return dot(d, d);
    }
    const V3 cp = sub(p, c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return dot(cp, cp);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        const V3 q = {a.x + ac.x * w, a.y + ac.y * w, a.z + ac.z * w};
        const V3 d = sub(p, q);
        return dot(d, d);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const V3 q = {b.x + (c.x - b.x) * w, b.y + (c.y - b.y) * w, b.z + (c.z - b.z) * w};
        const V3 d = sub(p, q);
        return dot(d, d);
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    const V3 q = {a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w};
    const V3 d = sub(p, q);
    return dot(d, d);
}

struct CellKey {
    int64_t x, y, z;
    bool operator==(const CellKey& o) const { return x == o.x && y == o.y && z == o.z; }
};

inline CellKey key_of(V3 p, double inv_cell) {
    return {(int64_t)std::floor(p.x * inv_cell), (int64_t)std::floor(p.y * inv_cell),
            (int64_t)std::floor(p.z * inv_cell)};
}

inline uint64_t hash_key(const CellKey& k) {
    uint64_t h = (uint64_t)k.x * 0x9E3779B97F4A7C15ull ^ (uint64_t)k.y * 0xC2B2AE3D27D4EB4Full ^
                 (uint64_t)k.z * 0x165667B19E3779F9ull;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    return h;
}

// Triangles binned into every cell their AABB touches.
class TriGrid {
  public:
    TriGrid(const Mesh& m, double cell) : m_(m), cell_(cell), inv_cell_(1.0 / cell) {
        const int64_t nt = m.n_tri();
        // pass 1: occupied cells and their triangle counts
        size_t cap = 1024;  // grows to keep the table at most half full
        while (cap < (size_t)nt) cap <<= 1;
        keys_.assign(cap, CellKey{0, 0, 0});
        used_.assign(cap, 0);
        start_.assign(cap, 0);
        count_.assign(cap, 0);
        for (int pass = 0; pass < 2; ++pass) {
            for (int64_t t = 0; t < nt; ++t) {
                CellKey lo, hi;
                aabb_cells(t, lo, hi);
                for (int64_t x = lo.x; x <= hi.x; ++x)
                    for (int64_t y = lo.y; y <= hi.y; ++y)
                        for (int64_t z = lo.z; z <= hi.z; ++z) {
                            if (pass == 0) {
                                count_[insert({x, y, z})] += 1;
                            } else {
                                const size_t s = find({x, y, z});
                                tri_[start_[s] + fill_[s]++] = (uint32_t)t;
                            }
                        }
            }
            if (pass == 0) {  // the table may have grown: walk its current size
                uint32_t acc = 0;
                for (size_t s = 0; s < keys_.size(); ++s) {
                    start_[s] = acc;
                    acc += count_[s];
                }
                tri_.assign(acc, 0);
                fill_.assign(keys_.size(), 0);
            }
        }
    }

    double nearest_sq(V3 q) const {
        const CellKey k = key_of(q, inv_cell_);
        double best = INFINITY;
        for (int64_t r = 0;; ++r) {
            for (int64_t dx = -r; dx <= r; ++dx)
                for (int64_t dy = -r; dy <= r; ++dy) {
                    const bool full = std::llabs(dx) == r || std::llabs(dy) == r;
                    for (int64_t dz = -r; dz <= r; dz += full ? 1 : 2 * r) {
                        visit({k.x + dx, k.y + dy, k.z + dz}, q, best);
                    }
                }
            // stop once the best hit is provably closer than any unscanned shell
            if (std::isfinite(best) && std::sqrt(best) <= (double)r * cell_) break;
            if (r + 1 > 64) break;
        }
        return best;
    }

  private:
    void aabb_cells(int64_t t, CellKey& lo, CellKey& hi) const {
        V3 mn{INFINITY, INFINITY, INFINITY}, mx{-INFINITY, -INFINITY, -INFINITY};
        for (int j = 0; j < 3; ++j) {
            const V3 v = m_.vertex(t * 3 + j);
            mn = {std::min(mn.x, v.x), std::min(mn.y, v.y), std::min(mn.z, v.z)};
            mx = {std::max(mx.x, v.x), std::max(mx.y, v.y), std::max(mx.z, v.z)};
        }
        lo = key_of(mn, inv_cell_);
        hi = key_of(mx, inv_cell_);
    }
    size_t insert(const CellKey& k) {
        size_t s = hash_key(k) & (keys_.size() - 1);
        for (;;) {
            if (!used_[s]) {
                used_[s] = 1;
                keys_[s] = k;
                if (++n_used_ * 2 > keys_.size()) grow();
                return find(k);
            }
            if (keys_[s] == k) return s;
            s = (s + 1) & (keys_.size() - 1);
        }
    }
    size_t find(const CellKey& k) const {
        size_t s = hash_key(k) & (keys_.size() - 1);
        for (;;) {
            if (!used_[s]) return SIZE_MAX;
            if (keys_[s] == k) return s;
            s = (s + 1) & (keys_.size() - 1);
        }
    }
    void grow() {  // pass 1 only: counts move with their keys
        std::vector<CellKey> ok;
        std::vector<uint32_t> oc;
        for (size_t s = 0; s < keys_.size(); ++s)
            if (used_[s]) { ok.push_back(keys_[s]); oc.push_back(count_[s]); }
        const size_t cap = keys_.size() * 2;
        keys_.assign(cap, CellKey{0, 0, 0});
        used_.assign(cap, 0);
        start_.assign(cap, 0);
        count_.assign(cap, 0);
        n_used_ = 0;
        for (size_t i = 0; i < ok.size(); ++i) {
            size_t s = hash_key(ok[i]) & (cap - 1);
            while (used_[s]) s = (s + 1) & (cap - 1);
            used_[s] = 1;
            keys_[s] = ok[i];
            count_[s] = oc[i];
            ++n_used_;
        }
    }
    void visit(const CellKey& c, V3 q, double& best) const {
        const size_t s = find(c);
        if (s == SIZE_MAX) return;
        for (uint32_t i = start_[s], e = start_[s] + count_[s]; i < e; ++i) {
            const int64_t t = tri_[i];
            const double d = point_triangle_dist2(q, m_.vertex(t * 3), m_.vertex(t * 3 + 1),
                                                  m_.vertex(t * 3 + 2));
            if (d < best) best = d;
        }
    }

    const Mesh m_;
    double cell_, inv_cell_;
    std::vector<CellKey> keys_;
    std::vector<uint8_t> used_;
    std::vector<uint32_t> start_, count_, fill_;  // per cell: first entry, entries
    std::vector<uint32_t> tri_;
    size_t n_used_ = 0;
};

int worker_count(int32_t threads, size_t work) {
    int n = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    if (n < 1) n = 1;
    if ((size_t)n > work / 256 + 1) n = (int)(work / 256 + 1);
    return n;
}

void nearest_all(const TriGrid& g, const V3* q, size_t n, int32_t threads, double* out) {
    const int nw = worker_count(threads, n);
    std::vector<std::thread> ws;
    for (int w = 0; w < nw; ++w)
        ws.emplace_back([&, w] {
            for (size_t i = (size_t)w; i < n; i += (size_t)nw) out[i] = g.nearest_sq(q[i]);
        });
    for (auto& t : ws) t.join();
}

OsnDistanceStats stats(std::vector<double> d) {
    OsnDistanceStats s{};
    if (d.empty()) return s;
    std::sort(d.begin(), d.end());
    const size_t n = d.size();
    double sum = 0.0, sq = 0.0;
    for (double x : d) { sum += x; sq += x * x; }
    s.mean = sum / (double)n;
    s.rms = std::sqrt(sq / (double)n);
    s.p95 = d[std::min(n - 1, (size_t)((double)n * 0.95))];
    s.p99 = d[std::min(n - 1, (size_t)((double)n * 0.99))];
    s.max = d[n - 1];
    s.n = (int64_t)n;
    return s;
}

bool valid(const float* xyz, int64_t nv, int32_t stride) {
    return nv >= 0 && stride >= 3 && (nv == 0 || xyz);
}

}  // namespace

extern "C" {

OsnMeshTolerance osn_mesh_tolerance_for_voxel(double voxel_m) {
    return OsnMeshTolerance{0.1 * voxel_m, voxel_m, 0.02};
}

int32_t osn_mesh_sample_surface(const float* xyz, int64_t n_vertices, int32_t stride,
                                int64_t n_samples, uint64_t seed, double* out_xyz,
                                int64_t out_cap, int64_t* n_out) {
    if (!valid(xyz, n_vertices, stride) || n_samples < 0 || !out_xyz || !n_out) return 1;
    const auto s = sample_surface(Mesh{xyz, n_vertices, stride}, n_samples, seed);
    const int64_t n = std::min<int64_t>((int64_t)s.size(), out_cap);
    std::memcpy(out_xyz, s.data(), (size_t)n * sizeof(V3));
    *n_out = n;
    return 0;
}

int32_t osn_mesh_nearest_sq(const float* xyz, int64_t n_vertices, int32_t stride,
                            const double* queries, int64_t n_queries, double cell,
                            int32_t threads, double* out_sq) {
    if (!valid(xyz, n_vertices, stride) || n_queries < 0 || !(cell > 0.0) ||
        (n_queries > 0 && (!queries || !out_sq)))
        return 1;
    const TriGrid g(Mesh{xyz, n_vertices, stride}, cell);
    nearest_all(g, reinterpret_cast<const V3*>(queries), (size_t)n_queries, threads, out_sq);
    return 0;
}

int32_t osn_mesh_compare(const float* xyz_a, int64_t n_vertices_a, int32_t stride_a,
                         const float* xyz_b, int64_t n_vertices_b, int32_t stride_b,
                         int64_t n_samples, double cell, uint64_t seed, int32_t threads,
                         OsnMeshAgreement* out) {
    if (!valid(xyz_a, n_vertices_a, stride_a) || !valid(xyz_b, n_vertices_b, stride_b) ||
        n_samples < 0 || !(cell > 0.0) || !out)
        return 1;
    const Mesh a{xyz_a, n_vertices_a, stride_a}, b{xyz_b, n_vertices_b, stride_b};
    const auto sa = sample_surface(a, n_samples, seed);
    const auto sb = sample_surface(b, n_samples, seed ^ 0x9E3779B97F4A7C15ull);
    const TriGrid ga(a, cell), gb(b, cell);
    std::vector<double> dab(sa.size()), dba(sb.size());
    nearest_all(gb, sa.data(), sa.size(), threads, dab.data());
    nearest_all(ga, sb.data(), sb.size(), threads, dba.data());
    for (double& x : dab) x = std::sqrt(x);
    for (double& x : dba) x = std::sqrt(x);
    out->a_to_b = stats(std::move(dab));
    out->b_to_a = stats(std::move(dba));
    out->mean_surface = 0.5 * (out->a_to_b.mean + out->b_to_a.mean);
    out->hausdorff = std::max(out->a_to_b.max, out->b_to_a.max);
    out->hausdorff_p99 = std::max(out->a_to_b.p99, out->b_to_a.p99);
    return 0;
}

int32_t osn_mesh_gate(const float* xyz_a, int64_t n_vertices_a, int32_t stride_a,
                      const float* xyz_b, int64_t n_vertices_b, int32_t stride_b,
                      OsnMeshTolerance tol, int64_t n_samples, double cell, int32_t threads,
                      OsnGateResult* out) {
    if (!out) return 1;
    OsnGateResult r{};
    const int32_t rc = osn_mesh_compare(xyz_a, n_vertices_a, stride_a, xyz_b, n_vertices_b,
                                        stride_b, n_samples, cell, 0xC0FFEEull, threads,
                                        &r.agreement);
    if (rc) return rc;
    r.n_vertices_a = n_vertices_a;
    r.n_vertices_b = n_vertices_b;
    const int64_t mx = std::max(n_vertices_a, n_vertices_b);
    r.vertex_count_rel = mx == 0 ? 0.0
                                 : std::fabs((double)n_vertices_a - (double)n_vertices_b) / (double)mx;
    if (r.agreement.mean_surface > tol.mean_surface_m) r.failures |= OSN_GATE_FAIL_MEAN;
    if (r.agreement.hausdorff_p99 > tol.hausdorff_p99_m) r.failures |= OSN_GATE_FAIL_P99;
    if (r.vertex_count_rel > tol.vertex_count_rel) r.failures |= OSN_GATE_FAIL_COUNT;
    r.passed = r.failures == 0;
    *out = r;
    return 0;
}

}  // extern "C"
