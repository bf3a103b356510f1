// OSNP point clouds, OSN1 mesh dumps and the analytic point clouds (host
// code). Layouts and generators as the reference defines them; see
// include/osn_io.h for the byte layouts and citations.

#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/osn_io.h"

namespace {

constexpr uint32_t kOsn1Magic = 0x314E534Fu;  // bytes "OSN1"

struct File {
    std::FILE* f;
    explicit File(std::FILE* p) : f(p) {}
    ~File() { if (f) std::fclose(f); }
};

int32_t osnp_header(std::FILE* f, int32_t* n, float cam[3]) {
    char magic[4] = {0, 0, 0, 0};
    int32_t version = 0, count = 0;
    float c[3];
    if (std::fread(magic, 1, 4, f) != 4 || std::fread(&version, 4, 1, f) != 1 ||
        std::fread(&count, 4, 1, f) != 1 || std::fread(c, 4, 3, f) != 3)
        return 3;
    if (std::memcmp(magic, "OSNP", 4) != 0 || version != 1 || count <= 0) return 3;
    *n = count;
    if (cam) std::memcpy(cam, c, sizeof(c));
    return 0;
}

int32_t osn1_header(std::FILE* f, int32_t* n) {
    uint32_t magic = 0, count = 0;
    if (std::fread(&magic, 4, 1, f) != 1 || std::fread(&count, 4, 1, f) != 1) return 3;
    if (magic != kOsn1Magic || count > 0x7fffffffu) return 3;
    *n = (int32_t)count;
    return 0;
}

}  // namespace

extern "C" {

int32_t osn_osnp_info(const char* path, int32_t* n_points, float cam[3]) {
    if (!path || !n_points) return 1;
    File f(std::fopen(path, "rb"));
    if (!f.f) return 2;
    return osnp_header(f.f, n_points, cam);
}

int32_t osn_osnp_read(const char* path, float* points, int32_t cap_points, int32_t* n_points,
                      float cam[3]) {
    if (!path || !points || !n_points || cap_points < 0) return 1;
    File f(std::fopen(path, "rb"));
    if (!f.f) return 2;
    int32_t n = 0;
    if (const int32_t rc = osnp_header(f.f, &n, cam)) return rc;
    *n_points = n;
    if (n > cap_points) return 5;
    if (std::fread(points, sizeof(float), (size_t)n * 3, f.f) != (size_t)n * 3) return 4;
    return 0;
}

int32_t osn_osnp_write(const char* path, const float* points, int32_t n_points,
                       const float cam[3]) {
    if (!path || !points || !cam || n_points <= 0) return 1;
    File f(std::fopen(path, "wb"));
    if (!f.f) return 2;
    const int32_t version = 1;
    const bool ok = std::fwrite("OSNP", 1, 4, f.f) == 4 && std::fwrite(&version, 4, 1, f.f) == 1 &&
                    std::fwrite(&n_points, 4, 1, f.f) == 1 && std::fwrite(cam, 4, 3, f.f) == 3 &&
                    std::fwrite(points, sizeof(float), (size_t)n_points * 3, f.f) ==
                        (size_t)n_points * 3;
    return ok ? 0 : 2;
}

int32_t osn_osn1_info(const char* path, int32_t* n_vertices) {
    if (!path || !n_vertices) return 1;
    File f(std::fopen(path, "rb"));
    if (!f.f) return 2;
    return osn1_header(f.f, n_vertices);
}

int32_t osn_osn1_read(const char* path, float* posnor, int32_t cap_vertices,
                      int32_t* n_vertices) {
    if (!path || !n_vertices || cap_vertices < 0 || (cap_vertices > 0 && !posnor)) return 1;
    File f(std::fopen(path, "rb"));
    if (!f.f) return 2;
    int32_t n = 0;
    if (const int32_t rc = osn1_header(f.f, &n)) return rc;
    *n_vertices = n;
    if (n > cap_vertices) return 5;
    if (std::fread(posnor, sizeof(float), (size_t)n * 6, f.f) != (size_t)n * 6) return 4;
    return 0;
}

int32_t osn_osn1_write(const char* path, const float* posnor, int32_t n_vertices) {
    if (!path || n_vertices < 0 || (n_vertices > 0 && !posnor)) return 1;
    File f(std::fopen(path, "wb"));
    if (!f.f) return 2;
    const uint32_t nv = (uint32_t)n_vertices;
    const bool ok = std::fwrite(&kOsn1Magic, 4, 1, f.f) == 1 && std::fwrite(&nv, 4, 1, f.f) == 1 &&
                    std::fwrite(posnor, sizeof(float), (size_t)nv * 6, f.f) == (size_t)nv * 6;
    return ok ? 0 : 2;
}

// analytic_scenes.hpp:18-30: theta at the centre of each band, phi from 0.
void osn_points_sphere(float radius, int32_t n_theta, int32_t n_phi, float* out) {
    const float pi = 3.14159265358979323846f;
    float* p = out;
    for (int i = 0; i < n_theta; ++i) {
        const float theta = pi * ((float)i + 0.5f) / (float)n_theta;
        for (int j = 0; j < n_phi; ++j) {
            const float phi = 2.0f * pi * (float)j / (float)n_phi;
            *p++ = radius * sinf(theta) * cosf(phi);
            *p++ = radius * sinf(theta) * sinf(phi);
            *p++ = radius * cosf(theta);
        }
    }
}

// analytic_scenes.hpp:33-44: an n x n grid with both edges included.
void osn_points_plane(float z0, float half_extent, int32_t n, float* out) {
    float* p = out;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            *p++ = -half_extent + 2.0f * half_extent * (float)i / (float)(n - 1);
            *p++ = -half_extent + 2.0f * half_extent * (float)j / (float)(n - 1);
            *p++ = z0;
        }
}

}  // extern "C"
