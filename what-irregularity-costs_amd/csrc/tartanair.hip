// TartanAir V2 ingest (host code): NED pose composition in FP64, the frame
// descriptor for the depth path, trajectory files and depth samples. The
// conventions and their reference citations are in include/osn_tartanair.h.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/osn_tartanair.h"

namespace {

// Camera axes in NED body axes: body = (Zc, Xc, Yc).
constexpr double kNedRCam[3][3] = {{0.0, 0.0, 1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}};

// R(q) for q = (qx, qy, qz, qw), normalised first.
void quat_matrix(const double q[4], double r[3][3]) {
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const double x = q[0] / n, y = q[1] / n, z = q[2] / n, w = q[3] / n;
    r[0][0] = 1.0 - 2.0 * (y * y + z * z);
    r[0][1] = 2.0 * (x * y - z * w);
    r[0][2] = 2.0 * (x * z + y * w);
    r[1][0] = 2.0 * (x * y + z * w);
    r[1][1] = 1.0 - 2.0 * (x * x + z * z);
    r[1][2] = 2.0 * (y * z - x * w);
    r[2][0] = 2.0 * (x * z - y * w);
    r[2][1] = 2.0 * (y * z + x * w);
    r[2][2] = 1.0 - 2.0 * (x * x + y * y);
}

}  // namespace

extern "C" {

int32_t osn_ta_pose_rotation(const double pose[7], double R_wc[9]) {
    if (!pose || !R_wc) return 1;
    const double qn = pose[3] * pose[3] + pose[4] * pose[4] + pose[5] * pose[5] + pose[6] * pose[6];
    if (!(qn > 0.0) || !std::isfinite(qn)) return 1;
    double r[3][3];
    quat_matrix(pose + 3, r);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double c = 0.0;
            for (int k = 0; k < 3; ++k) c += r[i][k] * kNedRCam[k][j];
            R_wc[i * 3 + j] = c;
        }
    return 0;
}

int32_t osn_ta_frame(const double pose[7], int32_t width, int32_t height, double fx, double fy,
                     double cx, double cy, float max_range_m, OsnTsdfHipFrame* out) {
    if (!out || width <= 0 || height <= 0 || fx == 0.0 || fy == 0.0) return 1;
    double R[9];
    if (const int32_t rc = osn_ta_pose_rotation(pose, R)) return rc;
    std::memset(out, 0, sizeof(*out));
    out->width = width;
    out->height = height;
    out->fx = (float)fx;
    out->fy = (float)fy;
    out->cx = (float)cx;
    out->cy = (float)cy;
    for (int i = 0; i < 9; ++i) {
        out->R_wc64[i] = R[i];
        out->R_wc[i] = (float)R[i];
    }
    for (int i = 0; i < 3; ++i) {
        out->t_wc64[i] = pose[i];
        out->t_wc[i] = (float)pose[i];
    }
    out->max_range_m = max_range_m;
    return 0;
}

int32_t osn_ta_load_poses(const char* path, double* poses, int32_t cap, int32_t* n) {
    if (!path || !n || cap < 0 || (cap > 0 && !poses)) return 1;
    *n = 0;
    std::FILE* f = std::fopen(path, "r");
    if (!f) return 2;
    char line[1024];
    int32_t count = 0, rc = 0;
    while (std::fgets(line, sizeof line, f)) {
        // whitespace-separated tokens; those that parse whole as numbers
        // count, the rest are ignored, and a pose needs exactly 7 (lib.rs:392-398)
        double v[8];
        int got = 0, tokens = 0;
        char* p = line;
        for (;;) {
            while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
            if (!*p) break;
            char* tok_end = p;
            while (*tok_end && *tok_end != ' ' && *tok_end != '\t' && *tok_end != '\r' &&
                   *tok_end != '\n')
                ++tok_end;
            ++tokens;
            char* end = nullptr;
            const double x = std::strtod(p, &end);
            if (end == tok_end && got < 8) v[got++] = x;
            p = tok_end;
        }
        if (tokens == 0) continue;  // blank line
        if (got != 7) { rc = 3; break; }
        if (count < cap) std::memcpy(poses + (size_t)count * 7, v, 7 * sizeof(double));
        ++count;
    }
    std::fclose(f);
    *n = count;
    if (rc) return rc;
    return count > cap ? 5 : 0;
}

void osn_ta_depth_from_rgba(const uint8_t* rgba, int64_t n_pixels, float* depth) {
    for (int64_t i = 0; i < n_pixels; ++i) {
        const uint8_t* px = rgba + i * 4;
        const uint8_t b[4] = {px[2], px[1], px[0], px[3]};  // B, G, R, A, little-endian
        uint32_t u = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) |
                     ((uint32_t)b[3] << 24);
        std::memcpy(depth + i, &u, 4);
    }
}

}  // extern "C"
