// apo_cavity.cpp -- CPU restatement of the reference's cavity hit ratios
// (app/cavity/utils.cu:17-65 with the ray-generation program of
// app/cavity/shaders, and the directions of :201-242).  TEST INFRASTRUCTURE
// ONLY (part of the oracle library; never linked into the product).
//
// Each vertex centre (as FP32) casts one ray per direction; a ray counts when
// some site sphere has an intersection parameter in [0.01, 1e16].  The
// reference traces with OptiX sphere primitives; here every ray is tested
// against every site, in FP32 with the quadratic
//   b = (o - c).d, q = |o - c|^2 - r^2, t = -b -/+ sqrt(b^2 - q).
#include <cmath>
#include <cstdint>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "apo_oracle.h"

namespace {

struct V3 {
    float x, y, z;
};

// sphereHalton(n, 3) (utils.cu:201-236): base-2 radical inverse for the
// height, base-3 for the angle, with the reference's truncated pi; then
// normalised (the ray-generation program normalises each direction).
std::vector<V3> directions(uint32_t samples) {
    std::vector<V3> out(samples);
    for (uint32_t k = 0; k < samples; k++) {
        float t = 0.f;
        uint32_t kk = k;
        for (float p = 0.5f; kk; p *= .5f, kk >>= 1)
            if (kk & 1u) t += p;
        t = 2.f * t - 1.f;
        const float st = std::sqrt(1.f - t * t);
        float phi = 0.f;
        const float ip = 1.f / 3.f;
        kk = k;
        for (float p = ip; kk; p *= ip, kk /= 3u) {
            const uint32_t a = kk % 3u;
            if (a) phi += float(a) * p;
        }
        const float phirad = static_cast<float>(static_cast<double>(phi * 4.f) * 3.1415926535);
        const V3 d{st * std::cos(phirad), st * std::sin(phirad), t};
        const float inv = 1.f / std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        out[k] = {d.x * inv, d.y * inv, d.z * inv};
    }
    return out;
}

bool hits(const V3& o, const V3& d, const float* s) {
    const V3 oc{o.x - s[0], o.y - s[1], o.z - s[2]};
    const float b = oc.x * d.x + oc.y * d.y + oc.z * d.z;
    const float q = (oc.x * oc.x + oc.y * oc.y + oc.z * oc.z) - s[3] * s[3];
    const float disc = b * b - q;
    if (disc < 0.f) return false;
    const float r = std::sqrt(disc);
    const float t1 = -b - r, t2 = -b + r;
    return (t1 >= 0.01f && t1 <= 1e16f) || (t2 >= 0.01f && t2 <= 1e16f);
}

}  // namespace

extern "C" {

void orc_cavity_directions(uint32_t samples, float* xyz) {
    const std::vector<V3> d = directions(samples);
    for (uint32_t k = 0; k < samples; k++) {
        xyz[3 * k] = d[k].x;
        xyz[3 * k + 1] = d[k].y;
        xyz[3 * k + 2] = d[k].z;
    }
}

void orc_cavity_hit_ratios(const float* xyzr, uint32_t n, const double* xyzw, uint64_t count, uint32_t samples,
                           uint32_t threads, float* ratios) {
#ifdef _OPENMP
    if (threads) omp_set_num_threads(int(threads));
#else
    (void)threads;
#endif
    const std::vector<V3> dirs = directions(samples);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t v = 0; v < int64_t(count); v++) {
        const V3 o{float(xyzw[4 * v]), float(xyzw[4 * v + 1]), float(xyzw[4 * v + 2])};
        uint32_t h = 0;
        for (uint32_t s = 0; s < samples; s++) {
            for (uint32_t i = 0; i < n; i++)
                if (hits(o, dirs[s], xyzr + 4 * size_t(i))) {
                    h++;
                    break;
                }
        }
        ratios[v] = float(h) / float(samples);
    }
}

}  // extern "C"
