// wc_scenes.cpp -- seeded initial states (the step before the SPH path; the reference builds its
// dam break in src/core/Fluid.cpp:104-134 with srand(0)/rand()).  Host code, no GPU.
//
// Random numbers: SplitMix64.  A particle's jitter is drawn x, then y, then z:
//   u = (next() >> 40) * 2^-24  (24 random bits, exactly representable),  j = (2u - 1) * amplitude.
// Positions: origin + ((float)i + 0.5f) * spacing + j, evaluated left to right in binary32.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/watercube.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float jitter(float amplitude) {
        const float u = (float)(next() >> 40) * (1.0f / 16777216.0f);
        return (2.0f * u - 1.0f) * amplitude;
    }
};

void set_record(wc_particle* p, float x, float y, float z, const float* v) {
    p->position[0] = x;
    p->position[1] = y;
    p->position[2] = z;
    p->density = 0.0f;
    p->velocity[0] = v ? v[0] : 0.0f;
    p->velocity[1] = v ? v[1] : 0.0f;
    p->velocity[2] = v ? v[2] : 0.0f;
    p->pressure = 0.0f;
}

}  // namespace

extern "C" {

WC_API int wc_scene_block(wc_particle* out, uint32_t nx, uint32_t ny, uint32_t nz, const float origin[3],
                          float spacing, float jitter, const float velocity[3], uint64_t seed) {
    if (!out || !origin || !(spacing > 0.0f) || !(jitter >= 0.0f)) return WC_ERR_INVALID_ARG;
    SplitMix64 rng{seed};
    size_t idx = 0;
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                const float px = origin[0] + ((float)x + 0.5f) * spacing;
                const float py = origin[1] + ((float)y + 0.5f) * spacing;
                const float pz = origin[2] + ((float)z + 0.5f) * spacing;
                const float jx = rng.jitter(jitter);
                const float jy = rng.jitter(jitter);
                const float jz = rng.jitter(jitter);
                set_record(&out[idx++], px + jx, py + jy, pz + jz, velocity);
            }
    return WC_OK;
}

WC_API int wc_scene_sphere(wc_particle* out, uint32_t count, const float center[3], float spacing, float jitter,
                           const float velocity[3], uint64_t seed) {
    if (!out || !center || !(spacing > 0.0f) || !(jitter >= 0.0f)) return WC_ERR_INVALID_ARG;
    if (count == 0) return WC_OK;
    // Lattice points (i + 0.5) * spacing of the unit cube nearest to `center`: the `count`
    // smallest squared distances (ties by lattice index), emitted in x-fastest lattice order.
    const double r_est = cbrt(3.0 * (double)count / (4.0 * M_PI)) * spacing * 1.25 + 2.0 * spacing;
    const int64_t lo[3] = {(int64_t)floor((center[0] - r_est) / spacing), (int64_t)floor((center[1] - r_est) / spacing),
                           (int64_t)floor((center[2] - r_est) / spacing)};
    const int64_t hi[3] = {(int64_t)ceil((center[0] + r_est) / spacing), (int64_t)ceil((center[1] + r_est) / spacing),
                           (int64_t)ceil((center[2] + r_est) / spacing)};
    struct Cand {
        double d2;
        int64_t i, j, k;
    };
    std::vector<Cand> cand;
    for (int64_t k = lo[2]; k <= hi[2]; ++k)
        for (int64_t j = lo[1]; j <= hi[1]; ++j)
            for (int64_t i = lo[0]; i <= hi[0]; ++i) {
                const double x = (i + 0.5) * spacing - center[0];
                const double y = (j + 0.5) * spacing - center[1];
                const double z = (k + 0.5) * spacing - center[2];
                cand.push_back({x * x + y * y + z * z, i, j, k});
            }
    if (cand.size() < count) return WC_ERR_INVALID_ARG;
    auto less = [](const Cand& a, const Cand& b) {
        if (a.d2 != b.d2) return a.d2 < b.d2;
        if (a.k != b.k) return a.k < b.k;
        if (a.j != b.j) return a.j < b.j;
        return a.i < b.i;
    };
    std::nth_element(cand.begin(), cand.begin() + (count - 1), cand.end(), less);
    cand.resize(count);
    std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
        if (a.k != b.k) return a.k < b.k;
        if (a.j != b.j) return a.j < b.j;
        return a.i < b.i;
    });
    for (const Cand& c : cand) {
        if (c.i < 0 || c.j < 0 || c.k < 0 || (c.i + 1) * (double)spacing > 1.0 || (c.j + 1) * (double)spacing > 1.0 ||
            (c.k + 1) * (double)spacing > 1.0)
            return WC_ERR_INVALID_ARG;
    }
    SplitMix64 rng{seed};
    for (uint32_t n = 0; n < count; ++n) {
        const Cand& c = cand[n];
        const float px = ((float)c.i + 0.5f) * spacing;
        const float py = ((float)c.j + 0.5f) * spacing;
        const float pz = ((float)c.k + 0.5f) * spacing;
        const float jx = rng.jitter(jitter);
        const float jy = rng.jitter(jitter);
        const float jz = rng.jitter(jitter);
        set_record(&out[n], px + jx, py + jy, pz + jz, velocity);
    }
    return WC_OK;
}

WC_API int wc_scene_reference(wc_particle* out, uint32_t n, float particle_radius, uint64_t seed) {
    if (!out || !(particle_radius > 0.0f)) return WC_ERR_INVALID_ARG;
    // Fluid.cpp:110-131: distance = 1.75 r, d = ceil(cbrt(n)), jitter in [-distance/4, distance/4)
    // (getJitter: rand()/RAND_MAX * jitter - jitter/2 with jitter = distance/2).
    const float distance = particle_radius * 1.75f;
    const int d = (int)ceil(cbrt((double)n));
    const float amplitude = distance * 0.5f * 0.5f;
    SplitMix64 rng{seed};
    for (int z = 0; z < d; ++z)
        for (int y = 0; y < d; ++y)
            for (int x = 0; x < d; ++x) {
                const int64_t index = (int64_t)z * d * d + (int64_t)y * d + x;
                if (index >= (int64_t)n) return WC_OK;
                const float jx = rng.jitter(amplitude);
                const float jy = rng.jitter(amplitude);
                const float jz = rng.jitter(amplitude);
                set_record(&out[index], (float)x * distance + jx, (float)y * distance + jy, (float)z * distance + jz,
                           nullptr);
            }
    return WC_OK;
}

}  // extern "C"
