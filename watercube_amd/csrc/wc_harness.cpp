// wc_harness.cpp -- headless C++ driver of the SPH step through watercube::Fluid (the reference
// drives the same stage from its Cinder frame loop, src/WaterCubeApp.cpp:76-103).
//
//   wc_harness <workload> <steps> [--seed S] [--out FILE]
//
// workloads: dam_break_16k (16,384 particles, dt 2e-4), block_1m (1,000,000, dt 1e-4),
// reference_default (80,000, dt = 1/60 s * 0.012).  Prints one JSON line; --out writes the final
// particle buffer as raw 32-byte records.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/watercube.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <workload> <steps> [--seed S] [--out FILE]\n", argv[0]);
        return 2;
    }
    const std::string name = argv[1];
    const long steps = std::strtol(argv[2], nullptr, 10);
    uint64_t seed = 0;
    const char* out = nullptr;
    for (int i = 3; i + 1 < argc; i += 2) {
        if (!std::strcmp(argv[i], "--seed")) seed = std::strtoull(argv[i + 1], nullptr, 10);
        else if (!std::strcmp(argv[i], "--out")) out = argv[i + 1];
    }
    try {
        int preset;
        float dt;
        std::vector<wc_particle> init;
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        if (name == "dam_break_16k") {
            preset = WC_PRESET_DAM_BREAK_16K;
            dt = 2e-4f;
            init.resize(32 * 32 * 16);
            watercube::check(wc_scene_block(init.data(), 32, 32, 16, zero, 0.015f, 0.0015f, zero, seed), "scene");
        } else if (name == "block_1m") {
            preset = WC_PRESET_BLOCK_1M;
            dt = 1e-4f;
            init.resize(1000000);
            watercube::check(wc_scene_block(init.data(), 100, 100, 100, zero, 0.008f, 0.0008f, zero, seed), "scene");
        } else if (name == "reference_default") {
            preset = WC_PRESET_REFERENCE_DEFAULT;
            dt = (1.0f / 60.0f) * 0.012f;
        } else {
            std::fprintf(stderr, "unknown workload %s\n", name.c_str());
            return 2;
        }
        watercube::Fluid fluid(preset);
        if (!init.empty()) fluid.initialParticles(init);
        fluid.setup(seed);
        const auto t0 = std::chrono::steady_clock::now();
        fluid.step(dt, (uint32_t)steps);
        fluid.synchronize();
        const auto t1 = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(t1 - t0).count();
        const std::vector<wc_particle> state = fluid.particles(1);
        double checksum = 0.0;
        for (const wc_particle& p : state) checksum += (double)p.position[0] + p.position[1] + p.position[2];
        if (out) {
            FILE* f = std::fopen(out, "wb");
            if (!f || std::fwrite(state.data(), sizeof(wc_particle), state.size(), f) != state.size()) {
                std::fprintf(stderr, "cannot write %s\n", out);
                return 1;
            }
            std::fclose(f);
        }
        const double n = (double)fluid.numParticles();
        std::printf("{\"workload\": \"%s\", \"particles\": %.0f, \"steps\": %ld, \"dt\": %g, \"ms_per_step\": %.6f, "
                    "\"particle_steps_per_s\": %.6e, \"position_checksum\": %.9e}\n",
                    name.c_str(), n, steps, (double)dt, steps ? s * 1e3 / steps : 0.0, steps ? n * steps / s : 0.0,
                    checksum);
    } catch (const watercube::Error& e) {
        std::fprintf(stderr, "wc_harness: %s (code %d)\n", e.what(), e.code());
        return 1;
    }
    return 0;
}
