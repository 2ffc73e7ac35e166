// wc_harness.cpp -- headless C++ driver of the SPH step through watercube::Fluid (the reference
// drives the same stage from its Cinder frame loop, src/WaterCubeApp.cpp:76-103).
//
//   wc_harness <workload> <steps> [--seed S] [--out STEM] [--in STEM]
//
// workloads: dam_break_16k (16,384 particles, dt 2e-4), block_1m (1,000,000, dt 1e-4),
// reference_default (80,000, dt = 1/60 s * 0.012).  Prints one JSON line.  --out writes the final
// state as a snapshot: STEM.bin, the particle buffer as raw 32-byte records, and STEM.json, the
// parameter block (raw bytes in hex), particle count, step index and dt.  --in resumes from such
// a snapshot (its parameters, records, step and dt) instead of the workload's initial state.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/watercube.hpp"

namespace {

std::string stem_of(const char* path) {
    std::string s(path);
    for (const char* ext : {".bin", ".json"}) {
        const size_t n = std::strlen(ext);
        if (s.size() > n && s.compare(s.size() - n, n, ext) == 0) return s.substr(0, s.size() - n);
    }
    return s;
}

bool write_snapshot(const std::string& stem, const wc_params& p, const std::vector<wc_particle>& state, long step,
                    float dt) {
    FILE* f = std::fopen((stem + ".bin").c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(state.data(), sizeof(wc_particle), state.size(), f) == state.size();
    std::fclose(f);
    if (!ok) return false;
    f = std::fopen((stem + ".json").c_str(), "w");
    if (!f) return false;
    std::string hex;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&p);
    char h[3];
    for (size_t i = 0; i < sizeof(p); ++i) {
        std::snprintf(h, sizeof h, "%02x", b[i]);
        hex += h;
    }
    std::fprintf(f,
                 "{\"format\": \"watercube-snapshot-1\", \"particles\": %zu, \"record_bytes\": 32, \"fields\": "
                 "[\"px\", \"py\", \"pz\", \"density\", \"vx\", \"vy\", \"vz\", \"pressure\"], \"step\": %ld, "
                 "\"dt\": %.9g, \"params_hex\": \"%s\"}\n",
                 state.size(), step, (double)dt, hex.c_str());
    return std::fclose(f) == 0;
}

// The value text after "key": in a one-level JSON object (numbers and strings only).
std::string json_value(const std::string& text, const char* key) {
    const std::string k = std::string("\"") + key + "\":";
    size_t i = text.find(k);
    if (i == std::string::npos) return "";
    i += k.size();
    while (i < text.size() && text[i] == ' ') ++i;
    if (i < text.size() && text[i] == '"') {
        const size_t e = text.find('"', i + 1);
        return e == std::string::npos ? "" : text.substr(i + 1, e - i - 1);
    }
    size_t e = i;
    while (e < text.size() && text[e] != ',' && text[e] != '}' && text[e] != '\n') ++e;
    return text.substr(i, e - i);
}

bool read_snapshot(const std::string& stem, wc_params& p, std::vector<wc_particle>& state, long& step, float& dt) {
    FILE* f = std::fopen((stem + ".json").c_str(), "r");
    if (!f) return false;
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    const std::string hex = json_value(text, "params_hex");
    if (hex.size() != 2 * sizeof(p)) return false;
    unsigned char* b = reinterpret_cast<unsigned char*>(&p);
    for (size_t i = 0; i < sizeof(p); ++i) b[i] = (unsigned char)std::strtoul(hex.substr(2 * i, 2).c_str(), nullptr, 16);
    const long n = std::strtol(json_value(text, "particles").c_str(), nullptr, 10);
    step = std::strtol(json_value(text, "step").c_str(), nullptr, 10);
    dt = std::strtof(json_value(text, "dt").c_str(), nullptr);
    if (n <= 0 || (uint32_t)n != p.num_particles) return false;
    state.resize((size_t)n);
    f = std::fopen((stem + ".bin").c_str(), "rb");
    if (!f) return false;
    const bool ok = std::fread(state.data(), sizeof(wc_particle), state.size(), f) == state.size();
    std::fclose(f);
    return ok;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <workload> <steps> [--seed S] [--out STEM] [--in STEM]\n", argv[0]);
        return 2;
    }
    const std::string name = argv[1];
    const long steps = std::strtol(argv[2], nullptr, 10);
    uint64_t seed = 0;
    const char* out = nullptr;
    const char* in = nullptr;
    for (int i = 3; i + 1 < argc; i += 2) {
        if (!std::strcmp(argv[i], "--seed")) seed = std::strtoull(argv[i + 1], nullptr, 10);
        else if (!std::strcmp(argv[i], "--out")) out = argv[i + 1];
        else if (!std::strcmp(argv[i], "--in")) in = argv[i + 1];
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
        long step0 = 0;
        wc_params base;
        watercube::check(wc_params_preset(&base, preset), "wc_params_preset");
        if (in) {
            std::vector<wc_particle> state;
            if (!read_snapshot(stem_of(in), base, state, step0, dt)) {
                std::fprintf(stderr, "cannot read snapshot %s\n", in);
                return 1;
            }
            init = std::move(state);
        }
        watercube::Fluid fluid(base);
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
        if (out && !write_snapshot(stem_of(out), fluid.params(), state, step0 + steps, dt)) {
            std::fprintf(stderr, "cannot write snapshot %s\n", out);
            return 1;
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
