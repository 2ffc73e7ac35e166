// watercube.hpp -- C++ host classes over the C ABI of watercube.h, shaped like the reference's
// src/core stage classes so that a caller of core::Fluid / core::Sort finds the same methods.
//
//   core::Fluid (src/core/Fluid.h:32-141)   ->  watercube::Fluid
//   core::Sort  (src/core/Sort.h:20-71)     ->  watercube::Sort
//
// Differences from the reference, each deliberate:
//   - the fluent setters return a reference to the object (the reference's thisRef() returns a
//     copy, src/core/Fluid.h:85, so its chained setters configure a temporary);
//   - the GPU buffers live in HBM behind a wc_ctx, not in GL buffer objects; the getters return
//     host copies;
//   - failures throw watercube::Error carrying the C ABI's status code and message.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "watercube.h"

namespace watercube {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const char* where) {
    if (rc != WC_OK) throw Error(rc, std::string(where) + ": " + wc_last_error());
}

// core::Sort: the grid build of one context.
class Sort {
public:
    explicit Sort(wc_ctx* ctx, const wc_params& p) : ctx_(ctx), params_(p) {}

    // Sort::run(buffer1, buffer2) (src/core/Sort.cpp:254-267)
    void run() { check(wc_sort_run(ctx_), "wc_sort_run"); }

    // Sort::getCountBuffer / getOffsetBuffer / getSortedBuffer (src/core/Sort.h:35-37)
    std::vector<uint32_t> getCountBuffer() const { return cells(&wc_get_counts, "wc_get_counts"); }
    std::vector<uint32_t> getOffsetBuffer() const { return cells(&wc_get_offsets, "wc_get_offsets"); }
    std::vector<uint32_t> getSortedBuffer() const {
        std::vector<uint32_t> v(params_.num_particles);
        check(wc_get_sorted(ctx_, v.data(), (uint32_t)v.size()), "wc_get_sorted");
        return v;
    }
    uint32_t numBins() const {
        const uint32_t r = (uint32_t)params_.grid_res;
        return r * r * r;
    }

private:
    using Getter = int (*)(wc_ctx*, uint32_t*, uint32_t);
    std::vector<uint32_t> cells(Getter g, const char* name) const {
        std::vector<uint32_t> v(numBins());
        check(g(ctx_, v.data(), (uint32_t)v.size()), name);
        return v;
    }
    wc_ctx* ctx_;
    wc_params params_;
};

// core::Fluid, simulation half.
class Fluid {
public:
    // Fluid::Fluid (src/core/Fluid.cpp:9-27): the reference defaults, or a preset.
    explicit Fluid(int preset = WC_PRESET_REFERENCE_DEFAULT) {
        check(wc_params_preset(&params_, preset), "wc_params_preset");
        gravity_direction_[0] = 0.0f;
        gravity_direction_[1] = -1.0f;
        gravity_direction_[2] = 0.0f;
        gravity_strength_ = -params_.gravity[1];
    }
    explicit Fluid(const wc_params& p) : params_(p) {}
    Fluid(const Fluid&) = delete;
    Fluid& operator=(const Fluid&) = delete;
    ~Fluid() { wc_destroy(ctx_); }

    // Fluent setters (src/core/Fluid.cpp:31-84).
    Fluid& numParticles(int n) { params_.num_particles = (uint32_t)n; return *this; }
    int numParticles() const { return (int)params_.num_particles; }
    Fluid& gridRes(int r) { params_.grid_res = r; return *this; }
    Fluid& size(float s) { params_.size = s; return *this; }
    // h = 4 r and m = 8 r, as Fluid::setup derives them (src/core/Fluid.cpp:209-210)
    Fluid& particleRadius(float r) {
        params_.particle_radius = r;
        params_.kernel_radius = r * 4.0f;
        params_.particle_mass = r * 8.0f;
        return *this;
    }
    Fluid& viscosityCoefficient(float c) { params_.viscosity = c; return *this; }
    Fluid& stiffness(float s) { params_.stiffness = s; return *this; }
    Fluid& restDensity(float d) { params_.rest_density = d; return *this; }
    Fluid& restPressure(float p) { params_.rest_pressure = p; return *this; }
    Fluid& gravityStrength(float g) { gravity_strength_ = g; applyGravity(); return *this; }
    Fluid& gravityDirection(float x, float y, float z) {
        gravity_direction_[0] = x; gravity_direction_[1] = y; gravity_direction_[2] = z;
        applyGravity();
        return *this;
    }
    // Fluid::updateGravity (src/core/Fluid.cpp:259-263) with rotate_gravity_ on: the direction
    // from the fluid's orientation q = (w, x, y, z) times the strength, pushed to a set-up context.
    void updateGravity(const float q_wxyz[4]) {
        check(wc_gravity_direction(q_wxyz, gravity_direction_), "wc_gravity_direction");
        applyGravity();
        if (ctx_) check(wc_set_params(ctx_, &params_), "wc_set_params");
    }
    // Mouse ray in the fluid's frame (Fluid::setMouseRay, src/core/Fluid.h:52).
    void setMouseRay(const float origin[3], const float direction[3]) {
        for (int a = 0; a < 3; ++a) {
            params_.camera_position[a] = origin[a];
            params_.mouse_ray_direction[a] = direction[a];
        }
        params_.mouse_force = 1;
        if (ctx_) check(wc_set_params(ctx_, &params_), "wc_set_params");
    }
    Fluid& initialParticles(std::vector<wc_particle> p) { initial_ = std::move(p); return *this; }

    // Fluid::setup (src/core/Fluid.cpp:203-235): derive constants, build the initial state
    // (generateInitialParticles unless one was given), allocate and upload.
    Fluid& setup(uint64_t seed = 0) {
        check(wc_params_derive(&params_), "wc_params_derive");
        if (initial_.empty()) generateInitialParticles(seed);
        if (initial_.size() != params_.num_particles)
            throw Error(WC_ERR_INVALID_ARG, "initial state size differs from numParticles");
        wc_destroy(ctx_);
        ctx_ = nullptr;
        check(wc_create(&params_, &ctx_), "wc_create");
        check(wc_upload(ctx_, initial_.data(), (uint32_t)initial_.size()), "wc_upload");
        sort_ = std::make_unique<Sort>(ctx_, params_);
        return *this;
    }

    // Fluid::generateInitialParticles (src/core/Fluid.cpp:104-134)
    void generateInitialParticles(uint64_t seed = 0) {
        initial_.assign(params_.num_particles, wc_particle{});
        check(wc_scene_reference(initial_.data(), params_.num_particles, params_.particle_radius, seed),
              "wc_scene_reference");
    }

    // Fluid::update (src/core/Fluid.cpp:342-354): dt = time * time_scale (Fluid.cpp:308).
    void update(double time) { step((float)time * params_.time_scale, 1); }

    // nsteps updates at a fixed step dt, replayed as captured graphs.
    void step(float dt, uint32_t nsteps) { check(wc_step(ctx_, dt, nsteps), "wc_step"); }

    // The passes of one update, for callers that interleave their own work.
    void runDensityProg() { check(wc_density(ctx_), "wc_density"); }
    void runUpdateProg(float dt) { check(wc_update(ctx_, dt), "wc_update"); }
    // Fluid::runAdvectProg (src/core/Fluid.cpp:326-336), on buffer 1; dt is the step itself
    void runAdvectProg(float dt) { check(wc_advect(ctx_, dt), "wc_advect"); }

    void synchronize() { check(wc_synchronize(ctx_), "wc_synchronize"); }

    std::vector<wc_particle> particles(int which = 1) const {
        std::vector<wc_particle> v(params_.num_particles);
        check(wc_download(ctx_, which, v.data(), (uint32_t)v.size()), "wc_download");
        return v;
    }

    Sort& sort() { return *sort_; }
    const wc_params& params() const { return params_; }
    wc_ctx* context() { return ctx_; }

private:
    void applyGravity() {
        for (int a = 0; a < 3; ++a) params_.gravity[a] = gravity_direction_[a] * gravity_strength_;
    }
    wc_params params_{};
    float gravity_direction_[3] = {0.0f, -1.0f, 0.0f};
    float gravity_strength_ = 0.0f;
    std::vector<wc_particle> initial_;
    wc_ctx* ctx_ = nullptr;
    std::unique_ptr<Sort> sort_;
};

}  // namespace watercube
