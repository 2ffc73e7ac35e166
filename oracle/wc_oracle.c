/*
 * wc_oracle.c -- CPU restatement of WaterCube's per-step SPH pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is compared against and
 * the CPU baseline the benchmark times beside it.  Only tests/, the smoke check and the
 * benchmark's cpu_baseline leg load it; the library never links or calls it.
 *
 * It transcribes the five live GLSL compute shaders of the reference and the order in which
 * src/core/Fluid.cpp and src/core/Sort.cpp dispatch them:
 *   count.comp:25-36        -> wco_count
 *   linearScan.comp:15-28   -> wco_linear_scan
 *   reorder.comp:32-45      -> wco_reorder
 *   density.comp:40-137     -> wco_density
 *   update.comp:48-232      -> wco_update   (file lines of the shader, not of this file)
 *   Sort::run  (src/core/Sort.cpp:254-267)     -> wco_sort_run
 *   Fluid::update (src/core/Fluid.cpp:342-354) -> wco_step
 * and the sixth, advect.comp:20-59 (compiled by the reference but never dispatched), as
 * wco_advect.
 *
 * GLSL leaves some arithmetic to the driver.  This restatement fixes it as follows, and the
 * GPU library reproduces the same choices:
 *   - IEEE-754 binary32, round to nearest even, correctly rounded '/' and sqrt, subnormals kept;
 *   - no FMA contraction (built with -ffp-contract=off), left-to-right association as written;
 *   - pow(x, 3) = (x*x)*x and pow(x, 2) = x*x;
 *   - length(v) = sqrt((v.x*v.x + v.y*v.y) + v.z*v.z); cross as in the GLSL specification;
 *   - ivec3(p / binSize) truncates toward zero, clamp(.., 0, R-1); NaN and negative quotients
 *     map to cell 0, quotients >= R to R-1 (the clamp's result for every finite input);
 *   - min/max/clamp follow C fminf/fmaxf;
 *   - the reorder's atomicAdd (reorder.comp:42) hands out slots in ascending particle order,
 *     so particles keep their input order inside each cell (a stable counting sort).
 *
 * Parity pins (what checks each function; all in tests/):
 *   wco_cell_key, wco_count      closed form on lattice points, brute-force histogram, sum = N
 *   wco_linear_scan              numpy cumsum, wraparound-free sizes
 *   wco_reorder                  numpy stable argsort of keys; every slot range holds its cell
 *   wco_density                  isolated particle = m*W(0); pair closed form; O(N^2)
 *                                brute force in NEIGHBORHOOD order on small scenes (bit-exact);
 *                                lattice interior within a few % of rest density
 *   wco_update                   single free-falling particle (closed form); pair forces
 *                                (analytic spiky / viscosity values); uniform velocity gives
 *                                zero viscosity; wall clamp and damping single-step outcomes;
 *                                O(N^2) brute-force force sum in NEIGHBORHOOD order
 *   wco_wall_density, wco_wall_forces, wco_mouse_force   hand-derived values at chosen points
 *   wco_advect                   x + v dt per axis; clamp and damping single-step outcomes at
 *                                the shader's border 0.01 and damping 0.3
 *   kernel constants             closed forms of src/core/Fluid.cpp:214-216
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/watercube.h"

#define WCO_API __attribute__((visibility("default")))

/* ---------------------------------------------------------------- scalar helpers --------- */

/* density.comp:53-55  poly6Kernel(r) = pow(h*h - r*r, 3) * poly6KernelConst */
WCO_API float wco_poly6(const wc_params* P, float r) {
    const float h = P->kernel_radius;
    const float a = h * h - r * r;
    return ((a * a) * a) * P->poly6_const;
}

/* update.comp:62-64  spikyKernel(r, d) = pow(h - d, 2) * (r / d) * spikyKernelConst */
WCO_API void wco_spiky(const wc_params* P, const float r[3], float d, float out[3]) {
    const float a = P->kernel_radius - d;
    const float s = a * a;
    for (int c = 0; c < 3; ++c) out[c] = (s * (r[c] / d)) * P->spiky_const;
}

/* update.comp:67-69  Wvis(r) = (h - r) * viscosityKernelConst */
WCO_API float wco_wvis(const wc_params* P, float r) { return (P->kernel_radius - r) * P->visc_const; }

static inline float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }

static inline int cell_axis(float x, float bin, int R) {
    const float q = x / bin;
    if (!(q >= 0.0f)) return 0;      /* negative or NaN: int() <= 0, clamped to 0 */
    if (q >= (float)R) return R - 1; /* int() >= R, clamped to R-1 */
    int c = (int)q;                  /* truncation toward zero, as GLSL ivec3(float) */
    return c > R - 1 ? R - 1 : c;
}

/* count.comp:31-33, reorder.comp:38-40, density.comp:88, update.comp:141 */
WCO_API void wco_cell(const wc_params* P, const float pos[3], int c[3]) {
    c[0] = cell_axis(pos[0], P->bin_size, P->grid_res);
    c[1] = cell_axis(pos[1], P->bin_size, P->grid_res);
    c[2] = cell_axis(pos[2], P->bin_size, P->grid_res);
}

WCO_API uint32_t wco_cell_key(const wc_params* P, const float pos[3]) {
    int c[3];
    wco_cell(P, pos, c);
    const int R = P->grid_res;
    return (uint32_t)(c[2] * R * R + c[1] * R + c[0]);
}

/* density.comp:57-79 wallDensity, including the z branch that tests p.y (line 74). */
WCO_API float wco_wall_density(const wc_params* P, const float p[3]) {
    const float h = P->kernel_radius, m = P->particle_mass, size = P->size;
    float density = 0.0f;
    if (p[0] < h) {
        density += m * wco_poly6(P, p[0]);
    } else if (p[0] > size - h) {
        density += m * wco_poly6(P, size - p[0]);
    }
    if (p[1] < h) {
        density += m * wco_poly6(P, p[1]);
    } else if (p[1] > size - h) {
        density += m * wco_poly6(P, size - p[1]);
    }
    if (p[2] < h) {
        density += m * wco_poly6(P, p[2]);
    } else if (p[1] > size - h) {
        density += m * wco_poly6(P, size - p[2]);
    }
    return density * 4.0f;
}

/* update.comp:71-100 wallForces.  GLSL vec3.length() is the component count, 3, so every call
 * passes d = 3 to spikyKernel. */
WCO_API void wco_wall_forces(const wc_params* P, const float p[3], float out[3]) {
    const float h = P->kernel_radius, size = P->size;
    float f[3] = {0.0f, 0.0f, 0.0f};
    float r[3], w[3];
    if (p[0] < h) {
        r[0] = 0.0f - p[0]; r[1] = p[1] - p[1]; r[2] = p[2] - p[2];
        wco_spiky(P, r, 3.0f, w);
        for (int c = 0; c < 3; ++c) f[c] += w[c];
    } else if (p[0] > size - h) {
        r[0] = size - p[0]; r[1] = p[1] - p[1]; r[2] = p[2] - p[2];
        wco_spiky(P, r, 3.0f, w);
        for (int c = 0; c < 3; ++c) f[c] += w[c];
    }
    if (p[1] < h) {
        r[0] = p[0] - p[0]; r[1] = 0.0f - p[1]; r[2] = p[2] - p[2];
        wco_spiky(P, r, 3.0f, w);
        for (int c = 0; c < 3; ++c) f[c] += w[c];
    } else if (p[1] > size - h) {
        r[0] = p[0] - p[0]; r[1] = size - p[1]; r[2] = p[2] - p[2];
        wco_spiky(P, r, 3.0f, w);
        for (int c = 0; c < 3; ++c) f[c] += w[c];
    }
    if (p[2] < h) {
        r[0] = p[0] - p[0]; r[1] = p[1] - p[1]; r[2] = 0.0f - p[2];
        wco_spiky(P, r, 3.0f, w);
        for (int c = 0; c < 3; ++c) f[c] += w[c];
    } else if (p[2] > size - h) {
        r[0] = p[0] - p[0]; r[1] = p[1] - p[1]; r[2] = size - p[2];
        wco_spiky(P, r, 3.0f, w);
        for (int c = 0; c < 3; ++c) f[c] += w[c];
    }
    for (int c = 0; c < 3; ++c) out[c] = f[c] * 0.01f;
}

/* update.comp:105-132 intersectAABB + mouseForce for one particle. */
WCO_API void wco_mouse_force(const wc_params* P, const float pos[3], float pressure, float out[3]) {
    const float* o = P->camera_position;
    const float* dir = P->mouse_ray_direction;
    const float size = P->size, h = P->kernel_radius;
    float t1[3], t2[3];
    for (int c = 0; c < 3; ++c) {
        const float tmin = (0.0f - o[c]) / dir[c];
        const float tmax = (size - o[c]) / dir[c];
        t1[c] = fminf(tmin, tmax);
        t2[c] = fmaxf(tmin, tmax);
    }
    const float tnear = fmaxf(fmaxf(t1[0], t1[1]), t1[2]);
    const float tfar = fminf(fminf(t2[0], t2[1]), t2[2]);
    out[0] = out[1] = out[2] = 0.0f;
    if (tnear > tfar) return;
    float tm[3] = {pos[0] - o[0], pos[1] - o[1], pos[2] - o[2]};
    /* cross(x, y) = (x1*y2 - y1*x2, x2*y0 - y2*x0, x0*y1 - y0*x1) */
    const float cx = dir[1] * tm[2] - tm[1] * dir[2];
    const float cy = dir[2] * tm[0] - tm[2] * dir[0];
    const float cz = dir[0] * tm[1] - tm[0] * dir[1];
    const float dist = len3(cx, cy, cz);
    if (dist > h) return;
    float w[3];
    wco_spiky(P, tm, dist + 1e-16f, w);
    const float s = (-P->particle_mass) * pressure;
    for (int c = 0; c < 3; ++c) out[c] = (s * w[c]) * 0.00001f;
}

/* density.comp:132-133 (cubic, live) or :129 (linear, commented). */
WCO_API float wco_eos(const wc_params* P, float density) {
    if (P->eos == WC_EOS_LINEAR) return P->rest_pressure + P->stiffness * (density - P->rest_density);
    const float x = density / P->rest_density;
    return P->rest_pressure + P->stiffness * (((x * x) * x) - 1.0f);
}

/* ---------------------------------------------------------------- grid build ------------ */

/* count.comp:25-36 over all particles; keys[i] receives the cell index of particle i.
 * counts must hold R^3 zeros on entry (Sort::clearCountBuffer, Sort.cpp:135-141). */
WCO_API void wco_count(const wc_params* P, const wc_particle* in, uint32_t n, uint32_t* keys, uint32_t* counts) {
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t k = wco_cell_key(P, in[i].position);
        keys[i] = k;
        counts[k] += 1u;
    }
}

/* linearScan.comp:15-28: exclusive prefix sum in cell-index order. */
WCO_API void wco_linear_scan(const uint32_t* counts, uint32_t ncells, uint32_t* offsets) {
    uint32_t prefix = 0;
    for (uint32_t i = 0; i < ncells; ++i) {
        offsets[i] = prefix;
        prefix += counts[i];
    }
}

/* reorder.comp:32-45 with slots handed out in ascending particle order.  `cursor` must hold
 * R^3 zeros on entry and ends equal to the per-cell counts (Sort.cpp:263-264).  sorted[slot]
 * receives the input index of the particle placed at `slot` (may be NULL). */
WCO_API void wco_reorder(const wc_params* P, const wc_particle* in, uint32_t n, const uint32_t* keys,
                         const uint32_t* offsets, uint32_t* cursor, wc_particle* out, uint32_t* sorted) {
    (void)P;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t k = keys[i];
        const uint32_t slot = offsets[k] + cursor[k]++;
        out[slot] = in[i];
        if (sorted) sorted[slot] = i;
    }
}

/* Sort::run(in, out) (Sort.cpp:254-267): clear counts, count, clear offsets, scan, clear
 * counts, reorder. */
WCO_API void wco_sort_run(const wc_params* P, const wc_particle* in, wc_particle* out, uint32_t* keys,
                          uint32_t* counts, uint32_t* offsets, uint32_t* sorted) {
    const uint32_t R = (uint32_t)P->grid_res;
    const size_t ncells = (size_t)R * R * R;
    memset(counts, 0, ncells * sizeof(uint32_t));
    wco_count(P, in, P->num_particles, keys, counts);
    memset(offsets, 0, ncells * sizeof(uint32_t));
    wco_linear_scan(counts, (uint32_t)ncells, offsets);
    memset(counts, 0, ncells * sizeof(uint32_t));
    wco_reorder(P, in, P->num_particles, keys, offsets, counts, out, sorted);
}

/* ---------------------------------------------------------------- fluid passes ---------- */

static inline void neighbour_cell(int b, int d[3]) {
    /* density.comp:40-50 NEIGHBORHOOD: dx outermost, dz innermost, each -1..1 */
    d[0] = b / 9 - 1;
    d[1] = (b / 3) % 3 - 1;
    d[2] = b % 3 - 1;
}

static void density_one(const wc_params* P, wc_particle* ps, const uint32_t* counts, const uint32_t* offsets,
                        uint32_t i) {
    const int R = P->grid_res;
    const float h = P->kernel_radius, m = P->particle_mass;
    float pos[3] = {ps[i].position[0], ps[i].position[1], ps[i].position[2]};
    int coord[3];
    wco_cell(P, pos, coord);
    float density = m * wco_poly6(P, 0.0f);
    for (int b = 0; b < 27; ++b) {
        int d[3];
        neighbour_cell(b, d);
        const int nx = coord[0] + d[0], ny = coord[1] + d[1], nz = coord[2] + d[2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= R || ny >= R || nz >= R) continue;
        const uint32_t index = (uint32_t)(nz * R * R + ny * R + nx);
        const uint32_t count = counts[index];
        const uint32_t offset = offsets[index];
        for (uint32_t l = 0; l < count; ++l) {
            const uint32_t j = offset + l;
            if (j == i) continue;
            const float rx = pos[0] - ps[j].position[0];
            const float ry = pos[1] - ps[j].position[1];
            const float rz = pos[2] - ps[j].position[2];
            const float dist = len3(rx, ry, rz);
            if (dist >= h) continue;
            density += m * wco_poly6(P, dist);
        }
    }
    const float wall = P->wall_density ? wco_wall_density(P, pos) : 0.0f;
    ps[i].density = density + wall;
    ps[i].pressure = wco_eos(P, density);
}

/* density.comp:81-137 for every particle, in place.  Each particle's result depends only on
 * positions, which the pass does not write, so any thread count gives identical bits. */
WCO_API void wco_density(const wc_params* P, wc_particle* ps, const uint32_t* counts, const uint32_t* offsets,
                         int nthreads) {
    const int64_t n = (int64_t)P->num_particles;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static, 1024) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; ++i) density_one(P, ps, counts, offsets, (uint32_t)i);
    (void)nthreads;
}

static void update_one(const wc_params* P, const wc_particle* in, const uint32_t* counts, const uint32_t* offsets,
                       wc_particle* out, float dt, uint32_t i) {
    const int R = P->grid_res;
    const float h = P->kernel_radius, m = P->particle_mass;
    const wc_particle p = in[i];
    int coord[3];
    wco_cell(P, p.position, coord);
    float pf[3] = {0.0f, 0.0f, 0.0f};
    float vf[3] = {0.0f, 0.0f, 0.0f};
    float ext[3];
    for (int c = 0; c < 3; ++c) ext[c] = P->gravity[c] * p.density;

    for (int b = 0; b < 27; ++b) {
        int d[3];
        neighbour_cell(b, d);
        const int nx = coord[0] + d[0], ny = coord[1] + d[1], nz = coord[2] + d[2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= R || ny >= R || nz >= R) continue;
        const uint32_t index = (uint32_t)(nz * R * R + ny * R + nx);
        const uint32_t count = counts[index];
        const uint32_t offset = offsets[index];
        for (uint32_t l = 0; l < count; ++l) {
            const uint32_t j = offset + l;
            if (j == i) continue;
            const wc_particle* o = &in[j];
            float r[3] = {p.position[0] - o->position[0], p.position[1] - o->position[1],
                          p.position[2] - o->position[2]};
            const float dist = len3(r[0], r[1], r[2]);
            if (dist >= h) continue;
            /* update.comp:178-183 */
            const float pressure = (p.pressure + o->pressure) / (2.0f * o->density);
            if (pressure > 0.0f) {
                float w[3];
                wco_spiky(P, r, dist + 1e-16f, w);
                const float mp = m * pressure;
                for (int c = 0; c < 3; ++c) pf[c] -= mp * w[c];
            }
            /* update.comp:186-187 */
            const float wv = wco_wvis(P, dist);
            for (int c = 0; c < 3; ++c) {
                const float dv = o->velocity[c] - p.velocity[c];
                vf[c] += (m * (dv / o->density)) * wv;
            }
        }
    }

    float mouse[3] = {0.0f, 0.0f, 0.0f};
    float wall[3] = {0.0f, 0.0f, 0.0f};
    if (P->mouse_force) wco_mouse_force(P, p.position, p.pressure, mouse);
    if (P->wall_force) wco_wall_forces(P, p.position, wall);
    for (int c = 0; c < 3; ++c) ext[c] += mouse[c] + wall[c];

    float vel[3], pos[3];
    const float denom = p.density + 1e-16f;
    for (int c = 0; c < 3; ++c) {
        vf[c] *= P->viscosity;
        const float force = (pf[c] + vf[c]) + ext[c];
        const float acc = force / denom;
        const float v = p.velocity[c] + acc * dt;
        vel[c] = fminf(fmaxf(v, -P->max_speed), P->max_speed);
        pos[c] = p.position[c] + vel[c] * dt;
    }
    /* update.comp:202-227 */
    const float lo = P->border, hi = P->size - P->border;
    for (int c = 0; c < 3; ++c) {
        if (pos[c] < lo) {
            vel[c] *= -P->wall_damping;
            pos[c] = lo;
        } else if (pos[c] > hi) {
            vel[c] *= -P->wall_damping;
            pos[c] = hi;
        }
    }
    wc_particle q = p;
    for (int c = 0; c < 3; ++c) {
        q.position[c] = pos[c];
        q.velocity[c] = vel[c];
    }
    out[i] = q;
}

/* update.comp:134-232 for every particle: in -> out at the same slot. */
WCO_API void wco_update(const wc_params* P, const wc_particle* in, const uint32_t* counts, const uint32_t* offsets,
                        wc_particle* out, float dt, int nthreads) {
    const int64_t n = (int64_t)P->num_particles;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static, 1024) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; ++i) update_one(P, in, counts, offsets, out, dt, (uint32_t)i);
    (void)nthreads;
}

/* advect.comp:20-59, in place: x = x + v dt, then per axis the clamp to [border, size - border]
 * with v *= -damping, border 0.01 and damping 0.3 fixed in the shader (advect.comp:30-31). */
WCO_API void wco_advect(const wc_params* P, wc_particle* ps, uint32_t n, float dt) {
    const float damping = 0.3f, border = 0.01f;
    const float hi = P->size - border;
    for (uint32_t i = 0; i < n; ++i) {
        float* x = ps[i].position;
        float* v = ps[i].velocity;
        for (int a = 0; a < 3; ++a) x[a] = x[a] + v[a] * dt;
        for (int a = 0; a < 3; ++a) {
            if (x[a] < border) {
                v[a] = v[a] * -damping;
                x[a] = border;
            } else if (x[a] > hi) {
                v[a] = v[a] * -damping;
                x[a] = hi;
            }
        }
    }
}

/* Fluid::update (Fluid.cpp:342-354) repeated nsteps times: sort buf1 -> buf2, density on buf2,
 * update buf2 -> buf1.  keys/sorted hold N entries, counts/offsets R^3. */
WCO_API void wco_step(const wc_params* P, wc_particle* buf1, wc_particle* buf2, uint32_t* keys, uint32_t* counts,
                      uint32_t* offsets, uint32_t* sorted, float dt, uint32_t nsteps, int nthreads) {
    for (uint32_t s = 0; s < nsteps; ++s) {
        wco_sort_run(P, buf1, buf2, keys, counts, offsets, sorted);
        wco_density(P, buf2, counts, offsets, nthreads);
        wco_update(P, buf2, counts, offsets, buf1, dt, nthreads);
    }
}

/* Per-pass wall time of wco_step, for the CPU baseline: ms[0] sort, ms[1] density, ms[2] update. */
#include <time.h>
static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}
WCO_API void wco_step_timed(const wc_params* P, wc_particle* buf1, wc_particle* buf2, uint32_t* keys,
                            uint32_t* counts, uint32_t* offsets, uint32_t* sorted, float dt, uint32_t nsteps,
                            int nthreads, double ms[3]) {
    ms[0] = ms[1] = ms[2] = 0.0;
    for (uint32_t s = 0; s < nsteps; ++s) {
        double t0 = now_ms();
        wco_sort_run(P, buf1, buf2, keys, counts, offsets, sorted);
        double t1 = now_ms();
        wco_density(P, buf2, counts, offsets, nthreads);
        double t2 = now_ms();
        wco_update(P, buf2, counts, offsets, buf1, dt, nthreads);
        double t3 = now_ms();
        ms[0] += t1 - t0;
        ms[1] += t2 - t1;
        ms[2] += t3 - t2;
    }
}

WCO_API int wco_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
