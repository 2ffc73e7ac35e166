/*
 * watercube.h -- C ABI of the MI355X (gfx950) SPH step library.
 *
 * The library replaces the GPU half of WaterCube's per-frame simulation: the grid build of
 * src/core/Sort.{h,cpp} (cell hash, per-cell count, prefix-sum offsets, reorder) and the two
 * fluid passes that src/core/Fluid.cpp dispatches (density/pressure, force + integration with
 * the cube boundary).  Every entry point below names the reference interface it stands in for
 * as file:line relative to the reference repository root.
 *
 * Conventions
 *   - Every call returns an int status: WC_OK (0) or a negative WC_ERR_* code.  The text of the
 *     last failure on the calling thread is returned by wc_last_error().  Nothing throws across
 *     the ABI.
 *   - A context owns all device memory it allocates, for its whole lifetime, and one HIP stream.
 *     Calls on one context are ordered on that stream; a context is not thread-safe.
 *   - Host arrays belong to the caller.  Sizes are element counts.
 *   - The particle record is the reference's 32-byte std430 struct, byte for byte.
 */
#ifndef WATERCUBE_H
#define WATERCUBE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WC_API __attribute__((visibility("default")))

#define WC_ABI_VERSION 1

/* Status codes. */
#define WC_OK 0
#define WC_ERR_INVALID_ARG (-1) /* bad pointer, size or parameter value */
#define WC_ERR_HIP (-2)         /* a HIP runtime call failed; text names the call */
#define WC_ERR_NO_MEMORY (-3)   /* device or host allocation failed */
#define WC_ERR_NO_DEVICE (-4)   /* no usable gfx950 device */
#define WC_ERR_CAPACITY (-5)    /* size above what the context was created for */
#define WC_ERR_DEVICE_FAULT (-6) /* a kernel reported a failed bounded wait (scan look-back) */

/*
 * One particle: struct Particle of src/core/util.h:29-35, laid out as the std430 GLSL struct
 * of assets/sort/count.comp:5-10 (vec3 + float, vec3 + float): offsets 0, 12, 16, 28; 32 bytes.
 */
typedef struct wc_particle {
    float position[3];
    float density;
    float velocity[3];
    float pressure;
} wc_particle;

/* Equation of state used by the density pass. */
#define WC_EOS_CUBIC 0  /* p = p0 + k((rho/rho0)^3 - 1): assets/fluid/density.comp:132-133 (live) */
#define WC_EOS_LINEAR 1 /* p = p0 + k(rho - rho0): assets/fluid/density.comp:129 (commented alternative) */

/*
 * Simulation parameters: the uniforms the reference passes per dispatch
 * (src/core/Sort.cpp:173-175, src/core/Fluid.cpp:276-285 and 305-317) plus the constants its
 * shaders hard-code (assets/fluid/update.comp:5, 202-203), each made a field.
 * wc_params_derive() fills bin_size and the three kernel constants from the primary fields.
 */
typedef struct wc_params {
    uint32_t num_particles; /* num_particles_ (src/core/Fluid.cpp:12) */
    int32_t grid_res;       /* grid_res_ R (Fluid.cpp:14); the grid has R^3 cells */
    float size;             /* size_: the cube is [0,size]^3 (Fluid.cpp:10) */
    float bin_size;         /* derived: size / R (Fluid.cpp:208) */
    float kernel_radius;    /* h (Fluid.cpp:209) */
    float particle_mass;    /* m (Fluid.cpp:210) */
    float rest_density;     /* rho0 (Fluid.cpp:18) */
    float rest_pressure;    /* p0 (Fluid.cpp:21) */
    float stiffness;        /* k (Fluid.cpp:20) */
    float viscosity;        /* mu, viscosityCoefficient (Fluid.cpp:19) */
    float gravity[3];       /* gravity_direction_ * gravity_strength_ (Fluid.cpp:310) */
    float poly6_const;      /* derived: (float)(315 / (64 pi h^9)) in double (Fluid.cpp:214) */
    float spiky_const;      /* derived: (float)(-45 / (pi h^6)) (Fluid.cpp:215) */
    float visc_const;       /* derived: (float)(45 / (pi h^6)) (Fluid.cpp:216) */
    float wall_damping;     /* update.comp:202 (0.3 in the reference) */
    float border;           /* update.comp:203 (0.001) */
    float max_speed;        /* update.comp:5 MAX_SPEED (50) */
    int32_t eos;            /* WC_EOS_CUBIC or WC_EOS_LINEAR */
    int32_t wall_density;   /* 1: add density.comp:57-79 wallDensity (with its z-branch test of p.y) */
    int32_t wall_force;     /* 1: add update.comp:71-100 wallForces (with vec3.length() == 3) */
    int32_t mouse_force;    /* 1: add update.comp:105-132 mouseForce for the ray below */
    float camera_position[3];     /* mouse ray origin (Fluid.cpp:314) */
    float mouse_ray_direction[3]; /* mouse ray direction (Fluid.cpp:315) */
    float time_scale;       /* time_scale_ (Fluid.cpp:24): dt = frame seconds * time_scale */
    float particle_radius;  /* particle_radius_ (Fluid.cpp:17); the reference sets h = 4r, m = 8r */
    uint32_t reserved[4];   /* zero */
} wc_params;

/* Parameter presets (wc_params_preset). */
#define WC_PRESET_REFERENCE_DEFAULT 0 /* Fluid.cpp:9-27 defaults: 80,000 particles, R = 21, cubic EOS, wall terms on */
#define WC_PRESET_DAM_BREAK_16K 1     /* 16,384 particles, h = 0.03, m = 3.375e-3, linear EOS k = 250, mu = 0.2 */
#define WC_PRESET_BLOCK_1M 2          /* 1,000,000 particles, h = 0.016, m = 5.12e-4, R = 62 */
#define WC_PRESET_BLOCK_SPHERE_17M 3  /* 17,777,216 particles, h = 0.004, m = 8e-6, R = 250 */

/* Fill *p with a preset and derive its constants.  No reference counterpart beyond the
 * constructor defaults of src/core/Fluid.cpp:9-27 and src/core/Fluid.cpp:206-216. */
WC_API int wc_params_preset(wc_params* p, int preset);

/* Recompute bin_size = size / grid_res and the poly6 / spiky / viscosity constants from
 * kernel_radius, as Fluid::setup does (src/core/Fluid.cpp:206-216). */
WC_API int wc_params_derive(wc_params* p);

/* ------------------------------------------------------------------------------------------
 * Scene generators (the step before the path: src/core/Fluid.cpp:104-134).  Host-only, no GPU.
 * Random numbers come from SplitMix64 seeded by `seed`; each particle draws its jitter in the
 * order x, y, z, each as u = (next() >> 40) * 2^-24, jitter = (2u - 1) * amplitude.
 * ------------------------------------------------------------------------------------------ */

/* Jittered lattice block: nx*ny*nz particles at origin + (i + 0.5) * spacing + jitter,
 * x fastest, then y, then z (src/core/Fluid.cpp:116-131 walks the same order).  Velocity is
 * `velocity` (may be NULL for zero); density and pressure are zero.  Writes nx*ny*nz records. */
WC_API int wc_scene_block(wc_particle* out, uint32_t nx, uint32_t ny, uint32_t nz, const float origin[3],
                          float spacing, float jitter, const float velocity[3], uint64_t seed);

/* Jittered lattice ball: of the lattice points (i + 0.5) * spacing of the unit cube, the `count`
 * nearest to `center` (ties broken by lattice index), written in x-fastest lattice order with
 * jitter and `velocity`.  Returns WC_ERR_INVALID_ARG if such a ball leaves the unit cube. */
WC_API int wc_scene_sphere(wc_particle* out, uint32_t count, const float center[3], float spacing, float jitter,
                           const float velocity[3], uint64_t seed);

/* The reference's own initial state (src/core/Fluid.cpp:104-134): ceil(cbrt(n))^3 lattice at
 * spacing 1.75 * particle_radius from the origin, jitter uniform in +-0.25 * spacing, first n
 * points.  rand() is replaced by the SplitMix64 draw above. */
WC_API int wc_scene_reference(wc_particle* out, uint32_t n, float particle_radius, uint64_t seed);

/* ------------------------------------------------------------------------------------------
 * Simulation context: Fluid's particle buffers + Sort's grid buffers on one GPU.
 * ------------------------------------------------------------------------------------------ */
typedef struct wc_ctx wc_ctx;

/* Allocate device buffers for p->num_particles particles and grid_res^3 cells on the current
 * HIP device, create the context's stream.  Stands in for Fluid::setup's buffer half and
 * Sort::prepareBuffers (src/core/Fluid.cpp:203-235, 139-166; src/core/Sort.cpp:67-94).  Unlike
 * the reference (Sort.cpp:76-82 sizes them by particle count) counts and offsets hold R^3 cells. */
WC_API int wc_create(const wc_params* p, wc_ctx** out);

/* Free everything the context owns.  NULL is accepted. */
WC_API int wc_destroy(wc_ctx* ctx);

/* Replace the parameters; num_particles and grid_res must not exceed the creation values. */
WC_API int wc_set_params(wc_ctx* ctx, const wc_params* p);
WC_API int wc_get_params(const wc_ctx* ctx, wc_params* out);

/* Fluid::updateGravity (src/core/Fluid.cpp:259-263) with rotate_gravity_ on: the gravity
 * direction in fluid coordinates, rotateWorldSpacePosition(vec3(0, -1, 0)) =
 * transpose(toMat3(q)) * (0, -1, 0) (Fluid.cpp:239-242), for the fluid's orientation quaternion
 * q = (w, x, y, z).  The caller scales it by the gravity strength into wc_params.gravity
 * (Fluid.cpp:310).  Host arithmetic in float, as glm's. */
WC_API int wc_gravity_direction(const float q_wxyz[4], float out[3]);

/* Copy n records into particle buffer 1 (the buffer Fluid::update sorts from and the update
 * pass writes to: Fluid.cpp:146, 347-350).  Both buffers are filled, like Fluid.cpp:146 and 155.
 * n is the context's particle count (all scenes, concatenated, for a batch). */
WC_API int wc_upload(wc_ctx* ctx, const wc_particle* host, uint32_t n);

/* Copy n records out of particle buffer `which` (1 or 2); waits for the stream.  Stands in for
 * util::getParticles (src/core/util.cpp:42-57). */
WC_API int wc_download(wc_ctx* ctx, int which, wc_particle* host, uint32_t n);

/* Sort::run(buffer1, buffer2) (src/core/Sort.cpp:254-267): clear counts, count, exclusive
 * scan, reorder -- as count.comp:25-36, linearScan.comp:15-28, reorder.comp:32-45 -- but with a
 * stable within-cell order (particles keep their buffer-1 order inside a cell). */
WC_API int wc_sort_run(wc_ctx* ctx);

/* Fluid::runDensityProg(buffer2) (src/core/Fluid.cpp:268-289): density.comp:81-137 in place. */
WC_API int wc_density(wc_ctx* ctx);

/* Fluid::runUpdateProg(buffer2, buffer1, dt) (src/core/Fluid.cpp:294-321): update.comp:134-232.
 * `dt` is the integration step itself (the reference passes frame time * time_scale_).  Reads
 * the densities and pressures of the preceding wc_density on the same sort, as the reference's
 * pass order does.  The sort and the density pass also check that every position, velocity,
 * density and pressure is 0 or within [2^-40, 2^40] in magnitude; when they are, the pass forms
 * each neighbour's quotients from shared reciprocals (same bits as IEEE division), otherwise,
 * and when no density pass followed the last sort, it divides as IEEE does. */
WC_API int wc_update(wc_ctx* ctx, float dt);

/* Fluid::runAdvectProg(buffer, time) (src/core/Fluid.cpp:326-336, advect.comp:20-59), whose call
 * in Fluid::update is commented out (Fluid.cpp:351): a position-only step on buffer 1, in place.
 * Per axis x = x + v dt, then if x < 0.01 or x > size - 0.01 it is clamped there and v = v * -0.3
 * (the shader's own border and damping, not the parameter block's).  `dt` is the step itself
 * (the reference passes time * time_scale_). */
WC_API int wc_advect(wc_ctx* ctx, float dt);

/* nsteps of Fluid::update (src/core/Fluid.cpp:342-354): sort, density, update, replayed as one
 * captured HIP graph per step.  Asynchronous: returns once the work is queued. */
WC_API int wc_step(wc_ctx* ctx, float dt, uint32_t nsteps);

/* One step without allocating: the step of wc_step (sort, density, force + integrate, in the
 * reference's pass order) for one scene whose buffer 1 lives in caller-owned device memory, for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops).  No
 * reference counterpart: the reference allocates its buffers once at setup (Fluid.cpp:139-166).
 *   p        the scene's parameters (p->num_particles particles, p->grid_res^3 cells)
 *   posrho   float4 per particle, (x, y, z, density) -- buffer 1's first half, read and rewritten
 *   velp     float4 per particle, (vx, vy, vz, pressure) -- its second half (both 16-byte aligned)
 *   status   device word (may be NULL): set to 0 at the start of the step, nonzero if a kernel's
 *            bounded wait failed (WC_ERR_DEVICE_FAULT's condition); read it after the stream syncs
 *   stream   hipStream_t to queue on (NULL: the default stream); the call returns once queued
 * Every intermediate (the sorted buffer 2, cell keys, ranges, scan state) lives in one fixed
 * device array of the library, WC_NOALLOC_MAX_PARTICLES particles and WC_NOALLOC_MAX_CELLS cells
 * in size (about 17 MB, held on every GPU the library runs on until the process ends); larger
 * scenes are refused with WC_ERR_CAPACITY.  The call makes no allocation of any kind.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller.  The neighbour passes search (no neighbour lists);
 * results are the same bits as wc_step's. */
#define WC_NOALLOC_MAX_PARTICLES 262144u
#define WC_NOALLOC_MAX_CELLS 262144u
WC_API int wc_step_noalloc(const wc_params* p, float dt, void* posrho, void* velp, uint32_t* status, void* stream);

/* ------------------------------------------------------------------------------------------
 * Batched scenes.  A context may hold many independent scenes that advance together: every pass
 * is one launch over all scenes (particles and cell grids concatenated in HBM behind a
 * per-scene table of constants), and a step is one captured graph.  No reference counterpart:
 * the reference simulates a single fluid (src/core/Scene.cpp:52-56).  wc_upload / wc_download
 * address the concatenation of all scenes; the cell and order getters above address scene 0.
 * ------------------------------------------------------------------------------------------ */
WC_API int wc_create_batch(const wc_params* scenes, uint32_t num_scenes, wc_ctx** out);
WC_API int wc_num_scenes(const wc_ctx* ctx);
WC_API int wc_scene_upload(wc_ctx* ctx, uint32_t scene, const wc_particle* host, uint32_t n);
WC_API int wc_scene_download(wc_ctx* ctx, uint32_t scene, int which, wc_particle* host, uint32_t n);
/* counts and offsets (relative to the scene's first particle) of one scene's R^3 cells; either
 * array may be NULL */
WC_API int wc_scene_get_cells(wc_ctx* ctx, uint32_t scene, uint32_t* counts, uint32_t* offsets, uint32_t ncells);
/* Per-scene step sizes used by wc_step_scenes (wc_step sets every scene to its dt). */
WC_API int wc_set_scene_dt(wc_ctx* ctx, const float* dt, uint32_t num_scenes);
WC_API int wc_step_scenes(wc_ctx* ctx, uint32_t nsteps);
/* For the last step: milliseconds from the start of its first pass until the scene's last
 * particle left the force pass (device wall clock), one entry per scene. */
WC_API int wc_scene_latency(wc_ctx* ctx, double* ms, uint32_t num_scenes);

/* Wait for the context's stream; reports a device-side failure flag as WC_ERR_DEVICE_FAULT. */
WC_API int wc_synchronize(wc_ctx* ctx);

/* Per-cell particle counts and exclusive offsets after the last sort (R^3 entries each): the
 * contents of Sort::getCountBuffer / getOffsetBuffer (src/core/Sort.h:35-36). */
WC_API int wc_get_counts(wc_ctx* ctx, uint32_t* host, uint32_t ncells);
WC_API int wc_get_offsets(wc_ctx* ctx, uint32_t* host, uint32_t ncells);

/* Cell index of every particle of buffer 1, as count.comp:31-33 computes it (n entries). */
WC_API int wc_get_cell_keys(wc_ctx* ctx, uint32_t* host, uint32_t n);

/* Sorted order of the last sort: out[slot] = index in buffer 1 of the particle now at `slot` of
 * buffer 2.  The permutation sort.comp:32-45 writes into Sort::getSortedBuffer (Sort.h:37). */
WC_API int wc_get_sorted(wc_ctx* ctx, uint32_t* host, uint32_t n);

/* Device addresses of the context's buffers, for callers that keep data in HBM.
 * posrho[b] / velp[b] are float4 arrays (x, y, z, density) and (vx, vy, vz, pressure) of
 * particle buffer b+1; cell_range is uint2 (offset, count) per cell.  Buffer 2 is the sort's
 * output: a caller that changes particles writes buffer 1 and sorts (wc_sort_run or wc_step)
 * before the next density and force pass, which rely on the sort's checks. */
typedef struct wc_device_view {
    void* posrho[2];
    void* velp[2];
    void* cell_range;
    void* sorted;
    void* stream; /* hipStream_t */
    uint32_t num_particles;
    uint32_t num_cells;
} wc_device_view;
WC_API int wc_device_view_get(wc_ctx* ctx, wc_device_view* out);

/* Timing of each kernel of the step, measured with HIP events recorded on the context's stream
 * around every launch (eager launches, no graph).  Runs nsteps full steps and returns in
 * ms_total[k] the summed duration of kernel k over them; wc_pass_name(k) names kernel k.
 * dt > 0 sets every scene's step size first; dt <= 0 keeps the per-scene ones. */
#define WC_NUM_PASSES 6
WC_API int wc_profile_steps(wc_ctx* ctx, float dt, uint32_t nsteps, double ms_total[WC_NUM_PASSES]);
WC_API const char* wc_pass_name(int k);

/* Tuning of the two neighbour passes (density, force).  Variant v in [0, wc_num_variants()) is
 * a (threads per workgroup, LDS list entries per lane, neighbour slots in flight per lane, drain
 * the list after every dx plane or only when full) choice; every variant computes bit-identical
 * results, only speed differs.  No reference
 * counterpart: the reference dispatches fixed 128-invocation work groups (src/core/util.h:18). */
typedef struct wc_tuning {
    int32_t density_variant;
    int32_t update_variant;
    /* 1: the density pass also stores each particle's accepted neighbours (up to 64, in
     * enumeration order) and the force pass reads them instead of searching again; a particle
     * with more neighbours is searched for as with 0.  Same results either way. */
    int32_t neighbour_list;
} wc_tuning;
WC_API int wc_num_variants(void);
WC_API int wc_variant_info(int v, int32_t* block, int32_t* list_cap, int32_t* chunk, int32_t* plane_drain);
WC_API int wc_get_tuning(const wc_ctx* ctx, wc_tuning* out);
WC_API int wc_set_tuning(wc_ctx* ctx, const wc_tuning* tuning);

/* Number of accepted neighbours of every sorted particle (slot order, n entries) from the last
 * density pass run with neighbour lists; 0xFFFFFFFF marks a particle with more than 64 (its list
 * was not kept).  WC_ERR_INVALID_ARG when the context has no neighbour lists. */
WC_API int wc_get_neighbour_counts(wc_ctx* ctx, uint32_t* host, uint32_t n);

/* Autotuner: sorts the current state once, then times every variant of the density pass and of
 * the force pass (reps launches each, HIP events on the context's stream), both without and with
 * neighbour lists, and keeps the fastest variant pair of the faster mode for later steps.  dt > 0
 * sets every scene's step size first; dt <= 0 keeps them.  Particle buffer 1 is restored
 * afterwards; buffer 2 (the sorted copy the next step rebuilds anyway) is left sorted.  Each
 * non-NULL array receives wc_num_variants() average launch times in milliseconds: the two passes
 * without lists (ms_density, ms_update) and with lists (ms_density_list, ms_update_list). */
WC_API int wc_tune(wc_ctx* ctx, float dt, uint32_t reps, wc_tuning* chosen, double* ms_density, double* ms_update,
                   double* ms_density_list, double* ms_update_list);

/* ------------------------------------------------------------------------------------------
 * Learned closure.  No reference counterpart: the reference has a single explicit step
 * (src/core/Fluid.cpp:342-354).  A small network predicts, per particle, the correction to the
 * acceleration that lets one step of K * dt follow K plain steps of dt; the plain step remains
 * the judge and the fallback.
 *
 * Network: 195 inputs -> 128 -> 128 -> 128 -> 3 (SiLU), bf16 weights and activations, fp32
 * accumulation on matrix cores.  Inputs per particle (bf16, rows of WC_MLP_IN_PAD values, the
 * last 13 zero): for each of its 32 nearest neighbours within the kernel radius (ascending
 * distance, missing ones zero) (x_j - x_i) / h and (v_j - v_i) / (1 m/s); then
 * 10 (rho_i / rho0 - 1), K / 4, and the filled fraction of the 32 places.  Output: the
 * acceleration correction in units of 9.81 m/s^2 (rows of 4 floats, the 4th zero).
 * ------------------------------------------------------------------------------------------ */
#define WC_MLP_IN 195
#define WC_MLP_IN_PAD 208
#define WC_MLP_HIDDEN 128
#define WC_MLP_OUT 3
typedef struct wc_mlp wc_mlp;

/* Network with Glorot-uniform weights drawn from SplitMix64(seed); max_batch bounds training
 * batches.  Parameters are one fp32 vector of wc_mlp_num_params() values: W1 [208][128], b1
 * [128], W2 [128][128], b2, W3 [128][128], b3, W4 [128][32], b4 [32] (row-major [input][output];
 * rows 195.. of W1 and columns 3.. of W4 / b4 are padding and stay zero). */
WC_API int wc_mlp_create(uint32_t max_batch, uint64_t seed, wc_mlp** out);
WC_API int wc_mlp_destroy(wc_mlp* net);
WC_API int wc_mlp_num_params(void);
WC_API int wc_mlp_get_params(wc_mlp* net, float* host, uint32_t n);
WC_API int wc_mlp_set_params(wc_mlp* net, const float* host, uint32_t n);
/* Inference with the fused kernel: x_dev n rows of WC_MLP_IN_PAD bf16, y_dev n rows of 4 floats. */
WC_API int wc_mlp_predict(wc_mlp* net, const void* x_dev, void* y_dev, uint32_t n);
/* Same network through the training path (layer-by-layer GEMMs); n <= max_batch. */
WC_API int wc_mlp_forward(wc_mlp* net, const void* x_dev, void* y_dev, uint32_t n);
/* Loss = mean over rows and the 3 outputs of (y - t)^2 for targets t_dev (n rows of 4 floats);
 * writes the gradient (wc_mlp_num_params() floats) and the loss. */
WC_API int wc_mlp_gradient(wc_mlp* net, const void* x_dev, const void* t_dev, uint32_t n, float* grad_host,
                           double* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; loss may be NULL. */
WC_API int wc_mlp_train_step(wc_mlp* net, const void* x_dev, const void* t_dev, uint32_t n, float lr, double* loss);
/* Average duration of the fused inference kernel over reps launches (HIP events). */
WC_API int wc_mlp_profile_predict(wc_mlp* net, const void* x_dev, void* y_dev, uint32_t n, uint32_t reps, double* ms);
/* Average duration of one training step (forward, backward, Adam; the same kernels as
 * wc_mlp_train_step, without the loss read-back) over reps steps (HIP events).  The steps update
 * the parameters. */
WC_API int wc_mlp_profile_train(wc_mlp* net, const void* x_dev, const void* t_dev, uint32_t n, float lr, uint32_t reps,
                                double* ms);

/* Device memory for the closure's inputs, targets and outputs. */
WC_API int wc_device_alloc(size_t bytes, void** out);
WC_API int wc_device_free(void* dev);
WC_API int wc_device_upload(void* dev, const void* host, size_t bytes);
WC_API int wc_device_download(void* host, const void* dev, size_t bytes);

/* Particle identities: once enabled, ids of buffer 1 start as 0..n-1 in its current order and
 * follow every particle through sorts, so runs can be compared particle by particle. */
WC_API int wc_track_ids(wc_ctx* ctx, int enable);
WC_API int wc_get_ids(wc_ctx* ctx, int which, uint32_t* host, uint32_t n);
/* Largest density written by the last density pass. */
WC_API int wc_max_density(wc_ctx* ctx, float* out);

/* Closure inputs of the sorted particles of buffer 2 (after wc_sort_run and wc_density): rows
 * lists the slots to describe (NULL: all n, in slot order); x_dev receives nrows (or n) rows.
 * With neighbour lists (wc_tuning.neighbour_list) the neighbours are read from the lists of the
 * last density pass, which must follow the last sort. */
WC_API int wc_closure_features(wc_ctx* ctx, float dt_mult, const uint32_t* rows, uint32_t nrows, void* x_dev);
/* The network on all n closure inputs of buffer 2 without forming the feature matrix: each row's
 * values are formed inside the inference kernel from the neighbour lists (required) of the last
 * density pass.  y_dev receives n rows of 4 floats, the same as wc_closure_features followed by
 * wc_mlp_predict. */
WC_API int wc_closure_infer(wc_ctx* ctx, wc_mlp* net, float dt_mult, void* y_dev);
/* Force pass buffer 2 -> buffer 1 with step dt; corr_dev (n rows of 4 floats, times corr_scale)
 * is added to the acceleration of particles at least 2h from every wall (NULL: none); acc_dev
 * (NULL: not written) receives each slot's plain acceleration (4th float: 1 where a correction
 * applies). */
WC_API int wc_closure_update(wc_ctx* ctx, float dt, const void* corr_dev, float corr_scale, void* acc_dev);

/* wc_closure_run applies, per particle, the average of the network corrections of the particle
 * and its accepted neighbours with weights h - r (h for the particle itself; neighbours within
 * 2h of a wall count with zero) (enable != 0, the default), or its own correction alone
 * (enable = 0). */
WC_API int wc_closure_set_smoothing(wc_ctx* ctx, int enable);

/* How wc_closure_run advances a learned step of K * dt with acceleration a (plain plus
 * correction): substeps = 0 (the default), one symplectic Euler step, v += K dt a, x += K dt v;
 * substeps != 0, the outcome of K symplectic Euler steps of dt under a frozen a,
 * v += K dt a, x += dt (K v_old + K (K+1)/2 dt a). */
WC_API int wc_closure_set_integrator(wc_ctx* ctx, int substeps);

typedef struct wc_closure_stats {
    uint32_t macro_steps;    /* learned steps of K * dt taken */
    uint32_t fallbacks;      /* of them replaced by K plain steps */
    double wall_ms;          /* device time of the whole run */
    double max_density_jump; /* largest rise of the density maximum over one learned step, / rho0 */
    double max_accepted_jump; /* largest rise of the density maximum over one accepted learned step,
                               * / rho0: the judged quantity, at most K * judge_tol by construction */
    double peak_density;      /* largest density maximum of the states the run kept, / rho0 */
} wc_closure_stats;

/* nmacro learned steps of K * dt.  Each: sort, density, closure inputs, network, force pass with
 * the correction.  Judge: after every learned step the new state's density maximum may exceed
 * the one before the step by at most K * judge_tol * rho0 (judge_tol per small step of dt the
 * learned step replaces); otherwise the step is replaced by K plain steps of dt from the saved
 * state.  net = NULL takes the same judged large steps with no correction (no closure inputs,
 * no network). */
WC_API int wc_closure_run(wc_ctx* ctx, wc_mlp* net, float dt, uint32_t K, uint32_t nmacro, float judge_tol,
                          wc_closure_stats* stats);

/* Text of the last failure on the calling thread ("" if none). */
WC_API const char* wc_last_error(void);

/* ABI version (WC_ABI_VERSION) and the device the library targets ("gfx950"). */
WC_API int wc_abi_version(void);
WC_API const char* wc_target_arch(void);

#ifdef __cplusplus
}
#endif

#endif /* WATERCUBE_H */
