// wc_api.hip -- host runtime behind include/watercube.h: parameters, contexts, device buffers,
// the captured per-step graph and per-kernel event timing.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "wc_internal.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define WC_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(WC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

namespace wc {

float dist2_cutoff(float h) {
    // Smallest float t with sqrtf(t) >= h.  sqrtf is monotonic, so for every d2 (NaN included)
    // "sqrtf(d2) >= h" and "d2 >= t" agree.
    float t = h * h;
    for (;;) {
        const float p = nextafterf(t, 0.0f);
        if (p >= 0.0f && p != t && sqrtf(p) >= h)
            t = p;
        else
            break;
    }
    while (sqrtf(t) < h) t = nextafterf(t, INFINITY);
    return t;
}

StepConsts make_consts(const wc_params& p) {
    StepConsts c;
    memset(&c, 0, sizeof(c));
    c.n = p.num_particles;
    c.R = p.grid_res;
    c.ncells = (uint32_t)p.grid_res * (uint32_t)p.grid_res * (uint32_t)p.grid_res;
    c.bin = p.bin_size;
    c.h = p.kernel_radius;
    c.hh = p.kernel_radius * p.kernel_radius;
    c.cut2 = dist2_cutoff(p.kernel_radius);
    c.m = p.particle_mass;
    c.poly6 = p.poly6_const;
    c.spiky = p.spiky_const;
    c.visc = p.visc_const;
    c.mu = p.viscosity;
    c.rho0 = p.rest_density;
    c.p0 = p.rest_pressure;
    c.k = p.stiffness;
    for (int a = 0; a < 3; ++a) {
        c.g[a] = p.gravity[a];
        c.cam[a] = p.camera_position[a];
        c.dir[a] = p.mouse_ray_direction[a];
    }
    c.size = p.size;
    c.size_minus_h = p.size - p.kernel_radius;
    c.border = p.border;
    c.hi = p.size - p.border;
    c.damping = p.wall_damping;
    c.max_speed = p.max_speed;
    c.eos = p.eos;
    c.wall_density = p.wall_density;
    c.wall_force = p.wall_force;
    c.mouse_force = p.mouse_force;
    return c;
}

}  // namespace wc

using wc::DeviceBuffers;
using wc::StepConsts;

struct wc_ctx {
    wc_params params;
    StepConsts consts;
    uint32_t n_cap = 0;
    uint32_t cells_cap = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffers b{};
    wc_particle* staging = nullptr;  // n_cap records for upload / download
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float graph_dt = 0.0f;
    bool graph_valid = false;
    int density_variant = 2;  // (64 threads, 32 list entries, 4 in flight, per-plane drain) until tuned
    int update_variant = 2;
};

namespace {

int check_params(const wc_params* p) {
    if (!p) return fail(WC_ERR_INVALID_ARG, "params is NULL");
    if (p->grid_res < 1 || p->grid_res > 1024)
        return fail(WC_ERR_INVALID_ARG, "grid_res %d outside [1, 1024]", p->grid_res);
    if (!(p->kernel_radius > 0.0f) || !(p->size > 0.0f) || !(p->bin_size > 0.0f))
        return fail(WC_ERR_INVALID_ARG, "kernel_radius, size and bin_size must be positive");
    if (p->eos != WC_EOS_CUBIC && p->eos != WC_EOS_LINEAR) return fail(WC_ERR_INVALID_ARG, "unknown eos %d", p->eos);
    return WC_OK;
}

int free_ctx_buffers(wc_ctx* c) {
    DeviceBuffers& b = c->b;
    void* ptrs[] = {b.posrho[0], b.posrho[1], b.velp[0], b.velp[1], b.key,         b.rank,
                    b.slot_tmp,  b.sorted,    b.skey,    b.counts,  b.range,       b.scan_status,
                    b.scan_ticket, b.error_flag, c->staging};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    memset(&b, 0, sizeof(b));
    c->staging = nullptr;
    return WC_OK;
}

void drop_graph(wc_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_valid = false;
}

int enqueue_sort(wc_ctx* c) {
    WC_HIP(wc::launch_hash(c->b, c->consts, c->stream));
    WC_HIP(wc::launch_scan(c->b, c->consts, c->stream));
    WC_HIP(wc::launch_place(c->b, c->consts, c->stream));
    WC_HIP(wc::launch_gather(c->b, c->consts, c->stream));
    return WC_OK;
}

int enqueue_step(wc_ctx* c, float dt) {
    int rc = enqueue_sort(c);
    if (rc) return rc;
    WC_HIP(wc::launch_density(c->b, c->consts, c->density_variant, c->stream));
    WC_HIP(wc::launch_update(c->b, c->consts, dt, c->update_variant, c->stream));
    return WC_OK;
}

int set_device(const wc_ctx* c) {
    WC_HIP(hipSetDevice(c->device));
    return WC_OK;
}

}  // namespace

extern "C" {

WC_API int wc_abi_version(void) { return WC_ABI_VERSION; }
WC_API const char* wc_target_arch(void) { return "gfx950"; }
WC_API const char* wc_last_error(void) { return g_last_error.c_str(); }

WC_API int wc_params_derive(wc_params* p) {
    if (!p) return fail(WC_ERR_INVALID_ARG, "params is NULL");
    if (p->grid_res < 1) return fail(WC_ERR_INVALID_ARG, "grid_res must be >= 1");
    if (!(p->kernel_radius > 0.0f)) return fail(WC_ERR_INVALID_ARG, "kernel_radius must be > 0");
    p->bin_size = p->size / (float)p->grid_res;
    const double h = (double)p->kernel_radius;
    p->poly6_const = (float)(315.0 / (64.0 * M_PI * pow(h, 9)));
    p->spiky_const = (float)(-45.0 / (M_PI * pow(h, 6)));
    p->visc_const = (float)(45.0 / (M_PI * pow(h, 6)));
    return WC_OK;
}

WC_API int wc_params_preset(wc_params* p, int preset) {
    if (!p) return fail(WC_ERR_INVALID_ARG, "params is NULL");
    memset(p, 0, sizeof(*p));
    p->size = 1.0f;
    p->rest_pressure = 0.0f;
    p->border = 0.001f;
    p->max_speed = 50.0f;
    p->time_scale = 1.0f;
    switch (preset) {
        case WC_PRESET_REFERENCE_DEFAULT:
            p->num_particles = 80000;
            p->grid_res = 21;
            p->particle_radius = 0.01f;
            p->kernel_radius = p->particle_radius * 4.0f;
            p->particle_mass = p->particle_radius * 8.0f;
            p->rest_density = 500.0f;
            p->viscosity = 200.0f;
            p->stiffness = 100.0f;
            p->gravity[1] = -1.0f * 900.0f;
            p->wall_damping = 0.3f;
            p->eos = WC_EOS_CUBIC;
            p->wall_density = 1;
            p->wall_force = 1;
            p->time_scale = 0.012f;
            break;
        case WC_PRESET_DAM_BREAK_16K:
        case WC_PRESET_BLOCK_1M:
        case WC_PRESET_BLOCK_SPHERE_17M: {
            float h, m;
            int R;
            uint32_t n;
            if (preset == WC_PRESET_DAM_BREAK_16K) {
                n = 16384; h = 0.03f; m = 3.375e-3f; R = 33;
            } else if (preset == WC_PRESET_BLOCK_1M) {
                n = 1000000; h = 0.016f; m = 5.12e-4f; R = 62;
            } else {
                n = 16777216u + 1000000u; h = 0.004f; m = 8e-6f; R = 250;
            }
            p->num_particles = n;
            p->grid_res = R;
            p->kernel_radius = h;
            p->particle_mass = m;
            p->particle_radius = h / 4.0f;
            p->rest_density = 1000.0f;
            p->stiffness = 250.0f;
            p->viscosity = 0.2f;
            p->gravity[1] = -9.81f;
            p->wall_damping = 0.5f;
            p->eos = WC_EOS_LINEAR;
            break;
        }
        default:
            return fail(WC_ERR_INVALID_ARG, "unknown preset %d", preset);
    }
    return wc_params_derive(p);
}

WC_API int wc_create(const wc_params* p, wc_ctx** out) {
    if (!out) return fail(WC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WC_ERR_NO_DEVICE, "no HIP device");
    wc_ctx* c = new (std::nothrow) wc_ctx();
    if (!c) return fail(WC_ERR_NO_MEMORY, "host allocation failed");
    c->params = *p;
    c->consts = wc::make_consts(*p);
    c->n_cap = p->num_particles;
    c->cells_cap = c->consts.ncells;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(WC_ERR_NO_DEVICE, "hipGetDevice failed");
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(WC_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    const size_t n = c->n_cap ? c->n_cap : 1;
    const size_t cells = c->cells_cap;
    const uint32_t tiles = (uint32_t)((cells + wc::kScanTile - 1) / wc::kScanTile);
    DeviceBuffers& b = c->b;
    b.num_scan_tiles = tiles;
    bool ok = true;
    auto alloc = [&](void** ptr, size_t bytes) {
        if (!ok) return;
        if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) ok = false;
    };
    alloc((void**)&b.posrho[0], n * sizeof(float4));
    alloc((void**)&b.posrho[1], n * sizeof(float4));
    alloc((void**)&b.velp[0], n * sizeof(float4));
    alloc((void**)&b.velp[1], n * sizeof(float4));
    alloc((void**)&b.key, n * 4);
    alloc((void**)&b.rank, n * 4);
    alloc((void**)&b.slot_tmp, n * 4);
    alloc((void**)&b.sorted, n * 4);
    alloc((void**)&b.skey, n * 4);
    alloc((void**)&b.counts, cells * 4 + 64);
    alloc((void**)&b.range, cells * 8 + 64);
    alloc((void**)&b.scan_status, (size_t)tiles * 8 + 8);
    alloc((void**)&b.scan_ticket, 64);
    alloc((void**)&b.error_flag, 64);
    alloc((void**)&c->staging, n * sizeof(wc_particle));
    if (!ok) {
        free_ctx_buffers(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(WC_ERR_NO_MEMORY, "device allocation failed (%u particles, %zu cells)", p->num_particles, cells);
    }
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMemsetAsync(b.counts, 0, cells * 4 + 64, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.range, 0, cells * 8 + 64, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.error_flag, 0, 64, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.sorted, 0, n * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.skey, 0, n * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.key, 0, n * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.posrho[0], 0, n * sizeof(float4), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.posrho[1], 0, n * sizeof(float4), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.velp[0], 0, n * sizeof(float4), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b.velp[1], 0, n * sizeof(float4), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        free_ctx_buffers(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(WC_ERR_HIP, "buffer initialisation failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return WC_OK;
}

WC_API int wc_destroy(wc_ctx* c) {
    if (!c) return WC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_ctx_buffers(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return WC_OK;
}

WC_API int wc_set_params(wc_ctx* c, const wc_params* p) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = check_params(p);
    if (rc) return rc;
    const uint64_t R = (uint64_t)p->grid_res;
    if (p->num_particles > c->n_cap || R * R * R > c->cells_cap)
        return fail(WC_ERR_CAPACITY, "num_particles %u / cells %llu exceed the context's %u / %u", p->num_particles,
                    (unsigned long long)(R * R * R), c->n_cap, c->cells_cap);
    c->params = *p;
    c->consts = wc::make_consts(*p);
    drop_graph(c);
    return WC_OK;
}

WC_API int wc_get_params(const wc_ctx* c, wc_params* out) {
    if (!c || !out) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    *out = c->params;
    return WC_OK;
}

WC_API int wc_upload(wc_ctx* c, const wc_particle* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n != c->params.num_particles)
        return fail(WC_ERR_INVALID_ARG, "upload of %u records into a context of %u particles", n,
                    c->params.num_particles);
    if (n == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipMemcpyAsync(c->staging, host, (size_t)n * sizeof(wc_particle), hipMemcpyHostToDevice, c->stream));
    WC_HIP(wc::launch_split_records(c->staging, c->b.posrho[0], c->b.velp[0], n, c->stream));
    WC_HIP(wc::launch_split_records(c->staging, c->b.posrho[1], c->b.velp[1], n, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

WC_API int wc_download(wc_ctx* c, int which, wc_particle* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (which != 1 && which != 2) return fail(WC_ERR_INVALID_ARG, "buffer must be 1 or 2, got %d", which);
    if (n > c->params.num_particles) return fail(WC_ERR_CAPACITY, "download of %u records from %u", n, c->params.num_particles);
    if (n == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    const int b = which - 1;
    WC_HIP(wc::launch_join_records(c->b.posrho[b], c->b.velp[b], c->staging, n, c->stream));
    WC_HIP(hipMemcpyAsync(host, c->staging, (size_t)n * sizeof(wc_particle), hipMemcpyDeviceToHost, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

WC_API int wc_sort_run(wc_ctx* c) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    return enqueue_sort(c);
}

WC_API int wc_density(wc_ctx* c) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(wc::launch_density(c->b, c->consts, c->density_variant, c->stream));
    return WC_OK;
}

WC_API int wc_update(wc_ctx* c, float dt) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(wc::launch_update(c->b, c->consts, dt, c->update_variant, c->stream));
    return WC_OK;
}

WC_API int wc_step(wc_ctx* c, float dt, uint32_t nsteps) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    if (nsteps == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    if (!c->graph_valid || c->graph_dt != dt) {
        drop_graph(c);
        WC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(c, dt);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(WC_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
        c->graph = g;
        e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(c);
            return fail(WC_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        c->graph_dt = dt;
        c->graph_valid = true;
    }
    for (uint32_t s = 0; s < nsteps; ++s) WC_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    return WC_OK;
}

WC_API int wc_synchronize(wc_ctx* c) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipStreamSynchronize(c->stream));
    uint32_t flag = 0;
    WC_HIP(hipMemcpy(&flag, c->b.error_flag, 4, hipMemcpyDeviceToHost));
    if (flag) return fail(WC_ERR_DEVICE_FAULT, "scan look-back exceeded its bounded wait (flag %u)", flag);
    return WC_OK;
}

static int download_range(wc_ctx* c, uint32_t* host, uint32_t ncells, int field) {
    if (!c || (!host && ncells)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (ncells != c->consts.ncells)
        return fail(WC_ERR_INVALID_ARG, "expected %u cells, got %u", c->consts.ncells, ncells);
    int rc = set_device(c);
    if (rc) return rc;
    std::vector<uint32_t> tmp((size_t)ncells * 2);
    WC_HIP(hipStreamSynchronize(c->stream));
    WC_HIP(hipMemcpy(tmp.data(), c->b.range, (size_t)ncells * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ncells; ++i) host[i] = tmp[2 * i + field];
    return WC_OK;
}

WC_API int wc_get_counts(wc_ctx* c, uint32_t* host, uint32_t ncells) { return download_range(c, host, ncells, 1); }
WC_API int wc_get_offsets(wc_ctx* c, uint32_t* host, uint32_t ncells) { return download_range(c, host, ncells, 0); }

static int download_u32(wc_ctx* c, const uint32_t* dev, uint32_t* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n > c->params.num_particles) return fail(WC_ERR_CAPACITY, "%u entries requested of %u", n, c->params.num_particles);
    if (n == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipStreamSynchronize(c->stream));
    WC_HIP(hipMemcpy(host, dev, (size_t)n * 4, hipMemcpyDeviceToHost));
    return WC_OK;
}

WC_API int wc_get_cell_keys(wc_ctx* c, uint32_t* host, uint32_t n) {
    return c ? download_u32(c, c->b.key, host, n) : fail(WC_ERR_INVALID_ARG, "ctx is NULL");
}
WC_API int wc_get_sorted(wc_ctx* c, uint32_t* host, uint32_t n) {
    return c ? download_u32(c, c->b.sorted, host, n) : fail(WC_ERR_INVALID_ARG, "ctx is NULL");
}

WC_API int wc_device_view_get(wc_ctx* c, wc_device_view* out) {
    if (!c || !out) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    out->posrho[0] = c->b.posrho[0];
    out->posrho[1] = c->b.posrho[1];
    out->velp[0] = c->b.velp[0];
    out->velp[1] = c->b.velp[1];
    out->cell_range = c->b.range;
    out->sorted = c->b.sorted;
    out->stream = (void*)c->stream;
    out->num_particles = c->params.num_particles;
    out->num_cells = c->consts.ncells;
    return WC_OK;
}

static const char* const kPassNames[WC_NUM_PASSES] = {"hash", "scan", "place", "gather", "density", "update"};

WC_API const char* wc_pass_name(int k) { return (k >= 0 && k < WC_NUM_PASSES) ? kPassNames[k] : ""; }

WC_API int wc_profile_steps(wc_ctx* c, float dt, uint32_t nsteps, double ms_total[WC_NUM_PASSES]) {
    if (!c || !ms_total) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    for (int k = 0; k < WC_NUM_PASSES; ++k) ms_total[k] = 0.0;
    if (nsteps == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    const size_t nev = (size_t)nsteps * (WC_NUM_PASSES + 1);
    std::vector<hipEvent_t> ev(nev, nullptr);
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            for (auto& x : ev)
                if (x) (void)hipEventDestroy(x);
            return fail(WC_ERR_HIP, "hipEventCreate failed");
        }
    hipError_t err = hipSuccess;
    for (uint32_t s = 0; s < nsteps && err == hipSuccess; ++s) {
        hipEvent_t* e = &ev[(size_t)s * (WC_NUM_PASSES + 1)];
        err = hipEventRecord(e[0], c->stream);
        if (err == hipSuccess) err = wc::launch_hash(c->b, c->consts, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[1], c->stream);
        if (err == hipSuccess) err = wc::launch_scan(c->b, c->consts, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[2], c->stream);
        if (err == hipSuccess) err = wc::launch_place(c->b, c->consts, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[3], c->stream);
        if (err == hipSuccess) err = wc::launch_gather(c->b, c->consts, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[4], c->stream);
        if (err == hipSuccess) err = wc::launch_density(c->b, c->consts, c->density_variant, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[5], c->stream);
        if (err == hipSuccess) err = wc::launch_update(c->b, c->consts, dt, c->update_variant, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[6], c->stream);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    if (err == hipSuccess) {
        for (uint32_t s = 0; s < nsteps; ++s) {
            hipEvent_t* e = &ev[(size_t)s * (WC_NUM_PASSES + 1)];
            for (int k = 0; k < WC_NUM_PASSES; ++k) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, e[k], e[k + 1]) == hipSuccess) ms_total[k] += ms;
            }
        }
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (err != hipSuccess) return fail(WC_ERR_HIP, "profiled step failed: %s", hipGetErrorString(err));
    return WC_OK;
}

WC_API int wc_num_variants(void) { return wc::num_variants(); }

WC_API int wc_variant_info(int v, int32_t* block, int32_t* list_cap, int32_t* chunk, int32_t* plane_drain) {
    if (v < 0 || v >= wc::num_variants()) return fail(WC_ERR_INVALID_ARG, "variant %d outside [0, %d)", v, wc::num_variants());
    const wc::PassVariant pv = wc::variant(v);
    if (block) *block = pv.block;
    if (list_cap) *list_cap = pv.list_cap;
    if (chunk) *chunk = pv.chunk;
    if (plane_drain) *plane_drain = pv.plane_drain ? 1 : 0;
    return WC_OK;
}

WC_API int wc_get_tuning(const wc_ctx* c, wc_tuning* out) {
    if (!c || !out) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    out->density_variant = c->density_variant;
    out->update_variant = c->update_variant;
    return WC_OK;
}

WC_API int wc_set_tuning(wc_ctx* c, const wc_tuning* t) {
    if (!c || !t) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    const int nv = wc::num_variants();
    if (t->density_variant < 0 || t->density_variant >= nv || t->update_variant < 0 || t->update_variant >= nv)
        return fail(WC_ERR_INVALID_ARG, "variant outside [0, %d)", nv);
    c->density_variant = t->density_variant;
    c->update_variant = t->update_variant;
    drop_graph(c);
    return WC_OK;
}

WC_API int wc_tune(wc_ctx* c, float dt, uint32_t reps, wc_tuning* chosen, double* ms_density, double* ms_update) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    if (reps == 0) reps = 1;
    int rc = set_device(c);
    if (rc) return rc;
    const int nv = wc::num_variants();
    const size_t n = c->params.num_particles;
    std::vector<double> td(nv, 0.0), tu(nv, 0.0);
    if (n > 0) {
        // Keep buffer 1 (the state the caller sees); the passes below rewrite it.
        float4 *save_pos = nullptr, *save_vel = nullptr;
        WC_HIP(hipMalloc(&save_pos, n * sizeof(float4)));
        if (hipMalloc(&save_vel, n * sizeof(float4)) != hipSuccess) {
            (void)hipFree(save_pos);
            return fail(WC_ERR_NO_MEMORY, "tuning copy of %zu particles", n);
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t err = hipEventCreate(&e0);
        if (err == hipSuccess) err = hipEventCreate(&e1);
        if (err == hipSuccess) err = hipMemcpyAsync(save_pos, c->b.posrho[0], n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess) err = hipMemcpyAsync(save_vel, c->b.velp[0], n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess) err = (enqueue_sort(c) == WC_OK) ? hipSuccess : hipErrorUnknown;
        // Both passes are idempotent on the sorted buffer: density rewrites the same densities,
        // the force pass rewrites buffer 1 from buffer 2 alone.
        for (int pass = 0; pass < 2 && err == hipSuccess; ++pass) {
            for (int v = 0; v < nv && err == hipSuccess; ++v) {
                auto launch = [&]() {
                    return pass == 0 ? wc::launch_density(c->b, c->consts, v, c->stream)
                                     : wc::launch_update(c->b, c->consts, dt, v, c->stream);
                };
                err = launch();  // warm-up
                if (err == hipSuccess) err = hipEventRecord(e0, c->stream);
                for (uint32_t r = 0; r < reps && err == hipSuccess; ++r) err = launch();
                if (err == hipSuccess) err = hipEventRecord(e1, c->stream);
                if (err == hipSuccess) err = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
                (pass == 0 ? td : tu)[v] = (double)ms / reps;
            }
            if (pass == 0 && err == hipSuccess) err = wc::launch_density(c->b, c->consts, c->density_variant, c->stream);
        }
        if (err == hipSuccess) err = hipMemcpyAsync(c->b.posrho[0], save_pos, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess) err = hipMemcpyAsync(c->b.velp[0], save_vel, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(save_pos);
        (void)hipFree(save_vel);
        if (err != hipSuccess) return fail(WC_ERR_HIP, "tuning failed: %s", hipGetErrorString(err));
        int bd = 0, bu = 0;
        for (int v = 1; v < nv; ++v) {
            if (td[v] < td[bd]) bd = v;
            if (tu[v] < tu[bu]) bu = v;
        }
        c->density_variant = bd;
        c->update_variant = bu;
        drop_graph(c);
    }
    if (chosen) {
        chosen->density_variant = c->density_variant;
        chosen->update_variant = c->update_variant;
    }
    for (int v = 0; v < nv; ++v) {
        if (ms_density) ms_density[v] = td[v];
        if (ms_update) ms_update[v] = tu[v];
    }
    return WC_OK;
}

}  // extern "C"
