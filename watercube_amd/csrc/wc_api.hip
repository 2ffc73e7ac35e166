// wc_api.hip -- host runtime behind include/watercube.h: parameters, contexts of one or many
// scenes, device buffers, the captured per-step graph, per-kernel event timing and the
// autotuner of the neighbour passes.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "wc_internal.hpp"
#include "wc_mlp.hpp"

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

uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// A start/stop pair of timing events, destroyed on every return path.
struct EventPair {
    hipEvent_t start = nullptr, stop = nullptr;
    hipError_t create() {
        hipError_t e = hipEventCreate(&start);
        if (e == hipSuccess) e = hipEventCreate(&stop);
        return e;
    }
    ~EventPair() {
        if (start) (void)hipEventDestroy(start);
        if (stop) (void)hipEventDestroy(stop);
    }
};

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
using wc::SceneDesc;

struct wc_ctx {
    std::vector<wc_params> params;   // per scene
    std::vector<SceneDesc> desc;     // per scene, mirrored in d_desc
    std::vector<uint32_t> tables;    // blk64 | blk128 | blk256 | tiles, each num_scenes + 1
    std::vector<uint32_t> scene_tab;  // several scenes: scene of each 64/128/256-thread block and tile
    uint32_t num_scenes = 0;
    uint32_t total_particles = 0;
    uint32_t total_cells = 0;        // padded to scan tiles per scene
    uint32_t particle_cap = 0;       // allocated particle slots (single-scene contexts may shrink)
    uint32_t cell_cap = 0;
    uint32_t tile_cap = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffers b{};
    SceneDesc* d_desc = nullptr;
    uint32_t* d_tables = nullptr;
    uint32_t* d_scene_tab = nullptr;
    wc_particle* staging = nullptr;  // particle_cap records for upload / download
    // captured steps: graph[0] one step, graph[1] kStepsPerGraph steps back to back (replayed
    // while enough steps remain, which saves the gap between graph launches)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    bool graph_valid = false;
    int density_variant = 2;  // (64 threads, 32 list entries, 4 in flight, per-plane drain) until tuned
    int update_variant = 2;
    bool neighbour_list = false;  // density writes neighbour lists, the force pass reads them
    double clock_khz = 100000.0;  // wall_clock64 rate
    // learned-closure workspace (allocated on first use)
    uint16_t* cl_x = nullptr;       // features, n x 208 bf16
    float* cl_y = nullptr;          // network output, n x 4
    uint32_t* cl_long = nullptr;    // slots whose list is longer than the selection in registers takes (n), count
    float4* cl_backup[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // two slots of (posrho, velp)
    uint32_t* cl_backup_ids[2] = {nullptr, nullptr};
    uint32_t* cl_host_rho = nullptr;  // pinned: the density maximum after each slot's learned step
    uint32_t* cl_rows = nullptr;    // row list for feature requests
    uint32_t cl_rows_cap = 0;
    bool cl_smooth = true;          // average the correction over neighbours in wc_closure_run
    bool cl_substeps = false;       // integrate a learned step as K frozen-force substeps
};

struct wc_mlp {
    wc::Mlp m;
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

// Lays the scenes out (particle and cell offsets, block and tile tables) from c->params.
int layout_scenes(wc_ctx* c) {
    const uint32_t S = c->num_scenes;
    c->desc.assign(S, SceneDesc{});
    c->tables.assign(4 * (size_t)(S + 1), 0u);
    uint32_t* blk64 = &c->tables[0];
    uint32_t* blk128 = blk64 + (S + 1);
    uint32_t* blk256 = blk128 + (S + 1);
    uint32_t* tiles = blk256 + (S + 1);
    uint64_t p_off = 0, c_off = 0;
    for (uint32_t s = 0; s < S; ++s) {
        const wc_params& p = c->params[s];
        SceneDesc& d = c->desc[s];
        d.c = wc::make_consts(p);
        d.p_off = (uint32_t)p_off;
        d.c_off = (uint32_t)c_off;
        const uint32_t nt = ceil_div(d.c.ncells, wc::kScanTile);
        blk64[s + 1] = blk64[s] + ceil_div(d.c.n, 64);
        blk128[s + 1] = blk128[s] + ceil_div(d.c.n, 128);
        blk256[s + 1] = blk256[s] + ceil_div(d.c.n, 256);
        tiles[s + 1] = tiles[s] + nt;
        p_off += d.c.n;
        c_off += (uint64_t)nt * wc::kScanTile;
        if (p_off > 0xFFFFFF00ull || c_off > 0xFFFFFF00ull)
            return fail(WC_ERR_CAPACITY, "scenes exceed 2^32 particles or cells in one context");
    }
    c->total_particles = (uint32_t)p_off;
    c->total_cells = (uint32_t)c_off;
    c->scene_tab.clear();
    if (S > 1) {
        for (const uint32_t* first : {blk64, blk128, blk256, tiles})
            for (uint32_t s = 0; s < S; ++s) c->scene_tab.insert(c->scene_tab.end(), first[s + 1] - first[s], s);
    }
    return WC_OK;
}

void fill_scene_views(wc_ctx* c) {
    const uint32_t S = c->num_scenes;
    wc::Scenes& v = c->b.sc;
    v.desc = c->d_desc;
    v.blk64 = c->d_tables;
    v.blk128 = c->d_tables + (S + 1);
    v.blk256 = c->d_tables + 2 * (S + 1);
    v.tiles = c->d_tables + 3 * (S + 1);
    v.scene64 = v.scene128 = v.scene256 = v.scene_tile = nullptr;
    if (S > 1 && c->d_scene_tab) {
        v.scene64 = c->d_scene_tab;
        v.scene128 = v.scene64 + c->tables[S];
        v.scene256 = v.scene128 + c->tables[2 * (size_t)(S + 1) - 1];
        v.scene_tile = v.scene256 + c->tables[3 * (size_t)(S + 1) - 1];
    }
    v.num_scenes = S;
    v.total_particles = c->total_particles;
    v.total_tiles = c->tables[4 * (size_t)(S + 1) - 1];
    v.blocks64 = c->tables[S];
    v.blocks128 = c->tables[2 * (size_t)(S + 1) - 1];
    v.blocks256 = c->tables[3 * (size_t)(S + 1) - 1];
}

int upload_tables(wc_ctx* c) {
    WC_HIP(hipMemcpyAsync(c->d_desc, c->desc.data(), c->desc.size() * sizeof(SceneDesc), hipMemcpyHostToDevice,
                          c->stream));
    WC_HIP(hipMemcpyAsync(c->d_tables, c->tables.data(), c->tables.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (!c->scene_tab.empty())
        WC_HIP(hipMemcpyAsync(c->d_scene_tab, c->scene_tab.data(), c->scene_tab.size() * 4, hipMemcpyHostToDevice,
                              c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    fill_scene_views(c);
    return WC_OK;
}

void free_ctx_buffers(wc_ctx* c) {
    DeviceBuffers& b = c->b;
    void* ptrs[] = {b.posrho[0],   b.posrho[1],   b.velp[0],     b.velp[1],     b.key,       b.rank,
                    b.slot_tmp,    b.sorted,      b.skey,        b.counts,      b.range,     b.scan_status,
                    b.error_flag,  b.scene_done,  b.step_start,  c->staging,  c->d_desc,
                    c->d_tables,   b.rho_max,     b.ids[0],      b.ids[1],      c->cl_x,     c->cl_y,
                    c->cl_backup[0][0], c->cl_backup[0][1], c->cl_backup[1][0], c->cl_backup[1][1],
                    c->cl_backup_ids[0], c->cl_backup_ids[1], c->cl_rows, b.nlist, b.ncount, c->d_scene_tab,
                    b.crowd,      b.crowd_count, c->cl_long};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->cl_host_rho) (void)hipHostFree(c->cl_host_rho);
    c->cl_host_rho = nullptr;
    memset(&b, 0, sizeof(b));
    c->staging = nullptr;
    c->d_desc = nullptr;
    c->d_tables = nullptr;
    c->d_scene_tab = nullptr;
    c->cl_x = nullptr;
    c->cl_y = nullptr;
    c->cl_long = nullptr;
    c->cl_backup[0][0] = c->cl_backup[0][1] = c->cl_backup[1][0] = c->cl_backup[1][1] = nullptr;
    c->cl_backup_ids[0] = c->cl_backup_ids[1] = nullptr;
    c->cl_rows = nullptr;
}

void drop_graph(wc_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        if (c->graph_exec[k]) (void)hipGraphExecDestroy(c->graph_exec[k]);
        if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
        c->graph_exec[k] = nullptr;
        c->graph[k] = nullptr;
    }
    c->graph_valid = false;
}

void destroy_ctx(wc_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_ctx_buffers(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Neighbour-list storage, allocated the first time the list mode is used.
int ensure_neighbour_lists(wc_ctx* c) {
    if (c->b.nlist && c->b.ncount) return WC_OK;
    const size_t n = c->particle_cap ? c->particle_cap : 1;
    if (!c->b.nlist && hipMalloc(&c->b.nlist, (size_t)wc::kNeighbourListMax * n * 4) != hipSuccess) {
        c->b.nlist = nullptr;
        return fail(WC_ERR_NO_MEMORY, "neighbour lists for %zu particles", n);
    }
    if (!c->b.ncount && hipMalloc(&c->b.ncount, n * 4) != hipSuccess) {
        c->b.ncount = nullptr;
        return fail(WC_ERR_NO_MEMORY, "neighbour counts for %zu particles", n);
    }
    return WC_OK;
}

int enqueue_sort(wc_ctx* c) {
    WC_HIP(wc::launch_hash(c->b, c->stream));
    WC_HIP(wc::launch_scan(c->b, c->stream));
    WC_HIP(wc::launch_place(c->b, c->stream));
    WC_HIP(wc::launch_gather(c->b, c->stream));
    return WC_OK;
}

int enqueue_step(wc_ctx* c) {
    int rc = enqueue_sort(c);
    if (rc) return rc;
    WC_HIP(wc::launch_density(c->b, c->density_variant, c->neighbour_list, c->stream));
    WC_HIP(wc::launch_update(c->b, c->update_variant, c->neighbour_list, c->stream));
    return WC_OK;
}

int set_device(const wc_ctx* c) {
    WC_HIP(hipSetDevice(c->device));
    return WC_OK;
}

int create_context(const wc_params* scenes, uint32_t num, wc_ctx** out) {
    if (!out) return fail(WC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!scenes || num == 0) return fail(WC_ERR_INVALID_ARG, "no scenes");
    for (uint32_t s = 0; s < num; ++s) {
        int rc = check_params(&scenes[s]);
        if (rc) return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WC_ERR_NO_DEVICE, "no HIP device");
    wc_ctx* c = new (std::nothrow) wc_ctx();
    if (!c) return fail(WC_ERR_NO_MEMORY, "host allocation failed");
    c->num_scenes = num;
    c->params.assign(scenes, scenes + num);
    int rc = layout_scenes(c);
    if (rc) {
        delete c;
        return rc;
    }
    c->particle_cap = c->total_particles;
    c->cell_cap = c->total_cells;
    c->tile_cap = c->tables.back();
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(WC_ERR_NO_DEVICE, "hipGetDevice failed");
    }
    int rate = 0;
    if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, c->device) == hipSuccess && rate > 0)
        c->clock_khz = rate;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(WC_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    const size_t n = c->particle_cap ? c->particle_cap : 1;
    const size_t cells = c->cell_cap ? c->cell_cap : 1;
    const size_t tiles = c->tile_cap ? c->tile_cap : 1;
    DeviceBuffers& b = c->b;
    bool ok = true;
    auto alloc = [&](void** ptr, size_t bytes) {
        if (!ok) return;
        if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) ok = false;
    };
    alloc((void**)&b.posrho[0], (n + wc::kRecordPad) * sizeof(float4));
    alloc((void**)&b.posrho[1], (n + wc::kRecordPad) * sizeof(float4));
    alloc((void**)&b.velp[0], (n + wc::kRecordPad) * sizeof(float4));
    alloc((void**)&b.velp[1], (n + wc::kRecordPad) * sizeof(float4));
    alloc((void**)&b.key, n * 4);
    alloc((void**)&b.rank, n * 4);
    alloc((void**)&b.slot_tmp, n * 4);
    alloc((void**)&b.sorted, n * 4);
    alloc((void**)&b.skey, n * 4);
    alloc((void**)&b.counts, cells * 4 + 64);
    alloc((void**)&b.range, cells * 8 + 64);
    alloc((void**)&b.scan_status, tiles * 8 + 8);
    alloc((void**)&b.error_flag, 64);
    alloc((void**)&b.scene_done, (size_t)num * 8 + 8);
    alloc((void**)&b.step_start, 64);
    alloc((void**)&b.rho_max, 64);
    b.crowd_cap = wc::crowd_capacity(n);
    alloc((void**)&b.crowd, (size_t)b.crowd_cap * sizeof(uint2));
    alloc((void**)&b.crowd_count, 64);
    alloc((void**)&c->staging, n * sizeof(wc_particle));
    alloc((void**)&c->d_desc, (size_t)num * sizeof(SceneDesc));
    alloc((void**)&c->d_tables, c->tables.size() * 4);
    if (!c->scene_tab.empty()) alloc((void**)&c->d_scene_tab, c->scene_tab.size() * 4);
    if (!ok) {
        destroy_ctx(c);
        return fail(WC_ERR_NO_MEMORY, "device allocation failed (%u particles, %u cells)", c->total_particles,
                    c->total_cells);
    }
    hipError_t e = hipSuccess;
    auto zero = [&](void* p, size_t bytes) {
        if (e == hipSuccess) e = hipMemsetAsync(p, 0, bytes, c->stream);
    };
    zero(b.counts, cells * 4 + 64);
    zero(b.range, cells * 8 + 64);
    zero(b.scan_status, tiles * 8 + 8);
    zero(b.error_flag, 64);
    zero(b.scene_done, (size_t)num * 8 + 8);
    zero(b.step_start, 64);
    zero(b.rho_max, 64);
    zero(b.crowd_count, 64);
    // rho_max[1] = 2: no density pass has checked the force pass's division operands yet
    if (e == hipSuccess) e = hipMemsetAsync(reinterpret_cast<char*>(b.rho_max) + 4, 2, 1, c->stream);
    zero(b.sorted, n * 4);
    zero(b.skey, n * 4);
    zero(b.key, n * 4);
    zero(b.slot_tmp, n * 4);
    zero(b.posrho[0], (n + wc::kRecordPad) * sizeof(float4));
    zero(b.posrho[1], (n + wc::kRecordPad) * sizeof(float4));
    zero(b.velp[0], (n + wc::kRecordPad) * sizeof(float4));
    zero(b.velp[1], (n + wc::kRecordPad) * sizeof(float4));
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        destroy_ctx(c);
        return fail(WC_ERR_HIP, "buffer initialisation failed: %s", hipGetErrorString(e));
    }
    rc = upload_tables(c);
    if (rc) {
        destroy_ctx(c);
        return rc;
    }
    *out = c;
    return WC_OK;
}

// Steps per replay of the longer captured graph.  A graph launch leaves the GPU idle for about
// 8 us before its first kernel; at one step per launch that is about 1.5 % of a 1M-particle step.
constexpr uint32_t kStepsPerGraph = 8;

int capture_steps(wc_ctx* c, uint32_t steps, hipGraph_t* graph, hipGraphExec_t* exec) {
    WC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = WC_OK;
    for (uint32_t s = 0; s < steps && !rc; ++s) rc = enqueue_step(c);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(WC_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    *graph = g;
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(WC_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return WC_OK;
}

int capture_step(wc_ctx* c) {
    drop_graph(c);
    int rc = capture_steps(c, 1, &c->graph[0], &c->graph_exec[0]);
    if (!rc) rc = capture_steps(c, kStepsPerGraph, &c->graph[1], &c->graph_exec[1]);
    if (rc) {
        drop_graph(c);
        return rc;
    }
    c->graph_valid = true;
    return WC_OK;
}

int run_steps(wc_ctx* c, uint32_t nsteps) {
    if (nsteps == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    if (!c->graph_valid) {
        rc = capture_step(c);
        if (rc) return rc;
    }
    uint32_t left = nsteps;
    for (; left >= kStepsPerGraph; left -= kStepsPerGraph) WC_HIP(hipGraphLaunch(c->graph_exec[1], c->stream));
    for (; left > 0; --left) WC_HIP(hipGraphLaunch(c->graph_exec[0], c->stream));
    return WC_OK;
}

// Sets every scene's dt.  Only the scene table changes (host copy and device copy); captured
// graphs read dt from the device table, so they stay valid.
int set_all_dt(wc_ctx* c, const float* dt, bool same) {
    bool changed = false;
    for (uint32_t s = 0; s < c->num_scenes; ++s) {
        const float v = same ? dt[0] : dt[s];
        if (c->desc[s].dt != v) {
            c->desc[s].dt = v;
            changed = true;
        }
    }
    if (!changed) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipMemcpyAsync(c->d_desc, c->desc.data(), c->desc.size() * sizeof(SceneDesc), hipMemcpyHostToDevice,
                          c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));  // the host table may change again before the copy runs
    return WC_OK;
}

int copy_records_in(wc_ctx* c, uint32_t first, const wc_particle* host, uint32_t n) {
    if (n == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipMemcpyAsync(c->staging + first, host, (size_t)n * sizeof(wc_particle), hipMemcpyHostToDevice,
                          c->stream));
    for (int b = 0; b < 2; ++b)
        WC_HIP(wc::launch_split_records(c->staging + first, c->b.posrho[b] + first, c->b.velp[b] + first, n,
                                        c->stream));
    // buffer 2 changed under the last sort: its cell keys, neighbour lists and division check
    WC_HIP(wc::launch_invalidate(c->b, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

int copy_records_out(wc_ctx* c, int which, uint32_t first, wc_particle* host, uint32_t n) {
    if (which != 1 && which != 2) return fail(WC_ERR_INVALID_ARG, "buffer must be 1 or 2, got %d", which);
    if (n == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    const int b = which - 1;
    WC_HIP(wc::launch_join_records(c->b.posrho[b] + first, c->b.velp[b] + first, c->staging + first, n, c->stream));
    WC_HIP(hipMemcpyAsync(host, c->staging + first, (size_t)n * sizeof(wc_particle), hipMemcpyDeviceToHost,
                          c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

int scene_cells(wc_ctx* c, uint32_t scene, uint32_t* counts, uint32_t* offsets, uint32_t ncells) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    if (scene >= c->num_scenes) return fail(WC_ERR_INVALID_ARG, "scene %u of %u", scene, c->num_scenes);
    const SceneDesc& d = c->desc[scene];
    if (ncells != d.c.ncells) return fail(WC_ERR_INVALID_ARG, "expected %u cells, got %u", d.c.ncells, ncells);
    int rc = set_device(c);
    if (rc) return rc;
    std::vector<uint32_t> tmp((size_t)ncells * 2);
    WC_HIP(hipStreamSynchronize(c->stream));
    if (ncells) WC_HIP(hipMemcpy(tmp.data(), c->b.range + d.c_off, (size_t)ncells * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ncells; ++i) {
        if (counts) counts[i] = tmp[2 * i + 1];
        if (offsets) offsets[i] = tmp[2 * i] - d.p_off;
    }
    return WC_OK;
}

int download_u32(wc_ctx* c, const uint32_t* dev, uint32_t* host, uint32_t n) {
    if (!host && n) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n > c->total_particles) return fail(WC_ERR_CAPACITY, "%u entries requested of %u", n, c->total_particles);
    if (n == 0) return WC_OK;
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipStreamSynchronize(c->stream));
    WC_HIP(hipMemcpy(host, dev, (size_t)n * 4, hipMemcpyDeviceToHost));
    return WC_OK;
}

const char* const kPassNames[WC_NUM_PASSES] = {"hash", "scan", "place", "gather", "density", "update"};

// ------------------------------------------------------------------------------------------
// Scratch of the allocation-free step (wc_step_noalloc): one fixed device array per GPU, laid
// out as the buffers a one-scene context allocates, for at most kNoAllocParticles particles and
// kNoAllocCells cells.  Host code never uses the array's name as an address: its per-device
// address comes from hipGetSymbolAddress on the first call on each device.
// ------------------------------------------------------------------------------------------
constexpr size_t kNoAllocParticles = WC_NOALLOC_MAX_PARTICLES;
constexpr size_t kNoAllocCells = WC_NOALLOC_MAX_CELLS;
static_assert(kNoAllocCells % wc::kScanTile == 0, "whole scan tiles");
constexpr size_t kNoAllocTiles = kNoAllocCells / wc::kScanTile;
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
// byte offsets in the array
constexpr size_t kNaHeader = 0;  // error word, rho_max[2], scene_done, step_start, crowded-cell count
constexpr size_t kNaDesc = kNaHeader + 256;
constexpr size_t kNaTables = kNaDesc + align256(sizeof(SceneDesc));
constexpr size_t kNaPosrho = kNaTables + 256;
constexpr size_t kNaVelp = kNaPosrho + align256((kNoAllocParticles + wc::kRecordPad) * sizeof(float4));
constexpr size_t kNaKey = kNaVelp + align256((kNoAllocParticles + wc::kRecordPad) * sizeof(float4));
constexpr size_t kNaRank = kNaKey + align256(kNoAllocParticles * 4);
constexpr size_t kNaSlot = kNaRank + align256(kNoAllocParticles * 4);
constexpr size_t kNaSorted = kNaSlot + align256(kNoAllocParticles * 4);
constexpr size_t kNaSkey = kNaSorted + align256(kNoAllocParticles * 4);
constexpr size_t kNaCounts = kNaSkey + align256(kNoAllocParticles * 4);
constexpr size_t kNaRange = kNaCounts + align256(kNoAllocCells * 4);
constexpr size_t kNaStatus = kNaRange + align256(kNoAllocCells * 8);
constexpr size_t kNaCrowd = kNaStatus + align256(kNoAllocTiles * 8);
constexpr size_t kNoAllocBytes = kNaCrowd + align256((kNoAllocParticles / (wc::kGatherDirectMax + 1) + 1) * 8);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// neighbour-pass variant of the allocation-free step: 64 threads, 32 list entries, 4 slots in
// flight, a drain per dx plane (a context's default before tuning)
constexpr int kNoAllocVariant = 2;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

// The scratch array's address on the current device `dev`, looked up once per device.
int scratch_on_device(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(WC_ERR_NO_DEVICE, "device %d outside [0, %d)", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(WC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
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
    if (!p) {
        if (out) *out = nullptr;
        return fail(WC_ERR_INVALID_ARG, "params is NULL");
    }
    return create_context(p, 1, out);
}

WC_API int wc_create_batch(const wc_params* scenes, uint32_t num_scenes, wc_ctx** out) {
    return create_context(scenes, num_scenes, out);
}

WC_API int wc_destroy(wc_ctx* c) {
    destroy_ctx(c);
    return WC_OK;
}

WC_API int wc_num_scenes(const wc_ctx* c) { return c ? (int)c->num_scenes : fail(WC_ERR_INVALID_ARG, "ctx is NULL"); }

WC_API int wc_gravity_direction(const float q[4], float out[3]) {
    if (!q || !out) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    // column 1 of glm::toMat3(q): (2(xy + wz), 1 - 2(xx + zz), 2(yz - wx)); transpose(M) * (0, -1, 0)
    // is minus that column
    const float xx = x * x, zz = z * z, xy = x * y, wz = w * z, yz = y * z, wx = w * x;
    out[0] = -(2.0f * (xy + wz));
    out[1] = -(1.0f - 2.0f * (xx + zz));
    out[2] = -(2.0f * (yz - wx));
    return WC_OK;
}

WC_API int wc_set_params(wc_ctx* c, const wc_params* p) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = check_params(p);
    if (rc) return rc;
    if (c->num_scenes != 1) return fail(WC_ERR_INVALID_ARG, "wc_set_params needs a single-scene context");
    const uint64_t R = (uint64_t)p->grid_res;
    const uint64_t padded = (R * R * R + wc::kScanTile - 1) / wc::kScanTile * wc::kScanTile;
    if (p->num_particles > c->particle_cap || padded > c->cell_cap)
        return fail(WC_ERR_CAPACITY, "num_particles %u / cells %llu exceed the context's %u / %u", p->num_particles,
                    (unsigned long long)(R * R * R), c->particle_cap, c->cell_cap);
    rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipStreamSynchronize(c->stream));
    const float dt = c->desc[0].dt;
    c->params[0] = *p;
    rc = layout_scenes(c);
    if (rc) return rc;
    c->desc[0].dt = dt;
    drop_graph(c);
    rc = upload_tables(c);
    if (rc) return rc;
    // keys, lists and the division check of buffer 2 were formed with the old constants
    WC_HIP(wc::launch_invalidate(c->b, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

WC_API int wc_get_params(const wc_ctx* c, wc_params* out) {
    if (!c || !out) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    *out = c->params[0];
    return WC_OK;
}

WC_API int wc_upload(wc_ctx* c, const wc_particle* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n != c->total_particles)
        return fail(WC_ERR_INVALID_ARG, "upload of %u records into a context of %u particles", n, c->total_particles);
    return copy_records_in(c, 0, host, n);
}

WC_API int wc_download(wc_ctx* c, int which, wc_particle* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n > c->total_particles) return fail(WC_ERR_CAPACITY, "download of %u records from %u", n, c->total_particles);
    return copy_records_out(c, which, 0, host, n);
}

WC_API int wc_scene_upload(wc_ctx* c, uint32_t scene, const wc_particle* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (scene >= c->num_scenes) return fail(WC_ERR_INVALID_ARG, "scene %u of %u", scene, c->num_scenes);
    if (n != c->desc[scene].c.n)
        return fail(WC_ERR_INVALID_ARG, "scene %u holds %u particles, got %u", scene, c->desc[scene].c.n, n);
    return copy_records_in(c, c->desc[scene].p_off, host, n);
}

WC_API int wc_scene_download(wc_ctx* c, uint32_t scene, int which, wc_particle* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (scene >= c->num_scenes) return fail(WC_ERR_INVALID_ARG, "scene %u of %u", scene, c->num_scenes);
    if (n > c->desc[scene].c.n)
        return fail(WC_ERR_CAPACITY, "scene %u holds %u particles, asked %u", scene, c->desc[scene].c.n, n);
    return copy_records_out(c, which, c->desc[scene].p_off, host, n);
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
    WC_HIP(wc::launch_density(c->b, c->density_variant, c->neighbour_list, c->stream));
    return WC_OK;
}

WC_API int wc_update(wc_ctx* c, float dt) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_all_dt(c, &dt, true);
    if (rc) return rc;
    rc = set_device(c);
    if (rc) return rc;
    WC_HIP(wc::launch_update(c->b, c->update_variant, c->neighbour_list, c->stream));
    return WC_OK;
}

WC_API int wc_advect(wc_ctx* c, float dt) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(wc::launch_advect(c->b, dt, c->stream));
    return WC_OK;
}

WC_API int wc_step(wc_ctx* c, float dt, uint32_t nsteps) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_all_dt(c, &dt, true);
    if (rc) return rc;
    return run_steps(c, nsteps);
}

WC_API int wc_step_noalloc(const wc_params* p, float dt, void* posrho, void* velp, uint32_t* status, void* stream) {
    int rc = check_params(p);
    if (rc) return rc;
    const uint64_t R = (uint64_t)p->grid_res;
    const uint64_t ncells = R * R * R;
    const uint32_t n = p->num_particles;
    if (n > kNoAllocParticles || ncells > kNoAllocCells)
        return fail(WC_ERR_CAPACITY, "wc_step_noalloc holds at most %zu particles and %zu cells; got %u and %llu",
                    kNoAllocParticles, kNoAllocCells, n, (unsigned long long)ncells);
    if (n == 0) return WC_OK;
    if (!posrho || !velp) return fail(WC_ERR_INVALID_ARG, "NULL particle buffer");
    if ((reinterpret_cast<uintptr_t>(posrho) | reinterpret_cast<uintptr_t>(velp)) & 15u)
        return fail(WC_ERR_INVALID_ARG, "particle buffers must be 16-byte aligned");
    int dev = 0;
    WC_HIP(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    rc = scratch_on_device(dev, &base);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);

    SceneDesc d;
    memset(&d, 0, sizeof(d));
    d.c = wc::make_consts(*p);
    d.dt = dt;
    const uint32_t tiles = ceil_div(ncells, wc::kScanTile);
    const uint32_t tables[8] = {0u, ceil_div(n, 64), 0u, ceil_div(n, 128), 0u, ceil_div(n, 256), 0u, tiles};
    DeviceBuffers b{};
    b.posrho[0] = static_cast<float4*>(posrho);
    b.velp[0] = static_cast<float4*>(velp);
    b.posrho[1] = reinterpret_cast<float4*>(base + kNaPosrho);
    b.velp[1] = reinterpret_cast<float4*>(base + kNaVelp);
    b.key = reinterpret_cast<uint32_t*>(base + kNaKey);
    b.rank = reinterpret_cast<uint32_t*>(base + kNaRank);
    b.slot_tmp = reinterpret_cast<uint32_t*>(base + kNaSlot);
    b.sorted = reinterpret_cast<uint32_t*>(base + kNaSorted);
    b.skey = reinterpret_cast<uint32_t*>(base + kNaSkey);
    b.counts = reinterpret_cast<uint32_t*>(base + kNaCounts);
    b.range = reinterpret_cast<uint2*>(base + kNaRange);
    b.scan_status = reinterpret_cast<unsigned long long*>(base + kNaStatus);
    b.error_flag = status ? status : reinterpret_cast<uint32_t*>(base + kNaHeader);
    b.rho_max = reinterpret_cast<uint32_t*>(base + kNaHeader + 8);
    b.scene_done = reinterpret_cast<unsigned long long*>(base + kNaHeader + 16);
    b.step_start = reinterpret_cast<unsigned long long*>(base + kNaHeader + 24);
    b.crowd_count = reinterpret_cast<uint32_t*>(base + kNaHeader + 32);
    b.crowd = reinterpret_cast<uint2*>(base + kNaCrowd);
    b.crowd_cap = wc::crowd_capacity(kNoAllocParticles);
    SceneDesc* d_desc = reinterpret_cast<SceneDesc*>(base + kNaDesc);
    uint32_t* d_tables = reinterpret_cast<uint32_t*>(base + kNaTables);
    wc::Scenes& v = b.sc;
    v.desc = d_desc;
    v.blk64 = d_tables;
    v.blk128 = d_tables + 2;
    v.blk256 = d_tables + 4;
    v.tiles = d_tables + 6;
    v.num_scenes = 1;
    v.total_particles = n;
    v.total_tiles = tiles;
    v.blocks64 = tables[1];
    v.blocks128 = tables[3];
    v.blocks256 = tables[5];
    WC_HIP(wc::launch_noalloc_setup(d, tables, d_desc, d_tables, b.counts, (uint32_t)ncells, b.error_flag, s));
    WC_HIP(wc::launch_hash(b, s));
    WC_HIP(wc::launch_scan(b, s));
    WC_HIP(wc::launch_place(b, s));
    WC_HIP(wc::launch_gather(b, s));
    WC_HIP(wc::launch_density(b, kNoAllocVariant, false, s));
    WC_HIP(wc::launch_update(b, kNoAllocVariant, false, s));
    return WC_OK;
}

WC_API int wc_set_scene_dt(wc_ctx* c, const float* dt, uint32_t num_scenes) {
    if (!c || !dt) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (num_scenes != c->num_scenes)
        return fail(WC_ERR_INVALID_ARG, "%u step sizes for %u scenes", num_scenes, c->num_scenes);
    return set_all_dt(c, dt, false);
}

WC_API int wc_step_scenes(wc_ctx* c, uint32_t nsteps) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    return run_steps(c, nsteps);
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

WC_API int wc_scene_latency(wc_ctx* c, double* ms, uint32_t num_scenes) {
    if (!c || !ms) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (num_scenes != c->num_scenes)
        return fail(WC_ERR_INVALID_ARG, "%u entries for %u scenes", num_scenes, c->num_scenes);
    int rc = set_device(c);
    if (rc) return rc;
    std::vector<unsigned long long> done(num_scenes);
    unsigned long long start = 0;
    WC_HIP(hipStreamSynchronize(c->stream));
    WC_HIP(hipMemcpy(done.data(), c->b.scene_done, (size_t)num_scenes * 8, hipMemcpyDeviceToHost));
    WC_HIP(hipMemcpy(&start, c->b.step_start, 8, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < num_scenes; ++s) ms[s] = done[s] > start ? (double)(done[s] - start) / c->clock_khz : 0.0;
    return WC_OK;
}

WC_API int wc_get_counts(wc_ctx* c, uint32_t* host, uint32_t ncells) { return scene_cells(c, 0, host, nullptr, ncells); }
WC_API int wc_get_offsets(wc_ctx* c, uint32_t* host, uint32_t ncells) { return scene_cells(c, 0, nullptr, host, ncells); }
WC_API int wc_scene_get_cells(wc_ctx* c, uint32_t scene, uint32_t* counts, uint32_t* offsets, uint32_t ncells) {
    return scene_cells(c, scene, counts, offsets, ncells);
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
    out->num_particles = c->total_particles;
    out->num_cells = c->total_cells;
    return WC_OK;
}

WC_API const char* wc_pass_name(int k) { return (k >= 0 && k < WC_NUM_PASSES) ? kPassNames[k] : ""; }

WC_API int wc_profile_steps(wc_ctx* c, float dt, uint32_t nsteps, double ms_total[WC_NUM_PASSES]) {
    if (!c || !ms_total) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    for (int k = 0; k < WC_NUM_PASSES; ++k) ms_total[k] = 0.0;
    if (nsteps == 0) return WC_OK;
    int rc = dt > 0.0f ? set_all_dt(c, &dt, true) : WC_OK;
    if (rc) return rc;
    rc = set_device(c);
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
        if (err == hipSuccess) err = wc::launch_hash(c->b, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[1], c->stream);
        if (err == hipSuccess) err = wc::launch_scan(c->b, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[2], c->stream);
        if (err == hipSuccess) err = wc::launch_place(c->b, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[3], c->stream);
        if (err == hipSuccess) err = wc::launch_gather(c->b, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[4], c->stream);
        if (err == hipSuccess) err = wc::launch_density(c->b, c->density_variant, c->neighbour_list, c->stream);
        if (err == hipSuccess) err = hipEventRecord(e[5], c->stream);
        if (err == hipSuccess) err = wc::launch_update(c->b, c->update_variant, c->neighbour_list, c->stream);
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
    if (v < 0 || v >= wc::num_variants())
        return fail(WC_ERR_INVALID_ARG, "variant %d outside [0, %d)", v, wc::num_variants());
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
    out->neighbour_list = c->neighbour_list ? 1 : 0;
    return WC_OK;
}

WC_API int wc_set_tuning(wc_ctx* c, const wc_tuning* t) {
    if (!c || !t) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    const int nv = wc::num_variants();
    if (t->density_variant < 0 || t->density_variant >= nv || t->update_variant < 0 || t->update_variant >= nv)
        return fail(WC_ERR_INVALID_ARG, "variant outside [0, %d)", nv);
    if (t->neighbour_list) {
        const int rc = set_device(c);
        if (rc) return rc;
        const int ra = ensure_neighbour_lists(c);
        if (ra) return ra;
    }
    c->density_variant = t->density_variant;
    c->update_variant = t->update_variant;
    c->neighbour_list = t->neighbour_list != 0;
    drop_graph(c);
    return WC_OK;
}

WC_API int wc_get_neighbour_counts(wc_ctx* c, uint32_t* host, uint32_t n) {
    if (!c || (!host && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (!c->b.ncount) return fail(WC_ERR_INVALID_ARG, "the context has no neighbour lists");
    if (n != c->total_particles) return fail(WC_ERR_INVALID_ARG, "n = %u, the context holds %u", n, c->total_particles);
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipMemcpyAsync(host, c->b.ncount, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

WC_API int wc_tune(wc_ctx* c, float dt, uint32_t reps, wc_tuning* chosen, double* ms_density, double* ms_update,
                   double* ms_density_list, double* ms_update_list) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    if (reps == 0) reps = 1;
    int rc = dt > 0.0f ? set_all_dt(c, &dt, true) : WC_OK;
    if (rc) return rc;
    rc = set_device(c);
    if (rc) return rc;
    const int nv = wc::num_variants();
    const size_t n = c->total_particles;
    // [mode][pass][variant]: mode 0 searches in both passes, mode 1 uses neighbour lists
    std::vector<double> t[2][2];
    for (auto& m : t)
        for (auto& v : m) v.assign(nv, 0.0);
    if (n > 0) {
        rc = ensure_neighbour_lists(c);
        if (rc) return rc;
        // Keep buffer 1 (the state the caller sees); the passes below rewrite it.
        float4 *save_pos = nullptr, *save_vel = nullptr;
        WC_HIP(hipMalloc(&save_pos, n * sizeof(float4)));
        if (hipMalloc(&save_vel, n * sizeof(float4)) != hipSuccess) {
            (void)hipFree(save_pos);
            return fail(WC_ERR_NO_MEMORY, "tuning copy of %zu particles", n);
        }
        EventPair ev;
        hipError_t err = ev.create();
        if (err == hipSuccess)
            err = hipMemcpyAsync(save_pos, c->b.posrho[0], n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess)
            err = hipMemcpyAsync(save_vel, c->b.velp[0], n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess) err = (enqueue_sort(c) == WC_OK) ? hipSuccess : hipErrorUnknown;
        // Both passes are idempotent on the sorted buffer: density rewrites the same densities
        // (and, with lists, the same lists), the force pass rewrites buffer 1 from buffer 2 alone.
        // In list mode the density variants run first, so every force-pass variant reads lists.
        for (int mode = 0; mode < 2 && err == hipSuccess; ++mode) {
            for (int pass = 0; pass < 2 && err == hipSuccess; ++pass) {
                for (int v = 0; v < nv && err == hipSuccess; ++v) {
                    auto launch = [&]() {
                        return pass == 0 ? wc::launch_density(c->b, v, mode == 1, c->stream)
                                         : wc::launch_update(c->b, v, mode == 1, c->stream);
                    };
                    err = launch();  // warm-up
                    if (err == hipSuccess) err = hipEventRecord(ev.start, c->stream);
                    for (uint32_t r = 0; r < reps && err == hipSuccess; ++r) err = launch();
                    if (err == hipSuccess) err = hipEventRecord(ev.stop, c->stream);
                    if (err == hipSuccess) err = hipEventSynchronize(ev.stop);
                    float ms = 0.0f;
                    if (err == hipSuccess) err = hipEventElapsedTime(&ms, ev.start, ev.stop);
                    t[mode][pass][v] = (double)ms / reps;
                }
            }
        }
        if (err == hipSuccess)
            err = hipMemcpyAsync(c->b.posrho[0], save_pos, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess)
            err = hipMemcpyAsync(c->b.velp[0], save_vel, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
        (void)hipFree(save_pos);
        (void)hipFree(save_vel);
        if (err != hipSuccess) return fail(WC_ERR_HIP, "tuning failed: %s", hipGetErrorString(err));
        int best[2][2] = {{0, 0}, {0, 0}};
        for (int mode = 0; mode < 2; ++mode)
            for (int pass = 0; pass < 2; ++pass)
                for (int v = 1; v < nv; ++v)
                    if (t[mode][pass][v] < t[mode][pass][best[mode][pass]]) best[mode][pass] = v;
        const double step0 = t[0][0][best[0][0]] + t[0][1][best[0][1]];
        const double step1 = t[1][0][best[1][0]] + t[1][1][best[1][1]];
        const int mode = step1 < step0 ? 1 : 0;
        c->density_variant = best[mode][0];
        c->update_variant = best[mode][1];
        c->neighbour_list = mode == 1;
        drop_graph(c);
    }
    if (chosen) {
        chosen->density_variant = c->density_variant;
        chosen->update_variant = c->update_variant;
        chosen->neighbour_list = c->neighbour_list ? 1 : 0;
    }
    for (int v = 0; v < nv; ++v) {
        if (ms_density) ms_density[v] = t[0][0][v];
        if (ms_update) ms_update[v] = t[0][1][v];
        if (ms_density_list) ms_density_list[v] = t[1][0][v];
        if (ms_update_list) ms_update_list[v] = t[1][1][v];
    }
    return WC_OK;
}

// ------------------------------------------------------------------------------------------
// Device memory helpers (for data the closure's training and inference read from HBM).
// ------------------------------------------------------------------------------------------
WC_API int wc_device_alloc(size_t bytes, void** out) {
    if (!out) return fail(WC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (hipMalloc(out, bytes ? bytes : 16) != hipSuccess) return fail(WC_ERR_NO_MEMORY, "hipMalloc of %zu bytes", bytes);
    return WC_OK;
}
WC_API int wc_device_free(void* p) {
    if (p) (void)hipFree(p);
    return WC_OK;
}
WC_API int wc_device_upload(void* dev, const void* host, size_t bytes) {
    if ((!dev || !host) && bytes) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (bytes) WC_HIP(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
    return WC_OK;
}
WC_API int wc_device_download(void* host, const void* dev, size_t bytes) {
    if ((!dev || !host) && bytes) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (bytes) WC_HIP(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
    return WC_OK;
}

// ------------------------------------------------------------------------------------------
// Particle identities and the density maximum.
// ------------------------------------------------------------------------------------------
WC_API int wc_track_ids(wc_ctx* c, int enable) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipStreamSynchronize(c->stream));
    if (!enable) {
        for (int b = 0; b < 2; ++b) {
            if (c->b.ids[b]) (void)hipFree(c->b.ids[b]);
            c->b.ids[b] = nullptr;
        }
        drop_graph(c);
        return WC_OK;
    }
    const size_t n = c->particle_cap ? c->particle_cap : 1;
    for (int b = 0; b < 2; ++b)
        if (!c->b.ids[b] && hipMalloc(&c->b.ids[b], n * 4) != hipSuccess)
            return fail(WC_ERR_NO_MEMORY, "identity buffers");
    std::vector<uint32_t> iota(n);
    for (size_t i = 0; i < n; ++i) iota[i] = (uint32_t)i;
    for (int b = 0; b < 2; ++b) WC_HIP(hipMemcpy(c->b.ids[b], iota.data(), n * 4, hipMemcpyHostToDevice));
    drop_graph(c);  // the captured kernels now carry the identity buffers
    return WC_OK;
}

WC_API int wc_get_ids(wc_ctx* c, int which, uint32_t* host, uint32_t n) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    if (which != 1 && which != 2) return fail(WC_ERR_INVALID_ARG, "buffer must be 1 or 2");
    if (!c->b.ids[0]) return fail(WC_ERR_INVALID_ARG, "identities are not tracked (wc_track_ids)");
    return download_u32(c, c->b.ids[which - 1], host, n);
}

WC_API int wc_max_density(wc_ctx* c, float* out) {
    if (!c || !out) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    int rc = set_device(c);
    if (rc) return rc;
    uint32_t bits = 0;
    WC_HIP(hipStreamSynchronize(c->stream));
    WC_HIP(hipMemcpy(&bits, c->b.rho_max, 4, hipMemcpyDeviceToHost));
    memcpy(out, &bits, 4);
    return WC_OK;
}

// ------------------------------------------------------------------------------------------
// The closure's network.
// ------------------------------------------------------------------------------------------
WC_API int wc_mlp_num_params(void) { return (int)wc::mlp_layout().total; }

WC_API int wc_mlp_create(uint32_t max_batch, uint64_t seed, wc_mlp** out) {
    if (!out) return fail(WC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WC_ERR_NO_DEVICE, "no HIP device");
    wc_mlp* h = new (std::nothrow) wc_mlp();
    if (!h) return fail(WC_ERR_NO_MEMORY, "host allocation failed");
    wc::Mlp& m = h->m;
    m.max_batch = max_batch ? max_batch : 1;
    (void)hipGetDevice(&m.device);
    const wc::MlpLayout L = wc::mlp_layout();
    const size_t B = m.max_batch;
    bool ok = hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking) == hipSuccess;
    auto alloc = [&](void** p, size_t bytes) {
        if (ok && hipMalloc(p, bytes ? bytes : 16) != hipSuccess) ok = false;
    };
    alloc((void**)&m.master, L.total * 4);
    alloc((void**)&m.grad, L.total * 4);
    alloc((void**)&m.adam_m, L.total * 4);
    alloc((void**)&m.adam_v, L.total * 4);
    alloc((void**)&m.w1, (size_t)wc::kMlpInPad * wc::kMlpHidden * 2);
    alloc((void**)&m.w1t, (size_t)wc::kMlpInPad * wc::kMlpHidden * 2);
    alloc((void**)&m.w2, (size_t)wc::kMlpHidden * wc::kMlpHidden * 2);
    alloc((void**)&m.w2t, (size_t)wc::kMlpHidden * wc::kMlpHidden * 2);
    alloc((void**)&m.w3, (size_t)wc::kMlpHidden * wc::kMlpHidden * 2);
    alloc((void**)&m.w3t, (size_t)wc::kMlpHidden * wc::kMlpHidden * 2);
    alloc((void**)&m.w4, (size_t)wc::kMlpHidden * wc::kMlpOutPad * 2);
    alloc((void**)&m.w4t, (size_t)wc::kMlpHidden * wc::kMlpOutPad * 2);
    for (int l = 0; l < 3; ++l) {
        alloc((void**)&m.z[l], B * wc::kMlpHidden * 4);
        alloc((void**)&m.h[l], B * wc::kMlpHidden * 2);
    }
    alloc((void**)&m.yacc, B * wc::kMlpOutPad * 4);
    alloc((void**)&m.d4, B * wc::kMlpOutPad * 2);
    alloc((void**)&m.gbuf, B * wc::kMlpHidden * 4);
    alloc((void**)&m.d, B * wc::kMlpHidden * 2);
    alloc((void**)&m.loss, 64);
    alloc((void**)&m.colpart, (B + wc::kColsumRows - 1) / wc::kColsumRows * wc::kMlpHidden * 4);
    if (!ok) {
        wc_mlp_destroy(h);
        return fail(WC_ERR_NO_MEMORY, "network buffers for a batch of %u", max_batch);
    }
    // Glorot-uniform weights from SplitMix64, zero biases; padded rows and columns stay zero.
    std::vector<float> w(L.total, 0.0f);
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto fill = [&](size_t off, int rows, int cols, int used_rows, int used_cols) {
        const double a = sqrt(6.0 / (double)(used_rows + used_cols));
        for (int r = 0; r < rows; ++r)
            for (int col = 0; col < cols; ++col) {
                const double u = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
                if (r < used_rows && col < used_cols) w[off + (size_t)r * cols + col] = (float)((2.0 * u - 1.0) * a);
            }
    };
    fill(L.w1, wc::kMlpInPad, wc::kMlpHidden, wc::kMlpIn, wc::kMlpHidden);
    fill(L.w2, wc::kMlpHidden, wc::kMlpHidden, wc::kMlpHidden, wc::kMlpHidden);
    fill(L.w3, wc::kMlpHidden, wc::kMlpHidden, wc::kMlpHidden, wc::kMlpHidden);
    fill(L.w4, wc::kMlpHidden, wc::kMlpOutPad, wc::kMlpHidden, wc::kMlpOut);
    hipError_t e = hipMemcpy(m.master, w.data(), L.total * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m.adam_m, 0, L.total * 4);
    if (e == hipSuccess) e = hipMemset(m.adam_v, 0, L.total * 4);
    if (e == hipSuccess) e = wc::mlp_refresh_bf16(m, m.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m.stream);
    if (e != hipSuccess) {
        wc_mlp_destroy(h);
        return fail(WC_ERR_HIP, "network initialisation: %s", hipGetErrorString(e));
    }
    *out = h;
    return WC_OK;
}

WC_API int wc_mlp_destroy(wc_mlp* h) {
    if (!h) return WC_OK;
    wc::Mlp& m = h->m;
    if (m.stream) (void)hipStreamSynchronize(m.stream);
    void* ptrs[] = {m.master, m.grad, m.adam_m, m.adam_v, m.w1, m.w1t, m.w2, m.w2t, m.w3, m.w3t, m.w4, m.w4t,
                    m.z[0], m.z[1], m.z[2], m.h[0], m.h[1], m.h[2], m.yacc, m.d4, m.gbuf, m.d, m.loss,
                    m.colpart};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (m.stream) (void)hipStreamDestroy(m.stream);
    delete h;
    return WC_OK;
}

WC_API int wc_mlp_get_params(wc_mlp* h, float* host, uint32_t n) {
    if (!h || !host) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n != wc::mlp_layout().total) return fail(WC_ERR_INVALID_ARG, "expected %zu parameters", wc::mlp_layout().total);
    WC_HIP(hipStreamSynchronize(h->m.stream));
    WC_HIP(hipMemcpy(host, h->m.master, (size_t)n * 4, hipMemcpyDeviceToHost));
    return WC_OK;
}

WC_API int wc_mlp_set_params(wc_mlp* h, const float* host, uint32_t n) {
    if (!h || !host) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n != wc::mlp_layout().total) return fail(WC_ERR_INVALID_ARG, "expected %zu parameters", wc::mlp_layout().total);
    WC_HIP(hipStreamSynchronize(h->m.stream));
    WC_HIP(hipMemcpy(h->m.master, host, (size_t)n * 4, hipMemcpyHostToDevice));
    WC_HIP(hipMemset(h->m.adam_m, 0, (size_t)n * 4));
    WC_HIP(hipMemset(h->m.adam_v, 0, (size_t)n * 4));
    h->m.adam_t = 0;
    WC_HIP(wc::mlp_refresh_bf16(h->m, h->m.stream));
    WC_HIP(hipStreamSynchronize(h->m.stream));
    return WC_OK;
}

WC_API int wc_mlp_predict(wc_mlp* h, const void* x_dev, void* y_dev, uint32_t n) {
    if (!h || ((!x_dev || !y_dev) && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    WC_HIP(wc::mlp_infer(h->m, (const uint16_t*)x_dev, (float*)y_dev, n, h->m.stream));
    WC_HIP(hipStreamSynchronize(h->m.stream));
    return WC_OK;
}

WC_API int wc_mlp_forward(wc_mlp* h, const void* x_dev, void* y_dev, uint32_t n) {
    if (!h || ((!x_dev || !y_dev) && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n > h->m.max_batch) return fail(WC_ERR_CAPACITY, "batch %u above the network's %u", n, h->m.max_batch);
    if (n == 0) return WC_OK;
    WC_HIP(wc::mlp_forward_train(h->m, (const uint16_t*)x_dev, nullptr, (float*)y_dev, n, h->m.stream));
    WC_HIP(hipStreamSynchronize(h->m.stream));
    return WC_OK;
}

static int mlp_grad(wc_mlp* h, const void* x_dev, const void* t_dev, uint32_t n, double* loss) {
    if (!h || !x_dev || !t_dev) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (n == 0 || n > h->m.max_batch) return fail(WC_ERR_CAPACITY, "batch %u outside [1, %u]", n, h->m.max_batch);
    wc::Mlp& m = h->m;
    WC_HIP(wc::mlp_forward_train(m, (const uint16_t*)x_dev, (const float*)t_dev, nullptr, n, m.stream));
    WC_HIP(wc::mlp_backward(m, (const uint16_t*)x_dev, n, m.stream));
    if (loss) {
        float sse = 0.0f;
        WC_HIP(hipMemcpyAsync(&sse, m.loss, 4, hipMemcpyDeviceToHost, m.stream));
        WC_HIP(hipStreamSynchronize(m.stream));
        *loss = (double)sse / (3.0 * n);
    }
    return WC_OK;
}

WC_API int wc_mlp_gradient(wc_mlp* h, const void* x_dev, const void* t_dev, uint32_t n, float* grad_host,
                           double* loss) {
    int rc = mlp_grad(h, x_dev, t_dev, n, loss);
    if (rc) return rc;
    if (grad_host) {
        WC_HIP(hipStreamSynchronize(h->m.stream));
        WC_HIP(hipMemcpy(grad_host, h->m.grad, wc::mlp_layout().total * 4, hipMemcpyDeviceToHost));
    }
    return WC_OK;
}

WC_API int wc_mlp_train_step(wc_mlp* h, const void* x_dev, const void* t_dev, uint32_t n, float lr, double* loss) {
    int rc = mlp_grad(h, x_dev, t_dev, n, loss);
    if (rc) return rc;
    WC_HIP(wc::mlp_adam(h->m, lr, h->m.stream));
    if (!loss) WC_HIP(hipStreamSynchronize(h->m.stream));
    return WC_OK;
}

WC_API int wc_mlp_profile_train(wc_mlp* h, const void* x_dev, const void* t_dev, uint32_t n, float lr, uint32_t reps,
                                double* ms) {
    if (!ms) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (reps == 0) reps = 1;
    int rc = wc_mlp_train_step(h, x_dev, t_dev, n, lr, nullptr);  // warm-up; checks the arguments
    if (rc) return rc;
    EventPair ev;
    hipError_t err = ev.create();
    if (err == hipSuccess) err = hipEventRecord(ev.start, h->m.stream);
    for (uint32_t r = 0; r < reps && err == hipSuccess; ++r) {
        err = wc::mlp_forward_train(h->m, (const uint16_t*)x_dev, (const float*)t_dev, nullptr, n, h->m.stream);
        if (err == hipSuccess) err = wc::mlp_backward(h->m, (const uint16_t*)x_dev, n, h->m.stream);
        if (err == hipSuccess) err = wc::mlp_adam(h->m, lr, h->m.stream);
    }
    if (err == hipSuccess) err = hipEventRecord(ev.stop, h->m.stream);
    if (err == hipSuccess) err = hipEventSynchronize(ev.stop);
    float t = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, ev.start, ev.stop);
    if (err != hipSuccess) return fail(WC_ERR_HIP, "timed training: %s", hipGetErrorString(err));
    *ms = (double)t / reps;
    return WC_OK;
}

WC_API int wc_mlp_profile_predict(wc_mlp* h, const void* x_dev, void* y_dev, uint32_t n, uint32_t reps, double* ms) {
    if (!h || !ms || ((!x_dev || !y_dev) && n)) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (reps == 0) reps = 1;
    EventPair ev;
    hipError_t err = ev.create();
    if (err == hipSuccess) err = wc::mlp_infer(h->m, (const uint16_t*)x_dev, (float*)y_dev, n, h->m.stream);
    if (err == hipSuccess) err = hipEventRecord(ev.start, h->m.stream);
    for (uint32_t r = 0; r < reps && err == hipSuccess; ++r)
        err = wc::mlp_infer(h->m, (const uint16_t*)x_dev, (float*)y_dev, n, h->m.stream);
    if (err == hipSuccess) err = hipEventRecord(ev.stop, h->m.stream);
    if (err == hipSuccess) err = hipEventSynchronize(ev.stop);
    float t = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, ev.start, ev.stop);
    if (err != hipSuccess) return fail(WC_ERR_HIP, "timed inference: %s", hipGetErrorString(err));
    *ms = (double)t / reps;
    return WC_OK;
}

// ------------------------------------------------------------------------------------------
// Learned closure on a single-scene context.
// ------------------------------------------------------------------------------------------
static int closure_workspace(wc_ctx* c) {
    if (c->num_scenes != 1) return fail(WC_ERR_INVALID_ARG, "the learned closure runs on single-scene contexts");
    const size_t n = c->particle_cap ? c->particle_cap : 1;
    bool ok = true;
    auto alloc = [&](void** p, size_t bytes) {
        if (ok && !*p && hipMalloc(p, bytes) != hipSuccess) ok = false;
    };
    alloc((void**)&c->cl_x, n * wc::kClosureFeatures * 2);
    alloc((void**)&c->cl_y, n * 16);
    alloc((void**)&c->cl_long, (n + 1) * sizeof(uint32_t));
    for (int slot = 0; slot < 2; ++slot) {
        alloc((void**)&c->cl_backup[slot][0], n * sizeof(float4));
        alloc((void**)&c->cl_backup[slot][1], n * sizeof(float4));
    }
    if (ok && !c->cl_host_rho && hipHostMalloc((void**)&c->cl_host_rho, 2 * sizeof(uint32_t)) != hipSuccess) {
        c->cl_host_rho = nullptr;
        ok = false;
    }
    if (!ok) return fail(WC_ERR_NO_MEMORY, "closure workspace for %zu particles", n);
    return WC_OK;
}

WC_API int wc_closure_features(wc_ctx* c, float dt_mult, const uint32_t* rows_host, uint32_t nrows, void* x_dev) {
    if (!c || !x_dev) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (c->num_scenes != 1) return fail(WC_ERR_INVALID_ARG, "the learned closure runs on single-scene contexts");
    int rc = set_device(c);
    if (rc) return rc;
    const uint32_t n = c->total_particles;
    const uint32_t* rows = nullptr;
    if (rows_host) {
        for (uint32_t r = 0; r < nrows; ++r)
            if (rows_host[r] >= n) return fail(WC_ERR_INVALID_ARG, "row %u names slot %u of %u", r, rows_host[r], n);
        if (nrows > c->cl_rows_cap) {
            if (c->cl_rows) (void)hipFree(c->cl_rows);
            c->cl_rows = nullptr;
            c->cl_rows_cap = 0;
            if (hipMalloc(&c->cl_rows, (size_t)nrows * 4) != hipSuccess) return fail(WC_ERR_NO_MEMORY, "row list");
            c->cl_rows_cap = nrows;
        }
        WC_HIP(hipMemcpyAsync(c->cl_rows, rows_host, (size_t)nrows * 4, hipMemcpyHostToDevice, c->stream));
        rows = c->cl_rows;
    } else {
        nrows = n;
    }
    WC_HIP(wc::launch_closure_features(c->b, rows, nrows, dt_mult, (uint16_t*)x_dev, c->neighbour_list,
                                       c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

// Inference on the closure inputs formed from the neighbour lists (no feature matrix): the
// selection of longer lists into cl_x's memory, then the network.
static_assert(wc::kSelRow == wc::kClosureNeighbours, "the selection writes rows the inference reads");
static_assert(wc::kSelRow <= wc::kClosureFeatures / 2, "selection rows fit in the feature buffer");
static hipError_t closure_infer_from_lists(wc_ctx* c, wc_mlp* net, float dt_mult, float* y) {
    const uint32_t n = c->total_particles;
    hipError_t e = wc::launch_closure_selection(c->b, n, reinterpret_cast<uint32_t*>(c->cl_x), c->cl_long,
                                                c->cl_long + c->particle_cap, c->stream);
    if (e != hipSuccess) return e;
    const wc::StepConsts& k = c->desc[0].c;
    const wc::ClosureRows rows{c->b.posrho[1], c->b.velp[1], c->b.nlist, c->b.ncount,
                               reinterpret_cast<const uint32_t*>(c->cl_x), (size_t)c->total_particles, k.h, k.rho0,
                               dt_mult};
    return wc::mlp_infer_closure(net->m, rows, y, n, c->stream);
}

WC_API int wc_closure_infer(wc_ctx* c, wc_mlp* net, float dt_mult, void* y_dev) {
    if (!c || !net || !y_dev) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (c->num_scenes != 1) return fail(WC_ERR_INVALID_ARG, "the learned closure runs on single-scene contexts");
    if (!c->neighbour_list) return fail(WC_ERR_INVALID_ARG, "closure inputs from the lists: enable neighbour lists");
    int rc = closure_workspace(c);
    if (rc) return rc;
    rc = set_device(c);
    if (rc) return rc;
    WC_HIP(hipStreamSynchronize(net->m.stream));
    WC_HIP(closure_infer_from_lists(c, net, dt_mult, (float*)y_dev));
    WC_HIP(hipStreamSynchronize(c->stream));
    return WC_OK;
}

WC_API int wc_closure_update(wc_ctx* c, float dt, const void* corr_dev, float corr_scale, void* acc_dev) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "ctx is NULL");
    if (c->num_scenes != 1) return fail(WC_ERR_INVALID_ARG, "the learned closure runs on single-scene contexts");
    int rc = set_device(c);
    if (rc) return rc;
    WC_HIP(wc::launch_closure_update(c->b, dt, (const float*)corr_dev, corr_scale, (float*)acc_dev, false, 0.0f,
                                     c->neighbour_list, c->stream));
    return WC_OK;
}

static int read_rho_max(wc_ctx* c, float* out) {
    uint32_t bits = 0;
    WC_HIP(hipMemcpyAsync(&bits, c->b.rho_max, 4, hipMemcpyDeviceToHost, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    memcpy(out, &bits, 4);
    return WC_OK;
}

WC_API int wc_closure_set_smoothing(wc_ctx* c, int enable) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "NULL context");
    c->cl_smooth = enable != 0;
    return WC_OK;
}

WC_API int wc_closure_set_integrator(wc_ctx* c, int substeps) {
    if (!c) return fail(WC_ERR_INVALID_ARG, "NULL context");
    c->cl_substeps = substeps != 0;
    return WC_OK;
}

WC_API int wc_closure_run(wc_ctx* c, wc_mlp* net, float dt, uint32_t K, uint32_t nmacro, float judge_tol,
                          wc_closure_stats* stats) {
    if (!c || !stats) return fail(WC_ERR_INVALID_ARG, "NULL argument");
    if (K == 0) return fail(WC_ERR_INVALID_ARG, "K must be >= 1");
    int rc = closure_workspace(c);
    if (rc) return rc;
    rc = set_all_dt(c, &dt, true);
    if (rc) return rc;
    rc = set_device(c);
    if (rc) return rc;
    memset(stats, 0, sizeof(*stats));
    // the judge's allowance for one learned step: judge_tol per small step of dt it replaces
    const float tol = judge_tol * (float)K;
    const uint32_t n = c->total_particles;
    const float rho0 = c->desc[0].c.rho0;
    const float dt_large = dt * (float)K;
    const float g = 9.81f;
    for (int slot = 0; slot < 2; ++slot)
        if (c->b.ids[0] && !c->cl_backup_ids[slot] &&
            hipMalloc(&c->cl_backup_ids[slot], (size_t)(n ? n : 1) * 4) != hipSuccess)
            return fail(WC_ERR_NO_MEMORY, "identity backup");
    EventPair ev;
    WC_HIP(ev.create());
    WC_HIP(hipEventRecord(ev.start, c->stream));
    auto sort_density = [&]() -> int {
        int r = enqueue_sort(c);
        if (r) return r;
        WC_HIP(wc::launch_density(c->b, c->density_variant, c->neighbour_list, c->stream));
        return WC_OK;
    };
    // One macro step: form the closure inputs, run the network, take the large step (whose force
    // pass also saves the sorted state it starts from into `slot`), sort and run the density pass
    // of the new state (the judge's input, and the next step's), and copy the density maximum to
    // the host.  Captured once per slot and replayed.
    auto enqueue_macro = [&](int slot) -> int {
        if (net && c->neighbour_list) {
            WC_HIP(closure_infer_from_lists(c, net, (float)K, c->cl_y));
        } else if (net) {
            WC_HIP(wc::launch_closure_features(c->b, nullptr, n, (float)K, c->cl_x, c->neighbour_list,
                                               c->stream));
            WC_HIP(wc::mlp_infer(net->m, c->cl_x, c->cl_y, n, c->stream));
        }
        wc::RecordBackup backup;
        backup.posrho = c->cl_backup[slot][0];
        backup.velp = c->cl_backup[slot][1];
        backup.ids = c->b.ids[0] ? c->cl_backup_ids[slot] : nullptr;
        WC_HIP(wc::launch_closure_update(c->b, dt_large, net ? c->cl_y : nullptr, g, nullptr, c->cl_smooth,
                                         c->cl_substeps ? (float)K : 0.0f, c->neighbour_list, c->stream, backup));
        int r = sort_density();
        if (r) return r;
        WC_HIP(hipMemcpyAsync(c->cl_host_rho + slot, c->b.rho_max, 4, hipMemcpyDeviceToHost, c->stream));
        return WC_OK;
    };
    struct MacroGraphs {
        hipStream_t stream = nullptr;
        hipGraph_t graph[2] = {nullptr, nullptr};
        hipGraphExec_t exec[2] = {nullptr, nullptr};
        hipEvent_t done[2] = {nullptr, nullptr};
        ~MacroGraphs() {
            // a speculative step may still be queued when an error ends the run
            if (stream) (void)hipStreamSynchronize(stream);
            for (int i = 0; i < 2; ++i) {
                if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
                if (graph[i]) (void)hipGraphDestroy(graph[i]);
                if (done[i]) (void)hipEventDestroy(done[i]);
            }
        }
    } mg;
    mg.stream = c->stream;
    // the network's own stream may still be writing its bf16 weights
    if (net) WC_HIP(hipStreamSynchronize(net->m.stream));
    float prev_max = 0.0f;
    rc = sort_density();
    if (!rc) rc = read_rho_max(c, &prev_max);
    stats->peak_density = prev_max / rho0;
    for (int slot = 0; slot < 2 && !rc && nmacro > 0; ++slot) {
        WC_HIP(hipEventCreateWithFlags(&mg.done[slot], hipEventDisableTiming));
        WC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_macro(slot);
        const hipError_t e = hipStreamEndCapture(c->stream, &mg.graph[slot]);
        if (!rc && e != hipSuccess) rc = fail(WC_ERR_HIP, "learned step capture: %s", hipGetErrorString(e));
        if (!rc) {
            const hipError_t ei = hipGraphInstantiate(&mg.exec[slot], mg.graph[slot], nullptr, nullptr, 0);
            if (ei != hipSuccess) rc = fail(WC_ERR_HIP, "learned step graph: %s", hipGetErrorString(ei));
        }
    }
    // Macro step m runs in slot m % 2.  Step m + 1 is queued before step m's verdict is read, so
    // the GPU does not wait on the host; it works on step m's result and saves it in the other
    // slot, leaving slot m % 2 (the state before step m) intact.  When step m fails its judge, the
    // state is restored from slot m % 2, K plain steps replace it, and step m + 1 is queued again
    // after them: the speculative one ran on the rejected state and is overwritten.  A step is
    // queued ahead only while the last density jump stayed under half the tolerance, so that a
    // likely rejection does not also throw away the step after it.
    auto launch = [&](uint32_t m) -> int {
        WC_HIP(hipGraphLaunch(mg.exec[m % 2], c->stream));
        WC_HIP(hipEventRecord(mg.done[m % 2], c->stream));
        return WC_OK;
    };
    uint32_t queued = 0;  // macro steps launched so far
    float last_jump = 0.0f;
    for (uint32_t m = 0; m < nmacro && !rc; ++m) {
        const int slot = (int)(m % 2);
        if (queued == m) {
            rc = launch(m);
            if (rc) break;
            queued = m + 1;
        }
        const bool ahead = m + 1 < nmacro && !(last_jump > 0.5f * tol);
        if (ahead) {
            rc = launch(m + 1);
            if (rc) break;
            queued = m + 2;
        }
        const hipError_t e = hipEventSynchronize(mg.done[slot]);
        if (e != hipSuccess) {
            rc = fail(WC_ERR_HIP, "learned step: %s", hipGetErrorString(e));
            break;
        }
        stats->macro_steps += 1;
        float now = 0.0f;
        memcpy(&now, c->cl_host_rho + slot, 4);
        const float jump = (now - prev_max) / rho0;
        last_jump = jump;
        if ((double)jump > stats->max_density_jump) stats->max_density_jump = jump;
        if (!(jump <= tol)) {  // NaN fails too
            stats->fallbacks += 1;
            WC_HIP(hipMemcpyAsync(c->b.posrho[0], c->cl_backup[slot][0], (size_t)n * sizeof(float4),
                                  hipMemcpyDeviceToDevice, c->stream));
            WC_HIP(hipMemcpyAsync(c->b.velp[0], c->cl_backup[slot][1], (size_t)n * sizeof(float4),
                                  hipMemcpyDeviceToDevice, c->stream));
            if (c->b.ids[0])
                WC_HIP(hipMemcpyAsync(c->b.ids[0], c->cl_backup_ids[slot], (size_t)n * 4, hipMemcpyDeviceToDevice,
                                      c->stream));
            rc = run_steps(c, K);
            if (!rc) rc = sort_density();
            if (!rc) rc = read_rho_max(c, &now);
            if (!rc && ahead) rc = launch(m + 1);
        }
        if (jump <= tol && (double)jump > stats->max_accepted_jump) stats->max_accepted_jump = jump;
        if ((double)now / rho0 > stats->peak_density) stats->peak_density = (double)now / rho0;
        prev_max = now;
    }
    if (!rc) {
        hipError_t e = hipEventRecord(ev.stop, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ev.stop);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev.start, ev.stop);
        if (e != hipSuccess) rc = fail(WC_ERR_HIP, "closure timing: %s", hipGetErrorString(e));
        stats->wall_ms = ms;
    }
    return rc;
}

}  // extern "C"
