// wc_internal.hpp -- types shared by the host runtime and the gfx950 kernels of the SPH step.
//
// A context holds one or more independent scenes.  Their particles are concatenated in the
// particle buffers and their cell grids in the cell arrays; a device-side table describes each
// scene (constants, first particle slot, first cell, step size).  Kernels launch once per pass
// over all scenes: a workgroup finds its scene with one scalar load from a per-block scene table
// (a context with one scene skips it).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/watercube.h"

namespace wc {

constexpr int kHashBlock = 256;       // threads per workgroup: hash, place, gather, copies
constexpr int kScanThreads = 256;     // scan workgroup
constexpr int kScanItems = 16;        // counts per scan thread (four 16-byte loads)
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kNoKey = 0xFFFFFFFFu;
constexpr uint32_t kLookbackSpinLimit = 1u << 24;  // bounded wait in the scan look-back
// Records allocated (and zeroed) past the last particle of the particle buffers: the neighbour
// passes load a cell's slots CHUNK at a time and may read up to CHUNK - 1 records beyond it.
constexpr size_t kRecordPad = 8;
// Neighbour lists (optional, chosen by the tuner): the density pass stores each particle's
// accepted neighbours in enumeration order, entry k of slot t at nlist[k * total + t], up to
// kNeighbourListMax of them; ncount[t] is their number, or kListOverflow when there were more
// (the force pass then searches for that particle itself).
constexpr uint32_t kNeighbourListMax = 64;
constexpr uint32_t kListOverflow = 0xFFFFFFFFu;
// The reorder ranks a particle inside its cell by counting the cell's particles with a smaller
// input index: at most kGatherDirectMax of them.  More crowded cells (a dense cluster, or every
// NaN position, which hashes to cell 0) go to a work list, and one of the gather's leading
// workgroups per listed cell sorts its input indices instead.  At most
// capacity / (kGatherDirectMax + 1) cells can be listed.
constexpr uint32_t kGatherDirectMax = 256;
inline uint32_t crowd_capacity(uint64_t particles) { return (uint32_t)(particles / (kGatherDirectMax + 1) + 1); }

// Per-scene constants, precomputed on the host in float so the device sees exactly the bits the
// CPU transcription uses.
struct StepConsts {
    uint32_t n;          // particles of the scene
    uint32_t ncells;     // R^3
    int32_t R;           // grid resolution
    float bin;           // binSize
    float h;             // kernel radius
    float hh;            // h * h, the left operand of poly6 (density.comp:54)
    float cut2;          // smallest float t with sqrtf(t) >= h: dist >= h  <=>  d2 >= cut2
    float m;             // particle mass
    float poly6, spiky, visc;
    float mu;            // viscosity coefficient
    float rho0, p0, k;   // rest density / pressure, stiffness
    float g[3];          // gravity vector
    float size;
    float size_minus_h;  // size - kernelRadius (density.comp:62, update.comp:78)
    float border;
    float hi;            // size - border (update.comp:208)
    float damping;
    float max_speed;
    int32_t eos;
    int32_t wall_density, wall_force, mouse_force;
    float cam[3], dir[3];
};

struct SceneDesc {
    StepConsts c;
    uint32_t p_off;  // first particle slot of the scene
    uint32_t c_off;  // first cell of the scene (a multiple of kScanTile)
    float dt;        // integration step
    uint32_t pad;
};

// Device tables of a context's scenes, passed by value to every kernel.
struct Scenes {
    const SceneDesc* desc;
    const uint32_t* blk64;   // first 64-thread workgroup of each scene; num_scenes + 1 entries
    const uint32_t* blk128;  // same for 128-thread workgroups
    const uint32_t* blk256;  // same for 256-thread workgroups
    const uint32_t* tiles;   // first scan tile of each scene; num_scenes + 1 entries
    // multi-scene contexts: the scene of every 64-, 128-, 256-thread block and of every scan tile
    // (NULL with one scene)
    const uint32_t* scene64;
    const uint32_t* scene128;
    const uint32_t* scene256;
    const uint32_t* scene_tile;
    uint32_t num_scenes;
    uint32_t total_particles;
    uint32_t total_tiles;
    uint32_t blocks64, blocks128, blocks256;
};

// Device buffers of one context.  posrho/velp are the 32-byte record split in its two
// 16-byte halves: (x, y, z, density) and (vx, vy, vz, pressure).
struct DeviceBuffers {
    float4* posrho[2];
    float4* velp[2];
    uint32_t* key;       // cell index (scene-local) per particle of buffer 1
    uint32_t* rank;      // arrival rank inside its cell (atomic order)
    uint32_t* slot_tmp;  // slot -> input index in arrival order
    uint32_t* sorted;    // slot -> input index in stable order
    uint32_t* skey;      // slot -> scene-local cell index of the sorted particle
    uint32_t* counts;    // per cell: atomic histogram; the scan reads and zeroes it
    uint2* range;        // per cell: (first slot, count); slots are context-global
    unsigned long long* scan_status;  // one look-back word per scan tile
    uint32_t* error_flag;             // nonzero after a failed bounded wait
    unsigned long long* scene_done;   // per scene: wall clock when its force pass finished
    unsigned long long* step_start;   // wall clock when the last step's hash pass started
    uint32_t* rho_max;                // [0] bits of the largest density of the last density pass;
                                      // [1] 0 when the force pass may share reciprocals
    uint32_t* ids[2];                 // particle identity per slot of buffer 1 / 2 (NULL: not tracked)
    uint32_t* nlist;                  // neighbour lists (NULL until the list mode is first used)
    uint32_t* ncount;
    uint2* crowd;                     // cells with more than kGatherDirectMax particles: (cell, scene)
    uint32_t* crowd_count;            // entries of `crowd` (reset by every sort)
    uint32_t crowd_cap;
    Scenes sc;
};

// Learned closure (single-scene contexts).
constexpr int kClosureNeighbours = 32;
constexpr int kClosureFeatures = 208;  // 195 used, zero padded
// Features of the sorted particles at `rows` (slots; NULL = every slot) into X (bf16 rows); with
// `lists`, from the neighbour lists of the last density pass instead of a search.
hipError_t launch_closure_features(const DeviceBuffers& b, const uint32_t* rows, uint32_t nrows, float dt_mult,
                                   uint16_t* X, bool lists, hipStream_t stream);
// Force pass with dt given directly; corr (4 floats per slot, units of g, may be NULL) is added to
// the acceleration of particles at least 2h from every wall, averaged over each particle's
// neighbours when `smooth`; acc_out (may be NULL) receives the plain acceleration.  substeps = K
// > 0 integrates as K steps of dt / K under the frozen force (0: one symplectic Euler step).
// With `lists` the neighbours come from the last density pass's neighbour lists.
// Where the force pass also copies each particle's sorted record (and identity) before
// integrating it: the learned closure's saved state, written without a separate copy.
struct RecordBackup {
    float4* posrho = nullptr;
    float4* velp = nullptr;
    uint32_t* ids = nullptr;
};
hipError_t launch_closure_update(const DeviceBuffers& b, float dt, const float* corr, float corr_scale,
                                 float* acc_out, bool smooth, float substeps, bool lists, hipStream_t stream,
                                 const RecordBackup& backup = RecordBackup{});

// Tuning variant of the neighbour passes; all variants produce identical results.
struct PassVariant {
    int block;         // threads per workgroup
    int list_cap;      // LDS list entries per lane
    int chunk;         // neighbour slots loaded per step of a lane's walk
    bool plane_drain;  // drain the list after every dx plane (else only when full and at the end)
};
int num_variants();
PassVariant variant(int v);

// Kernel launchers (wc_kernels.hip).  Each enqueues on `stream` and returns the launch status.
hipError_t launch_hash(const DeviceBuffers& b, hipStream_t stream);
hipError_t launch_scan(const DeviceBuffers& b, hipStream_t stream);
hipError_t launch_place(const DeviceBuffers& b, hipStream_t stream);
hipError_t launch_gather(const DeviceBuffers& b, hipStream_t stream);
// With `list`, the density pass also writes the neighbour lists and the force pass reads them.
hipError_t launch_density(const DeviceBuffers& b, int variant, bool list, hipStream_t stream);
// For the single scene's particles with more than kClosureNeighbours listed neighbours (or an
// overflowed list): the slots of the kClosureNeighbours nearest, in list order, at
// sel[t * kClosureFeatures / 2 + k].  Other rows are left alone.
// The kept neighbours' slots of every particle with more listed neighbours than the features keep
// (or an overflowed list), at sel[t * (kClosureFeatures / 2)]; heavy (n words) and heavy_count
// are workspace.
hipError_t launch_closure_selection(const DeviceBuffers& b, uint32_t n, uint32_t* sel, uint32_t* heavy,
                                    uint32_t* heavy_count, hipStream_t stream);
hipError_t launch_update(const DeviceBuffers& b, int variant, bool list, hipStream_t stream);
hipError_t launch_advect(const DeviceBuffers& b, float dt, hipStream_t stream);
// After the host changed particle buffer 2 or the parameters: recompute buffer 2's cell keys,
// mark the neighbour lists overflowed and the force pass's division operands unchecked.
hipError_t launch_invalidate(const DeviceBuffers& b, hipStream_t stream);
// Allocation-free step: writes one scene's table (d, and the 8 words blk64[0..1], blk128[0..1],
// blk256[0..1], tiles[0..1]) into scratch memory and zeroes counts[0, ncells) and *error_flag.
hipError_t launch_noalloc_setup(const SceneDesc& d, const uint32_t tables_host[8], SceneDesc* desc, uint32_t* tables,
                                uint32_t* counts, uint32_t ncells, uint32_t* error_flag, hipStream_t stream);
hipError_t launch_split_records(const wc_particle* rec, float4* posrho, float4* velp, uint32_t n,
                                hipStream_t stream);
hipError_t launch_join_records(const float4* posrho, const float4* velp, wc_particle* rec, uint32_t n,
                               hipStream_t stream);

StepConsts make_consts(const wc_params& p);
float dist2_cutoff(float h);

}  // namespace wc
