// wc_internal.hpp -- types shared by the host runtime and the gfx950 kernels of the SPH step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/watercube.h"

namespace wc {

constexpr int kBlock = 256;           // threads per workgroup for the per-particle kernels
constexpr int kScanThreads = 256;     // scan workgroup
constexpr int kScanItems = 16;        // counts per scan thread (four 16-byte loads)
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kNoKey = 0xFFFFFFFFu;
constexpr uint32_t kLookbackSpinLimit = 1u << 24;  // bounded wait in the scan look-back

// Everything one step's kernels read besides buffers, precomputed on the host in float so the
// device sees exactly the bits the CPU transcription uses.
struct StepConsts {
    uint32_t n;          // particles
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
    float size_minus_h;  // size - kernelRadius (density.comp:62, update.comp:215)
    float border;
    float hi;            // size - border (update.comp:345)
    float damping;
    float max_speed;
    int32_t eos;
    int32_t wall_density, wall_force, mouse_force;
    float cam[3], dir[3];
};

// Device buffers of one context.  posrho/velp are the 32-byte record split in its two
// 16-byte halves: (x, y, z, density) and (vx, vy, vz, pressure).
struct DeviceBuffers {
    float4* posrho[2];
    float4* velp[2];
    uint32_t* key;       // cell index per particle of buffer 1 (input order)
    uint32_t* rank;      // arrival rank inside its cell (atomic order)
    uint32_t* slot_tmp;  // slot -> input index in arrival order
    uint32_t* sorted;    // slot -> input index in stable order
    uint32_t* skey;      // slot -> cell index of the sorted particle
    uint32_t* counts;    // R^3 atomic histogram; the scan reads and zeroes it
    uint2* range;        // R^3 (exclusive offset, count)
    unsigned long long* scan_status;  // one look-back word per scan tile
    uint32_t* scan_ticket;            // dynamic tile counter
    uint32_t* error_flag;             // nonzero after a failed bounded wait
    uint32_t num_scan_tiles;
};

// Kernel launchers (wc_kernels.hip).  Each enqueues on `stream` and returns the launch status.
hipError_t launch_hash(const DeviceBuffers& b, const StepConsts& c, hipStream_t stream);
hipError_t launch_scan(const DeviceBuffers& b, const StepConsts& c, hipStream_t stream);
hipError_t launch_place(const DeviceBuffers& b, const StepConsts& c, hipStream_t stream);
hipError_t launch_gather(const DeviceBuffers& b, const StepConsts& c, hipStream_t stream);
// Tuning variant of the neighbour passes; all variants produce identical results.
struct PassVariant {
    int block;     // threads per workgroup
    int list_cap;  // LDS list entries per lane
    int chunk;     // neighbour slots loaded per step of a lane's walk
    bool plane_drain;  // drain the list after every dx plane (else only when full and at the end)
};
int num_variants();
PassVariant variant(int v);
hipError_t launch_density(const DeviceBuffers& b, const StepConsts& c, int variant, hipStream_t stream);
hipError_t launch_update(const DeviceBuffers& b, const StepConsts& c, float dt, int variant, hipStream_t stream);
hipError_t launch_split_records(const wc_particle* rec, float4* posrho, float4* velp, uint32_t n,
                                hipStream_t stream);
hipError_t launch_join_records(const float4* posrho, const float4* velp, wc_particle* rec, uint32_t n,
                               hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream);

StepConsts make_consts(const wc_params& p);
float dist2_cutoff(float h);

}  // namespace wc
