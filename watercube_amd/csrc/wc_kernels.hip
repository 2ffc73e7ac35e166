// wc_kernels.hip -- gfx950 kernels of the SPH step: grid build (hash + histogram, decoupled
// look-back scan, place, stable within-cell rank + gather), density/pressure, force + integrate.
// Every pass is one launch over all scenes of a context.
//
// Floating-point contract (shared with the CPU transcription): binary32, correctly rounded '/'
// and sqrtf (-fhip-fp32-correctly-rounded-divide-sqrt), subnormals kept, no FMA contraction
// (-ffp-contract=off), every expression associated left to right as the GLSL source writes it.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "wc_internal.hpp"

namespace wc {
namespace {

__device__ __forceinline__ int cell_axis(float x, float bin, int R) {
    // ivec3(p / binSize) truncates toward zero, then clamp(.., 0, R-1) (count.comp:32).
    const float q = x / bin;
    if (!(q >= 0.0f)) return 0;
    if (q >= (float)R) return R - 1;
    const int c = (int)q;
    return c > R - 1 ? R - 1 : c;
}

__device__ __forceinline__ uint32_t cell_key(const StepConsts& c, float x, float y, float z) {
    const int cx = cell_axis(x, c.bin, c.R);
    const int cy = cell_axis(y, c.bin, c.R);
    const int cz = cell_axis(z, c.bin, c.R);
    return (uint32_t)((cz * c.R + cy) * c.R + cx);
}

__device__ __forceinline__ float poly6(const StepConsts& c, float r) {
    const float a = c.hh - r * r;
    return ((a * a) * a) * c.poly6;
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// The dispatcher hands workgroup b to XCD b % 8.  Renumber workgroups so that every XCD works
// through one contiguous run of logical blocks: neighbouring particles (which read the same
// neighbour cells) then share one XCD's L2 instead of being fetched by all eight.
constexpr uint32_t kNumXcd = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
    const uint32_t q = nblocks / kNumXcd, r = nblocks % kNumXcd;
    const uint32_t x = b % kNumXcd, k = b / kNumXcd;
    return x * q + min(x, r) + k;
}

// Scene owning logical block (or tile) b: a table load when the context holds several scenes.
__device__ __forceinline__ uint32_t scene_of(const uint32_t* __restrict__ table, uint32_t num_scenes, uint32_t b) {
    return num_scenes == 1 ? 0u : table[b];
}
template <int BLOCK>
__device__ __forceinline__ const uint32_t* scene_table(const Scenes& S) {
    return BLOCK == 64 ? S.scene64 : (BLOCK == 128 ? S.scene128 : S.scene256);
}

template <int BLOCK>
__device__ __forceinline__ const uint32_t* first_blocks(const Scenes& S) {
    return BLOCK == 64 ? S.blk64 : (BLOCK == 128 ? S.blk128 : S.blk256);
}
template <int BLOCK>
__device__ __forceinline__ uint32_t total_blocks(const Scenes& S) {
    return BLOCK == 64 ? S.blocks64 : (BLOCK == 128 ? S.blocks128 : S.blocks256);
}

// Bits of rho_max[1], the force pass's division mode: 0 lets it share reciprocals (quot).
constexpr uint32_t kDivUnsafe = 1u;   // an operand of this step is outside div_operand's range
constexpr uint32_t kDivPending = 2u;  // sorted since the last density pass

// Walks the in-grid neighbour cells of scene-local cell `key` in NEIGHBORHOOD order
// (density.comp:40-50: dx outermost, dz innermost); visit(start, end, own) receives each cell's
// slot range (clipped to `limit`) and whether it is the particle's own cell (the 14th of the 27,
// the same iteration in every lane: std::true_type there, std::false_type elsewhere), and
// end_plane() runs after every dx plane.  The trip counts of both cell loops are wave-uniform.
template <typename Visit, typename EndPlane>
__device__ __forceinline__ void for_each_neighbour_cell(uint32_t key, const uint2* __restrict__ range,
                                                        const StepConsts& c, uint32_t limit, Visit& visit,
                                                        EndPlane& end_plane) {
    const int R = c.R;
    const int cx = (int)(key % (uint32_t)R);
    const int cy = (int)((key / (uint32_t)R) % (uint32_t)R);
    const int cz = (int)(key / ((uint32_t)R * (uint32_t)R));
#pragma unroll 1
    for (int dx = -1; dx <= 1; ++dx) {
        const int nx = cx + dx;
#pragma unroll 1
        for (int dyz = 0; dyz < 9; ++dyz) {
            const int ny = cy + (dyz / 3 - 1);
            const int nz = cz + (dyz % 3 - 1);
            if (nx < 0 || ny < 0 || nz < 0 || nx >= R || ny >= R || nz >= R) continue;
            const uint2 rg = range[(nz * R + ny) * R + nx];
            if (dx == 0 && dyz == 4)
                visit(rg.x, min(rg.x + rg.y, limit), std::true_type{});
            else
                visit(rg.x, min(rg.x + rg.y, limit), std::false_type{});
        }
        end_plane();
    }
}

// ------------------------------------------------------------------------------------------
// 1. cell hash + histogram (count.comp:25-36).  Consecutive lanes that fall in the same cell
//    form one run; the run's first lane does a single atomicAdd for the whole run, every lane
//    keeps its arrival rank.  Also resets the scan's look-back state and the per-scene clocks
//    for this step.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kHashBlock) void k_hash(Scenes S, const float4* __restrict__ posrho,
                                                     uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ counts,
                                                     unsigned long long* __restrict__ scan_status,
                                                     unsigned long long* __restrict__ scene_done,
                                                     unsigned long long* __restrict__ step_start,
                                                     uint32_t* __restrict__ rho_max,
                                                     uint32_t* __restrict__ crowd_count) {
    const uint32_t g = blockIdx.x * kHashBlock + threadIdx.x;
    if (g < S.total_tiles) scan_status[g] = 0ull;
    if (g < S.num_scenes) scene_done[g] = 0ull;
    if (g == 0) {
        *crowd_count = 0u;
        *step_start = wall_clock64();
        rho_max[0] = 0u;
        rho_max[1] = kDivPending;  // the force pass's operands are unchecked until a density pass runs
    }
    const uint32_t b = blockIdx.x;
    bool valid = false;
    uint32_t k = kNoKey, i = 0, c_off = 0;
    if (b < S.blocks256) {
        const uint32_t s = scene_of(S.scene256, S.num_scenes, b);
        const SceneDesc& d = S.desc[s];
        const uint32_t lt = (b - S.blk256[s]) * kHashBlock + threadIdx.x;
        valid = lt < d.c.n;
        if (valid) {
            i = d.p_off + lt;
            c_off = d.c_off;
            const float4 p = posrho[i];
            k = cell_key(d.c, p.x, p.y, p.z);
        }
    }
    const unsigned lane = lane_id();
    const uint32_t prev = __shfl_up(k, 1);
    const bool leader = (lane == 0) || (prev != k);
    const unsigned long long leaders = __ballot(leader);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
    const int first = 63 - __clzll((long long)(leaders & upto));
    const unsigned long long above = leaders & ~upto;
    const int next = above ? (__ffsll((long long)above) - 1) : 64;
    uint32_t base = 0;
    if (leader && valid) base = atomicAdd(&counts[c_off + k], (uint32_t)(next - first));
    base = __shfl(base, first);
    if (valid) {
        key[i] = k;
        rank[i] = base + (uint32_t)((int)lane - first);
    }
}

// ------------------------------------------------------------------------------------------
// 2. exclusive scan of the per-cell counts (linearScan.comp:15-28), single pass with decoupled
//    look-back, each scene on its own chain of tiles.  Writes range[cell] = (first slot, count)
//    and zeroes the counts for the next step.
//    Look-back word: (flag << 32) | value; flag 1 = tile aggregate, 2 = inclusive prefix.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long ld_status(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan(Scenes S, uint32_t* __restrict__ counts,
                                                       uint2* __restrict__ range,
                                                       unsigned long long* __restrict__ status,
                                                       uint32_t* __restrict__ error_flag) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_prefix;
    const unsigned tid = threadIdx.x;
    const unsigned lane = lane_id();
    const unsigned wave = tid >> 6;
    // Tile = workgroup index.  Each XCD dispatches its workgroups in index order, so the lowest
    // unfinished tile has been dispatched, and it waits on nothing (its predecessors are done):
    // the look-back always makes progress.  (A ticket counter gives the same guarantee, but
    // every workgroup then queues on one atomic: about 25-30 ns per tile.)
    const uint32_t tile = blockIdx.x;
    const uint32_t s = scene_of(S.scene_tile, S.num_scenes, tile);
    const uint32_t first_tile = S.tiles[s];
    const SceneDesc& d = S.desc[s];
    const uint32_t ncells = d.c.ncells;
    const size_t local = (size_t)(tile - first_tile) * kScanTile + (size_t)tid * kScanItems;
    uint32_t* cnt = counts + d.c_off;
    uint2* rng = range + d.c_off;

    uint32_t v[kScanItems];
    if (local + kScanItems <= ncells) {
        const uint4* src = reinterpret_cast<const uint4*>(cnt + local);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const uint4 w = src[q];
            v[4 * q + 0] = w.x;
            v[4 * q + 1] = w.y;
            v[4 * q + 2] = w.z;
            v[4 * q + 3] = w.w;
        }
        uint4* dst = reinterpret_cast<uint4*>(cnt + local);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) dst[q] = make_uint4(0u, 0u, 0u, 0u);
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            v[j] = (local + j < ncells) ? cnt[local + j] : 0u;
            if (local + j < ncells) cnt[local + j] = 0u;
        }
    }
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) sum += v[j];

    uint32_t incl = sum;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const uint32_t t = __shfl_up(incl, dd);
        if ((int)lane >= dd) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_prefix = 0, tile_total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const uint32_t t = s_wave[w];
        if ((unsigned)w < wave) wave_prefix += t;
        tile_total += t;
    }
    const uint32_t thread_excl = wave_prefix + incl - sum;

    if (wave == 0) {
        uint32_t tile_prefix = 0;
        if (tile == first_tile) {
            if (lane == 0) st_status(&status[tile], (2ull << 32) | tile_total);
        } else {
            if (lane == 0) st_status(&status[tile], (1ull << 32) | tile_total);
            long long look = (long long)tile - 1;
            uint32_t acc = 0, spins = 0;
            for (;;) {
                const long long idx = look - (long long)lane;
                // tiles before the scene's first tile count as a zero prefix
                unsigned long long w = (idx >= (long long)first_tile) ? ld_status(&status[idx]) : (2ull << 32);
                const uint32_t flag = (uint32_t)(w >> 32);
                const unsigned long long not_ready = __ballot(flag == 0u);
                const unsigned long long is_prefix = __ballot(flag == 2u);
                const int first_prefix = is_prefix ? (__ffsll((long long)is_prefix) - 1) : 64;
                const unsigned long long window =
                    (first_prefix >= 63) ? ~0ull : ((2ull << first_prefix) - 1ull);
                if (not_ready & window) {
                    if (++spins > kLookbackSpinLimit) {
                        if (lane == 0) atomicOr(error_flag, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const uint32_t contrib = ((int)lane <= first_prefix) ? (uint32_t)w : 0u;
                acc += wave_sum(contrib);
                if (first_prefix < 64) break;
                look -= 64;
            }
            tile_prefix = acc;
            if (lane == 0) st_status(&status[tile], (2ull << 32) | (uint32_t)(acc + tile_total));
        }
        if (lane == 0) s_prefix = tile_prefix;
    }
    __syncthreads();
    uint32_t run = d.p_off + s_prefix + thread_excl;
    if (local + kScanItems <= ncells) {
        uint4* dst = reinterpret_cast<uint4*>(rng + local);
#pragma unroll
        for (int q = 0; q < kScanItems / 2; ++q) {
            const uint32_t a = run;
            run += v[2 * q];
            const uint32_t b2 = run;
            run += v[2 * q + 1];
            dst[q] = make_uint4(a, v[2 * q], b2, v[2 * q + 1]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            if (local + j < ncells) rng[local + j] = make_uint2(run, v[j]);
            run += v[j];
        }
    }
}

// ------------------------------------------------------------------------------------------
// 3. place: slot = first slot of the cell + arrival rank (reorder.comp:38-43 without the copy).
// ------------------------------------------------------------------------------------------
// The first particle to arrive in a cell of more than kGatherDirectMax also lists the cell for the
// leading workgroups of k_gather, which sort its input indices instead of ranking them one by one.
__global__ __launch_bounds__(kHashBlock) void k_place(Scenes S, const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ rank,
                                                      const uint2* __restrict__ range,
                                                      uint32_t* __restrict__ slot_tmp, uint2* __restrict__ crowd,
                                                      uint32_t* __restrict__ crowd_count, uint32_t crowd_cap,
                                                      uint32_t* __restrict__ error_flag) {
    const uint32_t b = blockIdx.x;
    const uint32_t s = scene_of(S.scene256, S.num_scenes, b);
    const SceneDesc& d = S.desc[s];
    const uint32_t lt = (b - S.blk256[s]) * kHashBlock + threadIdx.x;
    if (lt >= d.c.n) return;
    const uint32_t i = d.p_off + lt;
    const uint32_t k = key[i];
    const uint2 rg = range[d.c_off + k];
    const uint32_t r = rank[i];
    const uint32_t slot = rg.x + r;
    if (slot - d.p_off < d.c.n) slot_tmp[slot] = i;  // always true when the scan succeeded
    if (rg.y > kGatherDirectMax && r == 0u) {       // one entry per crowded cell
        const uint32_t at = atomicAdd(crowd_count, 1u);
        if (at < crowd_cap)
            crowd[at] = make_uint2(d.c_off + k, s);
        else
            atomicOr(error_flag, 2u);  // cannot happen when the scan succeeded
    }
}

// Correctly rounded a / b from y = RN(1 / b): q0 = RN(a y), r = a - b q0 (exact), RN(q0 + r y)
// (Markstein's theorem).  Equal to the IEEE quotient, except that a = -0 gives +0, when b and y
// are normal and a is 0 or at least 2^-100 in magnitude with a normal quotient, so that no step
// underflows or overflows.  The force pass uses it only on steps whose operands the density pass
// found in that range (div_operand); then three quotients share one division.
__device__ __forceinline__ float quot(float a, float b, float y) {
    const float q0 = a * y;
    const float r = fmaf(-b, q0, a);
    return fmaf(r, y, q0);
}

// True for 0 and for 2^-40 <= |x| <= 2^40 (false for NaN and infinity).  Sums and differences of
// two such values are 0 or at least 2^-63 and at most 2^41 in magnitude, which keeps every
// quotient of the force pass inside the range where quot() is exact.
__device__ __forceinline__ bool div_operand(float x) {
    const uint32_t u = __float_as_uint(x) & 0x7fffffffu;
    return u == 0u || u - 0x2b800000u <= 0x53800000u - 0x2b800000u;
}

// ------------------------------------------------------------------------------------------
// 4. stable rank + gather: the particle that arrived at slot t moves to the cell's first slot +
//    (number of particles of its cell with a smaller input index), so every cell keeps input
//    order; the record is copied there (reorder.comp:44) and the cell index kept beside it.
//    Counting costs one pass over the cell per particle, so only cells of at most
//    kGatherDirectMax particles are ranked this way.  k_place lists each more crowded cell, and
//    the first kCrowdBlocks workgroups of k_gather sort its input indices instead.
// ------------------------------------------------------------------------------------------
// Moves the particle with input index i to slot dst of the sorted buffer (reorder.comp:44).
__device__ __forceinline__ void move_record(uint32_t i, uint32_t dst, uint32_t k, const float4* __restrict__ posrho_in,
                                            const float4* __restrict__ velp_in, float4* __restrict__ posrho_out,
                                            float4* __restrict__ velp_out, uint32_t* __restrict__ sorted,
                                            uint32_t* __restrict__ skey, const uint32_t* __restrict__ id_in,
                                            uint32_t* __restrict__ id_out, uint32_t* __restrict__ div_flags) {
    sorted[dst] = i;
    skey[dst] = k;
    const float4 v = velp_in[i];
    posrho_out[dst] = posrho_in[i];
    velp_out[dst] = v;
    if (id_out) id_out[dst] = id_in[i];
    // velocities are operands of the force pass's quotients (the density pass checks the rest)
    if (!(div_operand(v.x) && div_operand(v.y) && div_operand(v.z))) atomicOr(div_flags, kDivUnsafe);
}

// Crowded cells: one workgroup per listed cell (a loop over the list with stride nblocks, whose
// length k_place left in crowd_count) sorts the cell's input indices, slot_tmp[first .. first + c),
// in ascending order: chunks of kCrowdChunk by a bitonic sort in LDS, then merge passes between
// slot_tmp and sorted over the same slots (each element's place in the merged run is its index in
// its own run plus the number of smaller elements of the other run, by binary search; indices are
// unique).  The particle with the r-th smallest index moves to slot first + r, exactly the rank
// k_gather would give it.  Every loop is bounded by the cell's size: O(c log^2 c) work per cell.
constexpr uint32_t kCrowdChunk = 4096;
constexpr unsigned kCrowdBlocks = 32;

__device__ __forceinline__ uint32_t count_below(const uint32_t* __restrict__ run, uint32_t n, uint32_t x) {
    uint32_t a = 0, b = n;
    while (a < b) {
        const uint32_t m = (a + b) >> 1;
        if (run[m] < x)
            a = m + 1;
        else
            b = m;
    }
    return a;
}

__device__ __forceinline__ void gather_crowded_cells(
    const Scenes& S, const uint2* __restrict__ crowd, const uint32_t* __restrict__ crowd_count, uint32_t crowd_cap,
    uint32_t* __restrict__ slot_tmp, const uint2* __restrict__ range, const float4* __restrict__ posrho_in,
    const float4* __restrict__ velp_in, float4* __restrict__ posrho_out, float4* __restrict__ velp_out,
    uint32_t* __restrict__ sorted, uint32_t* __restrict__ skey, const uint32_t* __restrict__ id_in,
    uint32_t* __restrict__ id_out, uint32_t* __restrict__ div_flags, uint32_t* s_v, uint32_t block,
    uint32_t nblocks) {
    const uint32_t listed = min(*crowd_count, crowd_cap);
    for (uint32_t item = block; item < listed; item += nblocks) {
        const uint2 e = crowd[item];  // (context cell, scene)
        const SceneDesc& d = S.desc[e.y];
        const uint2 rg = range[e.x];
        const uint32_t first = rg.x, lim = d.p_off + d.c.n;
        if (first < d.p_off || first >= lim) continue;  // only after a failed scan
        const uint32_t c = min(rg.y, lim - first);
        uint32_t* src = slot_tmp + first;
        uint32_t* dst = sorted + first;
        for (uint32_t lo = 0; lo < c; lo += kCrowdChunk) {
            const uint32_t m = min(c - lo, kCrowdChunk);
            uint32_t P = 1;
            while (P < m) P <<= 1;
            for (uint32_t j = threadIdx.x; j < P; j += kHashBlock) s_v[j] = j < m ? src[lo + j] : 0xFFFFFFFFu;
            __syncthreads();
            for (uint32_t k = 2; k <= P; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t a = threadIdx.x; a < P; a += kHashBlock) {
                        const uint32_t b = a ^ j;
                        if (b > a) {
                            const uint32_t x = s_v[a], y = s_v[b];
                            if ((x > y) == ((a & k) == 0)) {
                                s_v[a] = y;
                                s_v[b] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (uint32_t j = threadIdx.x; j < m; j += kHashBlock) src[lo + j] = s_v[j];
            __syncthreads();
        }
        for (uint32_t w = kCrowdChunk; w < c; w <<= 1) {
            for (uint32_t j = threadIdx.x; j < c; j += kHashBlock) {
                const uint32_t run = j / w;
                const uint32_t lo = (run & ~1u) * w;
                const uint32_t mid = min(lo + w, c), hi = min(lo + 2 * w, c);
                const uint32_t x = src[j];
                const uint32_t at = (run & 1u) == 0u ? (j - lo) + count_below(src + mid, hi - mid, x)
                                                     : (j - mid) + count_below(src + lo, mid - lo, x);
                dst[lo + at] = x;
            }
            __syncthreads();
            uint32_t* tmp = src;
            src = dst;
            dst = tmp;
        }
        const uint32_t k = e.x - d.c_off;
        for (uint32_t r = threadIdx.x; r < c; r += kHashBlock)
            move_record(src[r], first + r, k, posrho_in, velp_in, posrho_out, velp_out, sorted, skey, id_in, id_out,
                        div_flags);
        __syncthreads();  // the next cell reuses the LDS chunk
    }
}

__global__ __launch_bounds__(kHashBlock) void k_gather(Scenes S, uint32_t* __restrict__ slot_tmp,
                                                       const uint32_t* __restrict__ key,
                                                       const uint2* __restrict__ range,
                                                       const float4* __restrict__ posrho_in,
                                                       const float4* __restrict__ velp_in,
                                                       float4* __restrict__ posrho_out, float4* __restrict__ velp_out,
                                                       uint32_t* __restrict__ sorted, uint32_t* __restrict__ skey,
                                                       const uint32_t* __restrict__ id_in, uint32_t* __restrict__ id_out,
                                                       uint32_t* __restrict__ div_flags,
                                                       const uint2* __restrict__ crowd,
                                                       const uint32_t* __restrict__ crowd_count, uint32_t crowd_cap) {
    __shared__ uint32_t s_v[kCrowdChunk];
    if (blockIdx.x < kCrowdBlocks) {  // the leading workgroups take the crowded cells k_place listed
        gather_crowded_cells(S, crowd, crowd_count, crowd_cap, slot_tmp, range, posrho_in,
                             velp_in, posrho_out, velp_out, sorted, skey, id_in, id_out, div_flags, s_v, blockIdx.x,
                             kCrowdBlocks);
        return;
    }
    // kCrowdBlocks is a multiple of the XCD count, so block b still runs on XCD b mod 8
    const uint32_t b = xcd_block(blockIdx.x - kCrowdBlocks, gridDim.x - kCrowdBlocks);
    const uint32_t s = scene_of(S.scene256, S.num_scenes, b);
    const SceneDesc& d = S.desc[s];
    const uint32_t lt = (b - S.blk256[s]) * kHashBlock + threadIdx.x;
    if (lt >= d.c.n) return;
    const uint32_t t = d.p_off + lt;
    const uint32_t lim = d.p_off + d.c.n;
    const uint32_t i = slot_tmp[t];
    if (i - d.p_off >= d.c.n) return;  // only after a failed scan (error flag set)
    const uint32_t k = key[i];
    const uint2 rg = range[d.c_off + k];
    if (rg.y > kGatherDirectMax) return;  // a crowded cell: sorted by the leading workgroups
    const uint32_t end = min(rg.x + rg.y, lim);
    uint32_t r = 0;
    for (uint32_t u = rg.x; u < end; ++u) r += (slot_tmp[u] < i) ? 1u : 0u;
    const uint32_t dst = rg.x + r;
    if (dst >= lim) return;
    move_record(i, dst, k, posrho_in, velp_in, posrho_out, velp_out, sorted, skey, id_in, id_out, div_flags);
}

// ------------------------------------------------------------------------------------------
// Neighbour search shared by the density and the force pass.
//
// Each lane owns one particle.  The 27 neighbour cells are visited in NEIGHBORHOOD order with a
// wave-uniform cell loop; inside a cell each lane walks its own slot range CHUNK slots at a time
// (CHUNK loads in flight).  Candidate tests are branch-free: every candidate is written to the
// lane's next entry of an LDS list and the entry count advances by the test's outcome.  The lists
// are drained at the end, (PLANE_DRAIN) after every dx plane, and whenever one lane's list could
// overflow on its next chunk; every lane drains then, not only the full ones.  The per-neighbour
// arithmetic thus runs over compacted entries with all lanes busy instead of under a mask where
// ~15% of the lanes accept a given candidate.  Entries are drained in enumeration order, so
// every sum is formed exactly as the reference forms it, and the result does not depend on
// BLOCK, CAP, CHUNK or PLANE_DRAIN (the tuning knobs of the passes) or on when the lists drain.
// ------------------------------------------------------------------------------------------
// Pointers into LDS (32-bit addresses, so list arithmetic stays 32-bit).
template <typename T>
using lds_ptr = __attribute__((address_space(3))) T*;

template <int BLOCK, int CAP, int CHUNK, bool PLANE_DRAIN, typename Entry, typename Drain>
__device__ __forceinline__ void search_neighbours(uint32_t key, uint32_t t, uint32_t limit, float px, float py,
                                                  float pz, const float4* __restrict__ pos,
                                                  const uint2* __restrict__ range, const StepConsts& c,
                                                  lds_ptr<Entry> list, Drain& drain_fn) {
    static_assert(CAP >= CHUNK, "list must hold a chunk");
    static_assert(CHUNK <= (int)kRecordPad + 1, "chunked loads may run CHUNK - 1 records past a cell");
    // `next` is the lane's next free list entry; the list holds (next - list) / BLOCK entries
    lds_ptr<Entry> next = list;
    const lds_ptr<Entry> full = list + (CAP - CHUNK) * BLOCK;  // past this a chunk might not fit
    auto drain = [&]() {
        drain_fn((int)((next - list) / BLOCK));
        next = list;
    };
    auto visit = [&](uint32_t start, uint32_t end, auto own) {
        // the chunks of a cell run in step across the wave: every lane visiting the cell takes part
        // in each iteration (with its own slots, or none once its range is done), so when one list
        // fills, every lane drains with it instead of waiting on a drain of a few lanes
        for (uint32_t u = start; __any(u < end); u += CHUNK) {
            if (__any(next > full)) drain();
            if (u >= end) continue;
            const uint32_t left = end - u;  // slots of the cell from u on (>= 1 here): u + j is one iff j < left
            // slots past the cell are loaded (the buffers are padded) and rejected below
            const float4* q = pos + u;
            float4 r[CHUNK];
#pragma unroll
            for (int j = 0; j < CHUNK; ++j) r[j] = q[j];
#pragma unroll
            for (int j = 0; j < CHUNK; ++j) {
                const float rx = px - r[j].x, ry = py - r[j].y, rz = pz - r[j].z;
                const float d2 = (rx * rx + ry * ry) + rz * rz;
                // dist < h  <=>  !(d2 >= cut2), NaN included (sqrtf(NaN) >= h is false)
                bool ok = ((uint32_t)j < left) & !(d2 >= c.cut2);
                if constexpr (decltype(own)::value) ok &= (u + (uint32_t)j != t);
                if constexpr (std::is_same<Entry, float>::value)
                    *next = d2;  // density: squared distance
                else
                    *next = u + (uint32_t)j;  // force pass: the neighbour's slot
                next += ok ? BLOCK : 0;
                // keep `next` an address in one register (not base + offset: one add fewer per candidate)
                asm("" : "+v"(next));
            }
        }
    };
    auto end_plane = [&]() {
        if (PLANE_DRAIN) drain();
    };
    for_each_neighbour_cell(key, range, c, limit, visit, end_plane);
    drain();
}

// ------------------------------------------------------------------------------------------
// 5. density + pressure (density.comp:81-137), in place on the sorted buffer.  The LDS list
//    holds each accepted neighbour's squared distance.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float wall_density(const StepConsts& c, float x, float y, float z) {
    // density.comp:57-79; the z branch's else-test reads p.y (line 74)
    float d = 0.0f;
    if (x < c.h) d += c.m * poly6(c, x);
    else if (x > c.size_minus_h) d += c.m * poly6(c, c.size - x);
    if (y < c.h) d += c.m * poly6(c, y);
    else if (y > c.size_minus_h) d += c.m * poly6(c, c.size - y);
    if (z < c.h) d += c.m * poly6(c, z);
    else if (y > c.size_minus_h) d += c.m * poly6(c, c.size - z);
    return d * 4.0f;
}

__device__ __forceinline__ float eos(const StepConsts& c, float rho) {
    if (c.eos == WC_EOS_LINEAR) return c.p0 + c.k * (rho - c.rho0);
    const float x = rho / c.rho0;
    return c.p0 + c.k * (((x * x) * x) - 1.0f);
}

// LIST: the LDS list holds slots instead; the drain recomputes each squared distance (the same
// expression, so the same bits) and also appends the slot to the particle's neighbour list in
// HBM, loading neighbour k + 1's position while neighbour k is summed.
template <int BLOCK, int CAP, int CHUNK, bool PLANE_DRAIN, bool LIST>
__global__ __launch_bounds__(BLOCK) void k_density(Scenes S, float4* __restrict__ posrho, float4* __restrict__ velp,
                                                   const uint32_t* __restrict__ skey,
                                                   const uint2* __restrict__ range, uint32_t* __restrict__ rho_max,
                                                   uint32_t* __restrict__ nlist, uint32_t* __restrict__ ncount) {
    using Entry = typename std::conditional<LIST, uint32_t, float>::type;
    __shared__ Entry s_list[CAP * BLOCK];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAnd(&rho_max[1], ~kDivPending);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = scene_of(scene_table<BLOCK>(S), S.num_scenes, b);
    const SceneDesc& d = S.desc[s];
    // the constants by value: read through a reference, every LDS store of the neighbour loop
    // could alias them (generic pointers may point into LDS), and they were reloaded per chunk
    const StepConsts c = d.c;
    const uint32_t lt = (b - first_blocks<BLOCK>(S)[s]) * BLOCK + threadIdx.x;
    if (lt - __lane_id() >= c.n) return;  // the whole wave is past the scene's last particle
    const bool active = lt < c.n;
    float density = 0.0f;
    if (active) {
        const uint32_t t = d.p_off + lt;
        const float4 p = posrho[t];
        const lds_ptr<Entry> list = (lds_ptr<Entry>)s_list + threadIdx.x;
        float rho = c.m * poly6(c, 0.0f);
        uint32_t written = 0;  // LIST: neighbours found so far
        auto drain = [&](int cnt) {
            if constexpr (LIST) {
                if (cnt <= 0) return;
                uint32_t u = list[0];
                float4 q = posrho[u];
                for (int k = 0; k < cnt; ++k) {
                    const uint32_t un = list[(k + 1 < cnt ? k + 1 : k) * BLOCK];
                    const float4 qn = posrho[un];
                    const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                    const float d2 = (rx * rx + ry * ry) + rz * rz;
                    rho += c.m * poly6(c, sqrtf(d2));
                    if (written < kNeighbourListMax) nlist[(size_t)written * S.total_particles + t] = u;
                    ++written;
                    u = un;
                    q = qn;
                }
            } else {
                for (int k = 0; k < cnt; ++k) rho += c.m * poly6(c, sqrtf(list[k * BLOCK]));
            }
        };
        search_neighbours<BLOCK, CAP, CHUNK, PLANE_DRAIN>(skey[t], t, d.p_off + c.n, p.x, p.y, p.z, posrho,
                                                          range + d.c_off, c, list, drain);
        if constexpr (LIST) ncount[t] = written <= kNeighbourListMax ? written : kListOverflow;
        const float wall = c.wall_density ? wall_density(c, p.x, p.y, p.z) : 0.0f;
        density = rho + wall;
        const float pressure = eos(c, rho);
        reinterpret_cast<float*>(posrho)[4 * (size_t)t + 3] = density;
        reinterpret_cast<float*>(velp)[4 * (size_t)t + 3] = pressure;
        // the other operands of the force pass's quotients (the sort checks the velocities)
        const bool ok = div_operand(p.x) && div_operand(p.y) && div_operand(p.z) && div_operand(pressure) &&
                        div_operand(density) && density != 0.0f;
        if (!ok) atomicOr(&rho_max[1], kDivUnsafe);
    }
    // largest density of the pass (positive floats order as their bit patterns; NaN ranks highest);
    // lanes past the scene's last particle take part as 0
    uint32_t bits = active ? __float_as_uint(density) & 0x7fffffffu : 0u;
    for (int k = 32; k >= 1; k >>= 1) bits = max(bits, (uint32_t)__shfl_xor((int)bits, k));
    if ((threadIdx.x & 63) == 0) atomicMax(rho_max, bits);
}

// ------------------------------------------------------------------------------------------
// 6. forces + symplectic Euler + cube boundary (update.comp:134-232): sorted buffer -> buffer 1.
// ------------------------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};

template <bool FAST = false>
__device__ __forceinline__ V3 spiky(const StepConsts& c, float rx, float ry, float rz, float d) {
    const float a = c.h - d;
    const float s = a * a;
    if constexpr (FAST) {
        const float y = 1.0f / d;
        return V3{(s * quot(rx, d, y)) * c.spiky, (s * quot(ry, d, y)) * c.spiky, (s * quot(rz, d, y)) * c.spiky};
    }
    return V3{(s * (rx / d)) * c.spiky, (s * (ry / d)) * c.spiky, (s * (rz / d)) * c.spiky};
}

__device__ __forceinline__ V3 wall_forces(const StepConsts& c, float x, float y, float z) {
    // update.comp:71-100; vec3.length() is 3
    V3 f{0.0f, 0.0f, 0.0f};
    if (x < c.h) {
        const V3 w = spiky(c, 0.0f - x, y - y, z - z, 3.0f);
        f.x += w.x; f.y += w.y; f.z += w.z;
    } else if (x > c.size_minus_h) {
        const V3 w = spiky(c, c.size - x, y - y, z - z, 3.0f);
        f.x += w.x; f.y += w.y; f.z += w.z;
    }
    if (y < c.h) {
        const V3 w = spiky(c, x - x, 0.0f - y, z - z, 3.0f);
        f.x += w.x; f.y += w.y; f.z += w.z;
    } else if (y > c.size_minus_h) {
        const V3 w = spiky(c, x - x, c.size - y, z - z, 3.0f);
        f.x += w.x; f.y += w.y; f.z += w.z;
    }
    if (z < c.h) {
        const V3 w = spiky(c, x - x, y - y, 0.0f - z, 3.0f);
        f.x += w.x; f.y += w.y; f.z += w.z;
    } else if (z > c.size_minus_h) {
        const V3 w = spiky(c, x - x, y - y, c.size - z, 3.0f);
        f.x += w.x; f.y += w.y; f.z += w.z;
    }
    return V3{f.x * 0.01f, f.y * 0.01f, f.z * 0.01f};
}

__device__ __forceinline__ V3 mouse_force(const StepConsts& c, float px, float py, float pz, float pressure) {
    // update.comp:105-132
    const float o[3] = {c.cam[0], c.cam[1], c.cam[2]};
    float t1[3], t2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float tmin = (0.0f - o[a]) / c.dir[a];
        const float tmax = (c.size - o[a]) / c.dir[a];
        t1[a] = fminf(tmin, tmax);
        t2[a] = fmaxf(tmin, tmax);
    }
    const float tnear = fmaxf(fmaxf(t1[0], t1[1]), t1[2]);
    const float tfar = fminf(fminf(t2[0], t2[1]), t2[2]);
    if (tnear > tfar) return V3{0.0f, 0.0f, 0.0f};
    const float tx = px - o[0], ty = py - o[1], tz = pz - o[2];
    const float cx = c.dir[1] * tz - ty * c.dir[2];
    const float cy = c.dir[2] * tx - tz * c.dir[0];
    const float cz = c.dir[0] * ty - tx * c.dir[1];
    const float dist = sqrtf((cx * cx + cy * cy) + cz * cz);
    if (dist > c.h) return V3{0.0f, 0.0f, 0.0f};
    const V3 w = spiky(c, tx, ty, tz, dist + 1e-16f);
    const float s = (-c.m) * pressure;
    return V3{(s * w.x) * 0.00001f, (s * w.y) * 0.00001f, (s * w.z) * 0.00001f};
}

__device__ __forceinline__ void integrate_axis(const StepConsts& c, float pf, float vf, float ext, float denom,
                                               float v0, float x0, float dt, float& vel, float& pos) {
    vf = vf * c.mu;
    const float force = (pf + vf) + ext;
    const float acc = force / denom;
    const float v = v0 + acc * dt;
    vel = fminf(fmaxf(v, -c.max_speed), c.max_speed);
    pos = x0 + vel * dt;
    if (pos < c.border) {
        vel = vel * -c.damping;
        pos = c.border;
    } else if (pos > c.hi) {
        vel = vel * -c.damping;
        pos = c.hi;
    }
}

// Pressure and viscosity contributions of one accepted neighbour (update.comp:178-187).
struct ForceSums {
    float pfx, pfy, pfz, vfx, vfy, vfz;

    // Returns the pair distance.  FAST: the quotients by rho_j, by 2 rho_j and by the distance
    // come from one reciprocal each (quot); the results are the same bits as with FAST false
    // whenever every operand passes div_operand.  A zero quotient may then carry the other sign,
    // which these sums cannot show: they start at +0, and x + (+-0) = x, +0 - (+-0) = +0.
    template <bool FAST>
    __device__ __forceinline__ float add(const StepConsts& c, const float4& p, const float4& pv, const float4& q,
                                         const float4& qv) {
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float dist = sqrtf((rx * rx + ry * ry) + rz * rz);
        float iw = 0.0f, pressure;
        if constexpr (FAST) {
            iw = 1.0f / q.w;  // RN(1 / rho_j); RN(1 / (2 rho_j)) is half of it
            pressure = quot(pv.w + qv.w, 2.0f * q.w, iw * 0.5f);
        } else {
            pressure = (pv.w + qv.w) / (2.0f * q.w);
        }
        if (pressure > 0.0f) {
            const V3 sw = spiky<FAST>(c, rx, ry, rz, dist + 1e-16f);
            const float mp = c.m * pressure;
            pfx = pfx - mp * sw.x;
            pfy = pfy - mp * sw.y;
            pfz = pfz - mp * sw.z;
        }
        const float wv = (c.h - dist) * c.visc;
        if constexpr (FAST) {
            vfx = vfx + (c.m * quot(qv.x - pv.x, q.w, iw)) * wv;
            vfy = vfy + (c.m * quot(qv.y - pv.y, q.w, iw)) * wv;
            vfz = vfz + (c.m * quot(qv.z - pv.z, q.w, iw)) * wv;
        } else {
            vfx = vfx + (c.m * ((qv.x - pv.x) / q.w)) * wv;
            vfy = vfy + (c.m * ((qv.y - pv.y) / q.w)) * wv;
            vfz = vfz + (c.m * ((qv.z - pv.z) / q.w)) * wv;
        }
        return dist;
    }
};

// Integration of the learned large step: the reference's update with the network's correction
// (units of g, times corr_scale) added to the acceleration of particles at least 2h from every
// wall, and the plain acceleration optionally written out (w = 1 for those particles).
// substeps = 0: one symplectic Euler step of dt (the reference's update with the correction).
// substeps = K > 0: the outcome of K symplectic Euler steps of dt / K under a force frozen at its
// current value, v = v0 + K (dt/K) a and x = x0 + (dt/K) (K v0 + K (K+1)/2 (dt/K) a), which
// removes the K (K-1)/2 (dt/K)^2 a position error of a single large step.
__device__ __forceinline__ void learned_axis(const StepConsts& c, float pf, float vf, float ext, float denom,
                                             float v0, float x0, float dt, float substeps, float corr,
                                             float& acc_plain, float& vel, float& pos) {
    vf = vf * c.mu;
    const float force = (pf + vf) + ext;
    acc_plain = force / denom;
    const float a = acc_plain + corr;
    const float v = v0 + a * dt;
    vel = fminf(fmaxf(v, -c.max_speed), c.max_speed);
    if (substeps > 0.0f) {
        const float dts = dt / substeps;
        pos = x0 + dts * (substeps * v0 + (substeps * (substeps + 1.0f) * 0.5f) * dts * a);
    } else {
        pos = x0 + vel * dt;
    }
    if (pos < c.border) {
        vel = vel * -c.damping;
        pos = c.border;
    } else if (pos > c.hi) {
        vel = vel * -c.damping;
        pos = c.hi;
    }
}

__device__ __forceinline__ bool closure_interior(const StepConsts& c, float x, float y, float z) {
    const float margin = 2.0f * c.h;
    return x > margin && y > margin && z > margin && x < c.size - margin && y < c.size - margin &&
           z < c.size - margin;
}

__device__ __forceinline__ void learned_integrate(const StepConsts& c, const ForceSums& f, const V3& ext, float denom,
                                                  const float4& p, const float4& pv, uint32_t t, float dt,
                                                  float substeps, bool have_corr, const V3& kc, float corr_scale,
                                                  float4* __restrict__ acc_out, float4& po, float4& vo) {
    float4 k = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool interior = closure_interior(c, p.x, p.y, p.z);
    if (have_corr && interior) k = make_float4(kc.x * corr_scale, kc.y * corr_scale, kc.z * corr_scale, 0.0f);
    float4 a;
    learned_axis(c, f.pfx, f.vfx, ext.x, denom, pv.x, p.x, dt, substeps, k.x, a.x, vo.x, po.x);
    learned_axis(c, f.pfy, f.vfy, ext.y, denom, pv.y, p.y, dt, substeps, k.y, a.y, vo.y, po.y);
    learned_axis(c, f.pfz, f.vfz, ext.z, denom, pv.z, p.z, dt, substeps, k.z, a.z, vo.z, po.z);
    a.w = interior ? 1.0f : 0.0f;
    if (acc_out) acc_out[t] = a;
}

// MODE 0: the reference's force pass.  MODE 1 (learned closure): dt comes from the argument and
// corr * corr_scale is added to the acceleration of particles at least 2h from every wall; with
// acc_out the plain acceleration is also written out.  MODE 2: as MODE 1 with the correction
// averaged over the particle and its accepted neighbours, weights h - distance (h for the
// particle itself); neighbours within 2h of a wall count with a zero correction.
// LIST: the neighbours come from the lists the density pass wrote (a particle whose list
// overflowed searches as without LIST).
template <int BLOCK, int CAP, int CHUNK, bool PLANE_DRAIN, int MODE, bool LIST>
__global__ __launch_bounds__(BLOCK) void k_update(Scenes S, const float4* __restrict__ posrho_in,
                                                  const float4* __restrict__ velp_in,
                                                  const uint32_t* __restrict__ skey,
                                                  const uint2* __restrict__ range, float4* __restrict__ posrho_out,
                                                  float4* __restrict__ velp_out,
                                                  unsigned long long* __restrict__ scene_done,
                                                  const uint32_t* __restrict__ id_sorted, uint32_t* __restrict__ id_out,
                                                  float dt_arg, const float4* __restrict__ corr, float corr_scale,
                                                  float4* __restrict__ acc_out, const uint32_t* __restrict__ nlist,
                                                  const uint32_t* __restrict__ ncount, float substeps,
                                                  const uint32_t* __restrict__ div_flags, RecordBackup backup) {
    // with neighbour lists the pass keeps no LDS (a particle whose list overflowed searches
    // without one, below), so that its occupancy is bound by registers alone
    __shared__ uint32_t s_list[LIST ? 1 : CAP * BLOCK];
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = scene_of(scene_table<BLOCK>(S), S.num_scenes, b);
    const SceneDesc& d = S.desc[s];
    const StepConsts c = d.c;  // by value, as in k_density: not reloaded after each LDS store
    const uint32_t lt = (b - first_blocks<BLOCK>(S)[s]) * BLOCK + threadIdx.x;
    if (lt >= c.n) return;
    const uint32_t t = d.p_off + lt;
    const float4 p = posrho_in[t];
    const float4 pv = velp_in[t];
    if (backup.posrho) {
        backup.posrho[t] = p;
        backup.velp[t] = pv;
        if (backup.ids && id_sorted) backup.ids[t] = id_sorted[t];
    }
    const lds_ptr<uint32_t> list = (lds_ptr<uint32_t>)s_list + threadIdx.x;
    ForceSums f{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    V3 csum{0.0f, 0.0f, 0.0f};  // MODE 2: weighted sum of neighbour corrections
    float wsum = 0.0f;
    // Every neighbour's pressure and viscosity terms.  FAST (a density pass ran since the last sort
    // and found all of this step's operands in div_operand's range; a wave-uniform flag) shares
    // one reciprocal between the quotients of a neighbour; the results are the same bits either
    // way.
    auto accumulate = [&](auto fast) {
        constexpr bool FAST = decltype(fast)::value;
        auto load = [&](uint32_t u, float4& q, float4& qv) {
            q = posrho_in[u];
            qv = velp_in[u];
        };
        // MODE 2: the neighbour's correction, loaded with its records (one neighbour ahead)
        auto load_corr = [&](uint32_t u) {
            if constexpr (MODE == 2) {
                if (corr) return corr[u];
            }
            return float4{0.0f, 0.0f, 0.0f, 0.0f};
        };
        auto add_corr = [&](float dist, const float4& q, const float4& cj) {
            if constexpr (MODE == 2) {
                const float w = c.h - dist;
                if (corr && closure_interior(c, q.x, q.y, q.z)) {
                    csum.x += w * cj.x;
                    csum.y += w * cj.y;
                    csum.z += w * cj.z;
                }
                wsum += w;
            }
        };
        // neighbour k + 1's records are loaded while neighbour k's terms are formed
        auto drain = [&](int cnt) {
            if (cnt <= 0) return;
            uint32_t u = list[0];
            float4 q, qv;
            load(u, q, qv);
            float4 cj = load_corr(u);
            for (int k = 0; k < cnt; ++k) {
                const uint32_t un = list[(k + 1 < cnt ? k + 1 : k) * BLOCK];
                float4 qn, qvn;
                load(un, qn, qvn);
                const float4 cjn = load_corr(un);
                const float dist = f.template add<FAST>(c, p, pv, q, qv);
                add_corr(dist, q, cj);
                q = qn;
                qv = qvn;
                cj = cjn;
            }
        };
        bool searched = true;
        if constexpr (LIST) {
            // lists are written by the density pass; after a sort with none since, they describe
            // the previous order and the particle searches instead
            const uint32_t cnt = (div_flags[0] & kDivPending) ? kListOverflow : ncount[t];
            if (cnt != kListOverflow) {
                searched = false;
                // the same loop as the drain, over the stored list (entry k at nlist[k * total + t]):
                // entry k + 2's index and entry k + 1's records load while entry k's terms are formed,
                // so no gather waits for the index it needs
                const size_t stride = S.total_particles;
                if (cnt > 0) {
                    const uint32_t last = cnt - 1;
                    const uint32_t u0 = nlist[t];
                    uint32_t u1 = nlist[(size_t)(last < 1u ? last : 1u) * stride + t];
                    float4 q, qv;
                    load(u0, q, qv);
                    float4 cj = load_corr(u0);
                    for (uint32_t k = 0; k < cnt; ++k) {
                        const uint32_t u2 = nlist[(size_t)(k + 2 < cnt ? k + 2 : last) * stride + t];
                        float4 qn, qvn;
                        load(u1, qn, qvn);
                        const float4 cjn = load_corr(u1);
                        const float dist = f.template add<FAST>(c, p, pv, q, qv);
                        add_corr(dist, q, cj);
                        q = qn;
                        qv = qvn;
                        cj = cjn;
                        u1 = u2;
                    }
                }
            }
        }
        if (searched) {
            if constexpr (LIST) {
                // a list that overflowed (more than kNeighbourListMax neighbours), or lists older than
                // the last sort: every candidate is tested and an accepted one added at once, in
                // enumeration order, one slot at a time (rare, so no LDS list is kept for it)
                auto visit = [&](uint32_t start, uint32_t end, auto own) {
                    for (uint32_t u = start; u < end; ++u) {
                        if (decltype(own)::value && u == t) continue;
                        const float4 q = posrho_in[u];
                        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                        const float d2 = (rx * rx + ry * ry) + rz * rz;
                        if (d2 >= c.cut2) continue;  // dist >= h; a NaN distance is accepted, as in the search
                        const float4 qv = velp_in[u];
                        const float4 cj = load_corr(u);
                        const float dist = f.template add<FAST>(c, p, pv, q, qv);
                        add_corr(dist, q, cj);
                    }
                };
                auto nothing = []() {};
                for_each_neighbour_cell(skey[t], range + d.c_off, c, d.p_off + c.n, visit, nothing);
            } else {
                search_neighbours<BLOCK, CAP, CHUNK, PLANE_DRAIN>(skey[t], t, d.p_off + c.n, p.x, p.y, p.z,
                                                                  posrho_in, range + d.c_off, c, list, drain);
            }
        }
    };
    if (div_flags[0] == 0u)
        accumulate(std::true_type{});
    else
        accumulate(std::false_type{});

    V3 ext{c.g[0] * p.w, c.g[1] * p.w, c.g[2] * p.w};
    V3 mouse{0.0f, 0.0f, 0.0f}, wall{0.0f, 0.0f, 0.0f};
    if (c.mouse_force) mouse = mouse_force(c, p.x, p.y, p.z, pv.w);
    if (c.wall_force) wall = wall_forces(c, p.x, p.y, p.z);
    ext.x = ext.x + (mouse.x + wall.x);
    ext.y = ext.y + (mouse.y + wall.y);
    ext.z = ext.z + (mouse.z + wall.z);

    const float denom = p.w + 1e-16f;
    float4 po, vo;
    if (MODE == 0) {
        const float dt = d.dt;
        integrate_axis(c, f.pfx, f.vfx, ext.x, denom, pv.x, p.x, dt, vo.x, po.x);
        integrate_axis(c, f.pfy, f.vfy, ext.y, denom, pv.y, p.y, dt, vo.y, po.y);
        integrate_axis(c, f.pfz, f.vfz, ext.z, denom, pv.z, p.z, dt, vo.z, po.z);
    } else {
        V3 kc{0.0f, 0.0f, 0.0f};
        if (corr) {
            const float4 ct = corr[t];
            if (MODE == 2) {
                const float inv = 1.0f / (c.h + wsum);
                kc = V3{(c.h * ct.x + csum.x) * inv, (c.h * ct.y + csum.y) * inv, (c.h * ct.z + csum.z) * inv};
            } else {
                kc = V3{ct.x, ct.y, ct.z};
            }
        }
        learned_integrate(c, f, ext, denom, p, pv, t, dt_arg, substeps, corr != nullptr, kc, corr_scale, acc_out, po,
                          vo);
    }
    po.w = p.w;
    vo.w = pv.w;
    posrho_out[t] = po;
    velp_out[t] = vo;
    if (id_out) id_out[t] = id_sorted[t];
    // completion time of this wave's particles; the scene's latest wave ends its step
    if ((threadIdx.x & 63) == 0) atomicMax(&scene_done[s], (unsigned long long)wall_clock64());
}


// ------------------------------------------------------------------------------------------
// advect.comp:20-59 (the reference compiles it but Fluid::update never dispatches it): a
// position-only step on buffer 1, x += v dt, then the cube clamp with the shader's own border
// 0.01 and damping 0.3 (advect.comp:30-31), in place.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kHashBlock) void k_advect(Scenes S, float4* __restrict__ posrho,
                                                       float4* __restrict__ velp, float dt) {
    const uint32_t b = blockIdx.x;
    const uint32_t s = scene_of(S.scene256, S.num_scenes, b);
    const SceneDesc& d = S.desc[s];
    const uint32_t lt = (b - S.blk256[s]) * kHashBlock + threadIdx.x;
    if (lt >= d.c.n) return;
    const uint32_t i = d.p_off + lt;
    float4 x = posrho[i], v = velp[i];
    const float damping = 0.3f, border = 0.01f;
    const float hi = d.c.size - border;
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    auto clamp_axis = [&](float& xa, float& va) {
        if (xa < border) {
            va = va * -damping;
            xa = border;
        } else if (xa > hi) {
            va = va * -damping;
            xa = hi;
        }
    };
    clamp_axis(x.x, v.x);
    clamp_axis(x.y, v.y);
    clamp_axis(x.z, v.z);
    posrho[i] = x;
    velp[i] = v;
}

// ------------------------------------------------------------------------------------------
// Learned closure features (single-scene contexts).  For every requested slot of the sorted
// buffer: its 32 nearest neighbours inside the kernel radius (the density pass's acceptance
// test; nearest by squared distance, among equal distances the earlier in NEIGHBORHOOD
// enumeration order), listed in enumeration order, each as ((x_j - x_i) / h, v_j - v_i);
// then rho_i / rho0 - 1 scaled by 10, the step multiplier / 4 and the fraction of the 32 places
// filled; zero padded to 208 bf16 values.
// ------------------------------------------------------------------------------------------
constexpr int kFeatBlock = 64;
constexpr int kFeatCap = 40;    // accepted neighbours held per particle between selections
constexpr int kFeatChunk = 4;   // candidate slots loaded per step of a lane's walk

// Removes the `drop` farthest of a lane's `cnt` listed neighbours (largest squared distance;
// of equal ones the later), keeping the rest in order.  Squared distances compare as their bit
// patterns: non-negative floats order like the integers, and NaN, which the acceptance test
// lets through, ranks above every number.
__device__ __forceinline__ int drop_farthest(lds_ptr<uint32_t> key, lds_ptr<uint32_t> slot, int cnt, int drop) {
    // the lane's keys into registers (all loads in flight), the `drop` largest marked one at a
    // time (of equal keys the later entry), then one pass that closes the gaps in order -- the
    // same entries, in the same order, as dropping and shifting one at a time
    uint32_t kv[kFeatCap];
#pragma unroll
    for (int k = 0; k < kFeatCap; ++k) kv[k] = k < cnt ? key[k * kFeatBlock] : 0u;
    uint64_t gone = 0;
    for (int m = 0; m < drop; ++m) {
        uint32_t best = 0;
        int at = 0;
#pragma unroll
        for (int k = 0; k < kFeatCap; ++k) {
            const bool ge = k < cnt && !((gone >> k) & 1ull) && kv[k] >= best;
            best = ge ? kv[k] : best;
            at = ge ? k : at;
        }
        gone |= 1ull << at;
    }
    int w = 0;
#pragma unroll
    for (int k = 0; k < kFeatCap; ++k) {
        if (k < cnt && !((gone >> k) & 1ull)) {  // w <= k: a move never overwrites an unread entry
            key[w * kFeatBlock] = kv[k];
            slot[w * kFeatBlock] = slot[k * kFeatBlock];
            ++w;
        }
    }
    return w;
}

// two floats -> two bf16 (round to nearest even) in one word, lo in the low half: one
// v_cvt_pk_bf16_f32 on gfx950
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    typedef __attribute__((__vector_size__(2 * sizeof(float)))) float f32x2;
    typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
    const f32x2 v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// One feature row: the kept neighbours' relative positions (/ h) and velocities, in
// bf16, then the scaled density deviation, dt multiplier and filled fraction.  slot_of(k) gives
// the k-th kept neighbour's slot, k < cnt.
// Software-pipelined: the gathers of group g + 1 are issued before group g is formed, and the
// scheduling barrier keeps the compiler from interleaving them with waits.  Positions k >= cnt
// load the particle's own record t and write zeros.
template <typename SlotOf>
__device__ __forceinline__ void write_feature_row(uint4* __restrict__ row, const StepConsts& c, uint32_t t, float4 p,
                                                  float4 pv, int cnt, SlotOf slot_of,
                                                  const float4* __restrict__ posrho, const float4* __restrict__ velp,
                                                  float dt_mult) {
    const float inv_h = 1.0f / c.h;
    float4 q[4], qv[4];
    auto load_group = [&](int g) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = g * 4 + j;
            const uint32_t u = k < cnt ? slot_of(k) : t;
            q[j] = posrho[u];
            qv[j] = velp[u];
        }
    };
    load_group(0);
    // eight groups of four neighbours: 24 values = three 16-byte stores
#pragma unroll
    for (int g = 0; g < kClosureNeighbours / 4; ++g) {
        float4 cq[4], cqv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cq[j] = q[j];
            cqv[j] = qv[j];
        }
        if (g + 1 < kClosureNeighbours / 4) load_group(g + 1);
        __builtin_amdgcn_sched_barrier(0);
        float v[24];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // zeros past cnt by masking the bits; the mask is opaque to the optimizer, which would
            // otherwise turn a select into a branch around the loads
            uint32_t m = g * 4 + j < cnt ? ~0u : 0u;
            asm("" : "+v"(m));
            auto keep = [m](float x) { return __uint_as_float(__float_as_uint(x) & m); };
            v[6 * j + 0] = keep((cq[j].x - p.x) * inv_h);
            v[6 * j + 1] = keep((cq[j].y - p.y) * inv_h);
            v[6 * j + 2] = keep((cq[j].z - p.z) * inv_h);
            v[6 * j + 3] = keep(cqv[j].x - pv.x);
            v[6 * j + 4] = keep(cqv[j].y - pv.y);
            v[6 * j + 5] = keep(cqv[j].z - pv.z);
        }
#pragma unroll
        for (int e = 0; e < 3; ++e)
            row[g * 3 + e] = make_uint4(pack_bf16(v[8 * e + 0], v[8 * e + 1]), pack_bf16(v[8 * e + 2], v[8 * e + 3]),
                                        pack_bf16(v[8 * e + 4], v[8 * e + 5]), pack_bf16(v[8 * e + 6], v[8 * e + 7]));
    }
    const float rho_rel = (p.w / c.rho0 - 1.0f) * 10.0f;
    row[24] = make_uint4(pack_bf16(rho_rel, dt_mult * 0.25f), pack_bf16((float)cnt * (1.0f / kClosureNeighbours), 0.0f),
                         0u, 0u);
    row[25] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_clear_word(uint32_t* __restrict__ w) {
    if (threadIdx.x == 0) *w = 0u;
}

// Lists a lane selects from by its threshold network (k_closure_select) hold up to kSelRegCap
// entries; longer ones are selected by bisection or by the wave, and lists that overflowed go
// through k_closure_features on kHeavyBlocks workgroups.
constexpr int kSelRegCap = 40;
constexpr int kHeavyBlocks = 64;
static_assert(kSelRegCap <= (int)kNeighbourListMax, "the selection's loads stay inside the lists");

// Loads at a wave-uniform base plus a per-lane 32-bit byte offset, the form global loads take with
// the base in scalar registers (one VGPR per address instead of two).  Slots are below 2^28.
__device__ __forceinline__ uint32_t load_at(const uint32_t* base, uint32_t slot) {
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + (slot << 2));
}
__device__ __forceinline__ float3 load_pos_at(const float4* base, uint32_t slot) {
    const float* f = reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (slot << 4));
    return float3{f[0], f[1], f[2]};
}

// The kClosureNeighbours nearest of the m listed neighbours of slot t, kClosureNeighbours < m <=
// CAP <= kClosureNeighbours + 8, in list order, to sel_out at the particle's slot.
// - The list entries load 8 at a time, two groups ahead of the neighbours' positions, which load 8
//   at a time, one group ahead of the distances formed from them.  The key of an entry is its
//   squared distance's bits, so NaN ranks farthest; entries past m take key 0.
// - The D = m - kClosureNeighbours largest keys, each as often as it occurs, come out of an
//   insertion network of 8 (two instructions per key and level, no indexing); the D-th largest, T,
//   is the threshold.  Every key above T is dropped, and of the keys equal to T the last
//   D - #{above} in list order: what dropping the largest key, of equal ones the later entry, D
//   times drops.
// - The entries are read again, the dropped ones removed in registers (each closes its gap, the
//   last first), and the kClosureNeighbours kept ones written in list order, entry j of slot t at
//   sel_out[t * kClosureNeighbours + j]: one whole 128-byte line per row, in 16-byte stores.
template <int CAP>
__device__ __forceinline__ void select_by_threshold(uint32_t t, uint32_t m, uint32_t n,
                                                    const float4* __restrict__ posrho,
                                                    const uint32_t* __restrict__ nlist, size_t stride,
                                                    uint32_t* __restrict__ sel_out) {
    constexpr int kTop = 8, kGroup = 8, kGroups = CAP / kGroup;
    static_assert(CAP <= (int)kNeighbourListMax && CAP - kClosureNeighbours <= kTop, "one network of 8");
    static_assert(CAP % kGroup == 0 && kGroups >= 2, "whole groups");
    const float4 p = posrho[t];
    auto entries = [&](int g, uint32_t (&u)[kGroup]) {
#pragma unroll
        for (int j = 0; j < kGroup; ++j) u[j] = load_at(nlist + (size_t)(g * kGroup + j) * stride, t);
    };
    // entries past m may be stale: one out of range is replaced by the particle's own slot
    auto gather = [&](int g, const uint32_t (&u)[kGroup], float3 (&q)[kGroup]) {
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            q[j] = load_pos_at(posrho, u[j] < n ? u[j] : t);
        }
    };
    uint32_t key[CAP];
    uint32_t ua[kGroup];
    float3 cur[kGroup], nxt[kGroup];
    entries(0, ua);
    gather(0, ua, cur);
    entries(1, ua);
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        if (g + 1 < kGroups) gather(g + 1, ua, nxt);
        if (g + 2 < kGroups) entries(g + 2, ua);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            const float rx = p.x - cur[j].x, ry = p.y - cur[j].y, rz = p.z - cur[j].z;
            uint32_t kk = __float_as_uint((rx * rx + ry * ry) + rz * rz);
            kk = (uint32_t)(g * kGroup + j) < m ? kk : 0u;
            // keeps the distance here, between the group's loads: left free, the instruction
            // selector moves all of them after the last group, with every position held meanwhile
            asm volatile("" : "+v"(kk));
            key[g * kGroup + j] = kk;
        }
#pragma unroll
        for (int j = 0; j < kGroup; ++j) cur[j] = nxt[j];
    }
    // top[0] >= top[1] >= ... >= top[7]: the 8 largest keys, each as often as it occurs
    uint32_t top[kTop];
#pragma unroll
    for (int i = 0; i < kTop; ++i) top[i] = 0u;
#pragma unroll
    for (int k = 0; k < CAP; ++k) {
        uint32_t x = key[k];
#pragma unroll
        for (int i = 0; i < kTop; ++i) {
            const uint32_t hi = max(top[i], x);
            x = min(top[i], x);
            top[i] = hi;
        }
    }
    const uint32_t drop = m - (uint32_t)kClosureNeighbours;  // 1 ... CAP - kClosureNeighbours
    uint32_t thr = top[0];
#pragma unroll
    for (int i = 1; i < kTop; ++i) thr = drop == (uint32_t)(i + 1) ? top[i] : thr;
    uint32_t above = 0;  // entries past m (key 0) are never above
#pragma unroll
    for (int k = 0; k < CAP; ++k) above += key[k] > thr ? 1u : 0u;
    // Keys equal to T are dropped from the end of the list.  With T = 0 the entries past m, last
    // in the list and of key 0, come first: they take CAP - m of the drops, on top of the real ones.
    uint32_t ties = drop - above + (thr == 0u ? (uint32_t)CAP - m : 0u);
    uint64_t gone = 0;
#pragma unroll
    for (int k = CAP - 1; k >= 0; --k) {
        const bool eq = key[k] == thr && ties != 0u;
        ties -= eq ? 1u : 0u;
        gone |= (key[k] > thr || eq) ? (1ull << k) : 0ull;
    }
    // the dropped entries before m (entries past m follow every kept one and need no removal)
    uint64_t dropped = gone & ((1ull << m) - 1ull);  // m <= CAP < 64
    __builtin_amdgcn_sched_barrier(0);
    // read the entries again rather than hold them through the selection (registers); the pointer
    // is passed through an empty asm statement so the loads are not merged with the first ones
    const uint32_t* again = nlist;
    asm volatile("" : "+s"(again));
    uint32_t u[CAP];
#pragma unroll
    for (int k = 0; k < CAP; ++k) u[k] = load_at(again + (size_t)k * stride, t);
    // at most CAP - kClosureNeighbours drops; a lane with fewer shifts nothing in the later rounds
#pragma unroll
    for (int r = 0; r < CAP - kClosureNeighbours; ++r) {
        if (__ballot(dropped != 0) == 0) break;  // wave-uniform
        const uint32_t g = dropped != 0 ? 63u - (uint32_t)__clzll((long long)dropped) : 64u;
        dropped &= ~(1ull << (g & 63u));
#pragma unroll
        for (int i = 0; i + 1 < CAP; ++i) u[i] = (uint32_t)i >= g ? u[i + 1] : u[i];
    }
    uint4* row = reinterpret_cast<uint4*>(sel_out + (size_t)t * kClosureNeighbours);
#pragma unroll
    for (int j = 0; j < kClosureNeighbours / 4; ++j) row[j] = make_uint4(u[4 * j], u[4 * j + 1], u[4 * j + 2], u[4 * j + 3]);
}

// The kClosureNeighbours nearest of the m listed neighbours of slot t, kSelRegCap < m <=
// kNeighbourListMax, in list order, to sel_out at the particle's slot (as select_by_threshold).
// - Keys are formed as there (entries past m take the largest key, 0xFFFFFFFF).
// - v, the kClosureNeighbours-th smallest key with multiplicity (the least v with at least
//   kClosureNeighbours keys <= v), is found by bisection over the 32-bit key range: 32 rounds of
//   one compare and one add per key, no indexing.
// - Kept: every key below v, and of the keys equal to v the first kClosureNeighbours - #{below} in
//   list order -- the kClosureNeighbours smallest by (key, list position), which is what dropping
//   the largest key, of equal ones the later entry, m - kClosureNeighbours times leaves.  Entries
//   past m follow every real one, so they are never among them.
// - The entries are read again and written to the lane's row of LDS at its next free word, which
//   advances only past a kept entry (a dropped entry is overwritten by the next kept one, or lands
//   in the row's spare last word); the row's kClosureNeighbours words then go out as 16-byte
//   stores, one 128-byte line.
constexpr int kSelStageWords = kClosureNeighbours + 1;  // per lane: odd, so lanes' rows fall in different banks
__device__ __forceinline__ void select_by_bisection(uint32_t t, uint32_t m, uint32_t n, const float4* __restrict__ posrho,
                                                    const uint32_t* __restrict__ nlist, size_t stride,
                                                    uint32_t* __restrict__ sel_out, lds_ptr<uint32_t> stage) {
    constexpr int CAP = (int)kNeighbourListMax, kGroup = 4, kGroups = CAP / kGroup;  // 64 keys held: smaller groups
    static_assert(CAP <= 64 && CAP % kGroup == 0, "one 64-bit mask, whole groups");
    const float4 p = posrho[t];
    // the list's columns in order, through one advancing scalar pointer (the empty asm keeps the
    // compiler from holding every column's address in scalar registers)
    const uint32_t* col = nlist;
    auto entries = [&](int, uint32_t (&u)[kGroup]) {
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            u[j] = load_at(col, t);
            col += stride;
            asm volatile("" : "+s"(col));
        }
    };
    auto gather = [&](const uint32_t (&u)[kGroup], float3 (&q)[kGroup]) {
#pragma unroll
        for (int j = 0; j < kGroup; ++j) q[j] = load_pos_at(posrho, u[j] < n ? u[j] : t);
    };
    uint32_t key[CAP];
    uint32_t ua[kGroup];
    float3 cur[kGroup], nxt[kGroup];
    entries(0, ua);
    gather(ua, cur);
    entries(1, ua);
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
        if (g + 1 < kGroups) gather(ua, nxt);
        if (g + 2 < kGroups) entries(g + 2, ua);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            const float rx = p.x - cur[j].x, ry = p.y - cur[j].y, rz = p.z - cur[j].z;
            uint32_t kk = __float_as_uint((rx * rx + ry * ry) + rz * rz);
            kk = (uint32_t)(g * kGroup + j) < m ? kk : 0xFFFFFFFFu;
            asm volatile("" : "+v"(kk));  // formed here, between the group's loads
            key[g * kGroup + j] = kk;
        }
#pragma unroll
        for (int j = 0; j < kGroup; ++j) cur[j] = nxt[j];
    }
    uint32_t lo = 0u, hi = 0xFFFFFFFFu;
#pragma unroll 1
    for (int round = 0; round < 32; ++round) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < CAP; ++k) c += key[k] <= mid ? 1u : 0u;
        const bool enough = c >= (uint32_t)kClosureNeighbours;
        hi = enough ? mid : hi;
        lo = enough ? lo : mid + 1u;
    }
    const uint32_t v = lo;
    uint32_t below = 0;
#pragma unroll
    for (int k = 0; k < CAP; ++k) below += key[k] < v ? 1u : 0u;
    uint32_t quota = (uint32_t)kClosureNeighbours - below;
    uint64_t keep = 0;
#pragma unroll
    for (int k = 0; k < CAP; ++k) {
        const bool eq = key[k] == v && quota != 0u;
        quota -= eq ? 1u : 0u;
        keep |= (key[k] < v || eq) ? (1ull << k) : 0ull;
    }
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t* again = nlist;  // read again, not merged with the first loads (registers)
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < CAP; ++k) {
        asm volatile("" : "+s"(again));
        stage[w] = load_at(again, t);
        again += stride;
        w += (uint32_t)(keep >> k) & 1u;
    }
    uint4* row = reinterpret_cast<uint4*>(sel_out + (size_t)t * kClosureNeighbours);
#pragma unroll
    for (int j = 0; j < kClosureNeighbours / 4; ++j)
        row[j] = make_uint4(stage[4 * j], stage[4 * j + 1], stage[4 * j + 2], stage[4 * j + 3]);
}

// One row of kSelRegCap + 1 ... kNeighbourListMax entries selected by the whole wave: slot t
// (uniform), its m listed neighbours (uniform), the lane's entry u (entry `lane`; the row's own
// slot past m) and that entry's position q.  Lane k ranks entry k among the row's m entries by
// (squared distance bits, list position), reading the others' keys lane by lane; the
// kClosureNeighbours lowest ranks are kept, in list order (the same entries as
// select_by_bisection keeps).  For waves with a few such rows: one row costs the wave about as
// much as the bisection costs each lane, but only the rows there are.
constexpr int kSelWaveRowsMax = 8;
struct WaveRow {
    uint32_t t, m, u;
    float3 q;
};
__device__ __forceinline__ void select_row_by_wave(const WaveRow& w, const float4* __restrict__ posrho,
                                                   uint32_t* __restrict__ sel_out) {
    static_assert(kNeighbourListMax == 64, "one entry per lane");
    const uint32_t lane = threadIdx.x & 63;
    const float4 p = posrho[w.t];
    const float rx = p.x - w.q.x, ry = p.y - w.q.y, rz = p.z - w.q.z;
    const uint32_t key = __float_as_uint((rx * rx + ry * ry) + rz * rz);
    uint32_t rank = 0;
    for (uint32_t j = 0; j < w.m; ++j) {
        const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)j);
        rank += (kj < key || (kj == key && j < lane)) ? 1u : 0u;
    }
    const bool keep = lane < w.m && rank < (uint32_t)kClosureNeighbours;
    const uint64_t km = __ballot(keep);
    if (keep) sel_out[(size_t)w.t * kClosureNeighbours + (uint32_t)__popcll(km & ((1ull << lane) - 1ull))] = w.u;
}

// Selection of the kept neighbours of every slot whose list is longer than kClosureNeighbours:
// lists of up to kSelRegCap entries by each lane's threshold network; longer ones up to
// kNeighbourListMax by each lane's bisection in a wave that has more than kSelWaveRowsMax of them,
// else by the whole wave one row at a time; overflowed lists are appended to `heavy` (one atomic
// per wave that has any) for k_closure_features' walk.  Rows of at most kClosureNeighbours need no
// selection.
__global__ __launch_bounds__(256) void k_closure_select(const float4* __restrict__ posrho, uint32_t n,
                                                        const uint32_t* __restrict__ nlist,
                                                        const uint32_t* __restrict__ ncount, size_t stride,
                                                        uint32_t* __restrict__ sel_out, uint32_t* __restrict__ heavy,
                                                        uint32_t* __restrict__ heavy_count) {
    __shared__ uint32_t s_stage[256 * kSelStageWords];
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const uint32_t m = t < n ? ncount[t] : 0u;
    const bool is_heavy = m > kNeighbourListMax;  // kListOverflow
    const uint64_t hm = __ballot(is_heavy);
    if (hm != 0) {  // wave-uniform
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)hm) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(heavy_count, (uint32_t)__popcll(hm));
        base = __shfl(base, (int)leader);
        if (is_heavy) heavy[base + (uint32_t)__popcll(hm & ((1ull << lane) - 1ull))] = t;
    }
    if (m > (uint32_t)kClosureNeighbours && m <= (uint32_t)kSelRegCap)
        select_by_threshold<kSelRegCap>(t, m, n, posrho, nlist, stride, sel_out);
    const bool longer = m > (uint32_t)kSelRegCap && !is_heavy;
    uint64_t rows = __ballot(longer);
    if (__popcll(rows) > kSelWaveRowsMax) {  // wave-uniform
        if (longer)
            select_by_bisection(t, m, n, posrho, nlist, stride, sel_out,
                                (lds_ptr<uint32_t>)s_stage + threadIdx.x * kSelStageWords);
        return;
    }
    // A few longer rows: every lane of the wave takes part in each, one after another, while the
    // next row's positions and the one after's list entries are in flight.
    const uint32_t lane = threadIdx.x & 63;
    auto next_row = [&](WaveRow& w) {  // wave-uniform; issues the row's list loads
        if (rows == 0) return false;
        const int l = __ffsll((unsigned long long)rows) - 1;
        rows &= rows - 1;
        w.t = (uint32_t)__builtin_amdgcn_readlane((int)t, l);
        w.m = (uint32_t)__builtin_amdgcn_readlane((int)m, l);
        w.u = lane < w.m ? nlist[(size_t)lane * stride + w.t] : w.t;
        return true;
    };
    auto positions = [&](WaveRow& w) {
        const float* f = reinterpret_cast<const float*>(posrho + w.u);
        w.q = float3{f[0], f[1], f[2]};
    };
    WaveRow a{}, b{}, c{};
    bool have_a = next_row(a), have_b = false;
    if (have_a) {
        positions(a);
        have_b = next_row(b);
    }
    while (have_a) {  // wave-uniform
        bool have_c = false;
        if (have_b) {
            positions(b);
            have_c = next_row(c);
        }
        select_row_by_wave(a, posrho, sel_out);
        a = b;
        b = c;
        have_a = have_b;
        have_b = have_c;
    }
}

// Rows whose particle has at most kClosureNeighbours listed neighbours keep all of them, in list
// order: written straight from the density pass's lists, with no distances and no LDS.  Other
// rows are left to k_closure_features.
__global__ __launch_bounds__(256) void k_closure_features_short(Scenes S, const float4* __restrict__ posrho,
                                                                const float4* __restrict__ velp,
                                                                const uint32_t* __restrict__ rows, uint32_t nrows,
                                                                float dt_mult, uint16_t* __restrict__ X,
                                                                const uint32_t* __restrict__ nlist,
                                                                const uint32_t* __restrict__ ncount) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nrows) return;
    const SceneDesc& d = S.desc[0];
    const StepConsts& c = d.c;
    const uint32_t t = rows ? rows[r] : d.p_off + r;
    uint4* row = reinterpret_cast<uint4*>(X + (size_t)r * kClosureFeatures);
    if (t - d.p_off >= c.n) {
        for (int q = 0; q < kClosureFeatures / 8; ++q) row[q] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const uint32_t listed = ncount[t];
    if (listed > (uint32_t)kClosureNeighbours) return;  // more than are kept, or overflowed
    // all listed slots first (coalesced across the wave, in flight together)
    const size_t stride = S.total_particles;
    uint32_t u[kClosureNeighbours];
#pragma unroll
    for (int k = 0; k < kClosureNeighbours; ++k) u[k] = (uint32_t)k < listed ? nlist[(size_t)k * stride + t] : t;
    write_feature_row(row, c, t, posrho[t], velp[t], (int)listed, [&](int k) { return u[k]; }, posrho, velp, dt_mult);
}

// One row of k_closure_features.
__device__ __forceinline__ void closure_row(const Scenes& S, const float4* __restrict__ posrho,
                                            const float4* __restrict__ velp, const uint32_t* __restrict__ skey,
                                            const uint2* __restrict__ range, const uint32_t* __restrict__ rows,
                                            uint32_t r, float dt_mult, uint16_t* __restrict__ X,
                                            const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                                            bool skip_short, uint32_t* __restrict__ sel_out, lds_ptr<uint32_t> key,
                                            lds_ptr<uint32_t> slot) {
    const SceneDesc& d = S.desc[0];
    const StepConsts& c = d.c;
    const uint32_t t = rows ? rows[r] : d.p_off + r;
    uint4* row = X ? reinterpret_cast<uint4*>(X + (size_t)r * kClosureFeatures) : nullptr;  // null: slots to sel_out
    if (t - d.p_off >= c.n) {
        if (row)
            for (int q = 0; q < kClosureFeatures / 8; ++q) row[q] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    // rows of short lists are written by k_closure_features_short
    if (skip_short && ncount[t] <= (uint32_t)kClosureNeighbours) return;
    const float4 p = posrho[t];
    const float4 pv = velp[t];
    int cnt = 0;
    auto visit = [&](uint32_t start, uint32_t end, auto own) {
        for (uint32_t u = start; u < end; u += kFeatChunk) {
            if (cnt > kFeatCap - kFeatChunk) cnt = drop_farthest(key, slot, cnt, cnt - kClosureNeighbours);
            const uint32_t left = end - u;
            const float4* q = posrho + u;  // the buffers are padded past the last particle
            float4 v[kFeatChunk];
#pragma unroll
            for (int j = 0; j < kFeatChunk; ++j) v[j] = q[j];
#pragma unroll
            for (int j = 0; j < kFeatChunk; ++j) {
                const float rx = p.x - v[j].x, ry = p.y - v[j].y, rz = p.z - v[j].z;
                const float d2 = (rx * rx + ry * ry) + rz * rz;
                bool ok = ((uint32_t)j < left) & !(d2 >= c.cut2);
                if constexpr (decltype(own)::value) ok &= (u + (uint32_t)j != t);
                key[cnt * kFeatBlock] = __float_as_uint(d2);
                slot[cnt * kFeatBlock] = u + (uint32_t)j;
                cnt += ok ? 1 : 0;
            }
        }
    };
    auto nothing = []() {};
    // With the density pass's neighbour lists the accepted neighbours are read in the same order
    // instead of searched for (a list that overflowed is searched for as without lists).
    const uint32_t listed = ncount ? ncount[t] : kListOverflow;
    if (listed != kListOverflow) {
        const size_t stride = S.total_particles;
        for (uint32_t k = 0; k < listed; k += kFeatChunk) {
            if (cnt > kFeatCap - kFeatChunk) cnt = drop_farthest(key, slot, cnt, cnt - kClosureNeighbours);
            uint32_t u[kFeatChunk];
            float4 v[kFeatChunk];
#pragma unroll
            for (int j = 0; j < kFeatChunk; ++j) u[j] = nlist[(size_t)min(k + (uint32_t)j, listed - 1) * stride + t];
#pragma unroll
            for (int j = 0; j < kFeatChunk; ++j) v[j] = posrho[u[j]];
#pragma unroll
            for (int j = 0; j < kFeatChunk; ++j) {
                const float rx = p.x - v[j].x, ry = p.y - v[j].y, rz = p.z - v[j].z;
                const float d2 = (rx * rx + ry * ry) + rz * rz;
                key[cnt * kFeatBlock] = __float_as_uint(d2);
                slot[cnt * kFeatBlock] = u[j];
                cnt += (k + (uint32_t)j < listed) ? 1 : 0;
            }
        }
    } else {
        for_each_neighbour_cell(skey[t], range + d.c_off, c, d.p_off + c.n, visit, nothing);
    }
    if (cnt > kClosureNeighbours) cnt = drop_farthest(key, slot, cnt, cnt - kClosureNeighbours);

    if (sel_out) {
        // the kept neighbours' slots instead of the row, entry k of slot t at
        // sel_out[t * kClosureNeighbours + k] as the selection writes them, for inference that
        // forms the values itself
        for (int k = 0; k < cnt; ++k) sel_out[(size_t)(t - d.p_off) * kClosureNeighbours + k] = slot[k * kFeatBlock];
        return;
    }
    write_feature_row(row, c, t, p, pv, cnt, [&](int k) { return slot[k * kFeatBlock]; }, posrho, velp, dt_mult);
}

__global__ __launch_bounds__(kFeatBlock) void k_closure_features(Scenes S, const float4* __restrict__ posrho,
                                                                 const float4* __restrict__ velp,
                                                                 const uint32_t* __restrict__ skey,
                                                                 const uint2* __restrict__ range,
                                                                 const uint32_t* __restrict__ rows, uint32_t nrows,
                                                                 float dt_mult, uint16_t* __restrict__ X,
                                                                 const uint32_t* __restrict__ nlist,
                                                                 const uint32_t* __restrict__ ncount,
                                                                 bool skip_short, uint32_t* __restrict__ sel_out,
                                                                 const uint32_t* __restrict__ nrows_dev) {
    static_assert(kFeatChunk <= (int)kRecordPad + 1, "chunked loads may run kFeatChunk - 1 records past a cell");
    static_assert(kFeatCap >= kClosureNeighbours + kFeatChunk, "a chunk must fit after a selection");
    __shared__ uint32_t s_key[kFeatCap * kFeatBlock];
    __shared__ uint32_t s_slot[kFeatCap * kFeatBlock];
    if (nrows_dev) nrows = *nrows_dev;  // a count an earlier kernel wrote
    // rows r, r + the grid, ...: one row per lane when the grid covers them all
    for (uint32_t r = blockIdx.x * kFeatBlock + threadIdx.x; r < nrows; r += gridDim.x * kFeatBlock)
        closure_row(S, posrho, velp, skey, range, rows, r, dt_mult, X, nlist, ncount, skip_short, sel_out,
                    (lds_ptr<uint32_t>)s_key + threadIdx.x, (lds_ptr<uint32_t>)s_slot + threadIdx.x);
}

// ------------------------------------------------------------------------------------------
// After the host rewrites the particle buffers or the parameters: the cell key of every slot of
// buffer 2 is recomputed from its position with the current constants (the neighbour passes
// read it where the reference recomputes clamp(ivec3(p / binSize)), density.comp:88,
// update.comp:141), the neighbour lists are marked overflowed (so every reader searches), and
// the force pass's division operands are marked unchecked.  The cell ranges stay those of the
// last sort, as the reference's counts and offsets do.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kHashBlock) void k_invalidate(Scenes S, const float4* __restrict__ posrho,
                                                           uint32_t* __restrict__ skey, uint32_t* __restrict__ ncount,
                                                           uint32_t* __restrict__ div_flags) {
    const uint32_t b = blockIdx.x;
    if (b == 0 && threadIdx.x == 0) atomicOr(div_flags, kDivPending | kDivUnsafe);
    if (b >= S.blocks256) return;
    const uint32_t s = scene_of(S.scene256, S.num_scenes, b);
    const SceneDesc& d = S.desc[s];
    const uint32_t lt = (b - S.blk256[s]) * kHashBlock + threadIdx.x;
    if (lt >= d.c.n) return;
    const uint32_t t = d.p_off + lt;
    const float4 p = posrho[t];
    skey[t] = cell_key(d.c, p.x, p.y, p.z);
    if (ncount) ncount[t] = kListOverflow;
}

// ------------------------------------------------------------------------------------------
// Allocation-free step: the scene's constants and block tables go into the library's fixed
// scratch array (from kernel arguments, so nothing is copied from host memory), and what a step
// expects to start at zero is cleared: the cell counts (a step whose scan failed leaves some
// behind) and the failure word.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kHashBlock) void k_noalloc_setup(SceneDesc d, uint4 t0, uint4 t1,
                                                              SceneDesc* __restrict__ desc,
                                                              uint32_t* __restrict__ tables,
                                                              uint32_t* __restrict__ counts, uint32_t ncells,
                                                              uint32_t* __restrict__ error_flag) {
    const uint32_t g = blockIdx.x * kHashBlock + threadIdx.x;
    if (g == 0) {
        *desc = d;
        reinterpret_cast<uint4*>(tables)[0] = t0;
        reinterpret_cast<uint4*>(tables)[1] = t1;
        *error_flag = 0u;
    }
    for (uint32_t i = g; i < ncells; i += gridDim.x * kHashBlock) counts[i] = 0u;
}

// ------------------------------------------------------------------------------------------
// Record <-> split-half conversion.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kHashBlock) void k_split(const float4* __restrict__ rec, float4* __restrict__ posrho,
                                                      float4* __restrict__ velp, uint32_t n) {
    const uint32_t i = blockIdx.x * kHashBlock + threadIdx.x;
    if (i >= n) return;
    posrho[i] = rec[2 * (size_t)i];
    velp[i] = rec[2 * (size_t)i + 1];
}

__global__ __launch_bounds__(kHashBlock) void k_join(const float4* __restrict__ posrho,
                                                     const float4* __restrict__ velp, float4* __restrict__ rec,
                                                     uint32_t n) {
    const uint32_t i = blockIdx.x * kHashBlock + threadIdx.x;
    if (i >= n) return;
    rec[2 * (size_t)i] = posrho[i];
    rec[2 * (size_t)i + 1] = velp[i];
}

inline unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

hipError_t launch_hash(const DeviceBuffers& b, hipStream_t s) {
    const Scenes& S = b.sc;
    size_t blocks = S.blocks256;
    const size_t reset = blocks_for(S.total_tiles > S.num_scenes ? S.total_tiles : S.num_scenes, kHashBlock);
    if (reset > blocks) blocks = reset;
    if (blocks == 0) blocks = 1;  // still resets the step clock
    hipLaunchKernelGGL(k_hash, dim3((unsigned)blocks), dim3(kHashBlock), 0, s, S, b.posrho[0], b.key, b.rank,
                       b.counts, b.scan_status, b.scene_done, b.step_start, b.rho_max, b.crowd_count);
    return hipGetLastError();
}

hipError_t launch_scan(const DeviceBuffers& b, hipStream_t s) {
    if (b.sc.total_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan, dim3(b.sc.total_tiles), dim3(kScanThreads), 0, s, b.sc, b.counts, b.range,
                       b.scan_status, b.error_flag);
    return hipGetLastError();
}

hipError_t launch_place(const DeviceBuffers& b, hipStream_t s) {
    if (b.sc.blocks256 == 0) return hipSuccess;
    hipLaunchKernelGGL(k_place, dim3(b.sc.blocks256), dim3(kHashBlock), 0, s, b.sc, b.key, b.rank, b.range,
                       b.slot_tmp, b.crowd, b.crowd_count, b.crowd_cap, b.error_flag);
    return hipGetLastError();
}

hipError_t launch_gather(const DeviceBuffers& b, hipStream_t s) {
    if (b.sc.blocks256 == 0) return hipSuccess;
    static_assert(kCrowdBlocks % kNumXcd == 0, "the leading workgroups keep the XCD mapping of the others");
    hipLaunchKernelGGL(k_gather, dim3(b.sc.blocks256 + kCrowdBlocks), dim3(kHashBlock), 0, s, b.sc, b.slot_tmp, b.key,
                       b.range, b.posrho[0], b.velp[0], b.posrho[1], b.velp[1], b.sorted, b.skey, b.ids[0],
                       b.ids[1], b.rho_max + 1, b.crowd, b.crowd_count, b.crowd_cap);
    return hipGetLastError();
}

// Tuning variants of the two neighbour passes: (block size, LDS list entries per lane, slots
// in flight per lane, drain after every dx plane).  Every variant computes identical bits.
#define WC_VARIANTS(X)     \
    X(256, 32, 4, true)    \
    X(128, 32, 4, true)    \
    X(64, 32, 4, true)     \
    X(64, 32, 8, true)     \
    X(64, 40, 4, false)    \
    X(64, 48, 4, false)    \
    X(64, 64, 4, false)    \
    X(128, 48, 4, false)   \
    X(128, 32, 2, true)    \
    X(64, 40, 8, false)    \
    X(128, 32, 8, true)

static const PassVariant kVariants[] = {
#define WC_ROW(B, C, K, P) {B, C, K, P},
    WC_VARIANTS(WC_ROW)
#undef WC_ROW
};

int num_variants() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }
PassVariant variant(int v) { return kVariants[(v >= 0 && v < num_variants()) ? v : 0]; }

static unsigned pass_blocks(const Scenes& S, int block) {
    return block == 64 ? S.blocks64 : (block == 128 ? S.blocks128 : S.blocks256);
}

hipError_t launch_density(const DeviceBuffers& b, int v, bool list, hipStream_t s) {
    const PassVariant pv = variant(v);
    const unsigned blocks = pass_blocks(b.sc, pv.block);
    if (blocks == 0) return hipSuccess;
    if (list && (!b.nlist || !b.ncount)) return hipErrorInvalidValue;
#define WC_LAUNCH(B, C, K, P)                                                                             \
    if (pv.block == B && pv.list_cap == C && pv.chunk == K && pv.plane_drain == P) {                     \
        if (list)                                                                                         \
            hipLaunchKernelGGL((k_density<B, C, K, P, true>), dim3(blocks), dim3(B), 0, s, b.sc, b.posrho[1], \
                               b.velp[1], b.skey, b.range, b.rho_max, b.nlist, b.ncount);                  \
        else                                                                                              \
            hipLaunchKernelGGL((k_density<B, C, K, P, false>), dim3(blocks), dim3(B), 0, s, b.sc, b.posrho[1],\
                               b.velp[1], b.skey, b.range, b.rho_max, nullptr, nullptr);                   \
        return hipGetLastError();                                                                         \
    }
    WC_VARIANTS(WC_LAUNCH)
#undef WC_LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_update(const DeviceBuffers& b, int v, bool list, hipStream_t s) {
    const PassVariant pv = variant(v);
    const unsigned blocks = pass_blocks(b.sc, pv.block);
    if (blocks == 0) return hipSuccess;
    if (list && (!b.nlist || !b.ncount)) return hipErrorInvalidValue;
#define WC_LAUNCH(B, C, K, P)                                                                             \
    if (pv.block == B && pv.list_cap == C && pv.chunk == K && pv.plane_drain == P) {                     \
        if (list)                                                                                         \
            hipLaunchKernelGGL((k_update<B, C, K, P, 0, true>), dim3(blocks), dim3(B), 0, s, b.sc,         \
                               b.posrho[1], b.velp[1], b.skey, b.range, b.posrho[0], b.velp[0],            \
                               b.scene_done, b.ids[1], b.ids[0], 0.0f, nullptr, 0.0f, nullptr, b.nlist,    \
                               b.ncount, 0.0f, b.rho_max + 1, RecordBackup{});                                  \
        else                                                                                              \
            hipLaunchKernelGGL((k_update<B, C, K, P, 0, false>), dim3(blocks), dim3(B), 0, s, b.sc,        \
                               b.posrho[1], b.velp[1], b.skey, b.range, b.posrho[0], b.velp[0],            \
                               b.scene_done, b.ids[1], b.ids[0], 0.0f, nullptr, 0.0f, nullptr, nullptr,    \
                               nullptr, 0.0f, b.rho_max + 1, RecordBackup{});                                   \
        return hipGetLastError();                                                                         \
    }
    WC_VARIANTS(WC_LAUNCH)
#undef WC_LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_closure_features(const DeviceBuffers& b, const uint32_t* rows, uint32_t nrows, float dt_mult,
                                   uint16_t* X, bool lists, hipStream_t s) {
    if (nrows == 0) return hipSuccess;
    if (lists && (!b.nlist || !b.ncount)) return hipErrorInvalidValue;
    if (lists) {
        hipLaunchKernelGGL(k_closure_features_short, dim3(blocks_for(nrows, 256)), dim3(256), 0, s, b.sc, b.posrho[1],
                           b.velp[1], rows, nrows, dt_mult, X, b.nlist, b.ncount);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_closure_features, dim3(blocks_for(nrows, kFeatBlock)), dim3(kFeatBlock), 0, s, b.sc,
                       b.posrho[1], b.velp[1], b.skey, b.range, rows, nrows, dt_mult, X,
                       lists ? b.nlist : nullptr, lists ? b.ncount : nullptr, lists, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_closure_selection(const DeviceBuffers& b, uint32_t n, uint32_t* sel, uint32_t* heavy,
                                    uint32_t* heavy_count, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!b.nlist || !b.ncount || !sel || !heavy || !heavy_count) return hipErrorInvalidValue;
    // a kernel clears the count (no memset node in a captured graph); the register selection runs
    // on every slot and lists the rare longer rows, which the walk then takes on a small grid
    hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(64), 0, s, heavy_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_closure_select, dim3(blocks_for(n, 256)), dim3(256), 0, s, b.posrho[1], n, b.nlist, b.ncount,
                       (size_t)b.sc.total_particles, sel, heavy, heavy_count);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_closure_features, dim3(kHeavyBlocks), dim3(kFeatBlock), 0, s, b.sc, b.posrho[1], b.velp[1],
                       b.skey, b.range, heavy, n, 0.0f, nullptr, b.nlist, b.ncount, true, sel, heavy_count);
    return hipGetLastError();
}

hipError_t launch_closure_update(const DeviceBuffers& b, float dt, const float* corr, float corr_scale,
                                 float* acc_out, bool smooth, float substeps, bool lists, hipStream_t s,
                                 const RecordBackup& backup) {
    if (b.sc.blocks64 == 0) return hipSuccess;
    if (lists && (!b.nlist || !b.ncount)) return hipErrorInvalidValue;
    if (lists) {
        if (smooth)
            hipLaunchKernelGGL((k_update<64, 32, 4, true, 2, true>), dim3(b.sc.blocks64), dim3(64), 0, s, b.sc,
                               b.posrho[1], b.velp[1], b.skey, b.range, b.posrho[0], b.velp[0], b.scene_done,
                               b.ids[1], b.ids[0], dt, reinterpret_cast<const float4*>(corr), corr_scale,
                               reinterpret_cast<float4*>(acc_out), b.nlist, b.ncount, substeps, b.rho_max + 1, backup);
        else
            hipLaunchKernelGGL((k_update<64, 32, 4, true, 1, true>), dim3(b.sc.blocks64), dim3(64), 0, s, b.sc,
                               b.posrho[1], b.velp[1], b.skey, b.range, b.posrho[0], b.velp[0], b.scene_done,
                               b.ids[1], b.ids[0], dt, reinterpret_cast<const float4*>(corr), corr_scale,
                               reinterpret_cast<float4*>(acc_out), b.nlist, b.ncount, substeps, b.rho_max + 1, backup);
        return hipGetLastError();
    }
    if (smooth)
        hipLaunchKernelGGL((k_update<64, 32, 4, true, 2, false>), dim3(b.sc.blocks64), dim3(64), 0, s, b.sc,
                           b.posrho[1], b.velp[1], b.skey, b.range, b.posrho[0], b.velp[0], b.scene_done, b.ids[1],
                           b.ids[0], dt, reinterpret_cast<const float4*>(corr), corr_scale,
                           reinterpret_cast<float4*>(acc_out), nullptr, nullptr, substeps, b.rho_max + 1, backup);
    else
        hipLaunchKernelGGL((k_update<64, 32, 4, true, 1, false>), dim3(b.sc.blocks64), dim3(64), 0, s, b.sc,
                           b.posrho[1], b.velp[1], b.skey, b.range, b.posrho[0], b.velp[0], b.scene_done, b.ids[1],
                           b.ids[0], dt, reinterpret_cast<const float4*>(corr), corr_scale,
                           reinterpret_cast<float4*>(acc_out), nullptr, nullptr, substeps, b.rho_max + 1, backup);
    return hipGetLastError();
}

hipError_t launch_noalloc_setup(const SceneDesc& d, const uint32_t tables_host[8], SceneDesc* desc,
                                uint32_t* tables, uint32_t* counts, uint32_t ncells, uint32_t* error_flag,
                                hipStream_t s) {
    const uint4 t0 = make_uint4(tables_host[0], tables_host[1], tables_host[2], tables_host[3]);
    const uint4 t1 = make_uint4(tables_host[4], tables_host[5], tables_host[6], tables_host[7]);
    unsigned blocks = blocks_for(ncells, kHashBlock);
    blocks = blocks < 1 ? 1 : (blocks > 256 ? 256 : blocks);
    hipLaunchKernelGGL(k_noalloc_setup, dim3(blocks), dim3(kHashBlock), 0, s, d, t0, t1, desc, tables, counts, ncells,
                       error_flag);
    return hipGetLastError();
}

hipError_t launch_invalidate(const DeviceBuffers& b, hipStream_t s) {
    const unsigned blocks = b.sc.blocks256 ? b.sc.blocks256 : 1;  // block 0 always marks the flags
    hipLaunchKernelGGL(k_invalidate, dim3(blocks), dim3(kHashBlock), 0, s, b.sc, b.posrho[1], b.skey, b.ncount,
                       b.rho_max + 1);
    return hipGetLastError();
}

hipError_t launch_advect(const DeviceBuffers& b, float dt, hipStream_t s) {
    if (b.sc.blocks256 == 0) return hipSuccess;
    hipLaunchKernelGGL(k_advect, dim3(b.sc.blocks256), dim3(kHashBlock), 0, s, b.sc, b.posrho[0], b.velp[0], dt);
    return hipGetLastError();
}

hipError_t launch_split_records(const wc_particle* rec, float4* posrho, float4* velp, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_split, dim3(blocks_for(n, kHashBlock)), dim3(kHashBlock), 0, s,
                       reinterpret_cast<const float4*>(rec), posrho, velp, n);
    return hipGetLastError();
}

hipError_t launch_join_records(const float4* posrho, const float4* velp, wc_particle* rec, uint32_t n,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_join, dim3(blocks_for(n, kHashBlock)), dim3(kHashBlock), 0, s, posrho, velp,
                       reinterpret_cast<float4*>(rec), n);
    return hipGetLastError();
}

}  // namespace wc
