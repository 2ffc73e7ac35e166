// wc_mlp.hip -- the learned closure's network on gfx950 matrix cores.
//
// MLP 195 -> 128 -> 128 -> 128 -> 3 with SiLU, bf16 weights and activations, fp32 accumulation,
// all products on v_mfma_f32_32x32x16_bf16.  The input is padded to 208 features and the output
// to 32 columns (zero weights), so every product is a whole number of 32x32x16 MFMA tiles.
//
// Training (forward with stored activations, backward, Adam) is a sequence of one generic MFMA
// GEMM kernel with strided operands and a few elementwise kernels.  Inference is one fused
// kernel: a workgroup of four waves takes 128 particles, each wave 32 of them through all four
// layers, activations staying in LDS between layers.
//
// MFMA 32x32x16 bf16 lane layout (wave64, lane l): A operand = A[l % 32][8 (l / 32) + 0..7],
// B operand = B[8 (l / 32) + 0..7][l % 32], accumulator element r = C[8 (r / 4) + 4 (l / 32) +
// r % 4][l % 32].
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "wc_mlp.hpp"

namespace wc {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// float -> bf16, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// SiLU with the hardware reciprocal (about 1 ulp): the network computes in bf16, so the correctly
// rounded division the rest of the library is built with would only cost time here.
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ f32x16 mfma(uint4 a, uint4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

// ------------------------------------------------------------------------------------------
// Generic GEMM: C[m][n] (+)= sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak], B(k,n) =
// B[k*sbk + n*sbn] (bf16), C fp32 row-major with leading dimension ldc.  64x64 tile per
// workgroup (2x2 waves of 32x32 on v_mfma_f32_32x32x16_bf16); K in stages of 64 staged through
// LDS as [row][k] (k contiguous, rows padded by 8), the next stage's global loads in flight while
// the current one is multiplied; split over gridDim.z with atomic adds.
//
// Each operand tile (64 rows x 64 k) loads as 512 runs of 8 values along the operand's contiguous
// dimension -- k, or the row (m for A, n for B) when ROW_CONTIG -- two runs per thread, as one
// 16-byte load where the run lies inside the matrix and the addresses are aligned (`vec`), value
// by value (zeros outside) elsewhere.  Row-contiguous runs are transposed as they go into LDS.
// ------------------------------------------------------------------------------------------
constexpr int kGemmTile = 64;
constexpr int kGemmKStage = 64;
constexpr int kGemmLdk = kGemmKStage + 8;

template <bool ROW_CONTIG>
__device__ __forceinline__ void gemm_load(uint4 (&v)[2], const uint16_t* __restrict__ X, long s_row, long s_k, int row0,
                                          int nrows, int k0, int kend, bool vec, int tid) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int q = tid * 2 + i;
        const int a = q >> 3, b = (q & 7) * 8;
        const int gr = row0 + (ROW_CONTIG ? b : a), gk = k0 + (ROW_CONTIG ? a : b);
        const bool full = ROW_CONTIG ? (gr + 8 <= nrows && gk < kend) : (gr < nrows && gk + 8 <= kend);
        if (vec && full) {
            v[i] = *reinterpret_cast<const uint4*>(&X[(long)gr * s_row + (long)gk * s_k]);
        } else {
            uint32_t w[4];
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                uint32_t e[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int rr = ROW_CONTIG ? gr + j + h : gr, kk = ROW_CONTIG ? gk : gk + j + h;
                    e[h] = (rr < nrows && kk < kend) ? X[(long)rr * s_row + (long)kk * s_k] : 0u;
                }
                w[j / 2] = e[0] | (e[1] << 16);
            }
            v[i] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

template <bool ROW_CONTIG>
__device__ __forceinline__ void gemm_store(const uint4 (&v)[2], uint16_t* tile, int tid) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int q = tid * 2 + i;
        const int a = q >> 3, b = (q & 7) * 8;
        if (!ROW_CONTIG) {
            *reinterpret_cast<uint4*>(&tile[a * kGemmLdk + b]) = v[i];
        } else {
            const uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
            for (int j = 0; j < 8; ++j) tile[(b + j) * kGemmLdk + a] = (uint16_t)(w[j / 2] >> (16 * (j & 1)));
        }
    }
}

template <bool A_ROWS, bool B_ROWS>
__global__ __launch_bounds__(256) void k_gemm(const uint16_t* __restrict__ A, long sam, long sak,
                                              const uint16_t* __restrict__ B, long sbk, long sbn,
                                              float* __restrict__ C, long ldc, int M, int N, int K, int k_split,
                                              int atomic, int vec_a, int vec_b) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kGemmTile * kGemmLdk];  // [m][k]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kGemmTile * kGemmLdk];  // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * kGemmTile, n0 = blockIdx.x * kGemmTile;
    const int kbeg = blockIdx.z * k_split;
    const int kend = min(K, kbeg + k_split);
    f32x16 acc = {};
    uint4 ra[2], rb[2];
    if (kbeg < kend) {
        gemm_load<A_ROWS>(ra, A, sam, sak, m0, M, kbeg, kend, vec_a != 0, tid);
        gemm_load<B_ROWS>(rb, B, sbn, sbk, n0, N, kbeg, kend, vec_b != 0, tid);
    }
    for (int k0 = kbeg; k0 < kend; k0 += kGemmKStage) {
        __syncthreads();  // the previous stage's operand reads are done
        gemm_store<A_ROWS>(ra, As, tid);
        gemm_store<B_ROWS>(rb, Bs, tid);
        __syncthreads();
        if (k0 + kGemmKStage < kend) {
            gemm_load<A_ROWS>(ra, A, sam, sak, m0, M, k0 + kGemmKStage, kend, vec_a != 0, tid);
            gemm_load<B_ROWS>(rb, B, sbn, sbk, n0, N, k0 + kGemmKStage, kend, vec_b != 0, tid);
        }
#pragma unroll
        for (int ks = 0; ks < kGemmKStage / 16; ++ks) {
            const int kq = ks * 16 + (lane >> 5) * 8;
            const uint4 a = *reinterpret_cast<const uint4*>(&As[(wm * 32 + (lane & 31)) * kGemmLdk + kq]);
            const uint4 b = *reinterpret_cast<const uint4*>(&Bs[(wn * 32 + (lane & 31)) * kGemmLdk + kq]);
            acc = mfma(a, b, acc);
        }
    }
    const int col = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row < M && col < N) {
            float* dst = &C[(long)row * ldc + col];
            if (atomic) atomicAdd(dst, acc[r]);
            else *dst = acc[r];
        }
    }
}

// z += bias (kept in Z), H = bf16(silu(z)); rows of Z and H have `ld` columns, n of them used.
__global__ void k_bias_silu(float* __restrict__ Z, const float* __restrict__ bias, uint16_t* __restrict__ H, int M,
                            int n, int ld) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)M * ld) return;
    const int c = (int)(i % ld);
    if (c >= n) {
        H[i] = 0;
        return;
    }
    const float z = Z[i] + bias[c];
    Z[i] = z;
    H[i] = f2bf(silu(z));
}

// Output layer: y = acc + b4 (first kOut of kOutPad columns).  With targets: D4 = bf16(scale (y - t)),
// loss += sum (y - t)^2.  Y receives y in 4-float rows.
__global__ void k_output(const float* __restrict__ acc, const float* __restrict__ b4, const float* __restrict__ T,
                         float* __restrict__ Y, uint16_t* __restrict__ D4, float scale, float* __restrict__ loss,
                         int M) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (m < M) {
        float y[kMlpOut];
#pragma unroll
        for (int c = 0; c < kMlpOut; ++c) y[c] = acc[(long)m * kMlpOutPad + c] + b4[c];
        if (Y) {
            float4 o = make_float4(y[0], y[1], y[2], 0.0f);
            reinterpret_cast<float4*>(Y)[m] = o;
        }
        if (T) {
#pragma unroll
            for (int c = 0; c < kMlpOutPad; ++c) {
                float d = 0.0f;
                if (c < kMlpOut) {
                    d = y[c] - T[(long)m * 4 + c];
                    l += d * d;
                }
                D4[(long)m * kMlpOutPad + c] = f2bf(scale * d);
            }
        }
    }
    if (loss) {
        for (int d = 32; d >= 1; d >>= 1) l += __shfl_xor(l, d);
        if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
    }
}

// D = bf16(G * silu'(Z)), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_silu_backward(const float* __restrict__ G, const float* __restrict__ Z, uint16_t* __restrict__ D,
                                long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float z = Z[i];
    const float s = 1.0f / (1.0f + __expf(-z));
    D[i] = f2bf(G[i] * (s * (1.0f + z * (1.0f - s))));
}

// out[c] += sum over rows of D[m][c] (bias gradients), D bf16 with ld columns
// Bias gradients: out[c] += sum over the M rows of D[m][c], c < n.  Partial sums per chunk of
// kColsumRows rows (64 columns x 4 row groups per workgroup, coalesced along the row), then one
// pass over the chunks; every sum is taken in a fixed order, so the result does not depend on
// scheduling.
__global__ __launch_bounds__(256) void k_colsum_part(const uint16_t* __restrict__ D, float* __restrict__ part, int M,
                                                     int n, int ld) {
    const int l = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + l, r0 = blockIdx.y * kColsumRows;
    const int r1 = min(M, r0 + kColsumRows);
    float s = 0.0f;
    if (c < n) {
#pragma unroll 8
        for (int r = r0 + g; r < r1; r += 4) s += bf2f(D[(long)r * ld + c]);
    }
    __shared__ float red[4][64];
    red[g][l] = s;
    __syncthreads();
    if (g == 0 && c < n) part[(long)blockIdx.y * n + c] = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
}

__global__ void k_colsum_sum(const float* __restrict__ part, float* __restrict__ out, int chunks, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float s = 0.0f;
#pragma unroll 8
    for (int i = 0; i < chunks; ++i) s += part[(long)i * n + c];
    out[c] += s;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of the master weights: row-major [k][n] for the GEMMs and transposed [n][k] for
// the fused inference kernel.
__global__ void k_cast(const float* __restrict__ w, uint16_t* __restrict__ rowmajor, uint16_t* __restrict__ transposed,
                       int K, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * N) return;
    const int k = i / N, n = i % N;
    const uint16_t h = f2bf(w[i]);
    rowmajor[i] = h;
    transposed[(long)n * K + k] = h;
}

// ------------------------------------------------------------------------------------------
// Fused inference: Y[n x 4] = MLP(X[n x 208]).  Persistent: one workgroup of eight waves per CU
// copies all four layers' weights and biases into LDS once (bf16, transposed [out][in], each row
// padded so that the 16 rows an MFMA operand read touches spread over the banks)
// and then walks tiles of 128 rows; each wave takes 16 rows of a tile through the four layers on
// v_mfma_f32_16x16x32_bf16.
//
// MFMA 16x16x32 bf16 lane layout (wave64, lane l, c = l % 16, q = l / 16): A operand =
// A[c][8 q + 0..7], B operand = B[8 q + 0..7][c], accumulator element r = D[4 q + r][c].
//
// Each layer computes the transposed product Z^T = W^T H^T: the weights are the A operand (rows
// of W^T from LDS) and the activations the B operand, so lane (q, c) works on row c.  Output tile
// j leaves hidden neuron 16 j + 4 q + r of row c in lane (q, c); the next layer's B operand wants
// 8 values of row c at input positions 32 s + 8 q + t.  Defining position p = 32 s + 8 q + t to
// hold neuron 32 s + 16 (t / 4) + 4 q + t % 4 -- the lane's own accumulators acc[2 s + t / 4][t % 4]
// -- makes every activation stay in the lane that produced it: the input columns of layers 2-4 are
// stored in that order in LDS, and activations never leave registers.
// ------------------------------------------------------------------------------------------
struct InferWeights {
    const uint16_t* w1t;  // [128][208]
    const uint16_t* w2t;  // [128][128]
    const uint16_t* w3t;  // [128][128]
    const uint16_t* w4t;  // [32][128]
    const float *b1, *b2, *b3, *b4;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x4 mfma16(uint4 a, uint4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

constexpr int kLd1 = kMlpInPad + 8;   // LDS row length of layer 1's weights (208 used, 8 zero)
// LDS row length of layers 2-4's weights: 72 dwords.  Measured on gfx950 (SQ_LDS_BANK_CONFLICT
// over the kernel's operand reads): rows of 68, 76, 80, 84 or 92 dwords spend 45 % of the LDS's
// active cycles in bank conflicts, 96 dwords 63 %, and 72 or 88 dwords 27 %.
constexpr int kLdH = kMlpHidden + 16;
constexpr int kInferWaves = 8;
constexpr int kInferRows = 16;                         // rows per wave and tile
constexpr int kInferTile = kInferWaves * kInferRows;   // rows per tile
constexpr int kKSteps1 = (kMlpInPad + 31) / 32;        // layer 1's k steps of 32 (the last half used)
constexpr int kKStepsH = kMlpHidden / 32;              // k steps of layers 2-4
constexpr int kW4Rows = 16;                            // the output layer's A operand rows (3 used)
constexpr int kInferLdsHalves = kMlpHidden * kLd1 + 2 * kMlpHidden * kLdH + kW4Rows * kLdH;
constexpr int kInferBias = 3 * kMlpHidden + 4;         // b1 | b2 | b3 | b4[0..3], floats
static_assert(kInferLdsHalves % 8 == 0, "the biases follow the weights 16-byte aligned");
static_assert(kInferLdsHalves * 2 + kInferBias * 4 <= 160 * 1024, "weights must fit one CU's LDS");
static_assert(32 * kKSteps1 <= kLd1 + 8, "layer 1's last k step reads at most 8 values past a row");

struct InferLds {
    uint16_t *w1, *w2, *w3, *w4;  // [128][kLd1], [128][kLdH] x 2, [16][kLdH]
    float* bias;                  // [kInferBias]
};

__device__ __forceinline__ InferLds infer_lds(uint16_t* s_mem) {
    InferLds L;
    L.w1 = s_mem;
    L.w2 = L.w1 + kMlpHidden * kLd1;
    L.w3 = L.w2 + kMlpHidden * kLdH;
    L.w4 = L.w3 + kMlpHidden * kLdH;
    L.bias = reinterpret_cast<float*>(s_mem + kInferLdsHalves);
    return L;
}

__device__ __forceinline__ uint32_t pack2bf(float lo, float hi) {
    typedef __attribute__((__vector_size__(2 * sizeof(float)))) float f32x2;
    typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
    const f32x2 v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// Layer 1's input order.  Its first six k steps (192 positions) hold the 32 neighbours' values:
// lane quarter q takes neighbours 8q .. 8q + 7, whose 48 values are features 48q .. 48q + 47, as
// 8 consecutive features per k step: position p = 32 s + 8 q + t holds feature 48 q + 8 s + t.
// Every lane then starts step s at the same offset, (8 s) mod 6, into its first neighbour of the
// step, a constant.  Step 6 holds features 192 .. 223 in order (the three scalars, then zeros).
__device__ __forceinline__ int layer1_chunk(int p8) {  // source chunk of 8 features for chunk p8
    return p8 < 24 ? 6 * (p8 & 3) + (p8 >> 2) : p8;
}

// Weights and biases into LDS, 8 values per copy.  Layer 1's input columns are stored in the
// order above, and its row padding is zeroed (its last k step multiplies it, and the first 8
// values of the next row, by zero inputs).  Layers 2-4 store input position p = 32 s + 8 q + t of
// each row from neuron 32 s + 16 (t / 4) + 4 q + t % 4: the 8 positions of one copy are two runs
// of 4 neurons, 16 apart.
__device__ __forceinline__ void stage_weights(const InferWeights& W, const InferLds& L, int tid, int nthreads) {
    for (int i = tid; i < kMlpHidden * (kLd1 / 8); i += nthreads) {
        const int r = i / (kLd1 / 8), q = i % (kLd1 / 8);
        const int src = layer1_chunk(q);
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (src < kMlpInPad / 8) v = reinterpret_cast<const uint4*>(W.w1t)[r * (kMlpInPad / 8) + src];
        *reinterpret_cast<uint4*>(&L.w1[r * kLd1 + q * 8]) = v;
    }
    constexpr int kChunks = kMlpHidden / 8;
    for (int i = tid; i < (2 * kMlpHidden + kW4Rows) * kChunks; i += nthreads) {
        const int r = i / kChunks, p8 = i % kChunks;
        const uint16_t* src;
        uint16_t* dst;
        if (r < kMlpHidden) {
            src = W.w2t + (size_t)r * kMlpHidden;
            dst = L.w2 + r * kLdH;
        } else if (r < 2 * kMlpHidden) {
            src = W.w3t + (size_t)(r - kMlpHidden) * kMlpHidden;
            dst = L.w3 + (r - kMlpHidden) * kLdH;
        } else {
            src = W.w4t + (size_t)(r - 2 * kMlpHidden) * kMlpHidden;
            dst = L.w4 + (r - 2 * kMlpHidden) * kLdH;
        }
        const int s = p8 >> 2, q = p8 & 3;
        const uint2 lo = *reinterpret_cast<const uint2*>(src + 32 * s + 4 * q);
        const uint2 hi = *reinterpret_cast<const uint2*>(src + 32 * s + 16 + 4 * q);
        *reinterpret_cast<uint4*>(dst + 8 * p8) = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
    for (int i = tid; i < kInferBias; i += nthreads)
        L.bias[i] = i < kMlpHidden       ? W.b1[i]
                    : i < 2 * kMlpHidden ? W.b2[i - kMlpHidden]
                    : i < 3 * kMlpHidden ? W.b3[i - 2 * kMlpHidden]
                                         : W.b4[i - 3 * kMlpHidden];
}

// Layer 1's B operand for a wave's 16 rows from row0, in layer 1's input order (rows at or past
// n, and features at or past 208, read as zeros).
__device__ __forceinline__ void load_rows(uint4 (&a)[kKSteps1], const uint16_t* __restrict__ X, long row0, uint32_t n,
                                          int lane) {
    const int row = lane & 15, q4 = lane >> 4;
    const bool live = row0 + row < (long)n;
#pragma unroll
    for (int s = 0; s < kKSteps1; ++s) a[s] = make_uint4(0u, 0u, 0u, 0u);
    if (live) {
        const uint16_t* src = X + (row0 + row) * kMlpInPad;
#pragma unroll
        for (int s = 0; s < kKSteps1; ++s) {
            const int chunk = layer1_chunk(4 * s + q4);
            if (chunk < kMlpInPad / 8) a[s] = *reinterpret_cast<const uint4*>(src + 8 * chunk);
        }
    }
}

// hout = bf16(silu(W^T hin + b)) for a wave's 16 rows: W^T rows from LDS with stride LDW, the
// input in registers (KSTEPS B operands), the output in the next layer's B operand order.
template <int KSTEPS, int LDW, int GROUP = 8>
__device__ __forceinline__ void infer_layer(const uint4 (&hin)[KSTEPS], const uint16_t* wt, const float* bias,
                                            uint4 (&hout)[kKStepsH], int lane) {
    const int c = lane & 15, q4 = lane >> 4;
    // accumulators start at the bias: element r of tile j is neuron 16 j + 4 q + r
    f32x4 acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = *reinterpret_cast<const f32x4*>(&bias[16 * j + 4 * q4]);
    // the weights of one group of GROUP output tiles in flight while the previous group's
    // products run; the scheduling barriers keep the compiler from hoisting every group's loads
    // (and their registers) up front
    static_assert(8 % GROUP == 0, "groups split a k step's 8 output tiles");
    constexpr int kGroups = KSTEPS * (8 / GROUP);
    const uint16_t* wrow = wt + c * LDW + 8 * q4;
    uint4 w[2][GROUP];
#pragma unroll
    for (int g = 0; g < GROUP; ++g) w[0][g] = *reinterpret_cast<const uint4*>(&wrow[16 * g * LDW]);
#pragma unroll
    for (int i = 0; i < kGroups; ++i) {
        const int s = i / (8 / GROUP), j0 = (i % (8 / GROUP)) * GROUP;
        if (i + 1 < kGroups) {
            const int s1 = (i + 1) / (8 / GROUP), j1 = ((i + 1) % (8 / GROUP)) * GROUP;
#pragma unroll
            for (int g = 0; g < GROUP; ++g)
                w[(i + 1) & 1][g] = *reinterpret_cast<const uint4*>(&wrow[16 * (j1 + g) * LDW + 32 * s1]);
        }
#pragma unroll
        for (int g = 0; g < GROUP; ++g) acc[j0 + g] = mfma16(w[i & 1][g], hin[s], acc[j0 + g]);
        __builtin_amdgcn_sched_barrier(0);
    }
    // every write below depends on every read above, so hout may be hin
#pragma unroll
    for (int s = 0; s < kKStepsH; ++s) {
        const f32x4 a = acc[2 * s], b = acc[2 * s + 1];
        hout[s] = make_uint4(pack2bf(silu(a[0]), silu(a[1])), pack2bf(silu(a[2]), silu(a[3])),
                             pack2bf(silu(b[0]), silu(b[1])), pack2bf(silu(b[2]), silu(b[3])));
    }
}

// Output layer: one 16-row A operand tile (rows 0..2 used); lanes 0..15 hold outputs 0..3 of
// row c and write Y's 4 columns (column 3 zero).
__device__ __forceinline__ void infer_output(const uint4 (&h)[kKStepsH], const InferLds& L, float* __restrict__ Y,
                                             long row0, int rows_left, int lane) {
    const int c = lane & 15, q4 = lane >> 4;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < kKStepsH; ++s)
        acc = mfma16(*reinterpret_cast<const uint4*>(&L.w4[c * kLdH + 32 * s + 8 * q4]), h[s], acc);
    if (q4 == 0 && c < rows_left) {
        const float* b4 = L.bias + 3 * kMlpHidden;
        float* y = Y + (row0 + c) * 4;
        y[0] = acc[0] + b4[0];
        y[1] = acc[1] + b4[1];
        y[2] = acc[2] + b4[2];
        y[3] = 0.0f;
    }
}

__global__ __launch_bounds__(kInferWaves * 64) void k_mlp_infer(const uint16_t* __restrict__ X, InferWeights W,
                                                                 float* __restrict__ Y, uint32_t n) {
    __shared__ __attribute__((aligned(16))) uint16_t s_mem[kInferLdsHalves + 2 * kInferBias];
    const InferLds L = infer_lds(s_mem);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    const uint32_t tiles = (n + kInferTile - 1) / kInferTile;
    uint4 xa[kKSteps1];
    if (blockIdx.x < tiles) load_rows(xa, X, (long)blockIdx.x * kInferTile + wave * kInferRows, n, lane);
    stage_weights(W, L, tid, kInferWaves * 64);
    __syncthreads();

    // From here on LDS is read-only and the waves are independent; the next tile's input rows
    // load while the current tile is computed.
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long row0 = (long)tile * kInferTile + wave * kInferRows;
        const int rows_left = (int)max(0L, min((long)kInferRows, (long)n - row0));
        uint4 h[kKStepsH];
        infer_layer<kKSteps1, kLd1>(xa, L.w1, L.bias, h, lane);
        const uint32_t next = tile + gridDim.x;
        if (next < tiles) load_rows(xa, X, (long)next * kInferTile + wave * kInferRows, n, lane);
        infer_layer<kKStepsH, kLdH>(h, L.w2, L.bias + kMlpHidden, h, lane);
        infer_layer<kKStepsH, kLdH>(h, L.w3, L.bias + 2 * kMlpHidden, h, lane);
        infer_output(h, L, Y, row0, rows_left, lane);
    }
}

// ------------------------------------------------------------------------------------------
// Inference on the closure's inputs without the feature matrix: each lane forms its own layer-1
// operand from the neighbour lists.  Lane l (row l % 16, quarter q = l / 16) holds, in layer 1's
// input order, the values of neighbours 8q .. 8q + 7 of its row: k step s < 6 takes neighbours
// a_s = floor(8 s / 6) and a_s + 1 of the eight, from value (8 s) mod 6 of the first.  The next
// tile's rows are fetched in phases between the current tile's layers (the row counts a tile
// earlier still, then the list entries, the records, and last the values), so the loads land
// while the matrix cores work.
// ------------------------------------------------------------------------------------------
constexpr int kFeatSteps = 6;  // k-steps of neighbour values (192 columns)
constexpr int kLaneNeighbours = 8;

struct RowFetch {
    uint32_t row, listed;
    bool live;
    uint32_t listed_next;  // the count of the row one tile further on, loaded a tile early
    float4 p;              // own position and density
    float3 pv;             // own velocity
    uint32_t e[kLaneNeighbours];                          // phase 1: list entries (or selected slots)
    float3 q[kLaneNeighbours], qv[kLaneNeighbours];       // phase 2: the neighbours' records
};

__device__ __forceinline__ float3 load3(const float4* a, uint32_t i) {
    const float* f = reinterpret_cast<const float*>(a) + 4 * (size_t)i;
    return float3{f[0], f[1], f[2]};
}

__device__ __forceinline__ uint32_t row_count(const ClosureRows& cr, long row0, uint32_t n, int lane) {
    const long row = row0 + (lane & 15);
    return row < (long)n ? cr.ncount[row] : 0u;
}

// phase 1: the row's list entries and own record; `listed` arrived a tile earlier, so the loads
// go to the list itself or, past 32 neighbours, to the 32 slots its selection wrote (cr.sel)
__device__ __forceinline__ void fetch_entries(RowFetch& f, const ClosureRows& cr, long row0, uint32_t n, int lane) {
    const int r16 = lane & 15, q4 = lane >> 4;
    f.live = row0 + r16 < (long)n;
    f.row = f.live ? (uint32_t)(row0 + r16) : 0u;
    f.listed = f.listed_next;
    f.p = cr.posrho[f.row];
    f.pv = load3(cr.velp, f.row);
    // entry k of the row at base + k * step: the selection's row (step 1) or the list's column
    // (step stride), one selected address advanced per entry rather than a branch per source
    const bool sel = f.listed > 32u;
    const size_t step = sel ? (size_t)1 : cr.stride;
    const uint32_t* src = (sel ? cr.sel + (size_t)f.row * kSelRow : cr.nlist + f.row) + (size_t)(kLaneNeighbours * q4) * step;
#pragma unroll
    for (int i = 0; i < kLaneNeighbours; ++i) {
        f.e[i] = *src;
        src += step;
    }
}

// phase 2: the neighbours' records (entries past a list may be stale: a slot out of range is
// replaced by the row's own, so every load stays inside the buffers)
__device__ __forceinline__ void fetch_records(RowFetch& f, const ClosureRows& cr, uint32_t n, int lane) {
    // a scalar base plus one 32-bit byte offset for both records (slots are below 2^28)
#pragma unroll
    for (int i = 0; i < kLaneNeighbours; ++i) {
        const uint32_t off = (f.e[i] < n ? f.e[i] : f.row) << 4;
        const float* a = reinterpret_cast<const float*>(reinterpret_cast<const char*>(cr.posrho) + off);
        const float* b = reinterpret_cast<const float*>(reinterpret_cast<const char*>(cr.velp) + off);
        f.q[i] = float3{a[0], a[1], a[2]};
        f.qv[i] = float3{b[0], b[1], b[2]};
    }
}

// phase 3: the layer-1 operand, formed as the feature kernel forms its rows (same operations, so
// the same bf16 values)
__device__ __forceinline__ void form_rows(uint4 (&a)[kKSteps1], const RowFetch& f, const ClosureRows& cr, int lane) {
    const int q4 = lane >> 4;
    if (!f.live) {
#pragma unroll
        for (int s = 0; s < kKSteps1; ++s) a[s] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const uint32_t cnt = f.listed > 32u ? 32u : f.listed;  // longer lists keep their 32 nearest
    const float inv_h = 1.0f / cr.h;
    float v[kLaneNeighbours][6];
    uint32_t mask[kLaneNeighbours];
#pragma unroll
    for (int i = 0; i < kLaneNeighbours; ++i) {
        // zeros past the list by masking the bits (an opaque mask keeps the loads unconditional)
        mask[i] = (uint32_t)(kLaneNeighbours * q4 + i) < cnt ? ~0u : 0u;
        asm("" : "+v"(mask[i]));
        v[i][0] = (f.q[i].x - f.p.x) * inv_h;
        v[i][1] = (f.q[i].y - f.p.y) * inv_h;
        v[i][2] = (f.q[i].z - f.p.z) * inv_h;
        v[i][3] = f.qv[i].x - f.pv.x;
        v[i][4] = f.qv[i].y - f.pv.y;
        v[i][5] = f.qv[i].z - f.pv.z;
    }
    // each packed pair belongs to one neighbour (6 values each, pairs start at even positions), so
    // its mask applies to the pair: bf16 pairs of masked values, with half the masking
#pragma unroll
    for (int s = 0; s < kFeatSteps; ++s) {
        const int first = (8 * s) / 6, o = (8 * s) % 6;  // compile-time
        static_assert((8 * kFeatSteps) % 2 == 0 && 6 % 2 == 0, "pairs within a neighbour");
        uint32_t w[4];
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            const int i = o + e < 6 ? first : first + 1, k = o + e < 6 ? o + e : o + e - 6;
            w[e / 2] = pack2bf(v[i][k], v[i][k + 1]) & mask[i];
        }
        a[s] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    // columns 192..207: the density deviation, the step multiplier and the filled fraction
    const float rho_rel = (f.p.w / cr.rho0 - 1.0f) * 10.0f;
    a[kFeatSteps] = q4 == 0 ? make_uint4(pack2bf(rho_rel, cr.dt_mult * 0.25f),
                                         pack2bf((float)cnt * (1.0f / 32.0f), 0.0f), 0u, 0u)
                            : make_uint4(0u, 0u, 0u, 0u);
}

__global__ __launch_bounds__(kInferWaves * 64) void k_mlp_infer_closure(ClosureRows cr, InferWeights W,
                                                                         float* __restrict__ Y, uint32_t n) {
    __shared__ __attribute__((aligned(16))) uint16_t s_mem[kInferLdsHalves + 2 * kInferBias];
    const InferLds L = infer_lds(s_mem);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    const uint32_t tiles = (n + kInferTile - 1) / kInferTile;
    uint4 xa[kKSteps1];
    RowFetch f;
    if (blockIdx.x < tiles) {
        f.listed_next = row_count(cr, (long)blockIdx.x * kInferTile + wave * kInferRows, n, lane);
        fetch_entries(f, cr, (long)blockIdx.x * kInferTile + wave * kInferRows, n, lane);
        const uint32_t second = blockIdx.x + gridDim.x;
        f.listed_next = second < tiles ? row_count(cr, (long)second * kInferTile + wave * kInferRows, n, lane) : 0u;
        fetch_records(f, cr, n, lane);
        form_rows(xa, f, cr, lane);
    }
    stage_weights(W, L, tid, kInferWaves * 64);
    __syncthreads();

    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long row0 = (long)tile * kInferTile + wave * kInferRows;
        const int rows_left = (int)max(0L, min((long)kInferRows, (long)n - row0));
        const uint32_t next = tile + gridDim.x;
        const bool more = next < tiles;
        uint4 h[kKStepsH];
        infer_layer<kKSteps1, kLd1, 4>(xa, L.w1, L.bias, h, lane);
        if (more) {
            fetch_entries(f, cr, (long)next * kInferTile + wave * kInferRows, n, lane);
            const uint32_t after = next + gridDim.x;
            f.listed_next = after < tiles ? row_count(cr, (long)after * kInferTile + wave * kInferRows, n, lane) : 0u;
        }
        infer_layer<kKStepsH, kLdH, 4>(h, L.w2, L.bias + kMlpHidden, h, lane);
        if (more) fetch_records(f, cr, n, lane);
        infer_layer<kKStepsH, kLdH, 4>(h, L.w3, L.bias + 2 * kMlpHidden, h, lane);
        infer_output(h, L, Y, row0, rows_left, lane);
        if (more) form_rows(xa, f, cr, lane);
    }
}

inline unsigned nblk(long n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------
hipError_t gemm(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C, long ldc, int M,
                int N, int K, bool accumulate, hipStream_t s) {
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    // split K so that long reductions (weight gradients over a batch) spread over the chip: at
    // least 256 k per split, up to about 512 workgroups
    const long tiles = (long)((M + kGemmTile - 1) / kGemmTile) * ((N + kGemmTile - 1) / kGemmTile);
    int splits = 1;
    if (K >= 1024 && tiles < 256) splits = (int)std::min<long>((K + 255) / 256, (512 + tiles - 1) / tiles);
    int k_split = (K + splits - 1) / splits;
    k_split = (k_split + kGemmKStage - 1) / kGemmKStage * kGemmKStage;
    splits = (K + k_split - 1) / k_split;
    const bool atomic = accumulate || splits > 1;
    if (atomic && !accumulate) {
        hipError_t e = hipMemset2DAsync(C, ldc * sizeof(float), 0, N * sizeof(float), M, s);
        if (e != hipSuccess) return e;
    }
    // which dimension of each operand is contiguous, and whether its runs of 8 may load as one
    // 16-byte access (aligned base, the other stride a multiple of 8)
    const bool a_rows = sam == 1, b_rows = sbn == 1;
    auto aligned = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
    const int vec_a = aligned(A) && (a_rows ? sak % 8 == 0 : (sak == 1 && sam % 8 == 0));
    const int vec_b = aligned(B) && (b_rows ? sbk % 8 == 0 : (sbk == 1 && sbn % 8 == 0));
    const dim3 grid((N + kGemmTile - 1) / kGemmTile, (M + kGemmTile - 1) / kGemmTile, splits);
#define WC_GEMM(AR, BR)                                                                                            \
    hipLaunchKernelGGL((k_gemm<AR, BR>), grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, k_split, \
                       atomic ? 1 : 0, vec_a, vec_b)
    if (a_rows && b_rows) WC_GEMM(true, true);
    else if (a_rows) WC_GEMM(true, false);
    else if (b_rows) WC_GEMM(false, true);
    else WC_GEMM(false, false);
#undef WC_GEMM
    return hipGetLastError();
}

MlpLayout mlp_layout() {
    MlpLayout L;
    size_t o = 0;
    L.w1 = o; o += (size_t)kMlpInPad * kMlpHidden;
    L.b1 = o; o += kMlpHidden;
    L.w2 = o; o += (size_t)kMlpHidden * kMlpHidden;
    L.b2 = o; o += kMlpHidden;
    L.w3 = o; o += (size_t)kMlpHidden * kMlpHidden;
    L.b3 = o; o += kMlpHidden;
    L.w4 = o; o += (size_t)kMlpHidden * kMlpOutPad;
    L.b4 = o; o += kMlpOutPad;
    L.total = o;
    return L;
}

hipError_t mlp_refresh_bf16(Mlp& m, hipStream_t s) {
    const MlpLayout L = mlp_layout();
    struct {
        size_t off;
        int K, N;
        uint16_t *rm, *tr;
    } mats[4] = {{L.w1, kMlpInPad, kMlpHidden, m.w1, m.w1t},
                 {L.w2, kMlpHidden, kMlpHidden, m.w2, m.w2t},
                 {L.w3, kMlpHidden, kMlpHidden, m.w3, m.w3t},
                 {L.w4, kMlpHidden, kMlpOutPad, m.w4, m.w4t}};
    for (auto& x : mats) {
        hipLaunchKernelGGL(k_cast, dim3(nblk((long)x.K * x.N, 256)), dim3(256), 0, s, m.master + x.off, x.rm, x.tr, x.K,
                           x.N);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t mlp_infer(const Mlp& m, const uint16_t* X, float* Y, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const MlpLayout L = mlp_layout();
    InferWeights W{m.w1t, m.w2t, m.w3t, m.w4t, m.master + L.b1, m.master + L.b2, m.master + L.b3, m.master + L.b4};
    // one workgroup per CU (the kernel holds most of a CU's LDS), each walking several tiles
    const unsigned tiles = nblk(n, kInferTile);
    hipLaunchKernelGGL(k_mlp_infer, dim3(tiles < 256u ? tiles : 256u), dim3(kInferWaves * 64), 0, s, X, W, Y, n);
    return hipGetLastError();
}

hipError_t mlp_infer_closure(const Mlp& m, const ClosureRows& rows, float* Y, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const MlpLayout L = mlp_layout();
    InferWeights W{m.w1t, m.w2t, m.w3t, m.w4t, m.master + L.b1, m.master + L.b2, m.master + L.b3, m.master + L.b4};
    const unsigned tiles = nblk(n, kInferTile);
    hipLaunchKernelGGL(k_mlp_infer_closure, dim3(tiles < 256u ? tiles : 256u), dim3(kInferWaves * 64), 0, s, rows, W,
                       Y, n);
    return hipGetLastError();
}

// Forward through the GEMM path, keeping what the backward pass needs.  With T, also the loss
// and the output-layer delta.
hipError_t mlp_forward_train(Mlp& m, const uint16_t* X, const float* T, float* Y, uint32_t M, hipStream_t s) {
    const MlpLayout L = mlp_layout();
    const float* b = m.master;
    hipError_t e;
#define WC_TRY(x)                            \
    do {                                     \
        e = (x);                             \
        if (e != hipSuccess) return e;       \
    } while (0)
    // layer 1..3
    WC_TRY(gemm(X, kMlpInPad, 1, m.w1, kMlpHidden, 1, m.z[0], kMlpHidden, M, kMlpHidden, kMlpInPad, false, s));
    hipLaunchKernelGGL(k_bias_silu, dim3(nblk((long)M * kMlpHidden, 256)), dim3(256), 0, s, m.z[0], b + L.b1, m.h[0],
                       (int)M, kMlpHidden, kMlpHidden);
    WC_TRY(hipGetLastError());
    WC_TRY(gemm(m.h[0], kMlpHidden, 1, m.w2, kMlpHidden, 1, m.z[1], kMlpHidden, M, kMlpHidden, kMlpHidden, false, s));
    hipLaunchKernelGGL(k_bias_silu, dim3(nblk((long)M * kMlpHidden, 256)), dim3(256), 0, s, m.z[1], b + L.b2, m.h[1],
                       (int)M, kMlpHidden, kMlpHidden);
    WC_TRY(hipGetLastError());
    WC_TRY(gemm(m.h[1], kMlpHidden, 1, m.w3, kMlpHidden, 1, m.z[2], kMlpHidden, M, kMlpHidden, kMlpHidden, false, s));
    hipLaunchKernelGGL(k_bias_silu, dim3(nblk((long)M * kMlpHidden, 256)), dim3(256), 0, s, m.z[2], b + L.b3, m.h[2],
                       (int)M, kMlpHidden, kMlpHidden);
    WC_TRY(hipGetLastError());
    WC_TRY(gemm(m.h[2], kMlpHidden, 1, m.w4, kMlpOutPad, 1, m.yacc, kMlpOutPad, M, kMlpOutPad, kMlpHidden, false, s));
    if (T) WC_TRY(hipMemsetAsync(m.loss, 0, sizeof(float), s));
    const float scale = 2.0f / (3.0f * (float)M);
    hipLaunchKernelGGL(k_output, dim3(nblk(M, 256)), dim3(256), 0, s, m.yacc, b + L.b4, T, Y, m.d4, scale,
                       T ? m.loss : nullptr, (int)M);
    WC_TRY(hipGetLastError());
    return hipSuccess;
}

// out[c] += column sums of the M x n bf16 matrix D (leading dimension ld).
static hipError_t colsum(Mlp& m, const uint16_t* D, float* out, int M, int n, int ld, hipStream_t s) {
    const int chunks = (M + kColsumRows - 1) / kColsumRows;
    hipLaunchKernelGGL(k_colsum_part, dim3((n + 63) / 64, chunks), dim3(256), 0, s, D, m.colpart, M, n, ld);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_colsum_sum, dim3((n + 127) / 128), dim3(128), 0, s, m.colpart, out, chunks, n);
    return hipGetLastError();
}

// Gradients of the mean squared error into m.grad (after mlp_forward_train with targets).
hipError_t mlp_backward(Mlp& m, const uint16_t* X, uint32_t M, hipStream_t s) {
    const MlpLayout L = mlp_layout();
    float* g = m.grad;
    hipError_t e;
    WC_TRY(hipMemsetAsync(g, 0, L.total * sizeof(float), s));
    // layer 4: dW4 = H3^T D4, db4 = colsum D4
    WC_TRY(gemm(m.h[2], 1, kMlpHidden, m.d4, kMlpOutPad, 1, g + L.w4, kMlpOutPad, kMlpHidden, kMlpOutPad, M, true, s));
    WC_TRY(colsum(m, m.d4, g + L.b4, (int)M, kMlpOut, kMlpOutPad, s));
    // G3 = D4 W4^T, D3 = G3 * silu'(Z3)
    WC_TRY(gemm(m.d4, kMlpOutPad, 1, m.w4, 1, kMlpOutPad, m.gbuf, kMlpHidden, M, kMlpHidden, kMlpOutPad, false, s));
    hipLaunchKernelGGL(k_silu_backward, dim3(nblk((long)M * kMlpHidden, 256)), dim3(256), 0, s, m.gbuf, m.z[2], m.d,
                       (long)M * kMlpHidden);
    WC_TRY(hipGetLastError());
    const size_t wofs[3] = {L.w1, L.w2, L.w3}, bofs[3] = {L.b1, L.b2, L.b3};
    const uint16_t* wrm[3] = {m.w1, m.w2, m.w3};
    for (int l = 2; l >= 0; --l) {
        // dW_l = H_{l-1}^T D_l (H_{-1} = X), db_l = colsum D_l
        const uint16_t* Hin = l == 0 ? X : m.h[l - 1];
        const int kin = l == 0 ? kMlpInPad : kMlpHidden;
        WC_TRY(gemm(Hin, 1, kin, m.d, kMlpHidden, 1, g + wofs[l], kMlpHidden, kin, kMlpHidden, M, true, s));
        WC_TRY(colsum(m, m.d, g + bofs[l], (int)M, kMlpHidden, kMlpHidden, s));
        if (l == 0) break;
        // G_{l-1} = D_l W_l^T, D_{l-1} = G * silu'(Z_{l-1})
        WC_TRY(gemm(m.d, kMlpHidden, 1, wrm[l], 1, kMlpHidden, m.gbuf, kMlpHidden, M, kMlpHidden, kMlpHidden, false,
                    s));
        hipLaunchKernelGGL(k_silu_backward, dim3(nblk((long)M * kMlpHidden, 256)), dim3(256), 0, s, m.gbuf,
                           m.z[l - 1], m.d, (long)M * kMlpHidden);
        WC_TRY(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t mlp_adam(Mlp& m, float lr, hipStream_t s) {
    const MlpLayout L = mlp_layout();
    m.adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)m.adam_t), c2 = 1.0f - powf(b2, (float)m.adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk((long)L.total, 256)), dim3(256), 0, s, m.master, m.grad, m.adam_m, m.adam_v,
                       (int)L.total, lr, b1, b2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return mlp_refresh_bf16(m, s);
#undef WC_TRY
}

}  // namespace wc
