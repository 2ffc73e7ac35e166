// wc_mlp.hpp -- the learned closure's network: sizes, device state, kernels' host entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace wc {

constexpr int kMlpIn = 195;      // 32 neighbours x (dx, dv) + 3 scalars
constexpr int kMlpInPad = 208;   // padded to whole 16-wide MFMA k-steps (zero features)
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 3;
constexpr int kMlpOutPad = 32;   // one 32-wide MFMA tile (zero weights past column 3)

// Offsets of the parameter blocks in the flat fp32 master vector (padded shapes):
// W1 [208][128], b1 [128], W2 [128][128], b2, W3 [128][128], b3, W4 [128][32], b4 [32].
// W_l is row-major [input][output].
struct MlpLayout {
    size_t w1, b1, w2, b2, w3, b3, w4, b4, total;
};
MlpLayout mlp_layout();

constexpr int kColsumRows = 1024;  // rows per partial column sum of the bias gradients

struct Mlp {
    uint32_t max_batch = 0;
    float* master = nullptr;  // fp32 parameters (MlpLayout)
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    long adam_t = 0;
    uint16_t *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *w4 = nullptr;      // bf16 row-major [in][out]
    uint16_t *w1t = nullptr, *w2t = nullptr, *w3t = nullptr, *w4t = nullptr;  // bf16 [out][in]
    // training workspace, max_batch rows each
    float* z[3] = {nullptr, nullptr, nullptr};     // pre-activations (fp32, 128 columns)
    uint16_t* h[3] = {nullptr, nullptr, nullptr};  // activations (bf16, 128 columns)
    float* yacc = nullptr;                         // output layer before bias (32 columns)
    uint16_t* d4 = nullptr;                        // output delta (bf16, 32 columns)
    float* gbuf = nullptr;                         // back-propagated gradient (fp32, 128 columns)
    uint16_t* d = nullptr;                         // hidden delta (bf16, 128 columns)
    float* loss = nullptr;                         // sum of squared errors of the last batch
    float* colpart = nullptr;                      // column sums per chunk of kColsumRows rows
    hipStream_t stream = nullptr;
    int device = 0;
};

hipError_t gemm(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C, long ldc, int M,
                int N, int K, bool accumulate, hipStream_t s);
hipError_t mlp_refresh_bf16(Mlp& m, hipStream_t s);
hipError_t mlp_infer(const Mlp& m, const uint16_t* X, float* Y, uint32_t n, hipStream_t s);
// The learned closure's inputs, for inference that forms each row's layer-1 operand itself, the
// values the closure's feature kernel would write: from a particle's neighbour list when it
// holds at most 32, else from the 32 nearest its selection wrote to `sel`.  Single scene, row r =
// slot r.
constexpr int kSelRow = 32;  // slots per row of the selection buffer (the closure's neighbours)
struct ClosureRows {
    const float4* posrho;    // sorted records: position, density
    const float4* velp;      // velocity, pressure
    const uint32_t* nlist;   // entry k of slot t at nlist[k * stride + t]
    const uint32_t* ncount;  // listed neighbours of slot t
    const uint32_t* sel;     // more than 32 listed: the 32 nearest, in list order, at sel[t * kSelRow + k]
    size_t stride;           // of the lists
    float h, rho0, dt_mult;
};
hipError_t mlp_infer_closure(const Mlp& m, const ClosureRows& rows, float* Y, uint32_t n, hipStream_t s);
hipError_t mlp_forward_train(Mlp& m, const uint16_t* X, const float* T, float* Y, uint32_t M, hipStream_t s);
hipError_t mlp_backward(Mlp& m, const uint16_t* X, uint32_t M, hipStream_t s);
hipError_t mlp_adam(Mlp& m, float lr, hipStream_t s);

}  // namespace wc
