// Learned radiance cache (neural radiance caching after Mueller, Rousselle,
// Novak and Keller 2021): a multiresolution hash-grid encoding of position
// (Mueller, Evans, Schied and Keller 2022) plus the vertex's direction,
// normal, roughness and albedo, fed to a 5 x 64 ReLU MLP with an RGB output.
// The MLP runs on BF16 MFMA with FP32 accumulation; parameters are FP32
// masters updated by Adam, with BF16 copies for the kernels.  New: the
// reference has no learned cache; the nearest behaviour is the irradiance
// cache ending a path at a hit (shaders/raytrace.rgen:1117-1123).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "rtx_amd.h"

namespace rtx {

constexpr int kNrcWidth = 64;                              // hidden width
constexpr int kNrcHidden = 5;                              // hidden layers (64 -> 64, ReLU)
constexpr int kNrcLayers = kNrcHidden + 1;                 // + the output layer (64 -> 3, padded to 16)
constexpr int kNrcOutPad = 16;                             // output rows of the last weight matrix
constexpr int kNrcMaxLevels = 16;
constexpr int kNrcFeatures = 2;                            // features per level
constexpr int kNrcInputs = 64;                             // network input width
constexpr int kNrcMlpParams = kNrcHidden * kNrcWidth * kNrcWidth + kNrcOutPad * kNrcWidth; // out-major
constexpr int kNrcRowsPerBlock = 128;                      // queries per workgroup (4 waves x 32 rows)

// Encoding constants: scene box, per-level resolution, table size.
struct NrcGrid {
    float lo[3], inv[3];        // position -> [0, 1]^3: (x - lo) * inv
    float res[kNrcMaxLevels];   // cells per axis of level l
    uint32_t dense[kNrcMaxLevels]; // 1: (res + 1)^3 <= table size, indexed densely
    uint32_t levels;
    uint32_t tableMask;         // entries per level - 1 (a power of two)
};

// Device state of one cache.
struct NrcBuffers {
    NrcGrid grid;
    // parameters: FP32 masters (grid [levels][T][2], then the MLP out-major),
    // their gradients and Adam moments; BF16 copies for the kernels
    float *params, *grads, *adamM, *adamV;
    size_t gridParams;          // levels * T * 2
    uint32_t *gridBf16;         // [levels][T] two bf16 features packed
    uint16_t *wFwd;             // [layer][out][64] bf16 (out-major: forward B operand)
    uint16_t *wBwd;             // [layer][in][64] bf16 (in-major: backward B operand; layer 5 out padded to 32)
    // inference's parameters: their moving average (FP32, from 0 at the
    // optimizer's reset) and its bias-corrected bf16 copies; with emaDecay 0
    // these alias params, gridBf16 and wFwd
    float *ema;
    uint32_t *gridBf16Ema;
    uint16_t *wFwdEma;
    float emaDecay;
    // training batch work (rows = trainBatch)
    uint16_t *acts;             // [6][rows][64] bf16: the input of every layer
    uint16_t *dz;               // [6][rows][64] bf16: d loss / d pre-activation of every layer
    float *pred;                // [rows][4]
    float *dFeat;               // [rows][32]: d loss / d encoded features
    float *partials;            // weight-gradient partials [chunks][6][64][64]
    uint32_t rowsCap;
    uint32_t step;              // Adam steps taken
    float lr, beta1, beta2, eps;
};

// Self-training suffixes (inference inside a render): a query whose path has
// endFlag in pathThr[pid].w also adds its term, over the throughput at the
// training vertex (trainPath[2 pid].xyz; .w = its target slot), to target.
struct NrcSelfTrain {
    const float4 *pathThr;
    const float4 *trainPath;
    float4 *target;   // NULL: no self-training
    uint32_t endFlag;
};
// n rows, or *count rows (at most n) when count is given
void launchNrcInfer(hipStream_t s, const NrcBuffers &nb, const float4 *in, uint32_t n, const uint32_t *count,
                    float4 *out, const float4 *aux, float4 *pathOut, int blocks, NrcSelfTrain st = NrcSelfTrain{});
// Rows (queries, training vertices) are rtx_nrc_input records, 4 float4 each:
// (pos, roughness), (dir, -), (normal, -), (albedo, -).
// Forward with saved activations, the loss gradient, the backward pass, the
// weight and grid gradients of rows [0, n) (n <= rowsCap); accumulates into
// nb.grads (which the caller zeroes) and writes the summed loss to *lossOut.
void launchNrcGradients(hipStream_t s, NrcBuffers nb, const float4 *in, const float4 *target, uint32_t n,
                        float invCount, float *lossOut, int cuCount);
void launchNrcAdam(hipStream_t s, NrcBuffers nb);
void launchNrcPackWeights(hipStream_t s, NrcBuffers nb);
// The moving average restarted at the parameters (then pack the weights).
hipError_t launchNrcEmaReset(hipStream_t s, NrcBuffers nb);
void launchNrcZeroGrads(hipStream_t s, NrcBuffers nb);
void launchNrcZero(hipStream_t s, float *p, size_t n); // a fill kernel

} // namespace rtx
