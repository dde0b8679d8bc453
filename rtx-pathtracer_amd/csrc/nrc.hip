// Learned radiance cache kernels: hash-grid encoding, the fused 5 x 64 MLP on
// BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation), its backward
// pass, weight and grid gradients, and Adam.
//
// MFMA tiling.  One wave owns 32 rows (queries) = two 16-row tiles; a layer
// is Y[rows][out] = X[rows][in] * W^T with the activations X as the A operand
// (row-major in LDS, 16 B per lane: lane i + 16 g reads row i, columns
// 8g..8g+7 of a 32-wide K step) and the weights as the B operand (stored
// out-major for the forward pass, so that lane j + 16 g reads output j,
// inputs 8g..8g+7).  A 16 x 16 FP32 result leaves lane j + 16 h holding
// column j of rows 4h..4h+3; it goes back to LDS as BF16 after the ReLU and
// becomes the next layer's A operand.  The weight gradients use the batch as
// the K dimension (activations and deltas transposed into LDS).
#include <hip/hip_runtime.h>

#include <cmath>

#include "nrc.h"

namespace rtx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kPitch = 72;        // bf16 per LDS row of a 64-wide tile (+8 spreads the banks)
constexpr int kBlock = 256;       // 4 waves
constexpr int kRowsPerWave = 32;
constexpr int kChunk = 256;       // batch rows per weight-gradient workgroup
constexpr int kChunkPitch = kChunk + 8;
constexpr int kWLayer = kNrcWidth * kNrcWidth; // bf16 per layer in the packed weight arrays

__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Hash-grid lookup of level l at the normalised position p: the trilinear
// blend of the 8 surrounding entries' two features.  Dense levels index
// x + (res+1)(y + (res+1) z); finer levels hash with the primes of Teschner
// et al. (2003) as instant-ngp does.  No contraction, so the weights and sums
// round the same way as a plain restatement.
__device__ void gridCorner(const NrcGrid &g, int l, const float p[3], int c, uint32_t &idx, float &w) {
#pragma clang fp contract(off)
    const float r = g.res[l];
    uint32_t ci[3];
    float wt = 1.0f;
    for (int a = 0; a < 3; ++a) {
        const float x = p[a] * r;
        float f = fminf(floorf(x), r - 1.0f);
        f = fmaxf(f, 0.0f);
        const float t = x - f;
        const uint32_t bit = (uint32_t)(c >> a) & 1u;
        ci[a] = (uint32_t)f + bit;
        wt = wt * (bit ? t : 1.0f - t);
    }
    if (g.dense[l]) {
        const uint32_t n1 = (uint32_t)r + 1u;
        idx = ci[0] + n1 * (ci[1] + n1 * ci[2]);
    } else {
        idx = (ci[0] ^ (ci[1] * 2654435761u) ^ (ci[2] * 805459861u)) & g.tableMask;
    }
    w = wt;
}

__device__ void normalisedPos(const NrcGrid &g, float4 in0, float p[3]) {
#pragma clang fp contract(off)
    p[0] = fminf(fmaxf((in0.x - g.lo[0]) * g.inv[0], 0.0f), 1.0f);
    p[1] = fminf(fmaxf((in0.y - g.lo[1]) * g.inv[1], 0.0f), 1.0f);
    p[2] = fminf(fmaxf((in0.z - g.lo[2]) * g.inv[2], 0.0f), 1.0f);
}

__device__ void encodeLevel(const NrcGrid &g, const uint32_t *table, int l, const float p[3], float &f0, float &f1) {
#pragma clang fp contract(off)
    const uint32_t *tab = table + (size_t)l * (g.tableMask + 1u);
    f0 = 0.0f;
    f1 = 0.0f;
    for (int c = 0; c < 8; ++c) {
        uint32_t idx;
        float w;
        gridCorner(g, l, p, c, idx, w);
        const uint32_t v = tab[idx];
        f0 = f0 + w * bf2f(v & 0xFFFFu);
        f1 = f1 + w * bf2f(v >> 16);
    }
}

// The network input of rows [row0, row0 + 128) into the block's activation
// tile (bf16, row-major, kPitch): 2 features per level, then direction,
// normal, roughness, albedo, a constant 1 (the layers have no bias) and
// zeros.  Rows past n are all zero.
__device__ void encodeTile(const NrcGrid &g, const uint32_t *table, const float4 *in, uint32_t row0, uint32_t n,
                           uint16_t *act) {
    const int t = threadIdx.x;
    const int r = t & (kNrcRowsPerBlock - 1);
    const uint32_t row = row0 + (uint32_t)r;
    const bool valid = row < n;
    uint16_t *dst = act + r * kPitch;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float4 in0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (valid) {
        in0 = in[4 * (size_t)row];
        normalisedPos(g, in0, p);
    }
    for (int l = t >> 7; l < kNrcMaxLevels; l += 2) {
        float f0 = 0.0f, f1 = 0.0f;
        if (valid && l < (int)g.levels) encodeLevel(g, table, l, p, f0, f1);
        dst[2 * l] = f2bf(f0);
        dst[2 * l + 1] = f2bf(f1);
    }
    if (t < kNrcRowsPerBlock) {
        float v[32];
        for (int k = 0; k < 32; ++k) v[k] = 0.0f;
        if (valid) {
            const float4 a = in[4 * (size_t)row + 1], b = in[4 * (size_t)row + 2], c = in[4 * (size_t)row + 3];
            v[0] = a.x, v[1] = a.y, v[2] = a.z;   // direction
            v[3] = b.x, v[4] = b.y, v[5] = b.z;   // normal
            v[6] = in0.w;                         // roughness
            v[7] = c.x, v[8] = c.y, v[9] = c.z;   // albedo
            v[10] = 1.0f;
        }
        for (int k = 0; k < 32; ++k) dst[32 + k] = f2bf(v[k]);
    }
}

// Copies `rows` rows of 64 bf16 (global, dense) into LDS rows of kPitch.
__device__ void loadRows(const uint16_t *src, int rows, uint16_t *dst) {
    const int chunks = rows * 8;
    for (int c = threadIdx.x; c < chunks; c += kBlock) {
        const int row = c >> 3, k = c & 7;
        *reinterpret_cast<uint4 *>(dst + row * kPitch + 8 * k) = *reinterpret_cast<const uint4 *>(src + row * 64 + 8 * k);
    }
}

// The MLP on one wave's 32 rows (act: the wave's tile, holding the input).
// Hidden layers overwrite the tile with their ReLU output; with SAVE every
// layer's input goes to acts[layer][row0 + .] as well.  Returns the output
// layer's fragments: out[r][e] = output column (lane & 15) of row
// 16 r + 4 (lane >> 4) + e.
template <bool SAVE>
__device__ void mlpWave(const uint16_t *wf, uint16_t *act, uint16_t *acts, uint32_t rowsCap, uint32_t row0,
                        f32x4 out[2]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    for (int layer = 0; layer < kNrcLayers; ++layer) {
        if (SAVE) {
            uint16_t *g = acts + ((size_t)layer * rowsCap + row0) * 64;
            for (int c = lane; c < kRowsPerWave * 8; c += 64) {
                const int row = c >> 3, k = c & 7;
                *reinterpret_cast<uint4 *>(g + row * 64 + 8 * k) = *reinterpret_cast<const uint4 *>(act + row * kPitch + 8 * k);
            }
        }
        const int nt = layer < kNrcHidden ? 4 : 1;
        const uint16_t *w = wf + layer * (kNrcWidth * kPitch);
        f32x4 acc[2][4];
        for (int r = 0; r < 2; ++r)
            for (int t = 0; t < 4; ++t) acc[r][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int s = 0; s < 2; ++s) {
            bf16x8 a[2];
            for (int r = 0; r < 2; ++r) a[r] = *reinterpret_cast<const bf16x8 *>(act + (16 * r + li) * kPitch + 32 * s + 8 * lg);
            for (int t = 0; t < nt; ++t) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(w + (16 * t + li) * kPitch + 32 * s + 8 * lg);
                for (int r = 0; r < 2; ++r) acc[r][t] = mfma(a[r], b, acc[r][t]);
            }
        }
        if (layer < kNrcHidden) {
            for (int r = 0; r < 2; ++r)
                for (int t = 0; t < 4; ++t)
                    for (int e = 0; e < 4; ++e)
                        act[(16 * r + 4 * lg + e) * kPitch + 16 * t + li] = f2bf(fmaxf(acc[r][t][e], 0.0f));
        } else {
            out[0] = acc[0][0];
            out[1] = acc[1][0];
        }
    }
}

// ------------------------------------------------------------- inference
// Rows are queries (4 float4 of input each).  out (optional): the raw RGB
// prediction per row.  pathOut (optional): the render's per-path radiance,
// to which max(prediction, 0) times the query's throughput aux[row].xyz is
// added for path aux[row].w.  n = *count when count is given.
template <bool TRAIN>
__global__ __launch_bounds__(kBlock) void kNrcForward(NrcBuffers nb, const float4 *in, uint32_t nHost,
                                                      const uint32_t *count, float4 *out, const float4 *aux,
                                                      float4 *pathOut, NrcSelfTrain st) {
    __shared__ __align__(16) uint16_t wf[kNrcLayers * kNrcWidth * kPitch];
    __shared__ __align__(16) uint16_t act[kNrcRowsPerBlock * kPitch];
    const uint32_t n = count ? min(*count, nHost) : nHost;
    // training runs whole 256-row chunks (rows past n are zeros), inference n rows
    const uint32_t rowsEnd = TRAIN ? ((n + 255u) & ~255u) : n;
    if ((uint32_t)blockIdx.x * kNrcRowsPerBlock >= rowsEnd) return;
    loadRows(nb.wFwd, kNrcLayers * kNrcWidth, wf);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    uint16_t *wact = act + wave * kRowsPerWave * kPitch;
    float *scratch = reinterpret_cast<float *>(wact); // 32 rows x 4 floats after the last layer
    for (uint32_t row0 = blockIdx.x * kNrcRowsPerBlock; row0 < rowsEnd; row0 += gridDim.x * kNrcRowsPerBlock) {
        __syncthreads(); // the previous tile's readers are done with act
        encodeTile(nb.grid, nb.gridBf16, in, row0, n, act);
        __syncthreads();
        f32x4 o[2];
        const uint32_t wrow0 = row0 + wave * kRowsPerWave;
        mlpWave<TRAIN>(wf, wact, nb.acts, nb.rowsCap, wrow0, o);
        // the output columns 0..2 of each row, through LDS to one lane per row
        if (li < 4)
            for (int r = 0; r < 2; ++r)
                for (int e = 0; e < 4; ++e) scratch[(16 * r + 4 * lg + e) * 4 + li] = o[r][e];
        if (lane < kRowsPerWave) {
            const uint32_t row = wrow0 + lane;
            const float4 rgb = make_float4(scratch[lane * 4], scratch[lane * 4 + 1], scratch[lane * 4 + 2], 0.0f);
            if (row < n) {
                if (TRAIN) {
                    reinterpret_cast<float4 *>(nb.pred)[row] = rgb;
                } else {
                    if (out) out[row] = rgb;
                    if (pathOut) {
                        const float4 a = aux[row];
                        const uint32_t pid = __float_as_uint(a.w);
                        float4 &po = pathOut[pid];
                        const float px = fmaxf(rgb.x, 0.0f), py = fmaxf(rgb.y, 0.0f), pz = fmaxf(rgb.z, 0.0f);
                        po.x += a.x * px;
                        po.y += a.y * py;
                        po.z += a.z * pz;
                        // a self-training suffix's end: the target of its training
                        // vertex gets the same term over the throughput there
                        if (st.target && (__float_as_uint(st.pathThr[pid].w) & st.endFlag)) {
                            const float4 tp = st.trainPath[2 * (size_t)pid];
                            float4 &t = st.target[__float_as_int(tp.w)];
                            t.x += tp.x != 0.0f ? a.x * px / tp.x : 0.0f;
                            t.y += tp.y != 0.0f ? a.y * py / tp.y : 0.0f;
                            t.z += tp.z != 0.0f ? a.z * pz / tp.z : 0.0f;
                        }
                    }
                }
            }
        }
    }
}

// --------------------------------------------------------------- backward
// Per wave of 32 rows: the loss gradient, then d loss / d pre-activation of
// every layer (dz, bf16) down to the encoded features (dFeat, fp32).
// Relative L2 loss (Mueller et al. 2021): mean over rows and channels of
// (p - t)^2 / (sg(p)^2 + 0.01); target.w = 0 marks a row without a target.
__global__ __launch_bounds__(kBlock) void kNrcBackward(NrcBuffers nb, const float4 *target, uint32_t n,
                                                       float invCount, float *lossOut) {
    __shared__ __align__(16) uint16_t wb[kNrcLayers * kNrcWidth * kPitch];
    __shared__ __align__(16) uint16_t delta[kNrcRowsPerBlock * kPitch];
    __shared__ float lossWave[4];
    loadRows(nb.wBwd, kNrcLayers * kNrcWidth, wb);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    uint16_t *d = delta + wave * kRowsPerWave * kPitch;
    const uint32_t row0 = blockIdx.x * kNrcRowsPerBlock + wave * kRowsPerWave;
    const uint32_t cap = nb.rowsCap;
    // output layer: dz5, 32 columns (K of the first product), zero past 3
    float lossSum = 0.0f;
    if (lane < kRowsPerWave) {
        const uint32_t row = row0 + lane;
        float g[3] = {0.0f, 0.0f, 0.0f};
        if (row < n) {
            const float4 p = reinterpret_cast<const float4 *>(nb.pred)[row];
            const float4 t = target[row];
            if (t.w != 0.0f) {
                const float pv[3] = {p.x, p.y, p.z}, tv[3] = {t.x, t.y, t.z};
                for (int c = 0; c < 3; ++c) {
                    const float den = pv[c] * pv[c] + 0.01f;
                    const float e = pv[c] - tv[c];
                    lossSum += e * e / den;
                    g[c] = 2.0f * e / den * invCount;
                }
            }
        }
        uint16_t *dr = d + lane * kPitch;
        for (int k = 0; k < 64; ++k) dr[k] = f2bf(k < 3 ? g[k] : 0.0f);
    }
    for (int off = 32; off > 0; off >>= 1) lossSum += __shfl_down(lossSum, off);
    if (lane == 0) lossWave[wave] = lossSum;
    __syncthreads(); // weights loaded; every wave's dz5 written
    if (threadIdx.x == 0) atomicAdd(lossOut, (lossWave[0] + lossWave[1] + lossWave[2] + lossWave[3]) * invCount);
    for (int layer = kNrcLayers - 1; layer >= 0; --layer) {
        // this layer's dz (rows x out) to global, for the weight gradients
        uint16_t *gdz = nb.dz + ((size_t)layer * cap + row0) * 64;
        for (int c = lane; c < kRowsPerWave * 8; c += 64) {
            const int row = c >> 3, k = c & 7;
            *reinterpret_cast<uint4 *>(gdz + row * 64 + 8 * k) = *reinterpret_cast<const uint4 *>(d + row * kPitch + 8 * k);
        }
        // d loss / d input of this layer = dz * W: K = outputs (32 for the padded output layer)
        const int ks = layer == kNrcLayers - 1 ? 1 : 2;
        const int nt = layer == 0 ? 2 : 4; // the first layer: only the encoded features (columns 0..31)
        const uint16_t *w = wb + layer * (kNrcWidth * kPitch);
        f32x4 acc[2][4];
        for (int r = 0; r < 2; ++r)
            for (int t = 0; t < 4; ++t) acc[r][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int s = 0; s < ks; ++s) {
            bf16x8 a[2];
            for (int r = 0; r < 2; ++r) a[r] = *reinterpret_cast<const bf16x8 *>(d + (16 * r + li) * kPitch + 32 * s + 8 * lg);
            for (int t = 0; t < nt; ++t) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(w + (16 * t + li) * kPitch + 32 * s + 8 * lg);
                for (int r = 0; r < 2; ++r) acc[r][t] = mfma(a[r], b, acc[r][t]);
            }
        }
        if (layer > 0) {
            // through the ReLU that produced this layer's input
            const uint16_t *ga = nb.acts + ((size_t)layer * cap + row0) * 64;
            for (int r = 0; r < 2; ++r)
                for (int t = 0; t < 4; ++t)
                    for (int e = 0; e < 4; ++e) {
                        const int row = 16 * r + 4 * lg + e, col = 16 * t + li;
                        const bool on = bf2f(ga[row * 64 + col]) > 0.0f;
                        d[row * kPitch + col] = f2bf(on ? acc[r][t][e] : 0.0f);
                    }
        } else {
            for (int r = 0; r < 2; ++r)
                for (int t = 0; t < 2; ++t)
                    for (int e = 0; e < 4; ++e) {
                        const uint32_t row = row0 + 16 * r + 4 * lg + e;
                        nb.dFeat[(size_t)row * 32 + 16 * t + li] = acc[r][t][e];
                    }
        }
    }
}

// ------------------------------------------------------ weight gradients
// dW[layer][out][in] partial over rows [chunk * 256, +256): sum over rows of
// dz[row][out] * act[row][in], with the rows as the MFMA's K dimension.
__global__ __launch_bounds__(kBlock) void kNrcWeightGrad(NrcBuffers nb) {
    __shared__ __align__(16) uint16_t aT[kNrcWidth * kChunkPitch];
    __shared__ __align__(16) uint16_t dT[kNrcWidth * kChunkPitch];
    const uint32_t chunk = blockIdx.x, layer = blockIdx.y;
    const uint32_t cap = nb.rowsCap;
    const uint16_t *ga = nb.acts + ((size_t)layer * cap + chunk * kChunk) * 64;
    const uint16_t *gd = nb.dz + ((size_t)layer * cap + chunk * kChunk) * 64;
    const int outCols = layer == kNrcLayers - 1 ? kNrcOutPad : kNrcWidth;
    for (int c = threadIdx.x; c < kChunk * 8; c += kBlock) {
        const int row = c >> 3, k = c & 7;
        const uint4 va = *reinterpret_cast<const uint4 *>(ga + row * 64 + 8 * k);
        const uint16_t *pa = reinterpret_cast<const uint16_t *>(&va);
        for (int e = 0; e < 8; ++e) aT[(8 * k + e) * kChunkPitch + row] = pa[e];
        if (8 * k < outCols) {
            const uint4 vd = *reinterpret_cast<const uint4 *>(gd + row * 64 + 8 * k);
            const uint16_t *pd = reinterpret_cast<const uint16_t *>(&vd);
            for (int e = 0; e < 8; ++e) dT[(8 * k + e) * kChunkPitch + row] = pd[e];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int nOut = outCols / 16, tiles = 4 * nOut;
    float *part = nb.partials + ((size_t)chunk * kNrcLayers + layer) * (kNrcWidth * kNrcWidth);
    for (int tile = wave; tile < tiles; tile += 4) {
        const int ti = tile / nOut, to = tile % nOut; // input rows tile, output columns tile
        f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int s = 0; s < kChunk / 32; ++s) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(aT + (16 * ti + li) * kChunkPitch + 32 * s + 8 * lg);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(dT + (16 * to + li) * kChunkPitch + 32 * s + 8 * lg);
            acc = mfma(a, b, acc);
        }
        // acc[e]: input 16 ti + 4 lg + e, output 16 to + li; stored [in][out]
        for (int e = 0; e < 4; ++e) part[(16 * ti + 4 * lg + e) * kNrcWidth + 16 * to + li] = acc[e];
    }
}

// Sums the partials of every chunk in chunk order into the MLP gradients
// (out-major, after the grid's parameters).
__global__ __launch_bounds__(256) void kNrcReduceGrad(NrcBuffers nb, uint32_t chunks) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint32_t)kNrcMlpParams) return;
    const uint32_t layer = i / (kNrcWidth * kNrcWidth), rem = i % (kNrcWidth * kNrcWidth);
    const uint32_t o = rem / kNrcWidth, k = rem % kNrcWidth;
    if (layer == kNrcLayers - 1 && o >= 3) return; // padding rows of the output layer stay zero
    float s = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c)
        s += nb.partials[((size_t)c * kNrcLayers + layer) * (kNrcWidth * kNrcWidth) + k * kNrcWidth + o];
    nb.grads[nb.gridParams + i] += s;
}

// d loss / d table entries: each row's feature gradients spread over the 8
// corners of every level with the trilinear weights (FP32 atomics).
__global__ __launch_bounds__(256) void kNrcGridGrad(NrcBuffers nb, const float4 *in, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t row = i / kNrcMaxLevels, l = i % kNrcMaxLevels;
    if (row >= n || l >= nb.grid.levels) return;
    float p[3];
    normalisedPos(nb.grid, in[4 * (size_t)row], p);
    const float g0 = nb.dFeat[(size_t)row * 32 + 2 * l], g1 = nb.dFeat[(size_t)row * 32 + 2 * l + 1];
    if (g0 == 0.0f && g1 == 0.0f) return;
    float *gg = nb.grads + (size_t)l * (nb.grid.tableMask + 1u) * 2;
    for (int c = 0; c < 8; ++c) {
        uint32_t idx;
        float w;
        gridCorner(nb.grid, (int)l, p, c, idx, w);
        atomicAdd(gg + 2 * (size_t)idx, w * g0);
        atomicAdd(gg + 2 * (size_t)idx + 1, w * g1);
    }
}

// ----------------------------------------------------------------- Adam
// Kingma and Ba (2015), bias-corrected: p -= lr m_hat / (sqrt(v_hat) + eps).
// Clears the gradient and refreshes the table's bf16 copy; then the moving
// average inference reads, e = d e + (1 - d) p from e = 0, whose table copy
// is bias-corrected the same way: e / (1 - d^t) (c3) after t steps.
__global__ __launch_bounds__(256) void kNrcAdam(NrcBuffers nb, size_t total, float c1, float c2, float c3) {
    uint16_t *tab = reinterpret_cast<uint16_t *>(nb.gridBf16);
    uint16_t *tabEma = reinterpret_cast<uint16_t *>(nb.gridBf16Ema);
    const bool ema = nb.ema != nb.params;
    const float d = nb.emaDecay;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const float g = nb.grads[i];
        const float m = nb.beta1 * nb.adamM[i] + (1.0f - nb.beta1) * g;
        const float v = nb.beta2 * nb.adamV[i] + (1.0f - nb.beta2) * g * g;
        nb.adamM[i] = m;
        nb.adamV[i] = v;
        const float p = nb.params[i] - nb.lr * (m * c1) / (sqrtf(v * c2) + nb.eps);
        nb.params[i] = p;
        nb.grads[i] = 0.0f;
        if (i < nb.gridParams) tab[i] = f2bf(p);
        if (ema) {
            const float e = d * nb.ema[i] + (1.0f - d) * p;
            nb.ema[i] = e;
            if (i < nb.gridParams) tabEma[i] = f2bf(e * c3);
        }
    }
}

// The MLP's bf16 copies: forward [layer][out][in], backward [layer][in][out]
// (the output layer's 16 padded outputs extend to 32 zero columns there),
// and the table's bf16 copy when all is set (after rtx_nrc_set_params).
// The moving average's copies are ema * c3, or the parameters themselves
// before the first step (c3 = 0).
__global__ __launch_bounds__(256) void kNrcPack(NrcBuffers nb, int withTable, float c3) {
    const float *emaSrc = c3 > 0.0f ? nb.ema : nb.params;
    const float emaScale = c3 > 0.0f ? c3 : 1.0f;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)kNrcMlpParams) {
        const uint32_t layer = (uint32_t)(i / (kNrcWidth * kNrcWidth)), rem = (uint32_t)(i % (kNrcWidth * kNrcWidth));
        const uint32_t o = rem / kNrcWidth, k = rem % kNrcWidth;
        const uint16_t v = f2bf(nb.params[nb.gridParams + i]);
        nb.wFwd[layer * kWLayer + o * kNrcWidth + k] = v;
        nb.wBwd[layer * kWLayer + k * kNrcWidth + o] = v;
        if (nb.ema != nb.params)
            nb.wFwdEma[layer * kWLayer + o * kNrcWidth + k] = f2bf(emaSrc[nb.gridParams + i] * emaScale);
    }
    if (withTable)
        for (size_t j = i; j < nb.gridParams; j += (size_t)gridDim.x * 256) {
            reinterpret_cast<uint16_t *>(nb.gridBf16)[j] = f2bf(nb.params[j]);
            if (nb.ema != nb.params) reinterpret_cast<uint16_t *>(nb.gridBf16Ema)[j] = f2bf(emaSrc[j] * emaScale);
        }
}

__global__ __launch_bounds__(256) void kNrcZero(float *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0.0f;
}

} // namespace

void launchNrcInfer(hipStream_t s, const NrcBuffers &nb, const float4 *in, uint32_t n, const uint32_t *count,
                    float4 *out, const float4 *aux, float4 *pathOut, int blocks, NrcSelfTrain st) {
    if (n == 0) return;
    const int tiles = (int)((n + kNrcRowsPerBlock - 1) / kNrcRowsPerBlock);
    NrcBuffers v = nb; // inference reads the moving average's copies
    v.gridBf16 = nb.gridBf16Ema;
    v.wFwd = nb.wFwdEma;
    hipLaunchKernelGGL(kNrcForward<false>, dim3(tiles < blocks ? tiles : blocks), dim3(kBlock), 0, s, v, in, n,
                       count, out, aux, pathOut, st);
}

void launchNrcGradients(hipStream_t s, NrcBuffers nb, const float4 *in, const float4 *target, uint32_t n,
                        float invCount, float *lossOut, int cuCount) {
    if (n == 0) return;
    // whole 256-row chunks: the rows past n get zero inputs and zero deltas,
    // so the weight gradients' last chunk reads zeros, not stale rows
    const uint32_t chunks = (n + kChunk - 1) / kChunk;
    const uint32_t tiles = chunks * (kChunk / kNrcRowsPerBlock);
    hipLaunchKernelGGL(kNrcForward<true>, dim3(tiles), dim3(kBlock), 0, s, nb, in, n, (const uint32_t *)nullptr,
                       (float4 *)nullptr, (const float4 *)nullptr, (float4 *)nullptr, NrcSelfTrain{});
    hipLaunchKernelGGL(kNrcBackward, dim3(tiles), dim3(kBlock), 0, s, nb, target, n, invCount, lossOut);
    hipLaunchKernelGGL(kNrcWeightGrad, dim3(chunks, kNrcLayers), dim3(kBlock), 0, s, nb);
    hipLaunchKernelGGL(kNrcReduceGrad, dim3((kNrcMlpParams + 255) / 256), dim3(256), 0, s, nb, chunks);
    const uint32_t pairs = n * kNrcMaxLevels;
    hipLaunchKernelGGL(kNrcGridGrad, dim3((pairs + 255) / 256), dim3(256), 0, s, nb, in, n);
    (void)cuCount;
}

void launchNrcAdam(hipStream_t s, NrcBuffers nb) {
    const size_t total = nb.gridParams + kNrcMlpParams;
    const double t = (double)nb.step;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)nb.beta1, t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)nb.beta2, t)));
    const float c3 = nb.emaDecay > 0.0f ? (float)(1.0 / (1.0 - std::pow((double)nb.emaDecay, t))) : 1.0f;
    hipLaunchKernelGGL(kNrcAdam, dim3(4096), dim3(256), 0, s, nb, total, c1, c2, c3);
    hipLaunchKernelGGL(kNrcPack, dim3((kNrcMlpParams + 255) / 256), dim3(256), 0, s, nb, 0, c3);
}

void launchNrcPackWeights(hipStream_t s, NrcBuffers nb) {
    const double t = (double)nb.step;
    const float c3 = nb.emaDecay > 0.0f && nb.step > 0 ? (float)(1.0 / (1.0 - std::pow((double)nb.emaDecay, t))) : 0.0f;
    hipLaunchKernelGGL(kNrcPack, dim3(4096), dim3(256), 0, s, nb, 1, c3);
}

hipError_t launchNrcEmaReset(hipStream_t s, NrcBuffers nb) {
    if (nb.ema != nb.params) launchNrcZero(s, nb.ema, nb.gridParams + kNrcMlpParams);
    launchNrcPackWeights(s, nb); // nb.step is 0 here: the copies are the parameters'
    return hipGetLastError();
}

void launchNrcZero(hipStream_t s, float *p, size_t n) {
    hipLaunchKernelGGL(kNrcZero, dim3(n < 256 * 4096 ? (unsigned)((n + 255) / 256) : 4096u), dim3(256), 0, s, p, n);
}

void launchNrcZeroGrads(hipStream_t s, NrcBuffers nb) {
    hipLaunchKernelGGL(kNrcZero, dim3(4096), dim3(256), 0, s, nb.grads, nb.gridParams + kNrcMlpParams);
}

} // namespace rtx
