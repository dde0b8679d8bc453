// Path-guiding update on the device (gfx950): replaces the host loop of
// PathGuiding::update (src/PathGuiding.cpp:276-312) over lightpmm's MAP-EM
// (external/lightpmm/include/pmm/VMMFactory.h) and SampleCollector's
// download-and-sort (src/SampleCollector.cpp:76-131).
//
//   kCompact*   valid samples of the frame in slot order (region, slot)
//   kRadix*     stable LSD radix sort by region, 6-bit digits (the reference
//               sorts with an unstable parallel std::sort)
//   kSegments   first sample of every region
//   kPrefit     per region: parallax mean = box centre, lobes repositioned
//               (PathGuiding::preFit, 350-391)
//   kGather     samples moved to the mean position (preFit) and packed
//   kRegionEM   one workgroup per region: fit on the first update, updateFit
//               after (VMMFactory.h:157-216, 305-362); E-step in FP32 with
//               lightpmm's fastexp, statistics reduced in FP64
//   kDistances  harmonic-mean lobe distances (incrementaldistance.h:58-104)
//   kToVmm      VMM_Theta for rendering (syncPMMsToVMM_Thetas, 117-139)
// Component split and merge (postFit's splitAll/mergeAll) are not built.
#include <hip/hip_runtime.h>

#include "dev_math.h"
#include "internal.h"

namespace rtx {

constexpr int kK = RTX_MAX_DISTRIBUTIONS;
constexpr int kSortBlock = 1024;
constexpr int kDigitBits = 6;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kEmBlock = 256;

// fastapprox fastexp (P. Mineiro, 2011), as lightpmm evaluates exp
// (pmm-vcl.h:158-184): 2^p from the float bit pattern with a rational
// correction of the fraction.
RTX_DI float fastexpf(float p) {
    const float x = 1.442695040f * p;
    const float clipp = fmaxf(-126.0f, x);
    const float w = truncf(clipp);
    const float z = clipp - w + 1.0f;
    const float v = 8388608.0f * (clipp + 121.2740575f + 27.7280233f / (4.84252568f - z) - 1.49012907f * z);
    return __int_as_float((int)rintf(v));
}

RTX_DI float kappaToMeanCosine(float kappa) { // VMFKernel.h:46-56
    if (kappa > 5.0f) return 1.0f - 1.0f / kappa;
    if (!(kappa > 0.0f)) return 0.0f;
    return 1.0f / tanhf(kappa) - 1.0f / kappa;
}
RTX_DI float meanCosineToKappa(float r) { return (r * 3.0f - r * r * r) / (1.0f - r * r); } // Banerjee 2005

RTX_DI void calNorm(GuideRegion &g, int k) { // VMFKernel.h:651-659
    g.e2k[k] = fastexpf(-2.0f * g.kappa[k]);
    g.norm[k] = g.kappa[k] > 0.0f ? g.kappa[k] / (2.0f * kPi * (1.0f - g.e2k[k])) : 1.0f / (4.0f * kPi);
}

// ------------------------------------------------------------ block scans
template <int BLOCK> RTX_DI uint32_t scanExclusive(uint32_t v, uint32_t *sh, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            uint32_t t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[BLOCK / 64] = acc;
    }
    __syncthreads();
    const uint32_t r = sh[wave] + x - v;
    total = sh[BLOCK / 64];
    __syncthreads();
    return r;
}

// One block: exclusive scan of n values in place, base added; returns nothing.
__global__ __launch_bounds__(1024) void kScanInPlace(uint32_t *a, uint32_t n, uint32_t *totalOut) {
    __shared__ uint32_t sh[1024 / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? a[i] : 0u;
        uint32_t total;
        const uint32_t ex = scanExclusive<1024>(v, sh, total);
        if (i < n) a[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && totalOut) *totalOut = carry;
}

// ---------------------------------------------------------- compaction
RTX_DI bool validSlot(const rtx_directional_data *dd, uint32_t i, uint32_t slots, uint32_t numRegions) {
    return i < slots && dd[i].flags < numRegions;
}

__global__ __launch_bounds__(kSortBlock) void kCompactCount(const rtx_directional_data *dd, uint32_t slots,
                                                           uint32_t numRegions, uint32_t *blockSums) {
    __shared__ uint32_t sh[kSortBlock / 64 + 1];
    uint32_t total;
    scanExclusive<kSortBlock>(validSlot(dd, blockIdx.x * kSortBlock + threadIdx.x, slots, numRegions) ? 1u : 0u, sh,
                              total);
    if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kSortBlock) void kCompactScatter(const rtx_directional_data *dd, uint32_t slots,
                                                             uint32_t numRegions, const uint32_t *blockOffsets,
                                                             uint32_t *keys, uint32_t *vals) {
    __shared__ uint32_t sh[kSortBlock / 64 + 1];
    const uint32_t i = blockIdx.x * kSortBlock + threadIdx.x;
    const bool v = validSlot(dd, i, slots, numRegions);
    uint32_t total;
    const uint32_t pos = blockOffsets[blockIdx.x] + scanExclusive<kSortBlock>(v ? 1u : 0u, sh, total);
    if (v) {
        keys[pos] = dd[i].flags;
        vals[pos] = i;
    }
}

// ------------------------------------------------------------ radix sort
__global__ __launch_bounds__(kSortBlock) void kRadixHist(const uint32_t *keys, const uint32_t *count, int shift,
                                                        uint32_t nblk, uint32_t *hist) {
    __shared__ uint32_t h[kDigits];
    if (threadIdx.x < kDigits) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = *count;
    const uint32_t i = blockIdx.x * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kDigits - 1)], 1u);
    __syncthreads();
    if (threadIdx.x < kDigits) hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: an item's place among the block's items of its digit is
// the count in earlier waves plus the count in earlier lanes of its wave
// (lanes of a digit matched with one ballot per digit bit).
__global__ __launch_bounds__(kSortBlock) void kRadixScatter(const uint32_t *keysIn, const uint32_t *valsIn,
                                                           uint32_t *keysOut, uint32_t *valsOut,
                                                           const uint32_t *count, int shift, uint32_t nblk,
                                                           const uint32_t *hist) {
    constexpr int kWaves = kSortBlock / 64;
    __shared__ uint32_t cnt[kWaves][kDigits];
    for (int t = threadIdx.x; t < kWaves * kDigits; t += kSortBlock) (&cnt[0][0])[t] = 0;
    __syncthreads();
    const uint32_t n = *count;
    const uint32_t i = blockIdx.x * kSortBlock + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? keysIn[i] : 0u;
    const uint32_t d = (key >> shift) & (kDigits - 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kDigitBits; ++b) {
        const uint64_t bb = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint64_t below = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
    const uint32_t rankInWave = (uint32_t)__popcll(below);
    if (valid && rankInWave == 0) cnt[wave][d] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x < kDigits) { // exclusive prefix over the waves, per digit
        uint32_t acc = 0;
        for (int w = 0; w < kWaves; ++w) {
            uint32_t t = cnt[w][threadIdx.x];
            cnt[w][threadIdx.x] = acc;
            acc += t;
        }
    }
    __syncthreads();
    if (valid) {
        const uint32_t pos = hist[d * nblk + blockIdx.x] + cnt[wave][d] + rankInWave;
        keysOut[pos] = key;
        valsOut[pos] = valsIn[i];
    }
}

__global__ __launch_bounds__(256) void kSegmentCount(const uint32_t *keys, const uint32_t *count, uint32_t *regionCount) {
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) atomicAdd(&regionCount[keys[i]], 1u);
}

__global__ __launch_bounds__(256) void kZero(uint32_t *a, uint32_t n) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a[i] = 0;
}

// --------------------------------------------------------------- preFit
// PathGuiding::preFit (src/PathGuiding.cpp:350-391) for regions with samples:
// the parallax mean is the box centre; after the first update the lobes
// with a finite distance keep pointing at their sources
// (incrementaldistance.h:106-128).
__global__ __launch_bounds__(64) void kPrefit(GuideFitBuffers g, bool parallax, bool firstFit) {
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    if (r >= g.numRegions || !parallax) return;
    if (g.regionStart[r + 1] == g.regionStart[r]) return;
    GuideRegion &R = g.regions[r];
    const rtx_aabb b = g.boxes[r];
    for (int c = 0; c < 3; ++c) {
        R.lastMean[c] = R.mean[c];
        R.mean[c] = b.min[c] + 0.5f * (b.max[c] - b.min[c]);
    }
    if (firstFit) return;
    const float shift[3] = {R.lastMean[0] - R.mean[0], R.lastMean[1] - R.mean[1], R.lastMean[2] - R.mean[2]};
    for (int k = 0; k < R.K; ++k) {
        if (isinf(R.dist[k]) || R.distW[k] <= 0.0f) continue;
        V3 v = v3(R.mu[k][0], R.mu[k][1], R.mu[k][2]) * R.dist[k] + v3(shift[0], shift[1], shift[2]);
        const float len = length(v);
        R.dist[k] = len;
        v = v / len;
        R.mu[k][0] = v.x;
        R.mu[k][1] = v.y;
        R.mu[k][2] = v.z;
    }
}

__global__ __launch_bounds__(256) void kGather(GuideFitBuffers g, const rtx_directional_data *dd, bool parallax) {
    const uint32_t n = g.counters[0];
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const rtx_directional_data x = dd[g.valsA[i]];
        V3 dir = v3(x.direction);
        float dist = x.distance;
        if (parallax) {
            const GuideRegion &R = g.regions[g.keysA[i]];
            if (dist > 0.0f) {
                V3 v = v3(x.position) + dist * dir - v3(R.mean[0], R.mean[1], R.mean[2]);
                dist = length(v);
                dir = v / dist;
            } else {
                dist = INFINITY;
            }
        }
        g.samples[2 * i] = f4(dir, x.weight);
        g.samples[2 * i + 1] = make_float4(dist, 0.0f, 0.0f, 0.0f);
    }
}

// -------------------------------------------------------------------- EM
struct EmShared {
    float w[kK], kappa[kK], norm[kK], mu[kK][3];
    double stat[kEmBlock / 64][kK * 4 + 2]; // per wave: W_k, rW_k, sumWeight, logLikelihood
    double tot[kK * 4 + 2];
    int stop;
};

RTX_DI double waveSumD(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// E-step over the region's samples with the model in LDS; leaves the
// block totals in E.tot.
RTX_DI void eStep(EmShared &E, int K, const float4 *smp, uint32_t n) {
    float W[kK], RW[kK][3];
    for (int k = 0; k < kK; ++k) W[k] = RW[k][0] = RW[k][1] = RW[k][2] = 0.0f;
    float sumW = 0.0f;
    double ll = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += kEmBlock) {
        const float4 a4 = smp[2 * j];
        const V3 d = v3(a4);
        float a[kK];
        float total = 0.0f;
        for (int k = 0; k < K; ++k) {
            const float c = E.mu[k][0] * d.x + E.mu[k][1] * d.y + E.mu[k][2] * d.z;
            a[k] = E.w[k] * (E.norm[k] * fastexpf(E.kappa[k] * fminf(c - 1.0f, 0.0f)));
            total += a[k];
        }
        if (!(total > 1e-8f)) continue; // SoftAssignmentWeights::valid
        const float inv = 1.0f / total;
        for (int k = 0; k < K; ++k) {
            const float aw = a[k] * inv * a4.w;
            RW[k][0] += d.x * aw;
            RW[k][1] += d.y * aw;
            RW[k][2] += d.z * aw;
            W[k] += aw;
        }
        sumW += a4.w;
        ll += (double)a4.w * (double)logf(total);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < K; ++k) {
        double v0 = waveSumD(W[k]), v1 = waveSumD(RW[k][0]), v2 = waveSumD(RW[k][1]), v3_ = waveSumD(RW[k][2]);
        if (lane == 0) {
            E.stat[wave][4 * k] = v0;
            E.stat[wave][4 * k + 1] = v1;
            E.stat[wave][4 * k + 2] = v2;
            E.stat[wave][4 * k + 3] = v3_;
        }
    }
    double s = waveSumD((double)sumW), l = waveSumD(ll);
    if (lane == 0) {
        E.stat[wave][4 * kK] = s;
        E.stat[wave][4 * kK + 1] = l;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 4 * kK + 2; t += kEmBlock) {
        double acc = 0.0;
        for (int w = 0; w < kEmBlock / 64; ++w) acc += E.stat[w][t];
        E.tot[t] = acc;
    }
    __syncthreads();
}

RTX_DI float applyWeightPrior(const EmProps &P, float w, int K) {
    const float inv = P.vPrior * (float)K + 1.0f, nrm = 1.0f / inv, prior = P.vPrior * nrm;
    return w * nrm + prior;
}
RTX_DI float removeWeightPrior(const EmProps &P, float w, int K) {
    const float inv = P.vPrior * (float)K + 1.0f, nrm = 1.0f / inv, prior = P.vPrior * nrm;
    return w > prior ? (w - prior) * inv : 0.0f;
}

__global__ __launch_bounds__(kEmBlock) void kRegionEM(GuideFitBuffers g, EmProps P, int update) {
    __shared__ EmShared E;
    __shared__ float oldW[kK], oldRW[kK][3];
    const uint32_t r = blockIdx.x;
    const uint32_t s0 = g.regionStart[r], n = g.regionStart[r + 1] - s0;
    if (n == 0) return;
    GuideRegion &R = g.regions[r];
    const int K = R.K;
    const float4 *smp = g.samples + 2 * (size_t)s0;
    if (threadIdx.x < (unsigned)kK) {
        const int k = threadIdx.x;
        E.w[k] = R.w[k];
        E.kappa[k] = R.kappa[k];
        E.norm[k] = R.norm[k];
        E.mu[k][0] = R.mu[k][0];
        E.mu[k][1] = R.mu[k][1];
        E.mu[k][2] = R.mu[k][2];
        oldW[k] = oldRW[k][0] = oldRW[k][1] = oldRW[k][2] = 0.0f;
        if (update && k < K) { // computeSufficentStatsFromMixture, VMMFactory.h:543-555
            const float Wk = removeWeightPrior(P, R.w[k], K) * R.sampleWeight;
            const float nk = (float)R.totalNumSamples * R.w[k];
            const float a = (R.r[k] * (nk + P.rPriorWeight) - P.rPrior * P.rPriorWeight) / nk * Wk;
            oldW[k] = Wk;
            oldRW[k][0] = R.mu[k][0] * a;
            oldRW[k][1] = R.mu[k][1] * a;
            oldRW[k][2] = R.mu[k][2] * a;
        }
    }
    if (threadIdx.x == 0) E.stop = 0;
    __syncthreads();
    const uint32_t maxItr = n > (uint32_t)(2 * K) ? P.maxItr : 0u;
    const float maxMeanCosine = kappaToMeanCosine(P.maxKappa);
    // model totals and convergence state live in thread 0
    float lastLL = 0.0f, absInvLastLL = 0.0f;
    uint32_t i = 0;
    for (; i < maxItr; ++i) {
        eStep(E, K, smp, n);
        if (threadIdx.x == 0 && i == 0) {
            const float sumWeight = (float)E.tot[4 * kK];
            if (update) {
                R.sampleWeight += sumWeight;
                R.numSamples += (float)n;
                R.totalNumSamples += n;
            } else {
                R.sampleWeight = sumWeight;
                R.numSamples = (float)n;
                R.totalNumSamples = n;
            }
        }
        __syncthreads();
        if (threadIdx.x < (unsigned)K) { // parameterUpdate, VMMFactory.h:436-456
            const int k = threadIdx.x;
            const float Wk = (float)E.tot[4 * k] + oldW[k];
            const float rw0 = (float)E.tot[4 * k + 1] + oldRW[k][0], rw1 = (float)E.tot[4 * k + 2] + oldRW[k][1],
                        rw2 = (float)E.tot[4 * k + 3] + oldRW[k][2];
            const float mixtureWeight = applyWeightPrior(P, Wk / R.sampleWeight, K);
            const float rLength = sqrtf(rw0 * rw0 + rw1 * rw1 + rw2 * rw2);
            if (rLength > 0.0f) {
                R.mu[k][0] = rw0 / rLength;
                R.mu[k][1] = rw1 / rLength;
                R.mu[k][2] = rw2 / rLength;
            }
            const float avgCosine = rLength / Wk;
            const float nk = (float)R.totalNumSamples * mixtureWeight;
            float withPrior = (avgCosine * nk + P.rPrior * P.rPriorWeight) / (nk + P.rPriorWeight);
            withPrior = isnan(withPrior) ? maxMeanCosine : fminf(withPrior, maxMeanCosine);
            const float kap = meanCosineToKappa(withPrior);
            R.w[k] = mixtureWeight;
            const bool small = kap < 1e-3f;
            R.kappa[k] = small ? 0.0f : kap;
            R.r[k] = small ? 0.0f : withPrior;
            calNorm(R, k);
            E.w[k] = R.w[k];
            E.kappa[k] = R.kappa[k];
            E.norm[k] = R.norm[k];
            E.mu[k][0] = R.mu[k][0];
            E.mu[k][1] = R.mu[k][1];
            E.mu[k][2] = R.mu[k][2];
        }
        if (threadIdx.x == 0 && i >= P.minItr) {
            const float ll = (float)E.tot[4 * kK + 1];
            if (i > P.minItr && (ll - lastLL) * absInvLastLL < P.relLogLikelihoodThreshold) E.stop = 1;
            lastLL = ll;
            absInvLastLL = 1.0f / fabsf(ll);
        }
        __syncthreads();
        if (E.stop) break;
    }
    if (threadIdx.x == 0) R.numEMIterations += i + (E.stop ? 0u : 0u);
}

// incrementaldistance.h:58-104
__global__ __launch_bounds__(kEmBlock) void kDistances(GuideFitBuffers g) {
    __shared__ double part[kEmBlock / 64][2 * kK];
    const uint32_t r = blockIdx.x;
    const uint32_t s0 = g.regionStart[r], n = g.regionStart[r + 1] - s0;
    if (n == 0) return;
    GuideRegion &R = g.regions[r];
    const int K = R.K;
    const float4 *smp = g.samples + 2 * (size_t)s0;
    float bw[kK], bd[kK];
    for (int k = 0; k < kK; ++k) bw[k] = bd[k] = 0.0f;
    for (uint32_t j = threadIdx.x; j < n; j += kEmBlock) {
        const float4 a4 = smp[2 * j];
        const float dist = smp[2 * j + 1].x;
        if (!(dist > 0.0f)) continue;
        const V3 d = v3(a4);
        float p[kK], wp[kK], total = 0.0f;
        for (int k = 0; k < K; ++k) {
            const float c = R.mu[k][0] * d.x + R.mu[k][1] * d.y + R.mu[k][2] * d.z;
            p[k] = R.norm[k] * fastexpf(R.kappa[k] * fminf(c - 1.0f, 0.0f));
            wp[k] = R.w[k] * p[k];
            total += wp[k];
        }
        if (!(total > 1e-8f)) continue;
        const float f = a4.w / total;
        for (int k = 0; k < K; ++k) {
            const float v = wp[k] * p[k] * f;
            bw[k] += v;
            bd[k] += v / dist;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < K; ++k) {
        double a = waveSumD(bw[k]), b = waveSumD(bd[k]);
        if (lane == 0) {
            part[wave][2 * k] = a;
            part[wave][2 * k + 1] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)K) {
        const int k = threadIdx.x;
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kEmBlock / 64; ++w) {
            a += part[w][2 * k];
            b += part[w][2 * k + 1];
        }
        const float nw = R.distW[k] + (float)a;
        R.dist[k] = nw / (R.distW[k] / R.dist[k] + (float)b);
        R.distW[k] = nw;
    }
}

// syncPMMsToVMM_Thetas (src/PathGuiding.cpp:117-139) with VMF_Theta::setK
// (src/PathGuiding.h:44-50)
__global__ __launch_bounds__(64) void kToVmm(GuideFitBuffers g, bool parallax, rtx_vmm_theta *out) {
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    if (r >= g.numRegions) return;
    const GuideRegion &R = g.regions[r];
    rtx_vmm_theta v;
    v.usedDistributions = R.K;
    v.meanPosition[0] = v.meanPosition[1] = v.meanPosition[2] = 0.0f;
    for (int k = 0; k < kK; ++k) {
        rtx_vmf_theta &t = v.thetas[k];
        const bool on = k < R.K;
        v.pi[k] = on ? R.w[k] : 0.0f;
        float kap = on ? R.kappa[k] : 0.0f;
        kap = kap < 1e-3f ? 0.0f : kap;
        t.mu[0] = on ? R.mu[k][0] : 0.0f;
        t.mu[1] = on ? R.mu[k][1] : 0.0f;
        t.mu[2] = on ? R.mu[k][2] : 1.0f;
        t.k = kap;
        t.norm = kap / (2.0f * kPi * (1.0f - expf(-2.0f * kap)));
        t.eMin2K = expf(-2.0f * kap);
        t.distance = -1.0f;
        t.target[0] = t.target[1] = t.target[2] = 0.0f;
    }
    if (parallax) {
        for (int c = 0; c < 3; ++c) v.meanPosition[c] = R.mean[c];
        for (int k = 0; k < kK; ++k) {
            float d = k < R.K ? R.dist[k] : INFINITY;
            d = isinf(d) ? -1.0f : d;
            v.thetas[k].distance = d;
            for (int c = 0; c < 3; ++c) v.thetas[k].target[c] = v.meanPosition[c] + d * v.thetas[k].mu[c];
        }
    }
    out[r] = v;
}

// VMMFactory::initialize (VMMFactory.h:133-147): K lobes of kappa initKappa at
// spherical-Fibonacci directions (594-632), equal weights; no samples yet.
__global__ __launch_bounds__(64) void kGuideInit(GuideFitBuffers g, uint32_t K, float initKappa) {
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    if (r >= g.numRegions) return;
    GuideRegion R;
    R.K = (int32_t)K;
    R.numEMIterations = 0;
    R.sampleWeight = R.numSamples = 0.0f;
    R.totalNumSamples = 0;
    const float gr = 1.618033988749895f;
    for (int n = 0; n < kK; ++n) {
        R.dist[n] = INFINITY;
        R.distW[n] = 0.0f;
        if ((uint32_t)n < K) {
            const float phi = 2.0f * kPi * ((float)n / gr);
            const float z = 1.0f - ((2.0f * n + 1.0f) / (float)K);
            const float theta = acosf(z);
            const float st = sinf(theta);
            R.mu[n][0] = st * cosf(phi);
            R.mu[n][1] = st * sinf(phi);
            R.mu[n][2] = cosf(theta);
            R.w[n] = 1.0f / (float)K;
            R.kappa[n] = initKappa;
            R.r[n] = kappaToMeanCosine(initKappa);
        } else {
            R.mu[n][0] = R.mu[n][1] = 0.0f;
            R.mu[n][2] = 1.0f;
            R.w[n] = R.kappa[n] = R.r[n] = 0.0f;
        }
        calNorm(R, n);
    }
    for (int c = 0; c < 3; ++c) R.mean[c] = R.lastMean[c] = 0.0f;
    g.regions[r] = R;
}

// ------------------------------------------------------------- launchers
static uint32_t blocksOf(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launchGuideInit(hipStream_t s, GuideFitBuffers g, uint32_t numInitialComponents, float initKappa) {
    hipLaunchKernelGGL(kGuideInit, dim3(blocksOf(g.numRegions, 64)), dim3(64), 0, s, g, numInitialComponents,
                       initKappa);
}

void launchGuideToVmm(hipStream_t s, GuideFitBuffers g, bool parallax, rtx_vmm_theta *vmmsOut) {
    hipLaunchKernelGGL(kToVmm, dim3(blocksOf(g.numRegions, 64)), dim3(64), 0, s, g, parallax, vmmsOut);
}

void launchGuideUpdate(hipStream_t s, GuideFitBuffers g, const rtx_directional_data *dd, uint32_t slots,
                       EmProps props, bool parallax, bool firstFit, rtx_vmm_theta *vmmsOut) {
    const uint32_t nblk = blocksOf(slots, kSortBlock);
    // valid samples in slot order
    hipLaunchKernelGGL(kCompactCount, dim3(nblk), dim3(kSortBlock), 0, s, dd, slots, g.numRegions, g.blockSums);
    hipLaunchKernelGGL(kScanInPlace, dim3(1), dim3(1024), 0, s, g.blockSums, nblk, g.counters);
    hipLaunchKernelGGL(kCompactScatter, dim3(nblk), dim3(kSortBlock), 0, s, dd, slots, g.numRegions, g.blockSums,
                       g.keysA, g.valsA);
    // stable sort by region
    int passes = 0;
    for (uint32_t span = 1; span < g.numRegions; span <<= kDigitBits) ++passes;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kDigitBits;
        hipLaunchKernelGGL(kRadixHist, dim3(nblk), dim3(kSortBlock), 0, s, g.keysA, g.counters, shift, nblk, g.hist);
        hipLaunchKernelGGL(kScanInPlace, dim3(1), dim3(1024), 0, s, g.hist, nblk * kDigits, (uint32_t *)nullptr);
        hipLaunchKernelGGL(kRadixScatter, dim3(nblk), dim3(kSortBlock), 0, s, g.keysA, g.valsA, g.keysB, g.valsB,
                           g.counters, shift, nblk, g.hist);
        std::swap(g.keysA, g.keysB);
        std::swap(g.valsA, g.valsB);
    }
    // region segments
    hipLaunchKernelGGL(kZero, dim3(blocksOf(g.numRegions + 1, 256)), dim3(256), 0, s, g.regionStart, g.numRegions + 1);
    hipLaunchKernelGGL(kSegmentCount, dim3(1024), dim3(256), 0, s, g.keysA, g.counters, g.regionStart);
    hipLaunchKernelGGL(kScanInPlace, dim3(1), dim3(1024), 0, s, g.regionStart, g.numRegions + 1, (uint32_t *)nullptr);
    // preFit, EM, distances
    hipLaunchKernelGGL(kPrefit, dim3(blocksOf(g.numRegions, 64)), dim3(64), 0, s, g, parallax, firstFit);
    hipLaunchKernelGGL(kGather, dim3(1024), dim3(256), 0, s, g, dd, parallax);
    hipLaunchKernelGGL(kRegionEM, dim3(g.numRegions), dim3(kEmBlock), 0, s, g, props, firstFit ? 0 : 1);
    if (parallax) hipLaunchKernelGGL(kDistances, dim3(g.numRegions), dim3(kEmBlock), 0, s, g);
    launchGuideToVmm(s, g, parallax, vmmsOut);
}

} // namespace rtx
