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
//   kSplitMerge one workgroup per region: postFit's component merge and split
//               (src/PathGuiding.cpp:414-441, 494-788) with the chi-square and
//               covariance statistics they read; decisions in lane 0, sample
//               loops over the workgroup
//   kDistances  harmonic-mean lobe distances (incrementaldistance.h:58-104)
//   kToVmm      VMM_Theta for rendering (syncPMMsToVMM_Thetas, 117-139)
#include <hip/hip_runtime.h>

#include "dev_math.h"
#include "dev_sort.h"
#include "internal.h"

namespace rtx {

constexpr int kK = RTX_MAX_LOBES; // array bound; rtx_guiding_options.maxComponents caps the lobes in use
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
        g.samples[2 * i + 1] = make_float4(dist, x.pdf, 0.0f, 0.0f);
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

// ======================================================= split and merge
// PMM::removeWeightPrior / applyWeightPrior (ParametricMixtureModel.h:339-375):
// unclamped, over the K active components.
RTX_DI void pmmRemovePrior(GuideRegion &R, float vPrior) {
    const float inv = vPrior * (float)R.K + 1.0f, nrm = 1.0f / inv, prior = vPrior * nrm;
    for (int k = 0; k < R.K; ++k) R.w[k] = (R.w[k] - prior) * inv;
}
RTX_DI void pmmApplyPrior(GuideRegion &R, float vPrior) {
    const float inv = vPrior * (float)R.K + 1.0f, nrm = 1.0f / inv, prior = vPrior * nrm;
    for (int k = 0; k < R.K; ++k) R.w[k] = R.w[k] * nrm + prior;
}

// Nori's coordinateAxis as lightpmm's Frame uses it (pmm-glm.h:107-133).
RTX_DI void tangentFrame(const float n[3], float x[3], float y[3]) {
    if (fabsf(n[0]) > fabsf(n[1])) {
        const float inv = 1.0f / sqrtf(n[0] * n[0] + n[2] * n[2]);
        y[0] = n[2] * inv, y[1] = 0.0f, y[2] = -n[0] * inv;
    } else {
        const float inv = 1.0f / sqrtf(n[1] * n[1] + n[2] * n[2]);
        y[0] = 0.0f, y[1] = n[2] * inv, y[2] = -n[1] * inv;
    }
    x[0] = y[1] * n[2] - y[2] * n[1];
    x[1] = y[2] * n[0] - y[0] * n[2];
    x[2] = y[0] * n[1] - y[1] * n[0];
}

// One vMF lobe for the merge metric.
struct Lobe {
    float w, mu[3], kappa, r, norm, e2k;
};
RTX_DI Lobe lobeOf(const GuideRegion &R, int k) {
    Lobe l;
    l.w = R.w[k], l.kappa = R.kappa[k], l.r = R.r[k], l.norm = R.norm[k], l.e2k = R.e2k[k];
    for (int c = 0; c < 3; ++c) l.mu[c] = R.mu[k][c];
    return l;
}
// VMFKernel::product (315-355) and division (362-402)
RTX_DI Lobe lobeCombine(const Lobe &a, const Lobe &b, bool divide) {
    const float sgn = divide ? -1.0f : 1.0f;
    float nm[3];
    for (int c = 0; c < 3; ++c) nm[c] = a.mu[c] * a.kappa + sgn * b.mu[c] * b.kappa;
    float nk = sqrtf(nm[0] * nm[0] + nm[1] * nm[1] + nm[2] * nm[2]);
    const bool small = nk < 1e-3f;
    if (small) nk = 0.0f;
    for (int c = 0; c < 3; ++c) nm[c] = small ? a.mu[c] : nm[c] / nk;
    const float ne2k = fastexpf(-2.0f * nk);
    const float cij = small ? 1.0f / (4.0f * kPi) : (1.0f / (2.0f * kPi)) * nk / (1.0f - ne2k);
    const float ti = nm[0] * a.mu[0] + nm[1] * a.mu[1] + nm[2] * a.mu[2];
    const float tj = nm[0] * b.mu[0] + nm[1] * b.mu[1] + nm[2] * b.mu[2];
    const float e = fastexpf(a.kappa * (ti - 1.0f) + sgn * b.kappa * (tj - 1.0f));
    Lobe o;
    o.w = divide ? e * a.w * a.norm / (b.w * b.norm * cij) : e * a.w * b.w * (a.norm * b.norm) / cij;
    o.kappa = nk;
    o.r = kappaToMeanCosine(nk);
    for (int c = 0; c < 3; ++c) o.mu[c] = nm[c];
    o.norm = cij;
    o.e2k = ne2k;
    return o;
}
// VMFKernel::mergeComponent (478-532)
RTX_DI Lobe lobeMerge(const Lobe &a, const Lobe &b) {
    Lobe o;
    float kappa = 0.0f, norm = 1.0f / (4.0f * kPi), e2k = 1.0f;
    const float weight = a.w + b.w;
    float mu[3];
    for (int c = 0; c < 3; ++c) mu[c] = (a.w * a.r * a.mu[c] + b.w * b.r * b.mu[c]) / weight;
    float r = mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2];
    if (r > 0.0f) {
        r = sqrtf(r);
        kappa = meanCosineToKappa(r);
        if (kappa < 1e-3f) kappa = 0.0f;
        e2k = expf(-2.0f * kappa);
        norm = kappa / (2.0f * kPi * (1.0f - e2k));
        for (int c = 0; c < 3; ++c) mu[c] /= r;
    } else {
        for (int c = 0; c < 3; ++c) mu[c] = a.mu[c];
    }
    o.w = weight, o.kappa = kappa, o.r = r, o.norm = norm, o.e2k = e2k;
    for (int c = 0; c < 3; ++c) o.mu[c] = mu[c];
    return o;
}
// computePearsonChiSquaredMergeMetric (src/PathGuiding.cpp:712-762)
RTX_DI float mergeMetric(const GuideRegion &R, int i, int j) {
    const Lobe a = lobeOf(R, i), b = lobeOf(R, j);
    const Lobe aa = lobeCombine(a, a, false), bb = lobeCombine(b, b, false), ab = lobeCombine(a, b, false);
    const Lobe m = lobeMerge(a, b);
    return lobeCombine(aa, m, true).w + 2.0f * lobeCombine(ab, m, true).w + lobeCombine(bb, m, true).w - m.w;
}

// PathGuiding::mergeComponents (765-788): statistics first, then
// PMM::mergeComponents (321-333).
RTX_DI void mergeComponents(GuideRegion &R, int a, int b) {
    const int last = R.K - 1;
    const float wA = R.w[a], wB = R.w[b], wM = wA + wB;
    { // IncrementalDistance::merge: harmonic means
        const float sw = wM / (wA / R.distW[a] + wB / R.distW[b]);
        const float d = wM / (wA / R.dist[a] + wB / R.dist[b]);
        R.dist[a] = d;
        R.distW[a] = sw;
        if (b != last) {
            R.dist[b] = R.dist[last];
            R.distW[b] = R.distW[last];
        }
    }
    const float inv = 1.0f / wM;
    { // IncrementalPearsonChiSquared::merge
        if (!isinf(inv)) {
            const float sa = wA * inv, sb = wB * inv;
            R.chi[a] = sa * R.chi[a] + sb * R.chi[b];
            R.chiN[a] = sa * R.chiN[a] + sb * R.chiN[b];
        } else {
            R.chi[a] = R.chiN[a] = 0.0f;
        }
        if (b != last) {
            R.chi[b] = R.chi[last];
            R.chiN[b] = R.chiN[last];
        }
    }
    { // IncrementalCovariance2D::merge
        if (!isinf(inv)) {
            float mm[3];
            for (int c = 0; c < 3; ++c) mm[c] = wA * R.mu[a][c] + wB * R.mu[b][c];
            const float il = 1.0f / sqrtf(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
            for (int c = 0; c < 3; ++c) mm[c] *= il;
            float fx[3], fy[3];
            tangentFrame(mm, fx, fy);
            const float xa = fx[0] * R.mu[a][0] + fx[1] * R.mu[a][1] + fx[2] * R.mu[a][2];
            const float ya = fy[0] * R.mu[a][0] + fy[1] * R.mu[a][1] + fy[2] * R.mu[a][2];
            const float xb = fx[0] * R.mu[b][0] + fx[1] * R.mu[b][1] + fx[2] * R.mu[b][2];
            const float yb = fy[0] * R.mu[b][0] + fy[1] * R.mu[b][1] + fy[2] * R.mu[b][2];
            const float mix = wA * wB / wM;
            const float dv[3] = {(xa - xb) * (xa - xb), (ya - yb) * (ya - yb), (xa - xb) * (ya - yb)};
            for (int c = 0; c < 3; ++c) R.cov[a][c] = (wA * R.cov[a][c] + wB * R.cov[b][c] + mix * dv[c]) * inv;
            R.covW[a] = R.covW[a] + R.covW[b];
        } else {
            R.cov[a][0] = R.cov[a][1] = R.cov[a][2] = 0.0f;
            R.covW[a] = 0.0f;
        }
        if (b != last) {
            for (int c = 0; c < 3; ++c) R.cov[b][c] = R.cov[last][c];
            R.covW[b] = R.covW[last];
        }
    }
    const Lobe m = lobeMerge(lobeOf(R, a), lobeOf(R, b));
    R.w[a] = m.w, R.kappa[a] = m.kappa, R.r[a] = m.r, R.norm[a] = m.norm, R.e2k[a] = m.e2k;
    for (int c = 0; c < 3; ++c) R.mu[a][c] = m.mu[c];
    R.w[b] = 0.0f, R.kappa[b] = 0.0f, R.r[b] = 0.0f, R.mu[b][0] = 0.0f, R.mu[b][1] = 0.0f, R.mu[b][2] = 1.0f;
    R.norm[b] = 1.0f / (4.0f * kPi), R.e2k[b] = 1.0f;
    R.K -= 1;
    const int k1 = R.K;
    if (b != k1) { // swapComponents(b, K)
        float t;
        t = R.w[b], R.w[b] = R.w[k1], R.w[k1] = t;
        t = R.kappa[b], R.kappa[b] = R.kappa[k1], R.kappa[k1] = t;
        t = R.r[b], R.r[b] = R.r[k1], R.r[k1] = t;
        for (int c = 0; c < 3; ++c) t = R.mu[b][c], R.mu[b][c] = R.mu[k1][c], R.mu[k1][c] = t;
        t = R.norm[b], R.norm[b] = R.norm[k1], R.norm[k1] = t;
        t = R.e2k[b], R.e2k[b] = R.e2k[k1], R.e2k[k1] = t;
    }
}

// Eigen 3's EigenSolver on a real 2x2 matrix (RealSchur with one Givens
// step of Anderson's continuous plane rotations, eigenvectors by
// back-substitution, columns normalised); lam[i] with column V[.][i].
RTX_DI void eigen2x2(float a00, float a01, float a10, float a11, float lam[2], float V[2][2]) {
    float T[2][2] = {{a00, a01}, {a10, a11}};
    float U[2][2] = {{1.0f, 0.0f}, {0.0f, 1.0f}};
    const float scale = fmaxf(fmaxf(fabsf(a00), fabsf(a01)), fmaxf(fabsf(a10), fabsf(a11)));
    if (scale < 1.17549435e-38f) {
        lam[0] = lam[1] = 0.0f;
        V[0][0] = V[1][1] = 1.0f;
        V[0][1] = V[1][0] = 0.0f;
        return;
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) T[i][j] /= scale;
    const float eps = 1.1920929e-07f;
    float sd = fabsf(T[0][0]) + fabsf(T[1][1]);
    if (sd == 0.0f) sd = fabsf(T[0][0]) + fabsf(T[0][1]) + fabsf(T[1][0]) + fabsf(T[1][1]);
    if (!(fabsf(T[1][0]) < eps * sd)) {
        const float p = 0.5f * (T[0][0] - T[1][1]);
        const float q = p * p + T[1][0] * T[0][1];
        if (q >= 0.0f) {
            const float z = sqrtf(fabsf(q));
            const float gp = p >= 0.0f ? p + z : p - z, gq = T[1][0];
            float c, sn;
            if (gq == 0.0f) {
                c = gp < 0.0f ? -1.0f : 1.0f;
                sn = 0.0f;
            } else if (gp == 0.0f) {
                c = 0.0f;
                sn = gq < 0.0f ? 1.0f : -1.0f;
            } else if (fabsf(gp) > fabsf(gq)) {
                const float t = gq / gp;
                float u = sqrtf(1.0f + t * t);
                if (gp < 0.0f) u = -u;
                c = 1.0f / u;
                sn = -t * c;
            } else {
                const float t = gp / gq;
                float u = sqrtf(1.0f + t * t);
                if (gq < 0.0f) u = -u;
                sn = -1.0f / u;
                c = -t * sn;
            }
            for (int j = 0; j < 2; ++j) {
                const float x = T[0][j], y = T[1][j];
                T[0][j] = c * x - sn * y;
                T[1][j] = sn * x + c * y;
            }
            for (int i = 0; i < 2; ++i) {
                float x = T[i][0], y = T[i][1];
                T[i][0] = c * x - sn * y;
                T[i][1] = sn * x + c * y;
                x = U[i][0], y = U[i][1];
                U[i][0] = c * x - sn * y;
                U[i][1] = sn * x + c * y;
            }
            T[1][0] = 0.0f;
        }
    } else {
        T[1][0] = 0.0f;
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) T[i][j] *= scale;
    lam[0] = T[0][0];
    lam[1] = T[1][1];
    const float w = T[0][0] - lam[1];
    const float nrm = fabsf(T[0][0]) + fabsf(T[0][1]) + fabsf(T[1][1]);
    const float x01 = w != 0.0f ? -T[0][1] / w : -T[0][1] / (eps * nrm);
    const float v0[2] = {U[0][0], U[1][0]};
    const float v1[2] = {U[0][0] * x01 + U[0][1], U[1][0] * x01 + U[1][1]};
    const float n0 = sqrtf(v0[0] * v0[0] + v0[1] * v0[1]), n1 = sqrtf(v1[0] * v1[0] + v1[1] * v1[1]);
    V[0][0] = v0[0] / n0, V[1][0] = v0[1] / n0;
    V[0][1] = v1[0] / n1, V[1][1] = v1[1] / n1;
}

// splitComponentUsingPCA (src/PathGuiding.cpp:494-561); the principal axis
// is a row of Eigen's eigenvector matrix, as the reference reads it.
RTX_DI void splitComponent(GuideRegion &R, int comp, float maxKappa, int maxK) {
    if (R.K >= maxK) return;
    const int tgt = R.K;
    float fx[3], fy[3];
    tangentFrame(R.mu[comp], fx, fy);
    float vx = 0.0f, vy = 0.0f, cxy = 0.0f;
    if (R.covW[comp] > 0.0f) vx = R.cov[comp][0], vy = R.cov[comp][1], cxy = R.cov[comp][2];
    float lam[2], V[2][2];
    eigen2x2(vx, cxy, cxy, vy, lam, V);
    int o0 = 0, o1 = 1;
    if (lam[1] <= lam[0]) o0 = 1, o1 = 0;
    const int pick = lam[o1] > lam[o0] ? o1 : o0;
    const float off = fminf(1.0f, 0.5f * sqrtf(lam[pick])); // std::min(1, NaN) is 1
    const float offset = isnan(off) ? 1.0f : off;
    const float px = V[pick][0] * offset, py = V[pick][1] * offset;
    const float z = sqrtf(1.0f - offset * offset);
    float muA[3], muB[3];
    for (int c = 0; c < 3; ++c) {
        muA[c] = px * fx[c] + py * fy[c] + z * R.mu[comp][c];
        muB[c] = -px * fx[c] - py * fy[c] + z * R.mu[comp][c];
    }
    R.K = tgt + 1;
    const float srcR = R.r[comp], splitW = R.w[comp] * 0.5f;
    const float maxAvgCos = kappaToMeanCosine(maxKappa);
    const float splitR = z > 0.0f ? fminf(maxAvgCos, srcR / z) : maxAvgCos;
    const float splitK = meanCosineToKappa(splitR);
    const bool small = splitK < 1e-3f;
    for (int q = 0; q < 2; ++q) {
        const int k = q == 0 ? comp : tgt;
        R.kappa[k] = small ? 0.0f : splitK;
        R.r[k] = small ? 0.0f : splitR;
        calNorm(R, k);
        R.w[k] = splitW;
    }
    for (int c = 0; c < 3; ++c) {
        R.mu[comp][c] = muA[c];
        R.mu[tgt][c] = muB[c];
    }
    R.dist[tgt] = R.dist[comp];
    R.distW[comp] = R.distW[tgt] = 0.5f * R.distW[comp];
    R.chi[comp] = R.chi[tgt] = R.chiN[comp] = R.chiN[tgt] = 0.0f;
    for (int c = 0; c < 3; ++c) R.cov[comp][c] = R.cov[tgt][c] = 0.0f;
    R.covW[comp] = R.covW[tgt] = 0.0f;
}

struct SmShared {
    EmShared E;
    float fs[kK][3], ft[kK][3];
    double part[kEmBlock / 64][5 * kK];
    float metric[kK * (kK - 1) / 2];
    int pairA[kK * (kK - 1) / 2], pairB[kK * (kK - 1) / 2];
    int mask[kK];
    float oldActive, sumWeight, norm;
    int numSplits, stop;
};

RTX_DI void loadModel(EmShared &E, const GuideRegion &R) {
    __syncthreads();
    if (threadIdx.x < (unsigned)kK) {
        const int k = threadIdx.x;
        E.w[k] = R.w[k];
        E.kappa[k] = R.kappa[k];
        E.norm[k] = R.norm[k];
        for (int c = 0; c < 3; ++c) E.mu[k][c] = R.mu[k][c];
    }
    __syncthreads();
}

// IncrementalPearsonChiSquared::updateDivergence(Masked) (54-148) and
// IncrementalCovariance2D::updateStatistics(Masked) (54-166) over the
// region's samples, for the masked components (all when useMask is false).
RTX_DI void chiCovUpdate(SmShared &S, GuideRegion &R, const float4 *smp, uint32_t n, bool useMask) {
    EmShared &E = S.E;
    loadModel(E, R);
    const int K = R.K;
    if (threadIdx.x < (unsigned)K) { // the covariance's frame compares squares
        const int k = threadIdx.x;
        const float *nn = E.mu[k];
        const bool xg = nn[0] * nn[0] > nn[1] * nn[1];
        const float inv = 1.0f / sqrtf((xg ? nn[0] * nn[0] : nn[1] * nn[1]) + nn[2] * nn[2]);
        float t[3];
        if (xg) t[0] = nn[2] * inv, t[1] = 0.0f, t[2] = -nn[0] * inv;
        else t[0] = 0.0f, t[1] = nn[2] * inv, t[2] = -nn[1] * inv;
        S.ft[k][0] = t[0], S.ft[k][1] = t[1], S.ft[k][2] = t[2];
        S.fs[k][0] = t[1] * nn[2] - t[2] * nn[1];
        S.fs[k][1] = t[2] * nn[0] - t[0] * nn[2];
        S.fs[k][2] = t[0] * nn[1] - t[1] * nn[0];
    }
    __syncthreads();
    float chiB[kK], bw[kK], bv[kK][3];
    for (int k = 0; k < kK; ++k) chiB[k] = bw[k] = bv[k][0] = bv[k][1] = bv[k][2] = 0.0f;
    for (uint32_t j = threadIdx.x; j < n; j += kEmBlock) {
        const float4 a4 = smp[2 * j];
        const float pdf = smp[2 * j + 1].y;
        const V3 d = v3(a4);
        float p[kK], mix = 0.0f;
        for (int k = 0; k < K; ++k) {
            const float c = E.mu[k][0] * d.x + E.mu[k][1] * d.y + E.mu[k][2] * d.z;
            p[k] = E.norm[k] * fastexpf(E.kappa[k] * fminf(c - 1.0f, 0.0f));
            mix += E.w[k] * p[k];
        }
        if (!(mix > 1e-8f)) continue; // PMM_EPSILON
        const float f = a4.w * a4.w * pdf / (mix * mix);
        const float inv = 1.0f / mix;
        for (int k = 0; k < K; ++k) {
            if (useMask && !S.mask[k]) continue;
            chiB[k] += p[k] * f;
            const float ws = a4.w * (E.w[k] * p[k] * inv);
            bw[k] += ws;
            const float lx = S.fs[k][0] * d.x + S.fs[k][1] * d.y + S.fs[k][2] * d.z;
            const float ly = S.ft[k][0] * d.x + S.ft[k][1] * d.y + S.ft[k][2] * d.z;
            bv[k][0] += lx * lx * ws;
            bv[k][1] += ly * ly * ws;
            bv[k][2] += lx * ly * ws;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < K; ++k) {
        const double v0 = waveSumD(chiB[k]), v1 = waveSumD(bw[k]), v2 = waveSumD(bv[k][0]), v3_ = waveSumD(bv[k][1]),
                     v4 = waveSumD(bv[k][2]);
        if (lane == 0) {
            S.part[wave][5 * k] = v0;
            S.part[wave][5 * k + 1] = v1;
            S.part[wave][5 * k + 2] = v2;
            S.part[wave][5 * k + 3] = v3_;
            S.part[wave][5 * k + 4] = v4;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)K && (!useMask || S.mask[threadIdx.x])) {
        const int k = threadIdx.x;
        double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < kEmBlock / 64; ++w)
            for (int q = 0; q < 5; ++q) t[q] += S.part[w][5 * k + q];
        if (n > 0) {
            const float cn = R.chiN[k] + (float)n;
            R.chi[k] = (R.chi[k] * R.chiN[k] + (float)t[0]) / cn;
            R.chiN[k] = cn;
        }
        const float nw = R.covW[k] + (float)t[1];
        for (int c = 0; c < 3; ++c) R.cov[k][c] = nw > 0.0f ? (R.cov[k][c] * R.covW[k] + (float)t[2 + c]) / nw : 0.0f;
        R.covW[k] = nw;
    }
    __syncthreads();
}

// VMMFactory::maskedFit (225-297) with maskedParameterUpdate (458-486)
RTX_DI void maskedFit(SmShared &S, GuideRegion &R, const EmProps &P, const float4 *smp, uint32_t n) {
    EmShared &E = S.E;
    const int K = R.K;
    if (threadIdx.x == 0) {
        float old = 0.0f;
        for (int k = 0; k < K; ++k)
            if (S.mask[k]) old += removeWeightPrior(P, R.w[k], K);
        S.oldActive = old;
        S.stop = 0;
    }
    loadModel(E, R);
    const uint32_t maxItr = n > (uint32_t)(2 * K) ? P.maxItr : 0u;
    const float maxMeanCosine = kappaToMeanCosine(P.maxKappa);
    float lastLL = 0.0f, absInvLastLL = 0.0f;
    for (uint32_t i = 0; i < maxItr; ++i) {
        eStep(E, K, smp, n);
        if (threadIdx.x == 0) {
            if (i == 0) S.sumWeight = (float)E.tot[4 * kK];
            float active = 0.0f;
            for (int k = 0; k < K; ++k)
                if (S.mask[k]) active += (float)E.tot[4 * k];
            S.norm = S.oldActive * (S.sumWeight / active);
        }
        __syncthreads();
        if (threadIdx.x < (unsigned)K && S.mask[threadIdx.x]) {
            const int k = threadIdx.x;
            const float Wk = (float)E.tot[4 * k] * S.norm;
            const float rw0 = (float)E.tot[4 * k + 1] * S.norm, rw1 = (float)E.tot[4 * k + 2] * S.norm,
                        rw2 = (float)E.tot[4 * k + 3] * S.norm;
            const float mixtureWeight = applyWeightPrior(P, Wk / S.sumWeight, K);
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
            for (int c = 0; c < 3; ++c) E.mu[k][c] = R.mu[k][c];
        }
        if (threadIdx.x == 0 && i >= P.minItr) {
            const float ll = (float)E.tot[4 * kK + 1];
            if (i > P.minItr && (ll - lastLL) * absInvLastLL < P.relLogLikelihoodThreshold) S.stop = 1;
            lastLL = ll;
            absInvLastLL = 1.0f / fabsf(ll);
        }
        __syncthreads();
        if (S.stop) break;
    }
    __syncthreads();
}

// postFit with splitAndMerge (src/PathGuiding.cpp:414-441): mergeAll when
// enough samples came in, the statistics, then splitAll.
__global__ __launch_bounds__(kEmBlock) void kSplitMerge(GuideFitBuffers g, EmProps P, rtx_guiding_options SP) {
    __shared__ SmShared S;
    const uint32_t r = blockIdx.x;
    const uint32_t s0 = g.regionStart[r], n = g.regionStart[r + 1] - s0;
    if (n == 0) return;
    GuideRegion &R = g.regions[r];
    const float4 *smp = g.samples + 2 * (size_t)s0;
    // ---- mergeAll (635-709)
    const bool doMerge = R.samplesSinceLastMerge + n > SP.minSamplesForMerging;
    if (doMerge) {
        if (threadIdx.x == 0) pmmRemovePrior(R, P.vPrior);
        for (;;) {
            __syncthreads();
            const int K = R.K;
            if (K <= 1) break;
            const int np = K * (K - 1) / 2;
            for (int t = threadIdx.x; t < np; t += kEmBlock) {
                int i = 0, rem = t; // pair t in the order (0,1), (0,2), ..., (1,2), ...
                while (rem >= K - 1 - i) rem -= K - 1 - i, ++i;
                const int j = i + 1 + rem;
                S.metric[t] = mergeMetric(R, i, j);
                S.pairA[t] = i;
                S.pairB[t] = j;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                // std::partition (libstdc++, bidirectional) by metric <= mergeMaxDivergence
                int first = 0, last = np;
                for (;;) {
                    while (first != last && S.metric[first] <= SP.mergeMaxDivergence) ++first;
                    if (first == last) break;
                    --last;
                    while (first != last && !(S.metric[last] <= SP.mergeMaxDivergence)) --last;
                    if (first == last) break;
                    float tm = S.metric[first];
                    S.metric[first] = S.metric[last], S.metric[last] = tm;
                    int ta = S.pairA[first], tb = S.pairB[first];
                    S.pairA[first] = S.pairA[last], S.pairB[first] = S.pairB[last];
                    S.pairA[last] = ta, S.pairB[last] = tb;
                    ++first;
                }
                const int nValid = first;
                // stable insertion sort by metric, ascending
                for (int q = 1; q < nValid; ++q) {
                    const float m = S.metric[q];
                    const int a = S.pairA[q], b = S.pairB[q];
                    int p = q - 1;
                    while (p >= 0 && S.metric[p] > m) {
                        S.metric[p + 1] = S.metric[p], S.pairA[p + 1] = S.pairA[p], S.pairB[p + 1] = S.pairB[p];
                        --p;
                    }
                    S.metric[p + 1] = m, S.pairA[p + 1] = a, S.pairB[p + 1] = b;
                }
                uint32_t used = 0;
                int ma[kK], mb[kK], nm = 0;
                for (int q = 0; q < nValid; ++q) {
                    const int a = min(S.pairA[q], S.pairB[q]), b = max(S.pairA[q], S.pairB[q]);
                    if (((used >> a) & 1u) || ((used >> b) & 1u)) continue;
                    ma[nm] = a, mb[nm] = b, ++nm;
                    used |= (1u << a) | (1u << b);
                }
                // by second component, descending (distinct)
                for (int q = 1; q < nm; ++q) {
                    const int a = ma[q], b = mb[q];
                    int p = q - 1;
                    while (p >= 0 && mb[p] < b) {
                        ma[p + 1] = ma[p], mb[p + 1] = mb[p];
                        --p;
                    }
                    ma[p + 1] = a, mb[p + 1] = b;
                }
                for (int q = 0; q < nm; ++q) mergeComponents(R, ma[q], mb[q]);
                S.stop = nm == 0;
            }
            __syncthreads();
            if (S.stop) break;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            pmmApplyPrior(R, P.vPrior);
            R.samplesSinceLastMerge = 0;
        }
    } else if (threadIdx.x == 0) {
        R.samplesSinceLastMerge += n;
    }
    __syncthreads();
    // ---- statistics after merging
    chiCovUpdate(S, R, smp, n, false);
    // ---- splitAll (563-632)
    const bool firstFit = (uint64_t)n == R.totalNumSamples;
    const bool fit = firstFit || n > SP.minSamplesForPostSplitFitting;
    if (threadIdx.x == 0) pmmRemovePrior(R, P.vPrior);
    __syncthreads();
    const int maxK = (int)SP.maxComponents;
    if (R.K < maxK) {
        for (;;) {
            if (threadIdx.x == 0) {
                const int K = R.K;
                const float invAvgSqr = (R.numSamples * R.numSamples) / (R.sampleWeight * R.sampleWeight);
                float cv[kK];
                int ci[kK];
                for (int k = 0; k < K; ++k) {
                    const bool seen = R.chiN[k] > (float)SP.minSamplesForSplitting;
                    const float wd = (firstFit || seen) ? (R.chi[k] * invAvgSqr - 1.0f) * R.w[k] : 0.0f;
                    cv[k] = fabsf(wd);
                    ci[k] = k;
                }
                int first = 0, last = K; // std::partition by |wd| >= splitMinDivergence
                for (;;) {
                    while (first != last && cv[first] >= SP.splitMinDivergence) ++first;
                    if (first == last) break;
                    --last;
                    while (first != last && !(cv[last] >= SP.splitMinDivergence)) --last;
                    if (first == last) break;
                    float tv = cv[first];
                    cv[first] = cv[last], cv[last] = tv;
                    int ti = ci[first];
                    ci[first] = ci[last], ci[last] = ti;
                    ++first;
                }
                const int nCand = first;
                const int numSplits = min(nCand, maxK - K);
                if (nCand > numSplits) // the numSplits largest, descending (std::partial_sort)
                    for (int q = 0; q < numSplits; ++q) {
                        int best = q;
                        for (int t = q + 1; t < nCand; ++t)
                            if (cv[t] > cv[best]) best = t;
                        float tv = cv[q];
                        cv[q] = cv[best], cv[best] = tv;
                        int ti = ci[q];
                        ci[q] = ci[best], ci[best] = ti;
                    }
                for (int k = 0; k < kK; ++k) S.mask[k] = 0;
                for (int q = 0; q < numSplits; ++q) {
                    S.mask[ci[q]] = 1;
                    S.mask[R.K] = 1;
                    splitComponent(R, ci[q], P.maxKappa, maxK);
                }
                S.numSplits = numSplits;
            }
            __syncthreads();
            if (S.numSplits == 0) break;
            if (fit) {
                if (threadIdx.x == 0) pmmApplyPrior(R, P.vPrior);
                __syncthreads();
                maskedFit(S, R, P, smp, n);
                chiCovUpdate(S, R, smp, n, true);
                if (threadIdx.x == 0) pmmRemovePrior(R, P.vPrior);
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) pmmApplyPrior(R, P.vPrior);
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
__global__ __launch_bounds__(64) void kToVmm(GuideFitBuffers g, bool parallax, rtx_vmm_theta_wide *out) {
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    if (r >= g.numRegions) return;
    const GuideRegion &R = g.regions[r];
    rtx_vmm_theta_wide v;
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
    for (int n = 0; n < kK; ++n) {
        R.chi[n] = R.chiN[n] = R.covW[n] = 0.0f;
        R.cov[n][0] = R.cov[n][1] = R.cov[n][2] = 0.0f;
    }
    R.samplesSinceLastMerge = 0;
    g.regions[r] = R;
}

__global__ void kClampCount(uint32_t *counters, uint32_t cap) {
    if (threadIdx.x == 0 && counters[0] > cap) counters[0] = cap;
}

// ------------------------------------------------------------- launchers
static uint32_t blocksOf(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launchGuideInit(hipStream_t s, GuideFitBuffers g, uint32_t numInitialComponents, float initKappa) {
    hipLaunchKernelGGL(kGuideInit, dim3(blocksOf(g.numRegions, 64)), dim3(64), 0, s, g, numInitialComponents,
                       initKappa);
}

void launchGuideToVmm(hipStream_t s, GuideFitBuffers g, bool parallax, rtx_vmm_theta_wide *vmmsOut) {
    hipLaunchKernelGGL(kToVmm, dim3(blocksOf(g.numRegions, 64)), dim3(64), 0, s, g, parallax, vmmsOut);
}

void launchGuideUpdate(hipStream_t s, GuideFitBuffers g, const rtx_directional_data *dd, uint32_t slots,
                       EmProps props, const rtx_guiding_options &sm, bool parallax, bool firstFit,
                       rtx_vmm_theta_wide *vmmsOut) {
    const uint32_t nblk = blocksOf(slots, kSortBlock);
    // valid samples in slot order
    hipLaunchKernelGGL(kCompactCount, dim3(nblk), dim3(kSortBlock), 0, s, dd, slots, g.numRegions, g.blockSums);
    hipLaunchKernelGGL(kScanInPlace<0>, dim3(1), dim3(1024), 0, s, g.blockSums, nblk, g.counters);
    hipLaunchKernelGGL(kCompactScatter, dim3(nblk), dim3(kSortBlock), 0, s, dd, slots, g.numRegions, g.blockSums,
                       g.keysA, g.valsA);
    // at most maxSamplesPerUpdate of them: the first in slot order
    if (sm.maxSamplesPerUpdate) hipLaunchKernelGGL(kClampCount, dim3(1), dim3(64), 0, s, g.counters, sm.maxSamplesPerUpdate);
    // stable sort by region
    int passes = 0;
    for (uint32_t span = 1; span < g.numRegions; span <<= kDigitBits) ++passes;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kDigitBits;
        hipLaunchKernelGGL(kRadixHist<0>, dim3(nblk), dim3(kSortBlock), 0, s, g.keysA, g.counters, shift, nblk, g.hist);
        hipLaunchKernelGGL(kScanInPlace<0>, dim3(1), dim3(1024), 0, s, g.hist, nblk * kDigits, (uint32_t *)nullptr);
        hipLaunchKernelGGL(kRadixScatter<0>, dim3(nblk), dim3(kSortBlock), 0, s, g.keysA, g.valsA, g.keysB, g.valsB,
                           g.counters, shift, nblk, g.hist);
        std::swap(g.keysA, g.keysB);
        std::swap(g.valsA, g.valsB);
    }
    // region segments
    hipLaunchKernelGGL(kZero, dim3(blocksOf(g.numRegions + 1, 256)), dim3(256), 0, s, g.regionStart, g.numRegions + 1);
    hipLaunchKernelGGL(kSegmentCount, dim3(1024), dim3(256), 0, s, g.keysA, g.counters, g.regionStart);
    hipLaunchKernelGGL(kScanInPlace<0>, dim3(1), dim3(1024), 0, s, g.regionStart, g.numRegions + 1, (uint32_t *)nullptr);
    // preFit, EM, distances
    hipLaunchKernelGGL(kPrefit, dim3(blocksOf(g.numRegions, 64)), dim3(64), 0, s, g, parallax, firstFit);
    hipLaunchKernelGGL(kGather, dim3(1024), dim3(256), 0, s, g, dd, parallax);
    hipLaunchKernelGGL(kRegionEM, dim3(g.numRegions), dim3(kEmBlock), 0, s, g, props, firstFit ? 0 : 1);
    if (sm.splitAndMerge) hipLaunchKernelGGL(kSplitMerge, dim3(g.numRegions), dim3(kEmBlock), 0, s, g, props, sm);
    if (parallax) hipLaunchKernelGGL(kDistances, dim3(g.numRegions), dim3(kEmBlock), 0, s, g);
    launchGuideToVmm(s, g, parallax, vmmsOut);
}

} // namespace rtx
