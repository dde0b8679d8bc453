// em_oracle.cpp - CPU restatement of the path-guiding update (TEST
// INFRASTRUCTURE, part of liborcpt.so; see oracle.h).
//
// PathGuiding::update (src/PathGuiding.cpp:276-312) with its per-region steps
// preFit (350-391), fit / updateFit (337-348) and the distance statistics of
// postFit (394-431), built on lightpmm's MAP-EM for von Mises-Fisher mixtures
// (external/lightpmm/include/pmm/VMMFactory.h: initialize 133-147, fit
// 157-216, updateFit 305-362, priors 385-433, parameterUpdate 436-456,
// computeSufficentStatsFromSamples 497-536, computeSufficentStatsFromMixture
// 543-555, computeUniformMus 594-632; VMFKernel.h: pdf 85-94, calNormalization
// 651-659, kappa/mean-cosine maps 46-66) and the harmonic-mean distance
// statistics of external/guiding/incrementaldistance.h (updateDistances
// 58-104, reposition 106-128).  lightpmm evaluates exp with the fastapprox
// "fastexp" of Paul Mineiro (pmm-vcl.h:158-184); it is restated here from that
// published method so that the fits follow the same numbers.  postFit's
// component merge and split (src/PathGuiding.cpp:414-441, 494-788, on by
// default: splitAndMerge, src/PathGuiding.h:124-129) follow below with the
// statistics they read: external/guiding/incrementalpearsonchisquared.h
// (54-213) and incrementalcovariance2d.h (54-263), lightpmm's VMF product,
// division and mergeComponent (VMFKernel.h:315-407, 478-532), PMM
// mergeComponents / weight priors (ParametricMixtureModel.h:298-375) and
// VMMFactory::maskedFit (VMMFactory.h:225-297, 458-486).  The PCA split reads
// a 2x2 eigen-decomposition from Eigen (absent here): restated as the one
// Givens step of Eigen 3's RealSchur::splitOffTwoRows for a 2x2 matrix.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr int kMaxK = RTX_MAX_LOBES; // array bound; SplitMergeProps::maxComponents caps the lobes in use
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Factory properties of PathGuiding::configureVMMFactory (src/PathGuiding.cpp:33-40)
// over lightpmm's defaults (VMMFactory.h:45-61).
struct Props {
    uint32_t minItr = 1, maxItr = 100;
    float relLogLikelihoodThreshold = 0.005f;
    uint32_t numInitialComponents = 8;
    float initKappa = 5.0f, maxKappa = 50000.0f;
    float vPrior = 0.01f, rPrior = 0.0f, rPriorWeight = 1.0f;
};

// fastapprox fastpow2 / fastexp (P. Mineiro, 2011): 2^p from the float bit
// layout with a rational correction of the fractional part.
template <class R> R fastexp(R p) {
    float x = 1.442695040f * (float)p;
    float clipp = std::max(-126.0f, x);
    float w = std::trunc(clipp);
    float z = clipp - w + 1.0f;
    float v = (1 << 23) * (clipp + 121.2740575f + 27.7280233f / (4.84252568f - z) - 1.49012907f * z);
    int32_t bits = (int32_t)std::nearbyint(v);
    float out;
    std::memcpy(&out, &bits, 4);
    return (R)out;
}

template <class R> R kappaToMeanCosine(R kappa) { // VMFKernel.h:46-56
    if (kappa > R(5)) return R(1) - R(1) / kappa;
    if (!(kappa > R(0))) return R(0);
    return R(1) / std::tanh(kappa) - R(1) / kappa;
}
template <class R> R meanCosineToKappa(R r) { return (r * R(3) - r * r * r) / (R(1) - r * r); } // Banerjee 2005

template <class R> struct Pmm {
    int K = 0;
    R w[kMaxK] = {}, mu[kMaxK][3] = {}, kappa[kMaxK] = {}, r[kMaxK] = {}, norm[kMaxK] = {}, e2k[kMaxK] = {};
    R sampleWeight = 0, numSamples = 0;
    uint64_t totalNumSamples = 0;
    uint32_t numEMIterations = 0;

    void calNorm(int k) { // VMFKernel.h:651-659
        e2k[k] = fastexp(R(-2) * kappa[k]);
        norm[k] = kappa[k] > R(0) ? kappa[k] / (R(2) * R(kPi) * (R(1) - e2k[k])) : R(1) / (R(4) * R(kPi));
    }
    void setKappaAndR(int k, R kap, R rr) {
        const bool small = kap < R(1e-3f);
        kappa[k] = small ? R(0) : kap;
        r[k] = small ? R(0) : rr;
        calNorm(k);
    }
    R pdf(int k, const R d[3]) const { // VMFKernel.h:85-94
        R c = mu[k][0] * d[0] + mu[k][1] * d[1] + mu[k][2] * d[2];
        R t = std::min(c - R(1), R(0));
        return norm[k] * fastexp(kappa[k] * t);
    }
};

template <class R> R applyWeightPrior(const Props &P, R w, int K) {
    const R inv = (R)P.vPrior * (R)K + R(1), nrm = R(1) / inv, prior = (R)P.vPrior * nrm;
    return w * nrm + prior;
}
template <class R> R removeWeightPrior(const Props &P, R w, int K) {
    const R inv = (R)P.vPrior * (R)K + R(1), nrm = R(1) / inv, prior = (R)P.vPrior * nrm;
    return w > prior ? (w - prior) * inv : R(0);
}
template <class R> R applyAvgCosinePrior(const Props &P, R a, R n) {
    return (a * n + (R)P.rPrior * (R)P.rPriorWeight) / (n + (R)P.rPriorWeight);
}
template <class R> R removeAvgCosinePrior(const Props &P, R a, R n) {
    return (a * (n + (R)P.rPriorWeight) - (R)P.rPrior * (R)P.rPriorWeight) / n;
}

template <class R> void initialize(const Props &P, Pmm<R> &m) { // VMMFactory.h:133-147, 594-632
    m = Pmm<R>{};
    const int K = (int)P.numInitialComponents;
    m.K = K;
    const float gr = 1.618033988749895f;
    for (int n = 0; n < kMaxK; ++n) {
        if (n < K) {
            float phi = 2.0f * (float)kPi * ((float)n / gr);
            float z = 1.0f - ((2.0f * n + 1.0f) / (float)K);
            float theta = std::acos(z);
            float st = std::sin(theta);
            m.mu[n][0] = (R)(st * std::cos(phi));
            m.mu[n][1] = (R)(st * std::sin(phi));
            m.mu[n][2] = (R)std::cos(theta);
            m.w[n] = R(1) / (R)K;
            R kap = (R)P.initKappa;
            m.kappa[n] = kap;
            m.r[n] = kappaToMeanCosine(kap);
        } else {
            m.mu[n][2] = R(1);
        }
        m.calNorm(n);
    }
}

template <class R> struct Stats {
    R rW[kMaxK][3] = {}, W[kMaxK] = {};
};

struct Sample {
    double pos[3], dir[3];
    double weight, distance, pdf;
};

template <class R>
Stats<R> statsFromSamples(const Pmm<R> &m, const std::vector<Sample> &s, R *sumWeight, R *logLik) {
    Stats<R> st;
    if (sumWeight) *sumWeight = 0;
    if (logLik) *logLik = 0;
    for (const Sample &x : s) {
        R d[3] = {(R)x.dir[0], (R)x.dir[1], (R)x.dir[2]};
        R a[kMaxK];
        R total = 0;
        for (int k = 0; k < m.K; ++k) {
            a[k] = m.w[k] * m.pdf(k, d);
            total += a[k];
        }
        if (!(total > R(1e-8f))) continue; // SoftAssignmentWeights::valid
        const R inv = R(1) / total;
        for (int k = 0; k < m.K; ++k) {
            const R aw = a[k] * inv * (R)x.weight;
            for (int c = 0; c < 3; ++c) st.rW[k][c] += d[c] * aw;
            st.W[k] += aw;
        }
        if (sumWeight) *sumWeight += (R)x.weight;
        if (logLik) *logLik += (R)x.weight * std::log(total);
    }
    return st;
}

template <class R> Stats<R> statsFromMixture(const Props &P, const Pmm<R> &m) {
    Stats<R> st;
    for (int k = 0; k < m.K; ++k) {
        st.W[k] = removeWeightPrior(P, m.w[k], m.K) * m.sampleWeight;
        R a = removeAvgCosinePrior(P, m.r[k], (R)m.totalNumSamples * m.w[k]) * st.W[k];
        for (int c = 0; c < 3; ++c) st.rW[k][c] = m.mu[k][c] * a;
    }
    return st;
}

template <class R> void parameterUpdate(const Props &P, Pmm<R> &m, int k, const R rW[3], R W) {
    const R maxMeanCosine = kappaToMeanCosine((R)P.maxKappa);
    const R mixtureWeight = applyWeightPrior(P, W / m.sampleWeight, m.K);
    const R rLength = std::sqrt(rW[0] * rW[0] + rW[1] * rW[1] + rW[2] * rW[2]);
    if (rLength > R(0))
        for (int c = 0; c < 3; ++c) m.mu[k][c] = rW[c] / rLength;
    const R avgCosine = rLength / W;
    R withPrior = applyAvgCosinePrior(P, avgCosine, (R)m.totalNumSamples * mixtureWeight);
    // lightpmm::min keeps the second operand when the first is NaN (a component without samples)
    withPrior = std::isnan(withPrior) ? maxMeanCosine : std::min(withPrior, maxMeanCosine);
    m.w[k] = mixtureWeight;
    m.setKappaAndR(k, meanCosineToKappa(withPrior), withPrior);
}

// fit (update = false) and updateFit (update = true)
template <class R> void emFit(const Props &P, Pmm<R> &m, const std::vector<Sample> &s, bool update) {
    const Stats<R> old = update ? statsFromMixture(P, m) : Stats<R>{};
    const uint32_t maxItr = s.size() > (size_t)(m.K * 2) ? P.maxItr : 0;
    R sumWeight = 0, ll = 0, lastLL = 0, absInvLastLL = 0;
    uint32_t i = 0;
    for (; i < maxItr; ++i) {
        Stats<R> st = statsFromSamples(m, s, i == 0 ? &sumWeight : nullptr, i >= P.minItr ? &ll : nullptr);
        if (i == 0) {
            if (update) {
                m.sampleWeight += sumWeight;
                m.numSamples += (R)s.size();
                m.totalNumSamples += s.size();
            } else {
                m.sampleWeight = sumWeight;
                m.numSamples = (R)s.size();
                m.totalNumSamples = s.size();
            }
        }
        for (int k = 0; k < m.K; ++k) {
            R rW[3] = {st.rW[k][0] + old.rW[k][0], st.rW[k][1] + old.rW[k][1], st.rW[k][2] + old.rW[k][2]};
            parameterUpdate(P, m, k, rW, st.W[k] + old.W[k]);
        }
        if (i >= P.minItr) {
            if (i > P.minItr) {
                const R rel = (ll - lastLL) * absInvLastLL;
                if (rel < (R)P.relLogLikelihoodThreshold) break;
            }
            lastLL = ll;
            absInvLastLL = R(1) / std::fabs(ll);
        }
    }
    m.numEMIterations += i;
}

template <class R> struct Region {
    Pmm<R> pmm;
    R mean[3] = {}, lastMean[3] = {};
    R dist[kMaxK], distW[kMaxK];
    // IncrementalPearsonChiSquared: divergencePlusOneTimesIntegralSqr, numSamples
    R chi[kMaxK] = {}, chiN[kMaxK] = {};
    // IncrementalCovariance2D: (var x, var y, cov xy), sumWeights
    R cov[kMaxK][3] = {}, covW[kMaxK] = {};
    uint32_t samplesSinceLastMerge = 0;
    Region() {
        for (int k = 0; k < kMaxK; ++k) {
            dist[k] = std::numeric_limits<R>::infinity();
            distW[k] = 0;
        }
    }
};

// incrementaldistance.h:58-104 (harmonic mean of distances, weighted by the
// soft assignment times the component pdf)
template <class R> void updateDistances(Region<R> &g, const std::vector<Sample> &s) {
    const Pmm<R> &m = g.pmm;
    R bw[kMaxK] = {}, bd[kMaxK] = {};
    for (const Sample &x : s) {
        if (!(x.distance > 0.0)) continue;
        R d[3] = {(R)x.dir[0], (R)x.dir[1], (R)x.dir[2]};
        R p[kMaxK], wp[kMaxK], total = 0;
        for (int k = 0; k < m.K; ++k) {
            p[k] = m.pdf(k, d);
            wp[k] = m.w[k] * p[k];
            total += wp[k];
        }
        if (!(total > R(1e-8f))) continue;
        const R f = (R)x.weight / total;
        for (int k = 0; k < m.K; ++k) {
            const R v = wp[k] * p[k] * f;
            bw[k] += v;
            bd[k] += v / (R)x.distance;
        }
    }
    for (int k = 0; k < m.K; ++k) {
        const R nw = g.distW[k] + bw[k];
        g.dist[k] = nw / (g.distW[k] / g.dist[k] + bd[k]);
        g.distW[k] = nw;
    }
}

// incrementaldistance.h:106-128: components with a finite distance keep
// pointing at their source after the mean position moves.
template <class R> void reposition(Region<R> &g, const R newPosToOldPos[3]) {
    for (int k = 0; k < g.pmm.K; ++k) {
        if (std::isinf(g.dist[k]) || g.distW[k] <= R(0)) continue;
        R v[3], len = 0;
        for (int c = 0; c < 3; ++c) {
            v[c] = g.pmm.mu[k][c] * g.dist[k] + newPosToOldPos[c];
            len += v[c] * v[c];
        }
        len = std::sqrt(len);
        g.dist[k] = len;
        for (int c = 0; c < 3; ++c) g.pmm.mu[k][c] = v[c] / len;
    }
}

// ======================================================= split and merge
// PathGuiding's switches (src/PathGuiding.h:124-129)
struct SplitMergeProps {
    bool splitAndMerge = true;
    uint32_t minSamplesForMerging = 8192, minSamplesForSplitting = 4096, minSamplesForPostSplitFitting = 4096;
    float splitMinDivergence = 0.5f, mergeMaxDivergence = 0.025f;
    bool splitRegions = false; // src/PathGuiding.h:118-119
    uint32_t samplesForRegionSplit = 10000;
    // scale limits (rtx_guiding_options): lobes per mixture (limits.glsl:6's
    // 16 by default), regions (0: no cap), samples per update (0: all)
    int maxComponents = RTX_MAX_DISTRIBUTIONS;
    uint32_t maxRegions = 0, maxSamplesPerUpdate = 0;
};

// PMM::removeWeightPrior / applyWeightPrior (ParametricMixtureModel.h:339-375):
// unclamped, over the K active components.
template <class R> void pmmRemovePrior(Pmm<R> &m, float vPrior) {
    const R inv = (R)vPrior * (R)m.K + R(1), nrm = R(1) / inv, prior = (R)vPrior * nrm;
    for (int k = 0; k < m.K; ++k) m.w[k] = (m.w[k] - prior) * inv;
}
template <class R> void pmmApplyPrior(Pmm<R> &m, float vPrior) {
    const R inv = (R)vPrior * (R)m.K + R(1), nrm = R(1) / inv, prior = (R)vPrior * nrm;
    for (int k = 0; k < m.K; ++k) m.w[k] = m.w[k] * nrm + prior;
}

// Component pdfs and the mixture pdf of one direction; false when the soft
// assignment is invalid (mixture pdf <= 1e-8, SoftAssignmentWeights::valid).
template <class R> bool componentPdfs(const Pmm<R> &m, const double dir[3], R p[kMaxK], R &mix) {
    R d[3] = {(R)dir[0], (R)dir[1], (R)dir[2]};
    mix = 0;
    for (int k = 0; k < m.K; ++k) {
        p[k] = m.pdf(k, d);
        mix += m.w[k] * p[k];
    }
    return mix > R(1e-8f);
}

// IncrementalPearsonChiSquared::updateDivergence(Masked) (54-148): running
// mean over batches of sum_samples pdf_k * weight^2 * pdf / mixturePdf^2.
template <class R> void chiUpdate(Region<R> &g, const std::vector<Sample> &s, const bool *mask) {
    if (s.empty()) return;
    const Pmm<R> &m = g.pmm;
    R batch[kMaxK] = {};
    for (const Sample &x : s) {
        R p[kMaxK], mix;
        R d[3] = {(R)x.dir[0], (R)x.dir[1], (R)x.dir[2]};
        mix = 0;
        for (int k = 0; k < m.K; ++k) {
            p[k] = m.pdf(k, d);
            mix += m.w[k] * p[k];
        }
        if (!(mix > R(1e-8f))) continue; // PMM_EPSILON
        const R f = (R)x.weight * (R)x.weight * (R)x.pdf / (mix * mix);
        for (int k = 0; k < m.K; ++k)
            if (!mask || mask[k]) batch[k] += p[k] * f;
    }
    for (int k = 0; k < m.K; ++k) {
        if (mask && !mask[k]) continue;
        const R n = g.chiN[k] + (R)s.size();
        g.chi[k] = (g.chi[k] * g.chiN[k] + batch[k]) / n;
        g.chiN[k] = n;
    }
}

// computeDivergence (150-160)
template <class R> R chiDivergence(const Region<R> &g, int k) {
    const Pmm<R> &m = g.pmm;
    const R invAvgSampleWeightSqr = (m.numSamples * m.numSamples) / (m.sampleWeight * m.sampleWeight);
    return g.chi[k] * invAvgSampleWeightSqr - R(1);
}

// The tangent frame of a direction: Nori's coordinateAxis as lightpmm uses
// it (pmm-glm.h:107-133): frame (x, y, z) with z = n.
template <class R> void tangentFrame(const R n[3], R x[3], R y[3]) {
    if (std::fabs(n[0]) > std::fabs(n[1])) {
        R inv = R(1) / std::sqrt(n[0] * n[0] + n[2] * n[2]);
        y[0] = n[2] * inv, y[1] = 0, y[2] = -n[0] * inv;
    } else {
        R inv = R(1) / std::sqrt(n[1] * n[1] + n[2] * n[2]);
        y[0] = 0, y[1] = n[2] * inv, y[2] = -n[1] * inv;
    }
    x[0] = y[1] * n[2] - y[2] * n[1]; // cross(y, z)
    x[1] = y[2] * n[0] - y[0] * n[2];
    x[2] = y[0] * n[1] - y[1] * n[0];
}

// IncrementalCovariance2D::updateStatistics(Masked) (54-166): per component,
// the weighted second moments of the sample directions in the component's
// tangent frame (s = x, t = y), soft-assignment times sample weight.
template <class R> void covUpdate(Region<R> &g, const std::vector<Sample> &s, const bool *mask) {
    const Pmm<R> &m = g.pmm;
    R fs[kMaxK][3], ft[kMaxK][3];
    for (int k = 0; k < m.K; ++k) {
        // the covariance code compares squares where coordinateAxis compares abs values
        const R *n = m.mu[k];
        const bool xGreater = n[0] * n[0] > n[1] * n[1];
        const R inv = R(1) / std::sqrt((xGreater ? n[0] * n[0] : n[1] * n[1]) + n[2] * n[2]);
        if (xGreater) ft[k][0] = n[2] * inv, ft[k][1] = 0, ft[k][2] = -n[0] * inv;
        else ft[k][0] = 0, ft[k][1] = n[2] * inv, ft[k][2] = -n[1] * inv;
        fs[k][0] = ft[k][1] * n[2] - ft[k][2] * n[1];
        fs[k][1] = ft[k][2] * n[0] - ft[k][0] * n[2];
        fs[k][2] = ft[k][0] * n[1] - ft[k][1] * n[0];
    }
    R bw[kMaxK] = {}, bv[kMaxK][3] = {};
    for (const Sample &x : s) {
        R d[3] = {(R)x.dir[0], (R)x.dir[1], (R)x.dir[2]};
        R a[kMaxK], total = 0;
        for (int k = 0; k < m.K; ++k) {
            a[k] = m.w[k] * m.pdf(k, d);
            total += a[k];
        }
        if (!(total > R(1e-8f))) continue;
        const R inv = R(1) / total;
        for (int k = 0; k < m.K; ++k) {
            if (mask && !mask[k]) continue;
            const R ws = (R)x.weight * (a[k] * inv);
            bw[k] += ws;
            const R lx = fs[k][0] * d[0] + fs[k][1] * d[1] + fs[k][2] * d[2];
            const R ly = ft[k][0] * d[0] + ft[k][1] * d[1] + ft[k][2] * d[2];
            bv[k][0] += lx * lx * ws;
            bv[k][1] += ly * ly * ws;
            bv[k][2] += lx * ly * ws;
        }
    }
    for (int k = 0; k < m.K; ++k) {
        if (mask && !mask[k]) continue;
        const R nw = g.covW[k] + bw[k];
        for (int c = 0; c < 3; ++c) g.cov[k][c] = nw > R(0) ? (g.cov[k][c] * g.covW[k] + bv[k][c]) / nw : R(0);
        g.covW[k] = nw;
    }
}

// One vMF lobe as lightpmm's kernels hold it, for the merge metric.
template <class R> struct Lobe {
    R w, mu[3], kappa, r, norm, e2k;
};
template <class R> Lobe<R> lobeOf(const Pmm<R> &m, int k) {
    return Lobe<R>{m.w[k], {m.mu[k][0], m.mu[k][1], m.mu[k][2]}, m.kappa[k], m.r[k], m.norm[k], m.e2k[k]};
}
constexpr float kVmfMinKappa = 1e-3f;
// VMFKernel::product (315-355) and division (362-402): the product/quotient
// lobe; its weight is the integral (scaled by the mixture weights).
template <class R> Lobe<R> lobeCombine(const Lobe<R> &a, const Lobe<R> &b, bool divide) {
    const R sgn = divide ? R(-1) : R(1);
    R nm[3];
    for (int c = 0; c < 3; ++c) nm[c] = a.mu[c] * a.kappa + sgn * b.mu[c] * b.kappa;
    R nk = std::sqrt(nm[0] * nm[0] + nm[1] * nm[1] + nm[2] * nm[2]);
    const bool small = nk < R(kVmfMinKappa);
    if (small) nk = 0;
    for (int c = 0; c < 3; ++c) nm[c] = small ? a.mu[c] : nm[c] / nk;
    const R ne2k = fastexp(R(-2) * nk);
    const R cij = small ? (R)(1.0f / (4.0f * (float)kPi)) : (R)(1.0f / (2.0f * (float)kPi)) * nk / (R(1) - ne2k);
    const R ti = nm[0] * a.mu[0] + nm[1] * a.mu[1] + nm[2] * a.mu[2];
    const R tj = nm[0] * b.mu[0] + nm[1] * b.mu[1] + nm[2] * b.mu[2];
    const R t = a.kappa * (ti - R(1)) + sgn * b.kappa * (tj - R(1));
    const R e = fastexp(t);
    Lobe<R> o;
    o.w = divide ? e * a.w * a.norm / (b.w * b.norm * cij) : e * a.w * b.w * (a.norm * b.norm) / cij;
    o.kappa = nk;
    o.r = kappaToMeanCosine(nk);
    for (int c = 0; c < 3; ++c) o.mu[c] = nm[c];
    o.norm = cij;
    o.e2k = ne2k;
    return o;
}
// VMFKernel::mergeComponent (478-532): moment-matched merge of two lobes.
template <class R> Lobe<R> lobeMerge(const Lobe<R> &a, const Lobe<R> &b) {
    Lobe<R> o;
    R kappa = 0, norm = (R)(1.0f / (4.0f * (float)kPi)), e2k = 1;
    const R weight = a.w + b.w;
    R mu[3];
    for (int c = 0; c < 3; ++c) mu[c] = (a.w * a.r * a.mu[c] + b.w * b.r * b.mu[c]) / weight;
    R r = mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2];
    if (r > R(0)) {
        r = std::sqrt(r);
        kappa = meanCosineToKappa(r);
        if (kappa < R(kVmfMinKappa)) kappa = 0;
        e2k = std::exp(R(-2) * kappa);
        norm = kappa / (R(2) * (R)(float)kPi * (R(1) - e2k));
        for (int c = 0; c < 3; ++c) mu[c] /= r;
    } else {
        for (int c = 0; c < 3; ++c) mu[c] = a.mu[c];
    }
    o.w = weight;
    o.kappa = kappa;
    o.r = r;
    for (int c = 0; c < 3; ++c) o.mu[c] = mu[c];
    o.norm = norm;
    o.e2k = e2k;
    return o;
}

// computePearsonChiSquaredMergeMetric (src/PathGuiding.cpp:712-762): for
// each pair i < j the chi-square divergence between the pair and its
// merged lobe, A^2/M + 2 AB/M + B^2/M - w_M.
template <class R> R mergeMetric(const Pmm<R> &m, int i, int j) {
    const Lobe<R> a = lobeOf(m, i), b = lobeOf(m, j);
    const Lobe<R> aa = lobeCombine(a, a, false), bb = lobeCombine(b, b, false), ab = lobeCombine(a, b, false);
    const Lobe<R> mg = lobeMerge(a, b);
    return lobeCombine(aa, mg, true).w + R(2) * lobeCombine(ab, mg, true).w + lobeCombine(bb, mg, true).w - mg.w;
}

// std::partition as libstdc++ implements it for bidirectional iterators
// (the order of the kept elements decides which split gets which index).
template <class T, class P> size_t partitionLikeLibstdcxx(std::vector<T> &v, P pred) {
    size_t first = 0, last = v.size();
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (pred(v[first])) ++first;
            else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!pred(v[last])) --last;
            else break;
        }
        std::swap(v[first], v[last]);
        ++first;
    }
}

// PathGuiding::mergeComponents (765-788): the statistics first (they read
// the weights before the merge), then PMM::mergeComponents (321-333): b is
// merged into a, the last component moves into b's slot.
template <class R> void mergeComponents(Region<R> &g, int a, int b) {
    Pmm<R> &m = g.pmm;
    const int last = m.K - 1;
    const R wA = m.w[a], wB = m.w[b], wM = wA + wB;
    { // IncrementalDistance::merge (incrementaldistance.h:165-204): harmonic means
        const R sw = wM / (wA / g.distW[a] + wB / g.distW[b]);
        const R d = wM / (wA / g.dist[a] + wB / g.dist[b]);
        g.dist[a] = d;
        g.distW[a] = sw;
        if (b != last) {
            g.dist[b] = g.dist[last];
            g.distW[b] = g.distW[last];
        }
    }
    { // IncrementalPearsonChiSquared::merge (172-198)
        const R inv = R(1) / wM;
        if (!std::isinf(inv)) {
            const R sa = wA * inv, sb = wB * inv;
            g.chi[a] = sa * g.chi[a] + sb * g.chi[b];
            g.chiN[a] = sa * g.chiN[a] + sb * g.chiN[b];
        } else {
            g.chi[a] = g.chiN[a] = 0;
        }
        if (b != last) {
            g.chi[b] = g.chi[last];
            g.chiN[b] = g.chiN[last];
        }
    }
    { // IncrementalCovariance2D::merge (178-224)
        const R inv = R(1) / wM;
        if (!std::isinf(inv)) {
            R mm[3];
            for (int c = 0; c < 3; ++c) mm[c] = wA * m.mu[a][c] + wB * m.mu[b][c];
            const R il = R(1) / std::sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
            for (int c = 0; c < 3; ++c) mm[c] *= il;
            R fx[3], fy[3];
            tangentFrame(mm, fx, fy);
            const R xa = fx[0] * m.mu[a][0] + fx[1] * m.mu[a][1] + fx[2] * m.mu[a][2];
            const R ya = fy[0] * m.mu[a][0] + fy[1] * m.mu[a][1] + fy[2] * m.mu[a][2];
            const R xb = fx[0] * m.mu[b][0] + fx[1] * m.mu[b][1] + fx[2] * m.mu[b][2];
            const R yb = fy[0] * m.mu[b][0] + fy[1] * m.mu[b][1] + fy[2] * m.mu[b][2];
            const R mix = wA * wB / wM;
            const R dv[3] = {(xa - xb) * (xa - xb), (ya - yb) * (ya - yb), (xa - xb) * (ya - yb)};
            for (int c = 0; c < 3; ++c) g.cov[a][c] = (wA * g.cov[a][c] + wB * g.cov[b][c] + mix * dv[c]) * inv;
            g.covW[a] = g.covW[a] + g.covW[b];
        } else {
            g.cov[a][0] = g.cov[a][1] = g.cov[a][2] = 0;
            g.covW[a] = 0;
        }
        if (b != last) {
            for (int c = 0; c < 3; ++c) g.cov[b][c] = g.cov[last][c];
            g.covW[b] = g.covW[last];
        }
    }
    const Lobe<R> mg = lobeMerge(lobeOf(m, a), lobeOf(m, b));
    m.w[a] = mg.w;
    m.kappa[a] = mg.kappa;
    m.r[a] = mg.r;
    for (int c = 0; c < 3; ++c) m.mu[a][c] = mg.mu[c];
    m.norm[a] = mg.norm;
    m.e2k[a] = mg.e2k;
    // reset b, then swap it with the last component
    m.w[b] = 0, m.kappa[b] = 0, m.r[b] = 0, m.mu[b][0] = 0, m.mu[b][1] = 0, m.mu[b][2] = 1;
    m.norm[b] = (R)(1.0f / (4.0f * (float)kPi)), m.e2k[b] = 1;
    m.K -= 1;
    if (b != m.K) {
        std::swap(m.w[b], m.w[m.K]);
        std::swap(m.kappa[b], m.kappa[m.K]);
        std::swap(m.r[b], m.r[m.K]);
        for (int c = 0; c < 3; ++c) std::swap(m.mu[b][c], m.mu[m.K][c]);
        std::swap(m.norm[b], m.norm[m.K]);
        std::swap(m.e2k[b], m.e2k[m.K]);
    }
}

// PathGuiding::mergeAll (635-709): repeatedly merge the disjoint pairs with
// metric <= mergeMaxDivergence, smallest metric first.
template <class R> uint32_t mergeAll(const SplitMergeProps &SP, Region<R> &g) {
    uint32_t total = 0;
    for (;;) {
        const int K = g.pmm.K;
        if (K <= 1) return total;
        struct Cand {
            R metric;
            int a, b;
        };
        std::vector<Cand> c;
        for (int i = 0; i < K; ++i)
            for (int j = i + 1; j < K; ++j) c.push_back(Cand{mergeMetric(g.pmm, i, j), i, j});
        const size_t nValid = partitionLikeLibstdcxx(c, [&](const Cand &x) { return x.metric <= (R)SP.mergeMaxDivergence; });
        std::stable_sort(c.begin(), c.begin() + nValid, [](const Cand &x, const Cand &y) { return x.metric < y.metric; });
        uint32_t used = 0;
        std::vector<std::pair<int, int>> merges;
        for (size_t q = 0; q < nValid; ++q) {
            const int a = std::min(c[q].a, c[q].b), b = std::max(c[q].a, c[q].b);
            if ((used >> a & 1u) || (used >> b & 1u)) continue;
            merges.push_back({a, b});
            used |= (1u << a) | (1u << b);
        }
        if (merges.empty()) break;
        std::sort(merges.begin(), merges.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
            return x.second > y.second;
        });
        for (const auto &mg : merges) mergeComponents(g, mg.first, mg.second);
        total += (uint32_t)merges.size();
    }
    return total;
}

// Eigen 3's EigenSolver on a real 2x2 matrix, as RealSchur computes it: the
// matrix scaled by its largest entry, one Givens rotation (makeGivens of
// Anderson's continuous plane rotations) splitting off the 2x2 block unless
// the subdiagonal is already negligible, eigenvalues from the diagonal,
// eigenvectors by back-substitution, columns normalised.  lam[i] and the
// column V[.][i] belong together, in Eigen's order.
template <class R> void eigen2x2(R a00, R a01, R a10, R a11, R lam[2], R V[2][2]) {
    R T[2][2] = {{a00, a01}, {a10, a11}};
    R U[2][2] = {{1, 0}, {0, 1}};
    const R scale = std::max(std::max(std::fabs(a00), std::fabs(a01)), std::max(std::fabs(a10), std::fabs(a11)));
    if (scale < std::numeric_limits<R>::min()) {
        lam[0] = lam[1] = 0;
        V[0][0] = V[1][1] = 1;
        V[0][1] = V[1][0] = 0;
        return;
    }
    for (auto &row : T)
        for (R &x : row) x /= scale;
    const R eps = std::numeric_limits<R>::epsilon();
    R sd = std::fabs(T[0][0]) + std::fabs(T[1][1]);
    if (sd == R(0)) sd = std::fabs(T[0][0]) + std::fabs(T[0][1]) + std::fabs(T[1][0]) + std::fabs(T[1][1]);
    if (!(std::fabs(T[1][0]) < eps * sd)) { // splitOffTwoRows(1)
        const R p = R(0.5) * (T[0][0] - T[1][1]);
        const R q = p * p + T[1][0] * T[0][1];
        if (q >= R(0)) {
            const R z = std::sqrt(std::fabs(q));
            const R gp = p >= R(0) ? p + z : p - z, gq = T[1][0];
            R c, s;
            if (gq == R(0)) {
                c = gp < R(0) ? R(-1) : R(1);
                s = 0;
            } else if (gp == R(0)) {
                c = 0;
                s = gq < R(0) ? R(1) : R(-1);
            } else if (std::fabs(gp) > std::fabs(gq)) {
                R t = gq / gp, u = std::sqrt(R(1) + t * t);
                if (gp < R(0)) u = -u;
                c = R(1) / u;
                s = -t * c;
            } else {
                R t = gp / gq, u = std::sqrt(R(1) + t * t);
                if (gq < R(0)) u = -u;
                s = -R(1) / u;
                c = -t * s;
            }
            // rows: T <- G^T T (adjoint rotation on the left), columns: T <- T G, U <- U G
            for (int j = 0; j < 2; ++j) {
                const R x = T[0][j], y = T[1][j];
                T[0][j] = c * x - s * y;
                T[1][j] = s * x + c * y;
            }
            for (int i = 0; i < 2; ++i) {
                R x = T[i][0], y = T[i][1];
                T[i][0] = c * x - s * y;
                T[i][1] = s * x + c * y;
                x = U[i][0], y = U[i][1];
                U[i][0] = c * x - s * y;
                U[i][1] = s * x + c * y;
            }
            T[1][0] = 0;
        }
    } else {
        T[1][0] = 0;
    }
    for (auto &row : T)
        for (R &x : row) x *= scale;
    lam[0] = T[0][0];
    lam[1] = T[1][1];
    // back-substitution: X = [[1, x01], [0, 1]], x01 = -T01 / (T00 - lam1)
    const R w = T[0][0] - lam[1];
    const R nrm = std::fabs(T[0][0]) + std::fabs(T[0][1]) + std::fabs(T[1][1]);
    const R x01 = w != R(0) ? -T[0][1] / w : -T[0][1] / (eps * nrm);
    R v0[2] = {U[0][0], U[1][0]};
    R v1[2] = {U[0][0] * x01 + U[0][1], U[1][0] * x01 + U[1][1]};
    const R n0 = std::sqrt(v0[0] * v0[0] + v0[1] * v0[1]), n1 = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1]);
    V[0][0] = v0[0] / n0, V[1][0] = v0[1] / n0;
    V[0][1] = v1[0] / n1, V[1][1] = v1[1] / n1;
}

// splitComponentUsingPCA (src/PathGuiding.cpp:494-561): the component is
// replaced by two at +-(half a standard deviation along the principal axis
// of its directions' covariance), each with half the weight.  The reference
// takes the principal axis from a ROW of Eigen's eigenvector matrix
// (computeEigenValuesVectors, 453-486); kept.
template <class R> void splitComponent(Region<R> &g, int comp, float maxKappa, int maxK) {
    Pmm<R> &m = g.pmm;
    if (m.K >= maxK) return;
    const int tgt = m.K;
    R fx[3], fy[3];
    tangentFrame(m.mu[comp], fx, fy);
    R vx = 0, vy = 0, cxy = 0;
    if (g.covW[comp] > R(0)) vx = g.cov[comp][0], vy = g.cov[comp][1], cxy = g.cov[comp][2];
    R lam[2], V[2][2];
    eigen2x2(vx, cxy, cxy, vy, lam, V); // Matrix2f << c[0][0], c[1][0], c[0][1], c[1][1]
    // (value, row) pairs sorted by value with the comparator <=, then the larger
    int order[2] = {0, 1};
    if (lam[1] <= lam[0]) order[0] = 1, order[1] = 0;
    const int pick = lam[order[1]] > lam[order[0]] ? order[1] : order[0];
    const R ev = lam[pick];
    const R pc[2] = {V[pick][0], V[pick][1]};
    const R off = std::min(R(1), R(0.5) * std::sqrt(ev));
    const R px = pc[0] * off, py = pc[1] * off;
    const R z = std::sqrt(R(1) - off * off);
    R muA[3], muB[3];
    for (int c = 0; c < 3; ++c) {
        muA[c] = px * fx[c] + py * fy[c] + z * m.mu[comp][c];
        muB[c] = -px * fx[c] - py * fy[c] + z * m.mu[comp][c];
    }
    m.K = tgt + 1;
    const R srcR = m.r[comp], srcW = m.w[comp];
    const R splitW = srcW * R(0.5);
    const R maxAvgCos = (R)kappaToMeanCosine(maxKappa);
    const R splitR = z > R(0) ? std::min(maxAvgCos, srcR / z) : maxAvgCos;
    const R splitK = meanCosineToKappa(splitR);
    m.setKappaAndR(comp, splitK, splitR);
    for (int c = 0; c < 3; ++c) m.mu[comp][c] = muA[c];
    m.w[comp] = splitW;
    m.setKappaAndR(tgt, splitK, splitR);
    for (int c = 0; c < 3; ++c) m.mu[tgt][c] = muB[c];
    m.w[tgt] = splitW;
    // IncrementalDistance::split: the new component inherits the distance, half the weight each
    g.dist[tgt] = g.dist[comp];
    g.distW[comp] = g.distW[tgt] = R(0.5) * g.distW[comp];
    g.chi[comp] = g.chi[tgt] = g.chiN[comp] = g.chiN[tgt] = 0;
    for (int c = 0; c < 3; ++c) g.cov[comp][c] = g.cov[tgt][c] = 0;
    g.covW[comp] = g.covW[tgt] = 0;
}

// VMMFactory::maskedFit (225-297): MAP-EM on the masked components only;
// their total weight is kept, the others are left as they are.
template <class R> void maskedFit(const Props &P, Pmm<R> &m, const std::vector<Sample> &s, const bool *mask) {
    const int K = m.K;
    R oldActive = 0;
    for (int k = 0; k < K; ++k)
        if (mask[k]) oldActive += removeWeightPrior(P, m.w[k], K);
    const uint32_t maxItr = s.size() > (size_t)(K * 2) ? P.maxItr : 0;
    R sumWeight = 0, ll = 0, lastLL = 0, absInvLastLL = 0;
    const R maxMeanCosine = kappaToMeanCosine((R)P.maxKappa);
    for (uint32_t i = 0; i < maxItr; ++i) {
        R sw = 0;
        Stats<R> st = statsFromSamples(m, s, i == 0 ? &sw : nullptr, i >= P.minItr ? &ll : nullptr);
        if (i == 0) sumWeight = sw;
        R activeSum = 0;
        for (int k = 0; k < K; ++k)
            if (mask[k]) activeSum += st.W[k];
        const R norm = oldActive * (sumWeight / activeSum);
        for (int k = 0; k < K; ++k) {
            if (!mask[k]) continue;
            for (int c = 0; c < 3; ++c) st.rW[k][c] *= norm;
            st.W[k] *= norm;
        }
        for (int k = 0; k < K; ++k) { // maskedParameterUpdate (458-486)
            if (!mask[k]) continue;
            const R mixtureWeight = applyWeightPrior(P, st.W[k] / sumWeight, K);
            const R rLength = std::sqrt(st.rW[k][0] * st.rW[k][0] + st.rW[k][1] * st.rW[k][1] + st.rW[k][2] * st.rW[k][2]);
            if (rLength > R(0))
                for (int c = 0; c < 3; ++c) m.mu[k][c] = st.rW[k][c] / rLength;
            const R avgCosine = rLength / st.W[k];
            R withPrior = applyAvgCosinePrior(P, avgCosine, (R)m.totalNumSamples * mixtureWeight);
            withPrior = std::isnan(withPrior) ? maxMeanCosine : std::min(withPrior, maxMeanCosine);
            m.w[k] = mixtureWeight;
            m.setKappaAndR(k, meanCosineToKappa(withPrior), withPrior);
        }
        if (i >= P.minItr) {
            if (i > P.minItr) {
                const R rel = (ll - lastLL) * absInvLastLL;
                if (rel < (R)P.relLogLikelihoodThreshold) break;
            }
            lastLL = ll;
            absInvLastLL = R(1) / std::fabs(ll);
        }
    }
}

// PathGuiding::splitAll (563-632): split every component whose weighted
// chi-square divergence reaches splitMinDivergence (largest first, up to the
// free slots), refit the split components, repeat until none qualifies.
template <class R>
uint32_t splitAll(const Props &P, const SplitMergeProps &SP, Region<R> &g, const std::vector<Sample> &s, bool fit) {
    Pmm<R> &m = g.pmm;
    const int maxK = SP.maxComponents;
    if (m.K >= maxK) return 0;
    const bool firstFit = s.size() == m.totalNumSamples;
    uint32_t total = 0;
    for (;;) {
        std::vector<std::pair<R, int>> c;
        for (int k = 0; k < m.K; ++k) {
            const bool seenEnough = g.chiN[k] > (R)SP.minSamplesForSplitting;
            const R wd = (firstFit || seenEnough) ? chiDivergence(g, k) * m.w[k] : R(0);
            c.push_back({std::fabs(wd), k});
        }
        const size_t nCand = partitionLikeLibstdcxx(c, [&](const std::pair<R, int> &x) {
            return x.first >= (R)SP.splitMinDivergence;
        });
        const uint32_t numSplits = (uint32_t)std::min<size_t>(nCand, (size_t)(maxK - m.K));
        if (numSplits == 0) break;
        if (nCand > numSplits)
            std::partial_sort(c.begin(), c.begin() + numSplits, c.begin() + nCand,
                              [](const std::pair<R, int> &x, const std::pair<R, int> &y) { return x.first > y.first; });
        bool mask[kMaxK] = {};
        for (uint32_t i = 0; i < numSplits; ++i) {
            mask[c[i].second] = true;
            mask[m.K] = true;
            splitComponent(g, c[i].second, P.maxKappa, maxK);
        }
        if (fit) {
            pmmApplyPrior(m, P.vPrior);
            maskedFit(P, m, s, mask);
            chiUpdate(g, s, mask);
            covUpdate(g, s, mask);
            pmmRemovePrior(m, P.vPrior);
        }
        total += numSplits;
    }
    return total;
}

// postFit with splitAndMerge (src/PathGuiding.cpp:414-441)
template <class R> void splitAndMerge(const Props &P, const SplitMergeProps &SP, Region<R> &g, const std::vector<Sample> &s) {
    Pmm<R> &m = g.pmm;
    g.samplesSinceLastMerge += (uint32_t)s.size();
    if (g.samplesSinceLastMerge > SP.minSamplesForMerging) {
        pmmRemovePrior(m, P.vPrior);
        mergeAll(SP, g);
        pmmApplyPrior(m, P.vPrior);
        g.samplesSinceLastMerge = 0;
    }
    chiUpdate(g, s, nullptr);
    covUpdate(g, s, nullptr);
    const bool firstFit = s.size() == m.totalNumSamples;
    const bool fit = firstFit || s.size() > (size_t)SP.minSamplesForPostSplitFitting;
    pmmRemovePrior(m, P.vPrior);
    splitAll(P, SP, g, s, fit);
    pmmApplyPrior(m, P.vPrior);
}

// V: rtx_vmm_theta (16 lobes, mixtures of at most 16) or rtx_vmm_theta_wide
template <class R, class V> void toVmm(const Region<R> &g, bool parallax, V &v) { // syncPMMsToVMM_Thetas
    constexpr int N = (int)(sizeof(v.pi) / sizeof(v.pi[0]));
    v = V{};
    v.usedDistributions = g.pmm.K;
    for (int k = 0; k < N; ++k) v.thetas[k].distance = -1.0f;
    for (int k = 0; k < g.pmm.K; ++k) {
        rtx_vmf_theta &t = v.thetas[k];
        v.pi[k] = (float)g.pmm.w[k];
        for (int c = 0; c < 3; ++c) t.mu[c] = (float)g.pmm.mu[k][c];
        // VMF_Theta::setK (src/PathGuiding.h:44-50), exact exp as the host computes it
        float kap = (float)g.pmm.kappa[k];
        kap = kap < 1e-3f ? 0.0f : kap;
        t.k = kap;
        t.norm = (float)(kap / (2 * kPi * (1 - std::exp(-2.0 * kap))));
        t.eMin2K = (float)std::exp(-2.0 * kap);
    }
    if (parallax) {
        for (int c = 0; c < 3; ++c) v.meanPosition[c] = (float)g.mean[c];
        for (int k = 0; k < N; ++k) {
            float d = k < g.pmm.K ? (float)g.dist[k] : std::numeric_limits<float>::infinity();
            d = std::isinf(d) ? -1.0f : d;
            v.thetas[k].distance = d;
            for (int c = 0; c < 3; ++c) v.thetas[k].target[c] = v.meanPosition[c] + d * v.thetas[k].mu[c];
        }
    }
}

} // namespace

struct orc_guide {
    Props props;
    SplitMergeProps sm;
    int prec = 64;
    bool parallax = true;
    bool firstFit = true;
    std::vector<rtx_aabb> boxes;
    std::vector<Region<double>> r64;
    std::vector<Region<float>> r32;
};

namespace {
std::string g_guideErr;

template <class R> std::vector<Region<R>> &regionsOf(orc_guide *g);
template <> std::vector<Region<double>> &regionsOf<double>(orc_guide *g) { return g->r64; }
template <> std::vector<Region<float>> &regionsOf<float>(orc_guide *g) { return g->r32; }

// PathGuiding::update for every region with samples (src/PathGuiding.cpp:276-293)
template <class R> void updateAll(orc_guide *g, const rtx_directional_data *dd, size_t n) {
    std::vector<Region<R>> &regs = regionsOf<R>(g);
    std::vector<std::vector<Sample>> per(regs.size());
    size_t taken = 0; // maxSamplesPerUpdate: the first valid samples in slot order
    for (size_t i = 0; i < n; ++i) {
        const rtx_directional_data &d = dd[i];
        if (d.flags == kInvalid || d.flags >= regs.size()) continue;
        if (g->sm.maxSamplesPerUpdate && taken == g->sm.maxSamplesPerUpdate) break;
        taken++;
        Sample x;
        for (int c = 0; c < 3; ++c) {
            x.pos[c] = d.position[c];
            x.dir[c] = d.direction[c];
        }
        x.weight = d.weight;
        x.distance = d.distance;
        x.pdf = d.pdf;
        per[d.flags].push_back(x);
    }
    for (size_t r = 0; r < regs.size(); ++r) {
        if (per[r].empty()) continue;
        Region<R> &reg = regs[r];
        std::vector<Sample> &s = per[r];
        if (g->parallax) { // preFit, src/PathGuiding.cpp:350-391
            const rtx_aabb &b = g->boxes[r];
            R mean[3];
            for (int c = 0; c < 3; ++c) mean[c] = (R)(b.min[c] + 0.5f * (b.max[c] - b.min[c]));
            for (int c = 0; c < 3; ++c) {
                reg.lastMean[c] = reg.mean[c];
                reg.mean[c] = mean[c];
            }
            for (Sample &x : s) {
                if (x.distance > 0.0) {
                    double v[3], len = 0;
                    for (int c = 0; c < 3; ++c) {
                        v[c] = x.pos[c] + x.distance * x.dir[c] - (double)mean[c];
                        len += v[c] * v[c];
                    }
                    len = std::sqrt(len);
                    for (int c = 0; c < 3; ++c) x.dir[c] = v[c] / len;
                    x.distance = len;
                } else {
                    x.distance = std::numeric_limits<double>::infinity();
                }
                for (int c = 0; c < 3; ++c) x.pos[c] = (double)mean[c];
            }
            if (!g->firstFit) {
                R shift[3];
                for (int c = 0; c < 3; ++c) shift[c] = reg.lastMean[c] - reg.mean[c];
                reposition(reg, shift);
            }
        }
        emFit(g->props, reg.pmm, s, !g->firstFit);
        if (g->sm.splitAndMerge) splitAndMerge(g->props, g->sm, reg, s);
        if (g->parallax) updateDistances(reg, s);
    }
    g->firstFit = false;
    // region splitting (src/PathGuiding.cpp:295-301, splitRegion 328-348)
    if (g->sm.splitRegions) {
        const size_t count = regs.size();
        const size_t cap = g->sm.maxRegions ? std::min<size_t>(g->sm.maxRegions, 1u << 20) : 1u << 20;
        for (size_t r = 0; r < count; ++r) {
            if (!(regs[r].pmm.numSamples > (R)g->sm.samplesForRegionSplit)) continue;
            if (g->boxes.size() >= cap) break; // the cap: no further region is split
            const rtx_aabb b = g->boxes[r];
            const float size[3] = {b.max[0] - b.min[0], b.max[1] - b.min[1], b.max[2] - b.min[2]};
            const int axis = size[0] > size[1] ? (size[0] > size[2] ? 0 : 2) : (size[1] > size[2] ? 1 : 2);
            rtx_aabb left = b, right = b;
            left.max[axis] -= 0.5f * size[axis];
            right.min[axis] += 0.5f * size[axis];
            g->boxes[r] = left;
            g->boxes.push_back(right);
            regs[r].pmm.numSamples *= R(0.25f);
            regs[r].pmm.sampleWeight *= R(0.25f);
            const Region<R> copy = regs[r];
            regs.push_back(copy);
        }
    }
}
} // namespace

extern "C" {

const char *orc_guide_last_error(void) { return g_guideErr.c_str(); }

int orc_guide_create(const rtx_aabb *sceneBox, uint32_t splitCount, int parallax, int prec, orc_guide **out) {
    if (!sceneBox || !out || (prec != 32 && prec != 64) || splitCount > 20) {
        g_guideErr = "invalid argument";
        return RTX_E_INVALID;
    }
    orc_guide *g = new orc_guide;
    g->prec = prec;
    g->parallax = parallax != 0;
    // Aabb::addEpsilon then splitCount rounds of splitting every box in half
    // along its largest axis (src/Shapes.h:30-54, src/PathGuiding.cpp:82-106)
    rtx_aabb b = *sceneBox;
    rtx_aabb e;
    for (int c = 0; c < 3; ++c) {
        float ext = b.max[c] - b.min[c], ctr = b.min[c] + 0.5f * ext;
        e.min[c] = ctr - 0.50001f * ext;
        e.max[c] = ctr + 0.50001f * ext;
    }
    std::vector<rtx_aabb> boxes{e};
    for (uint32_t it = 0; it < splitCount; ++it) {
        std::vector<rtx_aabb> next;
        for (const rtx_aabb &x : boxes) {
            float size[3] = {x.max[0] - x.min[0], x.max[1] - x.min[1], x.max[2] - x.min[2]};
            int axis = size[0] > size[1] ? (size[0] > size[2] ? 0 : 2) : (size[1] > size[2] ? 1 : 2);
            rtx_aabb l = x, r = x;
            l.max[axis] -= 0.5f * size[axis];
            r.min[axis] += 0.5f * size[axis];
            next.push_back(l);
            next.push_back(r);
        }
        boxes.swap(next);
    }
    g->boxes = boxes;
    if (prec == 64) {
        g->r64.resize(boxes.size());
        for (auto &r : g->r64) initialize(g->props, r.pmm);
    } else {
        g->r32.resize(boxes.size());
        for (auto &r : g->r32) initialize(g->props, r.pmm);
    }
    *out = g;
    return RTX_OK;
}

void orc_guide_destroy(orc_guide *g) { delete g; }

uint32_t orc_guide_regions(const orc_guide *g, rtx_aabb *boxes, rtx_vmm_theta *vmms, uint32_t capacity) {
    const uint32_t n = (uint32_t)g->boxes.size();
    for (uint32_t i = 0; i < std::min(n, capacity); ++i) {
        if (boxes) boxes[i] = g->boxes[i];
        if (vmms) {
            if (g->prec == 64) toVmm(g->r64[i], g->parallax, vmms[i]);
            else toVmm(g->r32[i], g->parallax, vmms[i]);
        }
    }
    return n;
}

uint32_t orc_guide_regions_wide(const orc_guide *g, rtx_aabb *boxes, rtx_vmm_theta_wide *vmms, uint32_t capacity) {
    const uint32_t n = (uint32_t)g->boxes.size();
    for (uint32_t i = 0; i < std::min(n, capacity); ++i) {
        if (boxes) boxes[i] = g->boxes[i];
        if (vmms) {
            if (g->prec == 64) toVmm(g->r64[i], g->parallax, vmms[i]);
            else toVmm(g->r32[i], g->parallax, vmms[i]);
        }
    }
    return n;
}

int orc_guide_update(orc_guide *g, const rtx_directional_data *samples, size_t n) {
    if (!g || (n && !samples)) {
        g_guideErr = "NULL argument";
        return RTX_E_INVALID;
    }
    if (g->prec == 64) updateAll<double>(g, samples, n);
    else updateAll<float>(g, samples, n);
    return RTX_OK;
}

float orc_fastexp(float x) { return fastexp<float>(x); }

int orc_guide_set_options(orc_guide *g, const rtx_guiding_options *o) {
    if (!g || !o) {
        g_guideErr = "NULL argument";
        return RTX_E_INVALID;
    }
    g->sm.splitAndMerge = o->splitAndMerge != 0;
    g->sm.minSamplesForMerging = o->minSamplesForMerging;
    g->sm.minSamplesForSplitting = o->minSamplesForSplitting;
    g->sm.minSamplesForPostSplitFitting = o->minSamplesForPostSplitFitting;
    g->sm.splitMinDivergence = o->splitMinDivergence;
    g->sm.mergeMaxDivergence = o->mergeMaxDivergence;
    g->sm.splitRegions = o->splitRegions != 0;
    g->sm.samplesForRegionSplit = o->samplesForRegionSplit;
    if (o->maxComponents < 1 || o->maxComponents > RTX_MAX_LOBES) {
        g_guideErr = "maxComponents must be in [1, RTX_MAX_LOBES]";
        return RTX_E_INVALID;
    }
    g->sm.maxComponents = (int)o->maxComponents;
    g->sm.maxRegions = o->maxRegions;
    g->sm.maxSamplesPerUpdate = o->maxSamplesPerUpdate;
    return RTX_OK;
}

// Test hooks (FP64): lobes as 8 doubles (weight, mu xyz, kappa, r, norm, e^-2kappa).
static Lobe<double> lobeFrom(const double *v) {
    return Lobe<double>{v[0], {v[1], v[2], v[3]}, v[4], v[5], v[6], v[7]};
}
static void lobeTo(const Lobe<double> &l, double *v) {
    v[0] = l.w, v[1] = l.mu[0], v[2] = l.mu[1], v[3] = l.mu[2], v[4] = l.kappa, v[5] = l.r, v[6] = l.norm, v[7] = l.e2k;
}
void orc_vmf_combine(const double a[8], const double b[8], int divide, double out[8]) {
    lobeTo(lobeCombine(lobeFrom(a), lobeFrom(b), divide != 0), out);
}
void orc_vmf_merge(const double a[8], const double b[8], double out[8]) {
    lobeTo(lobeMerge(lobeFrom(a), lobeFrom(b)), out);
}
double orc_merge_metric(const double a[8], const double b[8]) {
    Pmm<double> m;
    m.K = 2;
    const Lobe<double> la = lobeFrom(a), lb = lobeFrom(b);
    const Lobe<double> *l[2] = {&la, &lb};
    for (int k = 0; k < 2; ++k) {
        m.w[k] = l[k]->w, m.kappa[k] = l[k]->kappa, m.r[k] = l[k]->r, m.norm[k] = l[k]->norm, m.e2k[k] = l[k]->e2k;
        for (int c = 0; c < 3; ++c) m.mu[k][c] = l[k]->mu[c];
    }
    return mergeMetric(m, 0, 1);
}
// ---- the MAP-EM's steps on one FP64 mixture (VMMFactory.h:157-216,
// 305-362, 385-486), for tests that pin them by closed forms
static Props propsFrom(const double p[7]) {
    Props P;
    P.minItr = (uint32_t)p[0];
    P.maxItr = (uint32_t)p[1];
    P.relLogLikelihoodThreshold = (float)p[2];
    P.maxKappa = (float)p[3];
    P.vPrior = (float)p[4];
    P.rPrior = (float)p[5];
    P.rPriorWeight = (float)p[6];
    return P;
}
// m[]: K, then per component (w, mu xyz, kappa, r), then sampleWeight,
// numSamples, totalNumSamples: 1 + 6 * kMaxK + 3 doubles
constexpr int kMixDoubles = 1 + 6 * kMaxK + 3;
static bool mixFrom(const double *v, Pmm<double> &m) {
    m = Pmm<double>{};
    m.K = (int)v[0];
    if (m.K < 1 || m.K > kMaxK) return false;
    for (int k = 0; k < kMaxK; ++k) {
        const double *c = v + 1 + 6 * k;
        m.w[k] = c[0];
        m.mu[k][0] = c[1], m.mu[k][1] = c[2], m.mu[k][2] = c[3];
        m.kappa[k] = c[4];
        m.r[k] = c[5];
        m.calNorm(k);
    }
    m.sampleWeight = v[1 + 6 * kMaxK];
    m.numSamples = v[2 + 6 * kMaxK];
    m.totalNumSamples = (uint64_t)v[3 + 6 * kMaxK];
    return true;
}
static void mixTo(const Pmm<double> &m, double *v) {
    v[0] = m.K;
    for (int k = 0; k < kMaxK; ++k) {
        double *c = v + 1 + 6 * k;
        c[0] = m.w[k], c[1] = m.mu[k][0], c[2] = m.mu[k][1], c[3] = m.mu[k][2], c[4] = m.kappa[k], c[5] = m.r[k];
    }
    v[1 + 6 * kMaxK] = m.sampleWeight;
    v[2 + 6 * kMaxK] = m.numSamples;
    v[3 + 6 * kMaxK] = (double)m.totalNumSamples;
}
static std::vector<Sample> samplesFrom(const double *dirs, const double *weights, size_t n) {
    std::vector<Sample> s(n);
    for (size_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) s[i].dir[c] = dirs[3 * i + c], s[i].pos[c] = 0;
        s[i].weight = weights[i];
        s[i].distance = 0;
        s[i].pdf = 1;
    }
    return s;
}
// E-step statistics: per component the responsibility-weighted resultant
// rW (3 doubles) and weight W; the samples' weight sum and log-likelihood
int orc_em_estep(const double *mix, const double *dirs, const double *weights, size_t n, double *rW, double *W,
                 double *sumWeight, double *logLik) {
    Pmm<double> m;
    if (!mixFrom(mix, m)) return RTX_E_INVALID;
    const Stats<double> st = statsFromSamples(m, samplesFrom(dirs, weights, n), sumWeight, logLik);
    for (int k = 0; k < m.K; ++k) {
        for (int c = 0; c < 3; ++c) rW[3 * k + c] = st.rW[k][c];
        W[k] = st.W[k];
    }
    return RTX_OK;
}
// M-step of every component from given statistics (parameterUpdate)
int orc_em_mstep(const double props[7], double *mix, const double *rW, const double *W) {
    Pmm<double> m;
    if (!mixFrom(mix, m)) return RTX_E_INVALID;
    const Props P = propsFrom(props);
    for (int k = 0; k < m.K; ++k) parameterUpdate(P, m, k, rW + 3 * k, W[k]);
    mixTo(m, mix);
    return RTX_OK;
}
// fit (update = 0) or updateFit (update = 1) of a mixture to weighted directions
int orc_em_fit(const double props[7], double *mix, const double *dirs, const double *weights, size_t n, int update) {
    Pmm<double> m;
    if (!mixFrom(mix, m)) return RTX_E_INVALID;
    emFit(propsFrom(props), m, samplesFrom(dirs, weights, n), update != 0);
    mixTo(m, mix);
    return RTX_OK;
}
int orc_em_mix_doubles(void) { return kMixDoubles; }
// The element order partitionLikeLibstdcxx leaves for keys[] and "key <= pivot"
// (out: the original indices in their new order); returns the partition point.
size_t orc_partition_order(const double *keys, size_t n, double pivot, uint32_t *out) {
    std::vector<std::pair<double, uint32_t>> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = {keys[i], (uint32_t)i};
    const size_t mid = partitionLikeLibstdcxx(v, [&](const std::pair<double, uint32_t> &x) { return x.first <= pivot; });
    for (size_t i = 0; i < n; ++i) out[i] = v[i].second;
    return mid;
}

void orc_eigen2x2(const double a[4], double lam[2], double vec[4]) {
    double V[2][2];
    eigen2x2(a[0], a[1], a[2], a[3], lam, V);
    vec[0] = V[0][0], vec[1] = V[0][1], vec[2] = V[1][0], vec[3] = V[1][1];
}

} // extern "C"
