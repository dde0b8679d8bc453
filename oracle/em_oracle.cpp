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
// published method so that the fits follow the same numbers.  Component
// split and merge (postFit's splitAll/mergeAll, 494-788) are not restated.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr int kMaxK = RTX_MAX_DISTRIBUTIONS;
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
    double weight, distance;
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

template <class R> void toVmm(const Region<R> &g, bool parallax, rtx_vmm_theta &v) { // syncPMMsToVMM_Thetas
    v = rtx_vmm_theta{};
    v.usedDistributions = g.pmm.K;
    for (int k = 0; k < kMaxK; ++k) v.thetas[k].distance = -1.0f;
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
        for (int k = 0; k < kMaxK; ++k) {
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
    for (size_t i = 0; i < n; ++i) {
        const rtx_directional_data &d = dd[i];
        if (d.flags == kInvalid || d.flags >= regs.size()) continue;
        Sample x;
        for (int c = 0; c < 3; ++c) {
            x.pos[c] = d.position[c];
            x.dir[c] = d.direction[c];
        }
        x.weight = d.weight;
        x.distance = d.distance;
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
        if (g->parallax) updateDistances(reg, s);
    }
    g->firstFit = false;
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

} // extern "C"
