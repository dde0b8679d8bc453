// Path-guiding sampling on the device: the region lookup that replaces the
// guiding acceleration structure (src/PathGuiding.cpp:153-181,
// raytrace.guiding.rint/.rchit), the vMF mixture of shaders/guiding.glsl and
// getNewDirection (shaders/raytrace.rgen:923-960).
#pragma once

#include "dev_math.h"
#include "dev_shading.h"
#include "internal.h"

namespace rtx {

constexpr float kInv4Pi = 0.07957747155f; // M_INV_4PI, guiding.glsl:7

RTX_DI bool inBox(const float *lo, const float *hi, V3 x) {
    return lo[0] <= x.x && x.x <= hi[0] && lo[1] <= x.y && x.y <= hi[1] && lo[2] <= x.z && x.z <= hi[2];
}

// getGuidingRegion (raytrace.rgen:904-921): the region whose box holds x,
// bounds inclusive as in raytrace.guiding.rint:6-10; of several (a point on a
// shared face) the smallest index, -1 outside every region.  Walks the
// host-built box hierarchy; regions are disjoint apart from shared faces, so
// a lookup visits about log2(regions) nodes.
RTX_DI int guidingRegion(const DevGuiding &G, V3 x) {
    if (G.numNodes == 0) return -1;
    int stack[32];
    int sp = 0, node = 0, best = -1;
    for (;;) {
        const GuideNode &nd = G.nodes[node];
        if (inBox(nd.lo, nd.hi, x)) {
            if (nd.region >= 0) {
                if (best < 0 || nd.region < best) best = nd.region;
            } else {
                if (sp < 32) stack[sp++] = nd.right;
                node = node + 1;
                continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    return best;
}

// vMF lobe (guiding.glsl:34-47); with parallax compensation a lobe that has
// a distance points at its target as seen from x.
RTX_DI float vmfPdf(V3 wo, const rtx_vmf_theta &t, V3 x, bool parallax) {
    if (t.k == 0.0f) return kInv4Pi;
    V3 mu = v3(t.mu);
    if (parallax && t.distance > 0.0f) mu = normalize(v3(t.target) - x);
    return t.norm * expf(t.k * (dot(mu, wo) - 1.0f));
}

RTX_DI float vmmPdf(V3 wo, const rtx_vmm_theta_wide &m, V3 x, bool parallax) { // guiding.glsl:57-65
    float res = 0.0f;
    for (int i = 0; i < m.usedDistributions; ++i) res += m.pi[i] * vmfPdf(wo, m.thetas[i], x, parallax);
    return res;
}

RTX_DI V3 sampleVmf(Rng &r, const rtx_vmf_theta &t, V3 x, bool parallax) { // guiding.glsl:67-89
    if (t.k > 0.0f) {
        float r1 = r.next();
        float r2 = r.next();
        float cosTheta = 1.0f + logf(1.0f + t.eMin2K * r1 - r1) / t.k;
        float sinTheta = 1.0f - cosTheta * cosTheta <= 0.0f ? 0.0f : sqrtf(1.0f - cosTheta * cosTheta);
        float phi = 2.0f * kPi * r2;
        V3 mu = v3(t.mu);
        if (parallax && t.distance > 0.0f) mu = normalize(v3(t.target) - x);
        return toWorld(v3(sinTheta * cosf(phi), sinTheta * sinf(phi), cosTheta), mu);
    }
    return randomOnUnitSphere(r);
}

RTX_DI V3 sampleVmm(Rng &r, const rtx_vmm_theta_wide &m, V3 x, bool parallax) { // guiding.glsl:91-104
    float u = r.next();
    int i = 0;
    const int last = m.usedDistributions - 1;
    float piSum = m.pi[0];
    while (piSum < u && i < last) {
        i++;
        piSum += m.pi[i];
    }
    return sampleVmf(r, m.thetas[i], x, parallax);
}

struct GuidingParams {
    bool on;
    bool parallax;
    float prob;
};

// getNewDirection (raytrace.rgen:923-960)
RTX_DI float getNewDirection(Rng &r, const DevGuiding &G, const GuidingParams &gp, const Mat &m, V3 x, V3 n, V3 wi,
                             bool frontFace, V3 &newDir, int *regionOut = nullptr) {
    if (!gp.on || hasDiscreteDirection(m.type)) return sampleBSDF(r, m, wi, n, frontFace, newDir);
    const int region = guidingRegion(G, x);
    if (regionOut) *regionOut = region;
    if (region < 0) return 0.0f;
    const rtx_vmm_theta_wide &vm = G.vmms[region];
    float pdfMat;
    if (r.next() < gp.prob) {
        newDir = sampleVmm(r, vm, x, gp.parallax);
        pdfMat = pdfBSDF(m, n, wi, newDir);
    } else {
        pdfMat = sampleBSDF(r, m, wi, n, frontFace, newDir);
    }
    if (dot(newDir, n) < 0.0f || pdfMat <= 0.0f) return 0.0f;
    const float pdfGuiding = vmmPdf(newDir, vm, x, gp.parallax);
    if (isnan(pdfGuiding)) return sampleBSDF(r, m, wi, n, frontFace, newDir);
    return pdfMat * (1.0f - gp.prob) + pdfGuiding * gp.prob; // mix(pdfMat, pdfGuiding, guidingProb)
}

} // namespace rtx
