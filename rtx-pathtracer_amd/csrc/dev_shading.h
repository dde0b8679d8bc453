// Device restatement of the reference's shading layer in FP32:
// samplers (shaders/random.glsl, transform.glsl), BSDFs, light sampling
// (shaders/raytrace.rgen:144-742) and the closest-hit shaders
// (raytrace.rchit, raytrace.sphere.rchit, raytrace.rmiss).
#pragma once

#include "dev_math.h"
#include "dev_texture.h"
#include "internal.h"

namespace rtx {

// ------------------------------------------------------- transform.glsl
// coordinateAxis (Nori), shaders/transform.glsl:7-17
RTX_DI void coordinateAxis(V3 z, V3 &x, V3 &y) {
    if (fabsf(z.x) > fabsf(z.y)) {
        float invLen = 1.0f / sqrtf(z.x * z.x + z.z * z.z);
        y = v3(z.z * invLen, 0.0f, -z.x * invLen);
    } else {
        float invLen = 1.0f / sqrtf(z.y * z.y + z.z * z.z);
        y = v3(0.0f, z.z * invLen, -z.y * invLen);
    }
    x = cross(y, z);
}
RTX_DI V3 toWorld(V3 v, V3 n) { // shaders/transform.glsl:30-35
    V3 x, y;
    coordinateAxis(n, x, y);
    return v.x * x + v.y * y + v.z * n;
}

// ---------------------------------------------------------- random.glsl
RTX_DI V3 randomOnUnitSphere(Rng &r) { // shaders/random.glsl:60-68
    V3 res = v3(0.0f);
    for (int it = 0; it < kMaxRejections; ++it) {
        float a = r.negPos();
        float b = r.negPos();
        float c = r.negPos();
        res = v3(a, b, c);
        if (!(length(res) > 1.0f)) break;
    }
    return normalize(res);
}
RTX_DI V3 randomInHemisphere(Rng &r, V3 n) { // shaders/random.glsl:71-83
    V3 res = randomOnUnitSphere(r);
    if (dot(n, res) < 0.0f) res = reflect(res, n);
    return res;
}
RTX_DI V3 randomInHemisphereCosine(Rng &r, V3 n) { // shaders/random.glsl:86-94
    float u = r.next();
    float sqrt_u = sqrtf(u);
    float phi = 2.0f * kPi * r.next();
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return toWorld(v3(sqrt_u * cp, sqrt_u * sp, sqrtf(1.0f - u)), n);
}
RTX_DI V3 randomInHemisphereCosinePower(Rng &r, V3 refl, float p) { // shaders/random.glsl:97-106
    float u = r.next();
    float cosTheta = powf(u, 1.0f / (p + 1.0f));
    float phi = 2.0f * kPi * r.next();
    float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return toWorld(v3(sinTheta * cp, sinTheta * sp, cosTheta), refl);
}
RTX_DI V3 randomOnSphere(Rng &r, const rtx_sphere &s, V3 &normal) { // shaders/random.glsl:108-115
    normal = v3(0.0f);
    for (int it = 0; it < kMaxRejections; ++it) {
        float a = r.negPos();
        float b = r.negPos();
        float c = r.negPos();
        normal = v3(a, b, c);
        if (!(length(normal) > 1.0f)) break;
    }
    normal = normalize(normal);
    return v3(s.center) + normal * s.radius;
}
RTX_DI V3 randomBeckmannNormal(Rng &r, float alpha, V3 n) { // shaders/random.glsl:127-136
    float thetaM = atanf(sqrtf(-alpha * alpha * logf(1.0f - r.next())));
    float phiM = 2.0f * kPi * r.next();
    float st, ct, sp, cp;
    sincosf(thetaM, &st, &ct);
    sincosf(phiM, &sp, &cp);
    return toWorld(v3(st * cp, st * sp, ct), n);
}

// -------------------------------------------------------------- BSDFs
RTX_DI float fresnel(float eta, float cosThetaI) { // shaders/raytrace.rgen:145-160
    float sinThetaSqr = eta * eta * (1.0f - cosThetaI * cosThetaI);
    if (sinThetaSqr > 1.0f) return 1.0f;
    float cosThetaT = sqrtf(1.0f - sinThetaSqr);
    float Rs = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    float Rp = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
    return (Rs * Rs + Rp * Rp) / 2.0f;
}
RTX_DI float fresnelConductor(float cosThetaI, float eta, float k) { // shaders/raytrace.rgen:163-175
    if (cosThetaI < 0.0f) cosThetaI = -cosThetaI;
    float ek = eta * eta + k * k;
    float c2 = cosThetaI * cosThetaI;
    float Rs2 = (ek * c2 - 2.0f * eta * cosThetaI + 1.0f) / (ek * c2 + 2.0f * eta * cosThetaI + 1.0f);
    float Rp2 = (ek - 2.0f * eta * cosThetaI + c2) / (ek + 2.0f * eta * cosThetaI + c2);
    return (Rs2 + Rp2) / 2.0f;
}
// Beckmann D (shaders/raytrace.rgen:179-202); pow(E, x) written as exp(x).
// The reference forms tan(acos(cos)); here tan^2 = |n x m|^2 / (n.m)^2,
// the same quantity without acos/tan's cancellation near the pole, which
// matters in FP32 at small alpha (exp(-tan^2 / alpha^2) magnifies errors
// in tan^2 by 1/alpha^2).  The reference's guards reduce to tan^2 >= 0.
RTX_DI float beckmannD(float alpha, V3 n, V3 m) {
    float cosTheta = dot(n, m);
    if (cosTheta <= 0.0f) return 0.0f;
    V3 x = cross(n, m);
    float c2 = cosTheta * cosTheta;
    float tan2 = fmaxf(dot(x, x) / c2, 0.0f);
    float alphaSqr = alpha * alpha;
    return expf(-tan2 / alphaSqr) / (kPi * alphaSqr * (c2 * c2));
}
// G1 with the reference's two typos kept (tan of a cosine; 2.577*a for
// 2.577*a^2), shaders/raytrace.rgen:204-222
RTX_DI float G1(float alpha, V3 n, V3 m, V3 v) {
    float thetaV = dot(v, n);
    float c = dot(v, m) / thetaV;
    if (c <= 0.0f) return 0.0f;
    float a = 1.0f / (alpha * tanf(thetaV));
    if (a >= 1.6f) return 1.0f;
    float a2 = a * a;
    return (3.535f * a + 2.181f * a2) / (1.0f + 2.276f * a + 2.577f * a);
}
RTX_DI float Gterm(float alpha, V3 i, V3 o, V3 n, V3 m) { // shaders/raytrace.rgen:224-230
    float r = G1(alpha, n, m, i) * G1(alpha, n, m, o);
    return r < 0.0f ? 0.0f : r;
}

// GGX (Walter, Marschner, Li and Torrance 2007), the rough conductor's
// distribution when rtx_material.distribution is RTX_DISTRIBUTION_GGX (new;
// the reference has Beckmann only).  D(m) = a^2 / (pi cos^4 (a^2 + tan^2)^2)
// = a^2 / (pi (a^2 cos^2 + sin^2)^2), with sin^2 = |n x m|^2 (no cancellation
// at small a near the pole).
RTX_DI float ggxD(float alpha, V3 n, V3 m) {
    float c = dot(n, m);
    if (c <= 0.0f) return 0.0f;
    V3 x = cross(n, m);
    float a2 = alpha * alpha;
    float den = a2 * c * c + dot(x, x);
    return a2 / (kPi * den * den);
}
// Smith G1 of GGX: chi+(v.m / v.n) 2 / (1 + sqrt(1 + a^2 tan^2 theta_v)).
RTX_DI float ggxG1(float alpha, V3 n, V3 m, V3 v) {
    float cv = dot(v, n);
    if (dot(v, m) / cv <= 0.0f) return 0.0f;
    V3 x = cross(v, n);
    float tan2 = dot(x, x) / (cv * cv);
    return 2.0f / (1.0f + sqrtf(1.0f + alpha * alpha * tan2));
}
// Microfacet normal with density D(m) cos(theta_m): tan(theta_m) =
// a sqrt(u1 / (1 - u1)), phi = 2 pi u2 (the same two draws as the Beckmann
// sampler, in the same order).
RTX_DI V3 randomGGXNormal(Rng &r, float alpha, V3 n) {
    float u1 = r.next();
    float t = alpha * sqrtf(u1 / (1.0f - u1));
    float phiM = 2.0f * kPi * r.next();
    float ct = 1.0f / sqrtf(1.0f + t * t);
    float st = t * ct;
    float sp, cp;
    sincosf(phiM, &sp, &cp);
    return toWorld(v3(st * cp, st * sp, ct), n);
}

// Material fields the shading reads, gathered once per vertex.  kd and ks
// are the diffuse and specular colours with the vertex's textures applied
// (diffuse() and phong(), raytrace.rgen:232-267); `diffuse` and `specular`
// stay untextured for the lobe choice (raytrace.rgen:302-303, 493-494).
struct Mat {
    V3 lightColor, diffuse, specular, kd, ks;
    float highlight, ior, iorInv, eta, k, roughness;
    int type, texD, texS;
    bool ggx; // rough conductor with the GGX distribution
};
RTX_DI Mat loadMat(const rtx_material *m) {
    const float4 *p = reinterpret_cast<const float4 *>(m);
    float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5];
    Mat r;
    r.lightColor = v3(a);
    r.diffuse = v3(b);
    r.specular = v3(c);
    r.highlight = c.w;
    // d = (transparency, refractionIndex, refractionIndexInv, eta)
    r.ior = d.y;
    r.iorInv = d.z;
    r.eta = d.w;
    // e = (k, roughness, texIdDiffuse, texIdSpecular)
    r.k = e.x;
    r.roughness = e.y;
    r.texD = __float_as_int(e.z);
    r.texS = __float_as_int(e.w);
    r.type = __float_as_int(f.x);
    r.ggx = __float_as_int(f.y) == RTX_DISTRIBUTION_GGX;
    r.kd = r.diffuse;
    r.ks = r.specular;
    return r;
}
// The texture look-ups of diffuse() / phong() at texture coordinates (u, v).
RTX_DI void applyTextures(const DevScene &S, Mat &m, float u, float v) {
    if (m.texD != -1) {
        float4 t = texSample(S, m.texD, u, v);
        m.kd = m.diffuse * v3(t.x, t.y, t.z);
    }
    if (m.texS != -1) {
        float4 t = texSample(S, m.texS, u, v);
        m.ks = m.specular * v3(t.x, t.y, t.z);
    }
}

RTX_DI bool hasDiscreteDirection(int type) { // shaders/raytrace.rgen:733-742
    return type == RTX_MAT_DIELECTRIC || type == RTX_MAT_SPECULAR || type == RTX_MAT_CONDUCTOR;
}
RTX_DI bool neeCapable(int type) { // shaders/raytrace.rgen:602-611
    return type == RTX_MAT_DIFFUSE || type == RTX_MAT_PHONG || type == RTX_MAT_LIGHT ||
           type == RTX_MAT_ROUGH_CONDUCTOR;
}

RTX_DI V3 phong(const Mat &m, V3 n, V3 wi, V3 wo) { // shaders/raytrace.rgen:242-267
    V3 res = v3(0.0f);
    float cosThetaWo = dot(wo, n);
    if (cosThetaWo > 0.0f) {
        res = res + m.kd / kPi;
        float dotReflDir = dot(reflect(-wo, n), wi);
        if (dotReflDir > 0.0f)
            res = res + ((m.highlight + 2.0f) / (2.0f * kPi)) * m.ks * powf(dotReflDir, m.highlight);
    }
    return cosThetaWo * res;
}
RTX_DI V3 roughConductor(const Mat &m, V3 n, V3 wi, V3 wo) { // shaders/raytrace.rgen:269-280
    V3 hr = normalize(wi + wo);
    float cosThetaIHr = dot(wi, hr);
    float cosThetaI = dot(wi, n);
    float cosThetaO = dot(wo, n);
    if (cosThetaI <= 0.0f) return v3(0.0f);
    float GD = m.ggx ? ggxG1(m.roughness, n, hr, wi) * ggxG1(m.roughness, n, hr, wo) * ggxD(m.roughness, n, hr)
                     : Gterm(m.roughness, wi, wo, n, hr) * beckmannD(m.roughness, n, hr);
    float val = fresnelConductor(cosThetaIHr, m.eta, m.k) * GD / (4.0f * cosThetaI * cosThetaO);
    return v3(cosThetaO * val);
}
RTX_DI float pdfBSDF(const Mat &m, V3 n, V3 wi, V3 wo) { // shaders/raytrace.rgen:284-332
    if (m.type == RTX_MAT_ROUGH_CONDUCTOR) {
        V3 hr = normalize(wi + wo);
        float pm = (m.ggx ? ggxD(m.roughness, n, hr) : beckmannD(m.roughness, n, hr)) * fabsf(dot(hr, n));
        if (pm <= 0.0f || dot(wo, hr) <= 0.0f) return 0.0f;
        return pm / (4.0f * fabsf(dot(wo, hr)));
    }
    if (m.type == RTX_MAT_PHONG) {
        if (dot(n, wo) < 0.0f) return 0.0f;
        float lD = length(m.diffuse), lS = length(m.specular);
        float sum = lD + lS;
        if (sum == 0.0f) return 0.0f;
        V3 reflected = reflect(-wi, n);
        float pdf = 0.0f;
        if (dot(reflected, wo) > 0.0f) {
            pdf = (m.highlight + 1.0f) * powf(dot(reflected, wo), m.highlight) / (2.0f * kPi);
            pdf *= lS / sum;
        }
        pdf += dot(wo, n) / kPi * lD / sum;
        return pdf;
    }
    return dot(wo, n) / kPi;
}
RTX_DI float sampleBSDF(Rng &r, const Mat &m, V3 wi, V3 n, bool frontFace, V3 &newDir) { // raytrace.rgen:483-552
    switch (m.type) {
    case RTX_MAT_ROUGH_CONDUCTOR: {
        V3 worldM = m.ggx ? randomGGXNormal(r, m.roughness, n) : randomBeckmannNormal(r, m.roughness, n);
        newDir = reflect(-wi, worldM);
        return pdfBSDF(m, n, wi, newDir);
    }
    case RTX_MAT_PHONG: {
        float lD = length(m.diffuse), lS = length(m.specular);
        float sum = lD + lS;
        if (sum == 0.0f) return 0.0f;
        if (r.next() * sum > lD) {
            V3 reflected = reflect(-wi, n);
            newDir = randomInHemisphereCosinePower(r, reflected, m.highlight);
            if (dot(n, newDir) < 0.0f) return 0.0f;
            return pdfBSDF(m, n, wi, newDir);
        }
        newDir = randomInHemisphereCosine(r, n);
        return pdfBSDF(m, n, wi, newDir);
    }
    case RTX_MAT_SPECULAR:
    case RTX_MAT_CONDUCTOR:
        newDir = reflect(-wi, n);
        return 1.0f;
    case RTX_MAT_DIELECTRIC: {
        float eta = frontFace ? m.iorInv : m.ior;
        float cosThetaI = dot(wi, n);
        float F = fresnel(eta, cosThetaI);
        if (r.next() > F) {
            newDir = refract(-wi, n, eta);
            return 1.0f - F;
        }
        newDir = reflect(-wi, n);
        return F;
    }
    default:
        newDir = randomInHemisphereCosine(r, n);
        return dot(newDir, n) / kPi;
    }
}
RTX_DI V3 evalBsdf(const Mat &m, V3 n, V3 wi, V3 wo, bool frontFace) { // shaders/raytrace.rgen:557-599
    switch (m.type) {
    case RTX_MAT_DIFFUSE:
    case RTX_MAT_LIGHT:
        return dot(wo, n) * (m.kd / kPi);
    case RTX_MAT_PHONG:
        return phong(m, n, wi, wo);
    case RTX_MAT_ROUGH_CONDUCTOR: {
        V3 res = roughConductor(m, n, wi, wo);
        if (isnan(res.x)) return v3(0.0f);
        return res;
    }
    case RTX_MAT_DIELECTRIC: {
        float cosThetaI = dot(n, wi);
        float eta = frontFace ? m.iorInv : m.ior;
        if (dot(n, wo) < 0.0f) return m.specular * (1.0f - fresnel(eta, cosThetaI));
        return m.specular * fresnel(eta, cosThetaI);
    }
    case RTX_MAT_CONDUCTOR:
        return v3(fresnelConductor(dot(wi, n), m.eta, m.k));
    case RTX_MAT_SPECULAR:
        return m.specular;
    default:
        return v3(0.0f);
    }
}
RTX_DI float misWeight(bool power, float p1, float p2) { // shaders/raytrace.rgen:341-348
    if (power) {
        float s = p1 * p1;
        return s / (s + p2 * p2);
    }
    return p1 / (p1 + p2);
}

// ------------------------------------------------------------- hit info
struct HitInfo {
    V3 worldPos, normal;
    float u, v;
    int matIndex;
    int instanceIndex;
    bool frontFace, isSphere;
};

// Closest-hit shaders for a hit on primitive `prim` at (t, bu, bv):
// triangles raytrace.rchit:16-58, spheres raytrace.sphere.rchit:13-35.
RTX_DI HitInfo hitInfo(const DevScene &S, int prim, float t, float bu, float bv, V3 o, V3 d) {
    HitInfo h;
    V3 normal;
    if ((uint32_t)prim >= S.numTris) {
        int si = prim - (int)S.numTris;
        const rtx_sphere &s = S.spheres[si];
        V3 pos = o + t * d;
        normal = normalize(pos - v3(s.center));
        h.worldPos = pos;
        h.u = h.v = 0.0f;
        h.matIndex = s.materialIndex;
        h.isSphere = true;
        h.instanceIndex = si;
    } else {
        const int4 ref = S.triRef[prim]; // one load instead of the instance -> index -> vertex chain
        const rtx_instance &in = S.instances[ref.x];
        const rtx_vertex &v0 = S.vertices[ref.y], &v1 = S.vertices[ref.z], &v2 = S.vertices[ref.w];
        float b0 = 1.0f - bu - bv;
        normal = v3(v0.normal) * b0 + v3(v1.normal) * bu + v3(v2.normal) * bv;
        normal = normalize(xform(in.normalTransform, normal, 0.0f));
        V3 wp = v3(v0.pos) * b0 + v3(v1.pos) * bu + v3(v2.pos) * bv;
        h.worldPos = xform(in.transform, wp, 1.0f);
        h.u = v0.texCoord[0] * b0 + v1.texCoord[0] * bu + v2.texCoord[0] * bv;
        h.v = v0.texCoord[1] * b0 + v1.texCoord[1] * bu + v2.texCoord[1] * bv;
        h.matIndex = v0.materialIndex;
        h.isSphere = false;
        h.instanceIndex = ref.x;
    }
    if (dot(d, normal) < 0.0f) {
        h.frontFace = true;
        h.normal = normal;
    } else {
        h.frontFace = false;
        h.normal = -normal;
    }
    return h;
}

// sampleLights, shaders/raytrace.rgen:358-473
RTX_DI float sampleLights(const DevScene &S, Rng &r, V3 origin, V3 normal, V3 &lightDir, V3 &lightColor,
                          float &lightDist) {
    int iRandomLight = r.integer(RTX_SIZE_LIGHT_RANDOM - 1);
    int iLight = S.lightIndex[iRandomLight];
    const rtx_light &light = S.lights[iLight];
    if (light.type == RTX_LIGHT_ENV_MAP) { // raytrace.rgen:406-421: uniform hemisphere, pdf 1 / 2pi
        lightDir = randomInHemisphere(r, normal);
        lightColor = envColor(S, lightDir);
        lightDist = kTMax;
        return light.sampleProb / (2.0f * kPi);
    }
    if (light.type == RTX_LIGHT_POINT) {
        V3 toLight = v3(light.pos) - origin;
        lightDist = length(toLight);
        lightDir = toLight / lightDist;
        lightColor = v3(light.color) / (lightDist * lightDist);
        return light.sampleProb;
    }
    if (light.type == RTX_LIGHT_SPHERE) {
        const rtx_sphere &s = S.spheres[light.instanceIndex];
        lightColor = v3(S.materials[s.materialIndex].lightColor);
        V3 sphereNormal;
        V3 position = randomOnSphere(r, s, sphereNormal);
        V3 toLight = position - origin;
        lightDist = length(toLight);
        lightDir = normalize(toLight);
        float cosThetaLight = dot(-lightDir, sphereNormal);
        return light.sampleProb * lightDist * lightDist / (cosThetaLight * light.area);
    }
    // area light
    int iRandomTri = r.integer(RTX_SIZE_TRI_RANDOM - 1);
    float rx = r.next();
    float ry = r.next();
    float sqrtx = sqrtf(rx);
    float b0 = 1.0f - sqrtx, b1 = sqrtx * (1.0f - ry), b2 = ry * sqrtx;
    V3 pos, nrm;
    if (S.emitTris) { // the face table's entry holds the face (one load level instead of five)
        const float4 *et = S.emitTris + 6 * ((size_t)iLight * RTX_SIZE_TRI_RANDOM + iRandomTri);
        const rtx_instance &in = S.lightInst[iLight];
        const float4 p0 = et[0], p1 = et[1], p2 = et[2];
        pos = v3(p0) * b0 + v3(p1) * b1 + v3(p2) * b2;
        nrm = v3(et[3]) * b0 + v3(et[4]) * b1 + v3(et[5]) * b2;
        pos = xform(in.transform, pos, 1.0f);
        nrm = normalize(xform(in.normalTransform, nrm, 0.0f));
        lightColor = v3(p0.w, p1.w, p2.w);
    } else {
        const rtx_instance &in = S.instances[light.instanceIndex];
        int model = in.modelIndex;
        const rtx_face_sample ts = S.triTable[(size_t)iLight * RTX_SIZE_TRI_RANDOM + iRandomTri];
        const uint32_t *idx = S.indices + S.indexBase[model] + 3 * ts.index;
        const rtx_vertex *vb = S.vertices + S.vertexBase[model];
        const rtx_vertex &v0 = vb[idx[0]], &v1 = vb[idx[1]], &v2 = vb[idx[2]];
        pos = v3(v0.pos) * b0 + v3(v1.pos) * b1 + v3(v2.pos) * b2;
        nrm = v3(v0.normal) * b0 + v3(v1.normal) * b1 + v3(v2.normal) * b2;
        pos = xform(in.transform, pos, 1.0f);
        nrm = normalize(xform(in.normalTransform, nrm, 0.0f));
        lightColor = v3(S.materials[v0.materialIndex].lightColor);
    }
    V3 toLight = pos - origin;
    lightDist = length(toLight);
    lightDir = toLight / lightDist;
    float cosThetaLight = dot(-lightDir, nrm);
    if (cosThetaLight < 0.0f) cosThetaLight = -cosThetaLight;
    return light.sampleProb * lightDist * lightDist / cosThetaLight / light.area;
}

} // namespace rtx
