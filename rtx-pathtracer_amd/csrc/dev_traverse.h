// BVH traversal on the device.  Replaces the RT-core traversal behind the
// reference's traceRayEXT calls (closest hit: shaders/raytrace.rgen:1011-1022;
// shadow any-hit with TerminateOnFirstHit: raytrace.rgen:622-638) and the
// sphere intersection shader (raytrace.sphere.rint:11-29).
//
// Stack-based traversal (Aila & Laine 2009) over four-wide nodes; the
// stack's first LDS_DEPTH entries live in LDS (one column per thread, so
// consecutive lanes hit consecutive banks), deeper entries in a private
// array.  The builder bounds the stack use (3 entries per level) by
// kStackCapacity.
#pragma once

#include "dev_math.h"
#include "dev_texture.h"
#include "internal.h"

namespace rtx {

struct TravRay {
    V3 o, d, inv;
    V3 oinv; // o * inv: slab plane distances are fma(l, inv, -oinv)
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

RTX_DI TravRay makeTravRay(V3 o, V3 d) {
    TravRay r;
    r.o = o;
    r.d = d;
    // zero direction components get a tiny stand-in so slab tests never see 0*inf
    float dx = d.x != 0.0f ? d.x : copysignf(1e-30f, d.x);
    float dy = d.y != 0.0f ? d.y : copysignf(1e-30f, d.y);
    float dz = d.z != 0.0f ? d.z : copysignf(1e-30f, d.z);
    r.inv = v3(1.0f / dx, 1.0f / dy, 1.0f / dz);
    r.oinv = v3(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    // watertight test set-up (Woop, Benthin, Wald, JCGT 2013)
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (sel(d, kz) < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    float dkz = sel(d, kz);
    r.Sx = sel(d, kx) / dkz;
    r.Sy = sel(d, ky) / dkz;
    r.Sz = 1.0f / dkz;
    return r;
}

// Slab test against [lo, hi]: plane distances as one FMA each,
// fma(l, inv, -o * inv); the absolute part of its error (o * inv rounded) is
// covered by the builder's box padding (bvh_build.cpp), the relative part by
// the 1 + 2*gamma_3 factor of Ize, "Robust BVH ray traversal" (JCGT 2013).
RTX_DI bool slab(float lx, float ly, float lz, float hx, float hy, float hz, const TravRay &r, float tmin,
                 float tmax, float &tnear) {
    float tx0 = fmaf(lx, r.inv.x, -r.oinv.x), tx1 = fmaf(hx, r.inv.x, -r.oinv.x);
    float ty0 = fmaf(ly, r.inv.y, -r.oinv.y), ty1 = fmaf(hy, r.inv.y, -r.oinv.y);
    float tz0 = fmaf(lz, r.inv.z, -r.oinv.z), tz1 = fmaf(hz, r.inv.z, -r.oinv.z);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.0000004f;
    tnear = tn;
    return tn <= fminf(tf, tmax);
}

// Watertight ray/triangle test.  Products are not contracted into FMAs so
// that the edge functions of a shared edge are exact negatives.  Returns
// the hit distance and the barycentrics of vertices 1 and 2 (the
// reference's hit attributes, raytrace.rchit:28).
RTX_DI bool isectTri(const TravRay &r, float4 a, float4 b, float4 c, float tmin, float tmax, float &t, float &u,
                     float &v) {
#pragma clang fp contract(off)
    V3 A = v3(a) - r.o, B = v3(b) - r.o, C = v3(c) - r.o;
    float Az0 = sel(A, r.kz), Bz0 = sel(B, r.kz), Cz0 = sel(C, r.kz);
    float Ax = sel(A, r.kx) - r.Sx * Az0, Ay = sel(A, r.ky) - r.Sy * Az0;
    float Bx = sel(B, r.kx) - r.Sx * Bz0, By = sel(B, r.ky) - r.Sy * Bz0;
    float Cx = sel(C, r.kx) - r.Sx * Cz0, Cy = sel(C, r.ky) - r.Sy * Cz0;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Az0, Bz = r.Sz * Bz0, Cz = r.Sz * Cz0;
    float T = U * Az + V * Bz + W * Cz;
    float tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    u = V / det;
    v = W / det;
    return true;
}

// raytrace.sphere.rint:11-29: both roots reported, nearest in [tmin, tmax] kept
RTX_DI bool isectSphere(const rtx_sphere &s, V3 o, V3 d, float tmin, float tmax, float &t) {
    V3 oc = o - v3(s.center);
    float dotDOC = dot(d, oc);
    float rootTerm = dotDOC * dotDOC - dot(oc, oc) + s.radius * s.radius;
    if (rootTerm < 0.0f) return false;
    float root = sqrtf(rootTerm);
    float t1 = -dotDOC + root, t2 = -dotDOC - root;
    bool hit = false;
    float best = tmax;
    if (t1 >= tmin && t1 <= best) {
        best = t1;
        hit = true;
    }
    if (t2 >= tmin && t2 <= best) {
        best = t2;
        hit = true;
    }
    t = best;
    return hit;
}

// Stack with LDS_DEPTH entries in LDS (column per thread, stride BLOCK)
// and the remainder in a private array of kStackCapacity - LDS_DEPTH
// entries.  The private array is a separate local of the caller, not a
// member: a dynamically indexed array inside the struct keeps the whole
// struct in scratch, sp included, and every push and pop then waits on a
// scratch load of sp.
// The two pointers carry their address spaces (3: LDS, 5: private), so that
// a pop, which reads one or the other, compiles to ds_read and scratch_load
// rather than to a flat load of a selected pointer.
typedef __attribute__((address_space(3))) int LdsInt;
typedef __attribute__((address_space(5))) int PrivInt;

template <int BLOCK, int LDS_DEPTH> struct HybridStack {
    static constexpr int kSpill = kStackCapacity - LDS_DEPTH;
    LdsInt *lds;
    PrivInt *priv;
    int sp;
    // ldsColumn: this thread's column of the block's LDS stack (stride BLOCK);
    // spill: kSpill private entries.
    RTX_DI void bind(int *ldsColumn, int *spill) {
        lds = (LdsInt *)ldsColumn;
        priv = (PrivInt *)spill;
    }
    RTX_DI void push(int x) {
        if (sp < LDS_DEPTH) lds[sp * BLOCK] = x;
        else priv[sp - LDS_DEPTH] = x;
        sp++;
    }
    RTX_DI int pop() {
        sp--;
        return sp < LDS_DEPTH ? lds[sp * BLOCK] : priv[sp - LDS_DEPTH];
    }
};

// Whole stack in a private array (kStackCapacity entries, the caller's).
struct PrivateStack {
    PrivInt *priv;
    int sp;
    RTX_DI PrivateStack(int *spill) : priv((PrivInt *)spill), sp(0) {}
    RTX_DI void push(int x) { priv[sp++] = x; }
    RTX_DI int pop() { return priv[--sp]; }
};

struct VisitCount {
    uint32_t nodes = 0, prims = 0;
};

constexpr int32_t kLinkDone = (int32_t)0x80000000; // no node: a leaf never encodes to this

template <class Stack> RTX_DI int popOrDone(Stack &st) { return st.sp == 0 ? kLinkDone : st.pop(); }

// One four-wide node visit: of the children whose boxes the ray enters within
// [tmin, tmax], the nearest is returned and the others are pushed, farthest
// first; with none entered, the next stack entry (kLinkDone when empty).
template <class SceneT, class Stack>
RTX_DI int nodeStep4(const SceneT &S, const TravRay &r, float tmin, float tmax, int link, Stack &st) {
    const float4 *np = S.nodes + 8 * (size_t)link;
    const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
    const int4 ln = *reinterpret_cast<const int4 *>(np + 6);
    const float inf = __int_as_float(0x7f800000);
    float t[4], tn;
    int c[4] = {ln.x, ln.y, ln.z, ln.w};
    t[0] = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, r, tmin, tmax, tn) ? tn : inf;
    t[1] = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, r, tmin, tmax, tn) ? tn : inf;
    t[2] = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, r, tmin, tmax, tn) ? tn : inf;
    t[3] = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, r, tmin, tmax, tn) ? tn : inf;
    // sorting network on (t, link), ascending t
    auto cs = [&](int i, int j) {
        if (t[j] < t[i]) {
            float tt = t[i];
            t[i] = t[j];
            t[j] = tt;
            int cc = c[i];
            c[i] = c[j];
            c[j] = cc;
        }
    };
    cs(0, 1);
    cs(2, 3);
    cs(0, 2);
    cs(1, 3);
    cs(1, 2);
    if (t[3] < inf) st.push(c[3]);
    if (t[2] < inf) st.push(c[2]);
    if (t[1] < inf) st.push(c[1]);
    if (t[0] < inf) return c[0];
    return popOrDone(st);
}

// Closest hit (ANY = false) or first hit (ANY = true) in [tmin, tmax].
// Returns the primitive id (triangles 0..numTris-1, then spheres) or -1.
// Closest-hit ties in t go to the smaller primitive id.  Triangles marked
// with kAlphaBit pass the any-hit alpha test first (randomUInt: the frame's).
template <bool ANY, bool COUNT, class Stack>
RTX_DI int traverse(const DevScene &S, const TravRay &r, float tmin, float &tmax, float &bu, float &bv,
                    Stack &st, VisitCount &vc, uint32_t randomUInt = 0) {
    int best = -1;
    for (uint32_t i = 0; i < S.numSpheres; ++i) {
        float t;
        if (isectSphere(S.spheres[i], r.o, r.d, tmin, tmax, t)) {
            int id = (int)(S.numTris + i);
            if (ANY) return id;
            if (best < 0 || t < tmax || id < best) {
                tmax = t;
                best = id;
                bu = bv = 0.0f;
            }
        }
    }
    st.sp = 0;
    int link = 0;
    for (;;) {
        if (link >= 0) {
            if (COUNT) vc.nodes++;
            link = nodeStep4(S, r, tmin, tmax, link, st);
            if (link == kLinkDone) break;
        } else {
            uint32_t leaf = ~(uint32_t)link;
            uint32_t start = leaf >> 4, cnt = (leaf & 15u) + 1u;
            const float4 *tp = S.tris + 3 * (size_t)start;
            for (uint32_t k = 0; k < cnt; ++k) {
                float4 v0 = tp[3 * k], v1 = tp[3 * k + 1], v2 = tp[3 * k + 2];
                if (COUNT) vc.prims++;
                float t, u, v;
                if (isectTri(r, v0, v1, v2, tmin, tmax, t, u, v)) {
                    int id = __float_as_int(v0.w);
                    if (id & kAlphaBit) {
                        id &= ~kAlphaBit;
                        if (!alphaAccept(S, id, r.o.x, t, u, v, randomUInt)) continue;
                    }
                    if (ANY) return id;
                    if (best < 0 || t < tmax || id < best) {
                        tmax = t;
                        best = id;
                        bu = u;
                        bv = v;
                    }
                }
            }
            if (st.sp == 0) break;
            link = st.pop();
        }
    }
    return best;
}

// The part of a scene that traversal reads, per lane: launches that trace
// rays of several scenes (jobs of the serving pool) point each lane at its
// ray's scene when the ray is fetched.
struct LaneScene {
    const float4 *nodes;
    const float4 *tris;
    const rtx_sphere *spheres;
    uint32_t numTris, numSpheres;
    const DevScene *full; // for the any-hit alpha test
};
RTX_DI void pointAt(LaneScene &ls, const DevScene &S, const DevScene *full) {
    ls.nodes = S.nodes;
    ls.tris = S.tris;
    ls.spheres = S.spheres;
    ls.numTris = S.numTris;
    ls.numSpheres = S.numSpheres;
    ls.full = full;
}
RTX_DI bool alphaAccept(const LaneScene &S, int prim, float ox, float t, float bu, float bv, uint32_t randomUInt) {
    return alphaAccept(*S.full, prim, ox, t, bu, bv, randomUInt);
}

// ------------------------------------------------------------------
// Persistent traversal with lane refill (Aila & Laine 2009, sections 4-5):
// each wave loops until the queue is drained; lanes whose ray finished
// fetch a new ray (one atomic per wave), and the "while-while" loop with
// postponed leaves keeps lanes in step: a lane that reaches a leaf parks
// it and keeps descending until every lane of the wave has a leaf, then
// the wave intersects leaves together.
struct LaneRay {
    TravRay r;
    float tmin, tmax, bu, bv;
    int best;
    int link, leaf;
    uint32_t idx;
    uint32_t randomUInt; // the frame's, for the any-hit alpha test
    bool any;            // kHitLane: this lane's ray wants any hit
};

// How a persistent launch ends its rays: at the closest hit, at the first
// hit, or per lane (LaneRay::any), so that one launch traces occlusion and
// closest-hit rays side by side.
constexpr int kHitClosest = 0, kHitAny = 1, kHitLane = 2;

// One inner-node visit: returns the next link (pushing the farther entered
// children), or the next stack entry / kLinkDone when none is entered.
template <bool COUNT, class SceneT, class Stack>
RTX_DI int innerStep(const SceneT &S, LaneRay &L, Stack &st, VisitCount &vc) {
    if (COUNT) vc.nodes++;
    return nodeStep4(S, L.r, L.tmin, L.tmax, L.link, st);
}

// Intersects the primitives of leaf link `leaf`; returns true when an any-hit
// ray is done.
template <int HIT, bool COUNT, bool ALPHA, class SceneT>
RTX_DI bool leafStep(const SceneT &S, LaneRay &L, int leaf, VisitCount &vc) {
    const bool ANY = HIT == kHitAny || (HIT == kHitLane && L.any);
    uint32_t data = ~(uint32_t)leaf;
    uint32_t start = data >> 4, cnt = (data & 15u) + 1u;
    const float4 *tp = S.tris + 3 * (size_t)start;
    for (uint32_t k = 0; k < cnt; ++k) {
        float4 v0 = tp[3 * k], v1 = tp[3 * k + 1], v2 = tp[3 * k + 2];
        if (COUNT) vc.prims++;
        float t, u, v;
        if (isectTri(L.r, v0, v1, v2, L.tmin, L.tmax, t, u, v)) {
            int id = __float_as_int(v0.w);
            if (ALPHA && (id & kAlphaBit)) {
                id &= ~kAlphaBit;
                if (!alphaAccept(S, id, L.r.o.x, t, u, v, L.randomUInt)) continue;
            }
            if (ANY) {
                L.best = id;
                return true;
            }
            if (L.best < 0 || t < L.tmax || id < L.best) {
                L.tmax = t;
                L.best = id;
                L.bu = u;
                L.bv = v;
            }
        }
    }
    return false;
}

// Starts a ray: spheres are tested up front (all workloads have few).
template <int HIT, class SceneT> RTX_DI void beginRay(const SceneT &S, LaneRay &L) {
    const bool ANY = HIT == kHitAny || (HIT == kHitLane && L.any);
    L.best = -1;
    L.bu = L.bv = 0.0f;
    L.link = 0;
    L.leaf = kLinkDone;
    for (uint32_t i = 0; i < S.numSpheres; ++i) {
        float t;
        if (isectSphere(S.spheres[i], L.r.o, L.r.d, L.tmin, L.tmax, t)) {
            int id = (int)(S.numTris + i);
            if (L.best < 0 || t < L.tmax || id < L.best) {
                L.tmax = t;
                L.best = id;
            }
            if (ANY) {
                L.link = kLinkDone;
                return;
            }
        }
    }
}

// Drives the persistent loop.  load(idx, L) fills L.r/L.tmin/L.tmax (and
// L.any under kHitLane); S is the scene (a DevScene for launches over one
// scene, or a per-lane LaneScene that load() points at the ray's scene);
// store(L) consumes a finished ray.  `fetch` must be zero at launch.  Each
// wave first takes its own 64 rays of the grid's static span (thread i of
// the grid: ray i) without an atomic, so a launch over a short or empty queue
// costs no contention; after that a wave refills when at least kRefillLanes
// of its lanes are idle (or all are), with one atomic for the whole wave on
// the rays past the static span.  Once the queue is drained the wave stops
// fetching, so `fetch` ends below n + 64 * waves.
constexpr int kRefillLanes = 16;

template <int HIT, bool COUNT, bool ALPHA, class SceneT, class Stack, class Load, class Store>
RTX_DI void persistentTrace(const SceneT &S, uint32_t n, uint32_t *fetch, Stack &st, VisitCount &vc, Load load,
                            Store store) {
    LaneRay L;
    L.link = L.leaf = kLinkDone;
    L.idx = 0;
    L.any = HIT == kHitAny;
    bool active = false;
    bool drained = false; // wave-uniform
    bool first = true;    // wave-uniform
    const uint32_t staticSpan = gridDim.x * blockDim.x; // blockDim.x is a multiple of 64
    const uint32_t waveBase = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (;;) {
        // refill idle lanes (the whole wave is converged here)
        const bool idle = !active;
        const uint64_t m = __ballot(idle);
        const int nIdle = __popcll(m);
        const bool anyActive = nIdle < 64; // gfx950 waves are 64 wide; blocks are whole waves
        if (!drained && (nIdle >= kRefillLanes || !anyActive)) {
            const int lane = (int)__lane_id();
            uint32_t base = 0;
            if (first) { // every lane idle: the wave's static slice
                first = false;
                base = waveBase;
                if (staticSpan >= n) drained = true;
            } else {
                const int leader = __ffsll((unsigned long long)m) - 1;
                if (lane == leader) base = atomicAdd(fetch, (uint32_t)nIdle);
                base = staticSpan + __shfl(base, leader);
                if (base + (uint32_t)nIdle >= n) drained = true;
            }
            if (idle) {
                const uint64_t below = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
                const uint32_t idx = base + (uint32_t)__popcll(below);
                if (idx < n) {
                    L.idx = idx;
                    load(idx, L);
                    st.sp = 0;
                    beginRay<HIT>(S, L);
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            // inner nodes until every lane of the wave holds a leaf (or is done)
            while (L.link >= 0) {
                L.link = innerStep<COUNT>(S, L, st, vc);
                if (L.link < 0 && L.link != kLinkDone && L.leaf == kLinkDone) {
                    L.leaf = L.link;
                    L.link = popOrDone(st);
                }
                if (__ballot(L.leaf == kLinkDone) == 0) break;
            }
            // leaves
            while (L.leaf != kLinkDone) {
                if (leafStep<HIT, COUNT, ALPHA>(S, L, L.leaf, vc)) {
                    L.link = kLinkDone;
                    L.leaf = kLinkDone;
                    break;
                }
                L.leaf = kLinkDone;
                if (L.link < 0 && L.link != kLinkDone) {
                    L.leaf = L.link;
                    L.link = popOrDone(st);
                }
            }
            if (L.link == kLinkDone && L.leaf == kLinkDone) {
                store(L);
                active = false;
            }
        }
    }
}

} // namespace rtx
