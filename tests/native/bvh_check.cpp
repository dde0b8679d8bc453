// Structural check of the host BVH builder (rtx::buildBvh, csrc/bvh_build.cpp)
// over awkward triangle sets: empty, tiny, duplicated, zero-area, NaN / inf /
// huge coordinates and scales spread over 2^120.  For each tree: every
// triangle id appears once in the leaf order, every leaf range lies inside it
// and is reached once from the root, and every finite vertex of a leaf's
// triangles lies inside that leaf's box in its parent.  Built and run by
// tests/test_bvh_build.py, with AddressSanitizer and UBSan where available.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "bvh.h"
using namespace rtx;
static int fails = 0;
static void check(const std::vector<float> &v, const char *name) {
    uint32_t n = (uint32_t)(v.size() / 9);
    BvhBuildResult r;
    try { buildBvh(v.data(), n, r); } catch (std::exception &e) { printf("%s: threw %s\n", name, e.what()); return; }
    std::vector<int> seen(n, 0);
    if (r.primOrder.size() != n) { printf("%s: primOrder %zu != %u\n", name, r.primOrder.size(), n); ++fails; }
    for (auto p : r.primOrder) if (p < n) seen[p]++; else { printf("%s: prim id %u\n", name, p); ++fails; break; }
    for (uint32_t i = 0; i < n; ++i) if (seen[i] != 1) { printf("%s: tri %u seen %d\n", name, i, seen[i]); ++fails; break; }
    // walk: every leaf range in bounds, every triangle inside its leaf's box (finite coords only)
    std::vector<int> cover(r.primOrder.size(), 0);
    std::vector<int> st{0};
    size_t visited = 0;
    while (!st.empty() && n) {
        int ni = st.back(); st.pop_back();
        if (ni < 0 || (size_t)ni >= r.nodes.size()) { printf("%s: node %d\n", name, ni); ++fails; return; }
        ++visited;
        const Bvh4Node &nd = r.nodes[ni];
        for (int c = 0; c < 4; ++c) {
            int32_t l = nd.link[c];
            if (nd.f[0][c] == 1e30f) continue;
            if (l >= 0) { st.push_back(l); continue; }
            uint32_t code = ~(uint32_t)l, start = code >> 4, cnt = (code & 15) + 1;
            for (uint32_t k = start; k < start + cnt; ++k) {
                if (k >= cover.size()) { printf("%s: leaf past primOrder\n", name); ++fails; return; }
                cover[k]++;
                const float *t = &v[9 * (size_t)r.primOrder[k]];
                for (int q = 0; q < 3; ++q) for (int a = 0; a < 3; ++a) {
                    float x = t[3 * q + a];
                    if (std::isfinite(x) && (x < nd.f[2 * a][c] || x > nd.f[2 * a + 1][c])) { printf("%s: tri outside box\n", name); ++fails; return; }
                }
            }
        }
    }
    for (size_t k = 0; k < cover.size(); ++k) if (cover[k] != 1) { printf("%s: slot %zu covered %d\n", name, k, cover[k]); ++fails; break; }
    printf("%s: %u tris, %zu nodes (%zu reached), depth %u/%u\n", name, n, r.nodes.size(), visited, r.maxDepth, r.binaryDepth);
}
int main() {
    std::mt19937 g(1);
    std::uniform_real_distribution<float> u(-1, 1);
    auto rnd = [&](uint32_t n, float s) { std::vector<float> v(9 * (size_t)n); for (auto &x : v) x = u(g) * s; return v; };
    check({}, "empty");
    check(rnd(1, 1), "one");
    check(rnd(2, 1), "two");
    check(rnd(17, 1), "seventeen");
    check(rnd(100000, 10), "random100k");
    { auto v = rnd(1, 1); std::vector<float> w; for (int i = 0; i < 5000; ++i) w.insert(w.end(), v.begin(), v.end()); check(w, "identical5000"); }
    { std::vector<float> w(9 * 3000, 0.5f); check(w, "points3000"); }
    { std::vector<float> w; for (int i = 0; i < 4000; ++i) { float x = i * 1e-3f; float t[9] = {x, 0, 0, x, 1, 0, x, 0, 1}; w.insert(w.end(), t, t + 9); } check(w, "planes4000"); }
    { auto v = rnd(2000, 1); for (size_t i = 0; i < v.size(); i += 97) v[i] = NAN; check(v, "nan"); }
    { auto v = rnd(2000, 1); for (size_t i = 0; i < v.size(); i += 89) v[i] = INFINITY; check(v, "inf"); }
    { auto v = rnd(3000, 1); for (size_t i = 0; i < v.size(); i += 3) v[i] *= 1e30f; check(v, "huge"); }
    { std::vector<float> w; for (int i = 0; i < 60000; ++i) { float x = std::ldexp(1.0f, -(i % 120)); float t[9] = {x, x, x, x * 2, x, x, x, x * 2, x}; w.insert(w.end(), t, t + 9); } check(w, "geometric60000"); }
    printf("fails %d\n", fails);
    return fails != 0;
}
