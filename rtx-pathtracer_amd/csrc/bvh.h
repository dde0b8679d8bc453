// Bounding-volume hierarchy layout and host builder (pure C++).
#pragma once

#include <cstdint>
#include <vector>

namespace rtx {

// ----------------------------------------------------------------- BVH
// Binary BVH node holding both children's boxes (64 B, four 16-byte loads):
//   a = (c0.lo.x, c0.lo.y, c0.lo.z, c0.hi.x)
//   b = (c0.hi.y, c0.hi.z, c1.lo.x, c1.lo.y)
//   c = (c1.lo.z, c1.hi.x, c1.hi.y, c1.hi.z)
//   d = (link0, link1, -, -)
// link >= 0: internal node index; link < 0: leaf, ~link = start << 4 | (count - 1).
struct BvhNode {
    float a[4], b[4], c[4];
    int32_t d[4];
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must be 64 bytes");

constexpr int kLeafMaxPrims = 16;
constexpr int kMaxBvhDepth = 48; // traversal stack capacity (24 in LDS + 24 in registers/scratch)

inline int32_t makeLeaf(uint32_t start, uint32_t count) {
    return ~(int32_t)((start << 4) | (count - 1));
}

struct BvhBuildResult {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> primOrder; // triangle ids in leaf order
    uint32_t maxDepth = 0;
    double buildMs = 0;
};

// Binned-SAH build over triangles given as 9 floats each (world space).
void buildBvh(const float *triVerts, uint32_t numTris, BvhBuildResult &out);

} // namespace rtx
