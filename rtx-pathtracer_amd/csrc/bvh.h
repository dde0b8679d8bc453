// Bounding-volume hierarchy layout and host builder (pure C++).
#pragma once

#include <cstdint>
#include <vector>

namespace rtx {

// ----------------------------------------------------------------- BVH
// The builder makes a binary SAH tree (64-byte two-box nodes) and collapses
// it into four-wide nodes, the layout the kernels read (128 B, eight 16-byte
// slots, the first seven read):
//   f[0] = lo.x of children 0-3   f[1] = hi.x
//   f[2] = lo.y                   f[3] = hi.y
//   f[4] = lo.z                   f[5] = hi.z
//   link[0..3] (slot 6), pad (slot 7)
// link >= 0: internal node index; link < 0: leaf, ~link = start << 4 | (count - 1).
// An unused child slot has a point box at 1e30 that no ray reaches.
struct BvhNode {
    float a[4], b[4], c[4];
    int32_t d[4];
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must be 64 bytes");

struct Bvh4Node {
    float f[6][4];
    int32_t link[4];
    int32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "Bvh4Node must be 128 bytes");
constexpr uint32_t kBvh4NodeBytesRead = 112; // the seven slots the traversal loads

constexpr int kLeafMaxPrims = 16;
constexpr int kMaxBvhDepth = 48;   // depth limit of the binary build
constexpr int kStackCapacity = 64; // traversal stack entries (20 in LDS, the rest in registers/scratch)

inline int32_t makeLeaf(uint32_t start, uint32_t count) {
    return ~(int32_t)((start << 4) | (count - 1));
}

struct BvhBuildResult {
    std::vector<Bvh4Node> nodes;
    std::vector<uint32_t> primOrder; // triangle ids in leaf order
    uint32_t maxDepth = 0;   // of the four-wide tree (levels of nodes)
    uint32_t binaryDepth = 0;
    double buildMs = 0;
};

// Binned-SAH build over triangles given as 9 floats each (world space).
void buildBvh(const float *triVerts, uint32_t numTris, BvhBuildResult &out);

} // namespace rtx
