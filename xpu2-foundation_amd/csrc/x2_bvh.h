// x2_bvh.h — the host BVH8 builder's interface (x2_bvh.cpp).
#pragma once
#include <string>
#include <vector>

#include "x2_layout.h"

namespace x2 {
struct Bvh8Result {
    std::vector<Node8> nodes;
    std::vector<Node8Q> qnodes;  // the same nodes in one line each (the device's)
    const void* deviceNodes() const { return qnodes.data(); }
    size_t deviceNodeBytes() const { return nodes.size() * sizeof(DevNode); }
    std::vector<TriRecord> tris;
    std::vector<float> triUV;  // 6 floats per triangle record when the scene has uvs
    std::vector<uint8_t> triCull;  // 1 per triangle record (leaf order) of a single-sided mesh; empty when none is
    uint32_t maxStack = 0, depth = 0;  // traversal stack entries a ray can need; wide-tree depth
};
bool buildBvh8(const x2_scene& s, Bvh8Result& out, std::string& err);
// The records' vertices packed 9 floats a triangle (the leaf tests' array).
inline std::vector<float> triVertices(const std::vector<TriRecord>& tris) {
    std::vector<float> v;
    v.reserve(tris.size() * 9);
    for (const TriRecord& t : tris) v.insert(v.end(), {t.ax, t.ay, t.az, t.bx, t.by, t.bz, t.cx, t.cy, t.cz});
    return v;
}
}  // namespace x2
