// x2_bvh.cpp — host build of the device's acceleration structure: a binary BVH by binned SAH (object splits,
// 32 bins, at most 3 triangles a leaf unless the SAH prefers up to 8), collapsed to 8-wide nodes by repeatedly
// opening the largest-area inner child (the usual BVH-N widening), written in the 256-B SoA node layout with
// triangles reordered into 48-B records in leaf order. It stands where the reference builds OptiX GAS/IAS
// structures (async_gpu.h:802-1114) and Embree scenes (async_cpu.h:196-560); the traversal result does not depend
// on it (the closest-hit key is a function of the scene), only the work per ray does.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "x2_bvh.h"

namespace x2 {
namespace {

struct Aabb {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Aabb& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        if (lo[0] > hi[0]) return 0.0f;
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BinNode {
    Aabb box;
    uint32_t left = 0, right = 0, first = 0, count = 0;  // count > 0: leaf
};

#ifndef X2_MIN_LEAF
#define X2_MIN_LEAF 3  // with one-line nodes 1.01x over 4 on the 1080p headline (2: 1.003x); the traversal cost 0.5-2: none
#endif
#ifndef X2_TRAVERSAL_COST
#define X2_TRAVERSAL_COST 1.0f
#endif
#ifndef X2_BINS
#define X2_BINS 128
#endif
// SAH bins a split tries. On 2.6 M triangles 128 renders 1.005x faster than 32 (16: 0.994x; 256 and 1024 no faster;
// 128 only at nodes of 4,096 triangles or more 1.002x) for 0.8 s more build.
constexpr int Bins = X2_BINS;
constexpr uint32_t MinLeaf = X2_MIN_LEAF;  // a leaf at or below this many triangles; the SAH decides up to 8
constexpr float TraversalCost = X2_TRAVERSAL_COST, IntersectCost = 1.0f;

struct Builder {
    const std::vector<Aabb>& boxes;
    const std::vector<float>& centroid;
    std::vector<uint32_t>& order;
    std::vector<BinNode> nodes;

    Builder(const std::vector<Aabb>& b, const std::vector<float>& c, std::vector<uint32_t>& o)
        : boxes(b), centroid(c), order(o) {}

    void build() {
        nodes.reserve(order.size() / 2 + 2);
        nodes.emplace_back();
        struct Task {
            uint32_t node, first, count;
        };
        std::vector<Task> tasks{{0, 0, uint32_t(order.size())}};
        std::vector<Aabb> bin(Bins);
        std::vector<uint32_t> cnt(Bins), rightCount(Bins);
        std::vector<float> rightArea(Bins);
        while (!tasks.empty()) {
            const Task t = tasks.back();
            tasks.pop_back();
            Aabb bounds, cb;
            for (uint32_t i = t.first; i < t.first + t.count; ++i) {
                bounds.grow(boxes[order[i]]);
                cb.grow(&centroid[size_t(order[i]) * 3]);
            }
            nodes[t.node].box = bounds;
            if (t.count <= MinLeaf) {
                nodes[t.node].first = t.first;
                nodes[t.node].count = t.count;
                continue;
            }
            int axis = -1, split = -1;
            float bestCost = FLT_MAX;
            constexpr int nb = Bins;
            for (int k = 0; k < 3; ++k) {
                const float extent = cb.hi[k] - cb.lo[k];
                if (!(extent > 0.0f)) continue;
                for (int q = 0; q < nb; ++q) {
                    bin[q] = Aabb{};
                    cnt[q] = 0;
                }
                const float scale = float(nb) * (1.0f - 1e-6f) / extent;
                for (uint32_t i = t.first; i < t.first + t.count; ++i) {
                    const uint32_t p = order[i];
                    const int b = std::min(nb - 1, int((centroid[size_t(p) * 3 + k] - cb.lo[k]) * scale));
                    ++cnt[b];
                    bin[b].grow(boxes[p]);
                }
                Aabb acc;
                uint32_t n = 0;
                for (int b = nb - 1; b > 0; --b) {
                    acc.grow(bin[b]);
                    n += cnt[b];
                    rightArea[b] = acc.area();
                    rightCount[b] = n;
                }
                acc = Aabb{};
                n = 0;
                for (int b = 0; b < nb - 1; ++b) {
                    acc.grow(bin[b]);
                    n += cnt[b];
                    if (!n || !rightCount[b + 1]) continue;
                    const float cost = acc.area() * float(n) + rightArea[b + 1] * float(rightCount[b + 1]);
                    if (cost < bestCost) {
                        bestCost = cost;
                        axis = k;
                        split = b + 1;
                    }
                }
            }
            const float leafCost = IntersectCost * float(t.count) * bounds.area();
            const float splitCost = TraversalCost * bounds.area() + IntersectCost * bestCost;
            if (t.count <= MaxLeafTriangles && (axis < 0 || leafCost <= splitCost)) {
                nodes[t.node].first = t.first;
                nodes[t.node].count = t.count;
                continue;
            }
            uint32_t mid;
            if (axis < 0) {
                mid = t.first + t.count / 2;  // coincident centroids: halve the list
            } else {
                const float extent = cb.hi[axis] - cb.lo[axis];
                const float scale = float(nb) * (1.0f - 1e-6f) / extent;
                const float lo = cb.lo[axis];
                auto* begin = order.data() + t.first;
                auto* it = std::partition(begin, begin + t.count, [&](uint32_t p) {
                    return std::min(nb - 1, int((centroid[size_t(p) * 3 + axis] - lo) * scale)) < split;
                });
                mid = t.first + uint32_t(it - begin);
                if (mid == t.first || mid == t.first + t.count) mid = t.first + t.count / 2;
            }
            const uint32_t l = uint32_t(nodes.size());
            nodes.emplace_back();
            nodes.emplace_back();
            nodes[t.node].left = l;
            nodes[t.node].right = l + 1;
            nodes[t.node].count = 0;
            tasks.push_back({l, t.first, mid - t.first});
            tasks.push_back({l + 1, mid, t.first + t.count - mid});
        }
    }
};

}  // namespace

// Builds the BVH8 and the leaf-ordered triangle records. Returns false with err set on bad input.
// A node's children in one line: per axis the origin is the node's lowest child bound and the step the least power of
// two whose 255 steps reach its highest; each bound is the nearest step outward, checked (and moved outward until it
// holds) in the device's own float arithmetic, origin + q * step, with each operation rounded on its own.
static float decodeStep(float origin, uint32_t q, float step) {
    const float m = float(q) * step;  // exact: q < 256 and step a power of two
    return origin + m;
}
static void quantizeNodes(const std::vector<Node8>& in, std::vector<Node8Q>& out) {
    out.assign(in.size(), Node8Q{});
    for (size_t n = 0; n < in.size(); ++n) {
        const Node8& w = in[n];
        Node8Q& o = out[n];
        const float* lo[3] = {w.lox, w.loy, w.loz};
        const float* hi[3] = {w.hix, w.hiy, w.hiz};
        float origin[3];
        uint8_t ebits[3];
        float step[3];
        for (int a = 0; a < 3; ++a) {
            float l = INFINITY, h = -INFINITY;
            for (int i = 0; i < 8; ++i)
                if (w.child[i] != EmptyChild) {
                    l = std::min(l, lo[a][i]);
                    h = std::max(h, hi[a][i]);
                }
            if (!(l <= h)) l = h = 0.0f;  // no child: any bounds
            origin[a] = l;
            int e = -126;  // the least normal step
            const double extent = double(h) - double(l);
            if (extent > 0) e = std::max(-126, int(std::ceil(std::log2(extent / 255.0))));
            while (e < 127 && decodeStep(l, 255u, std::ldexp(1.0f, e)) < h) ++e;
            step[a] = std::ldexp(1.0f, e);
            ebits[a] = uint8_t(e + 127);
        }
        o.ox = origin[0], o.oy = origin[1], o.oz = origin[2];
        o.ex = ebits[0], o.ey = ebits[1], o.ez = ebits[2];
        o.pad0 = 0;
        for (int i = 0; i < 8; ++i) {
            o.child[i] = w.child[i];
            o.pad[i] = 0;
            for (int a = 0; a < 3; ++a) {
                uint32_t ql = 0, qh = 255;
                if (w.child[i] != EmptyChild) {
                    const double fl = std::floor((double(lo[a][i]) - double(origin[a])) / double(step[a]));
                    const double fh = std::ceil((double(hi[a][i]) - double(origin[a])) / double(step[a]));
                    ql = uint32_t(std::min(255.0, std::max(0.0, fl)));
                    qh = uint32_t(std::min(255.0, std::max(0.0, fh)));
                    while (ql > 0 && decodeStep(origin[a], ql, step[a]) > lo[a][i]) --ql;
                    while (qh < 255 && decodeStep(origin[a], qh, step[a]) < hi[a][i]) ++qh;
                }
                o.q[2 * a][i] = uint8_t(ql);
                o.q[2 * a + 1][i] = uint8_t(qh);
            }
        }
    }
}

bool buildBvh8(const x2_scene& s, Bvh8Result& out, std::string& err) {
    const uint32_t n = s.triangle_count;
    if (n >= (1u << 28)) {
        err = "too many triangles for the 28-bit leaf reference";
        return false;
    }
    std::vector<uint32_t> geomOf(n), materialOf(n), primOf(n);
    bool anySingle = false;
    if (s.geom_flags)
        for (uint32_t g = 0; g < s.geom_count; ++g) anySingle = anySingle || (s.geom_flags[g] & X2_GEOM_SINGLE_SIDED);
    for (uint32_t g = 0; g < s.geom_count; ++g) {
        const uint32_t a = s.geom_first[g], b = s.geom_first[g + 1];
        if (a > b || b > n) {
            err = "geom_first is not a non-decreasing prefix within triangle_count";
            return false;
        }
        if (s.geom_material[g] >= s.material_count) {
            err = "geom_material out of range";
            return false;
        }
        for (uint32_t t = a; t < b; ++t) {
            geomOf[t] = g;
            primOf[t] = t - a;
            materialOf[t] = s.geom_material[g];
        }
    }
    if (s.geom_count == 0 || s.geom_first[0] != 0 || s.geom_first[s.geom_count] != n) {
        err = "meshes must cover every triangle";
        return false;
    }
    std::vector<Aabb> boxes(n);
    std::vector<float> centroid(size_t(n) * 3);
    for (uint32_t t = 0; t < n; ++t) {
        for (int v = 0; v < 3; ++v) {
            const uint32_t idx = s.indices[size_t(t) * 3 + v];
            if (idx >= s.vertex_count) {
                err = "triangle index out of range";
                return false;
            }
            boxes[t].grow(s.positions + size_t(idx) * 3);
        }
        for (int k = 0; k < 3; ++k) centroid[size_t(t) * 3 + k] = 0.5f * (boxes[t].lo[k] + boxes[t].hi[k]);
    }
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    Builder b(boxes, centroid, order);
    if (n) b.build();

    // Collapse to BVH8. Each wide node opens its children greedily by area until it has 8.
    out.nodes.clear();
    out.tris.clear();
    out.triUV.clear();
    out.triCull.clear();
    out.tris.reserve(n);
    out.maxStack = 0;
    out.depth = 0;
    auto emptyNode = []() {
        Node8 w;
        for (int i = 0; i < 8; ++i) {
            w.lox[i] = w.loy[i] = w.loz[i] = INFINITY;
            w.hix[i] = w.hiy[i] = w.hiz[i] = -INFINITY;
            w.child[i] = EmptyChild;
            w.pad[i] = 0;
        }
        return w;
    };
    auto emitLeaf = [&](const BinNode& leaf) -> uint32_t {
        const uint32_t first = uint32_t(out.tris.size());
        for (uint32_t i = leaf.first; i < leaf.first + leaf.count; ++i) {
            const uint32_t t = order[i];
            TriRecord r;
            const float* pa = s.positions + size_t(s.indices[size_t(t) * 3 + 0]) * 3;
            const float* pb = s.positions + size_t(s.indices[size_t(t) * 3 + 1]) * 3;
            const float* pc = s.positions + size_t(s.indices[size_t(t) * 3 + 2]) * 3;
            r.ax = pa[0], r.ay = pa[1], r.az = pa[2];
            r.bx = pb[0], r.by = pb[1], r.bz = pb[2];
            r.cx = pc[0], r.cy = pc[1], r.cz = pc[2];
            r.geom = geomOf[t];
            r.prim = primOf[t];
            r.material = materialOf[t];
            if (anySingle) out.triCull.push_back((s.geom_flags[geomOf[t]] & X2_GEOM_SINGLE_SIDED) ? 1u : 0u);
            out.tris.push_back(r);
            if (s.uvs)
                for (int v = 0; v < 3; ++v) {
                    const float* uv = s.uvs + size_t(s.indices[size_t(t) * 3 + v]) * 2;
                    out.triUV.push_back(uv[0]);
                    out.triUV.push_back(uv[1]);
                }
        }
        return LeafBit | (first << 3) | (leaf.count - 1);
    };
    struct Work {
        uint32_t bin, wide, depth, stackAbove;
    };
    out.nodes.push_back(emptyNode());
    std::vector<Work> work;
    if (n == 0) return true;  // the root stays empty: every ray misses
    work.push_back({0, 0, 1, 0});
    while (!work.empty()) {
        const Work w = work.back();
        work.pop_back();
        std::vector<uint32_t> kids;
        if (b.nodes[w.bin].count) {
            kids.push_back(w.bin);  // a leaf root: one leaf child
        } else {
            kids = {b.nodes[w.bin].left, b.nodes[w.bin].right};
            while (kids.size() < 8) {
                int pick = -1;
                float best = -1.0f;
                for (size_t i = 0; i < kids.size(); ++i) {
                    const BinNode& k = b.nodes[kids[i]];
                    if (k.count == 0 && k.box.area() > best) {
                        best = k.box.area();
                        pick = int(i);
                    }
                }
                if (pick < 0) break;
                const BinNode k = b.nodes[kids[pick]];
                kids[pick] = k.left;
                kids.push_back(k.right);
            }
        }
        out.depth = std::max(out.depth, w.depth);
        const uint32_t stackHere = w.stackAbove + uint32_t(kids.size()) - 1;
        out.maxStack = std::max(out.maxStack, stackHere + 1);
        for (size_t i = 0; i < kids.size(); ++i) {
            const BinNode& k = b.nodes[kids[i]];
            Node8& node = out.nodes[w.wide];
            node.lox[i] = k.box.lo[0], node.loy[i] = k.box.lo[1], node.loz[i] = k.box.lo[2];
            node.hix[i] = k.box.hi[0], node.hiy[i] = k.box.hi[1], node.hiz[i] = k.box.hi[2];
            if (k.count) {
                node.child[i] = emitLeaf(k);
            } else {
                const uint32_t child = uint32_t(out.nodes.size());
                out.nodes[w.wide].child[i] = child;
                out.nodes.push_back(emptyNode());
                work.push_back({kids[i], child, w.depth + 1, stackHere});
            }
        }
    }
    quantizeNodes(out.nodes, out.qnodes);
    return true;
}

}  // namespace x2
