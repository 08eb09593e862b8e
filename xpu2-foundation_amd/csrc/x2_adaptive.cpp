// x2_adaptive.cpp — the host side of an adaptive frame (x2_adaptive.h): MoonRay's pass schedule, pixel error and
// region error trees (after Dammertz et al. 2010, "A hierarchical automatic stopping condition for Monte Carlo
// global illumination"), as the reference restates them in async_adaptive.h / async_adaptive_core.h. Built without
// floating-point contraction: each product and sum rounds on its own, as the reference's mulRn/addRn do.
#include "x2_adaptive.h"

#include <math.h>

#include <algorithm>

namespace x2::adaptive {
namespace {

constexpr float MinNodeSize = 1.27201964951407f;  // sqrt of the golden ratio (async_adaptive_core.h:29)
constexpr float MaxNodeSize = 16.0f;
constexpr float SplitThreshold = 256.0f;
constexpr float NoSplit = -1.0f;
// MoonRay judges a node's error against a fixed image area, whatever the image (async_adaptive_core.h:53-56).
constexpr float ReferenceArea = (815.0f * 1967.0f) / 4.0f;

IBox outward(const FBox& b) { return {int(floorf(b.x0)), int(floorf(b.y0)), int(ceilf(b.x1)), int(ceilf(b.y1))}; }

float nodeError(float summed, float area) { return summed / sqrtf(area * ReferenceArea); }

// The offset along a node (a fraction of its length) for a split at `at`, kept MinNodeSize from either end.
float keepFromEdge(float at, float length) {
    if (length * at >= MinNodeSize) return at;
    if (length >= MinNodeSize * 2) return MinNodeSize / length;
    return NoSplit;
}
float splitOffset(float at, float length) {
    if (at <= 0.5f) return keepFromEdge(at, length);
    const float mirrored = keepFromEdge(1.0f - at, length);
    return mirrored == NoSplit ? NoSplit : 1.0f - mirrored;
}

// A region's summed-area table: sum(x0..x, y0..y) of errors at [(y - y0) * w1 + (x - x0)], exclusive corners.
struct Table {
    const double* s;
    int x0, y0, w1;
    double at(int x, int y) const { return s[size_t(y - y0) * size_t(w1) + size_t(x - x0)]; }
    // The summed error of [ax, bx) x [ay, by).
    double box(int ax, int ay, int bx, int by) const { return at(bx, by) - at(ax, by) - at(bx, ay) + at(ax, ay); }
};

// The error a node has accumulated in its first k + 1 columns (rows when it splits across y), as a float.
struct Prefix {
    const Table& t;
    IBox b;
    bool acrossX;
    int n;
    float operator()(int k) const {
        return acrossX ? float(t.box(b.x0, b.y0, b.x0 + k + 1, b.y1)) : float(t.box(b.x0, b.y0, b.x1, b.y0 + k + 1));
    }
};

// Where along the node half its error lies: the first column whose prefix passes half the total (bisection),
// interpolated within that column.
float halfErrorAt(const Prefix& p) {
    const float total = p.n > 0 ? p(p.n - 1) : 0.0f, half = total * 0.5f;
    int k = p.n / 2;
    if (total != 0.0f) {
        int lo = 0, hi = p.n;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (p(mid) > half) hi = mid;
            else lo = mid + 1;
        }
        k = lo;
    }
    k = std::min(std::max(k, 1), p.n - 1);
    const float below = p(k - 1), above = p(k), step = above - below;
    const float within = step > 0 ? (half - below) / step : 0.5f;
    return (float(k) + within) / float(p.n);
}

enum Fate { Open, Converged, Split };

// A node's fate (AdaptiveRegionTree::updateImpl): converged, split at the half-error point of its longer side, or
// left open.
Fate judge(const Table& t, const IBox& root, const FBox& b, float target, FBox& first, FBox& second) {
    const float w = b.x1 - b.x0, h = b.y1 - b.y0;
    const bool acrossX = w > h;
    const float length = acrossX ? w : h;
    IBox ib = outward(b);
    ib = {std::max(ib.x0, root.x0), std::max(ib.y0, root.y0), std::min(ib.x1, root.x1), std::min(ib.y1, root.y1)};
    const Prefix p{t, ib, acrossX, acrossX ? ib.x1 - ib.x0 : ib.y1 - ib.y0};
    const float error = nodeError(p.n > 0 ? p(p.n - 1) : 0.0f, w * h);
    if (error < target && length < MaxNodeSize) return Converged;
    if (!(error < target * SplitThreshold && length > MinNodeSize * 2)) return Open;
    const float offset = splitOffset(halfErrorAt(p), length);
    if (offset == NoSplit) return Open;
    first = second = b;
    if (acrossX) first.x1 = second.x0 = b.x0 + offset * w;
    else first.y1 = second.y0 = b.y0 + offset * h;
    return Split;
}

}  // namespace

std::vector<Pass> schedule(uint32_t maxSamples) {
    std::vector<Pass> out{{0, 1}};
    uint32_t at = 1, left = maxSamples > 0 ? maxSamples - 1 : 0;
    for (uint32_t size = 4; left > size && out.size() < 999; size += 2) {
        out.push_back({at, at + size});
        at += size;
        left -= size;
    }
    if (left) out.push_back({at, at + left});
    return out;
}

float pixelError(const double total[4], const double odd[4], uint32_t samples) {
    const float n = float(samples), nOdd = float(samples / 2), nEven = n - nOdd;
    if (nOdd <= 0 || nEven <= 0) return -1.0f;
    float gap[4], mean[3];
    for (int c = 0; c < 4; ++c) {
        const float t = float(total[c]), o = float(odd[c]);
        if (t != t || o != o) return 0.0f;
        // The even half's mean less the odd half's: (t - o) / nEven - o / nOdd over a common denominator.
        gap[c] = fabsf((t * nOdd + -(o * n)) / (nOdd * nEven));
        if (c < 3) mean[c] = t / n;
    }
    const float lumaGap = 0.299f * gap[0] + 0.587f * gap[1] + 0.114f * gap[2];
    const float lumaMean = 0.299f * mean[0] + 0.587f * mean[1] + 0.114f * mean[2];
    if (lumaMean <= 0) return gap[3];
    return lumaGap / sqrtf(lumaMean) + gap[3];
}

Tree::Tree(IBox root, float target)
    : root_(root), target_(target), sat_(size_t(root.x1 - root.x0 + 1) * size_t(root.y1 - root.y0 + 1), 0.0) {
    nodes_.push_back({{float(root.x0), float(root.y0), float(root.x1), float(root.y1)}, -1, 0});
}

void Tree::update(const float* errors, int stride) {
    nodes_.resize(1);
    nodes_[0].children = -1;
    nodes_[0].complete = 0;
    const int w = root_.x1 - root_.x0;
    const size_t w1 = size_t(w) + 1;
    for (int y = root_.y0; y < root_.y1; ++y)  // a pixel without both halves yet: the whole region samples
        for (int x = root_.x0; x < root_.x1; ++x)
            if (errors[size_t(y) * stride + x] < 0) return;
    for (int y = root_.y0; y < root_.y1; ++y) {
        const size_t r = size_t(y - root_.y0);
        double run = 0;
        for (int x = 0; x < w; ++x) {
            run += double(errors[size_t(y) * stride + root_.x0 + x]);
            sat_[(r + 1) * w1 + size_t(x) + 1] = sat_[r * w1 + size_t(x) + 1] + run;
        }
    }
    updateNode(0);
}

// Breadth first: every node of a level is judged, its children appended; then each inner node, deepest first, is
// complete when both children are.
void Tree::updateNode(int index) {
    const Table t{sat_.data(), root_.x0, root_.y0, root_.x1 - root_.x0 + 1};
    for (size_t i = size_t(index); i < nodes_.size(); ++i) {
        FBox a, b;
        const Fate f = judge(t, root_, nodes_[i].bounds, target_, a, b);
        if (f == Converged) nodes_[i].complete = 1;
        if (f != Split) continue;
        nodes_[i].children = int(nodes_.size());
        nodes_.push_back({a, -1, 0});
        nodes_.push_back({b, -1, 0});
    }
    for (size_t i = nodes_.size(); i-- > size_t(index);) {
        Node& n = nodes_[i];
        if (n.children >= 0) n.complete = nodes_[size_t(n.children)].complete && nodes_[size_t(n.children) + 1].complete;
    }
}

uint64_t Tree::sampleArea(int tx, int ty) const {
    uint64_t bits = 0;
    for (const Node& n : nodes_) {
        if (n.children >= 0 || n.complete) continue;  // open leaves only
        const FBox c{std::max(n.bounds.x0, float(tx)), std::max(n.bounds.y0, float(ty)), std::min(n.bounds.x1, float(tx + 8)),
                     std::min(n.bounds.y1, float(ty + 8))};
        if (c.x1 <= c.x0 || c.y1 <= c.y0) continue;
        const IBox r = outward(c);
        for (int y = std::max(r.y0, ty); y < std::min(r.y1, ty + 8); ++y)
            for (int x = std::max(r.x0, tx); x < std::min(r.x1, tx + 8); ++x) bits |= 1ull << ((y - ty) * 8 + (x - tx));
    }
    if (!bits) return 0;
    int rows[2] = {8, -1}, cols[2] = {8, -1};  // the active bits' bounding rectangle (ActivePixelMask::fillGaps)
    for (int q = 0; q < 64; ++q)
        if (bits >> q & 1) {
            rows[0] = std::min(rows[0], q / 8), rows[1] = std::max(rows[1], q / 8);
            cols[0] = std::min(cols[0], q % 8), cols[1] = std::max(cols[1], q % 8);
        }
    uint64_t filled = 0;
    for (int y = rows[0]; y <= rows[1]; ++y)
        for (int x = cols[0]; x <= cols[1]; ++x) filled |= 1ull << (y * 8 + x);
    return filled;
}

Regions::Regions(int width, int height, float target) : width_(width), height_(height) {
    constexpr int tile = 8, overlap = 8, most = 4;
    auto upTo8 = [](int n) { return (n + 7) / 8 * 8; };
    if (tile + overlap >= width || tile + overlap >= height) {
        regionW_ = width;
        regionH_ = height;
    } else {
        perDim_ = std::max(1, std::min({most, width / (tile + overlap), height / (tile + overlap)}));
        regionW_ = upTo8(1 + (width - 1) / perDim_);
        regionH_ = upTo8(1 + (height - 1) / perDim_);
    }
    for (int i = 0; i < perDim_ * perDim_; ++i) {
        const int rx = i % perDim_, ry = i / perDim_;
        const IBox b{std::max(rx * regionW_ - overlap, 0), std::max(ry * regionH_ - overlap, 0),
                     std::min((rx + 1) * regionW_ + overlap, width), std::min((ry + 1) * regionH_ + overlap, height)};
        trees_.emplace_back(b, target);
    }
}

void Regions::update(const float* errors) {
    for (Tree& t : trees_) t.update(errors, width_);
}

bool Regions::done() const {
    return std::all_of(trees_.begin(), trees_.end(), [](const Tree& t) { return t.done(); });
}

uint64_t Regions::sampleArea(int tx, int ty) const {
    const int rx = std::min(tx / regionW_, perDim_ - 1), ry = std::min(ty / regionH_, perDim_ - 1);
    return trees_[size_t(ry * perDim_ + rx)].sampleArea(tx, ty);
}

}  // namespace x2::adaptive
