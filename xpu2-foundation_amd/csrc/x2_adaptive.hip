// x2_adaptive.hip — the host side of an adaptive frame (x2_adaptive.h): MoonRay's pass schedule, pixel error and
// region error trees (after Dammertz et al. 2010, "A hierarchical automatic stopping condition for Monte Carlo
// global illumination"), as the reference restates them in async_adaptive.h / async_adaptive_core.h. The node
// arithmetic is x2_adaptive_core.h's, which the device's pass chain shares. Built without floating-point
// contraction: each product and sum rounds on its own, as the reference's mulRn/addRn do.
//
// The schedule, the region layout and the tree (Tree, Regions) and cellular (Cellular) rules are adapted from MoonRay
// rndr/adaptive/AdaptiveRegionTree.{h,cc}, OverlappingRegions.h, ActivePixelMask.h and RenderDriver.cc, by way of the
// reference's restatement (xpu2/interactive/async_adaptive.h): Apache License 2.0, copyright 2023-2024 DreamWorks
// Animation LLC (http://www.apache.org/licenses/LICENSE-2.0). Changes: restated in this build's own structure and
// names; the rules and arithmetic are kept so that the same errors give the same masks.
#include "x2_adaptive.h"

#include <math.h>

#include <algorithm>
#include <thread>

#include "x2_adaptive_core.h"

namespace x2::adaptive {

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
    return pixelErrorOf(total, odd, samples);
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

// Depth first from the root, skipping converged subtrees and those that miss the tile; each open leaf's overlap
// with the tile, rounded outward, is sampled.
uint64_t Tree::sampleArea(int tx, int ty) const {
    const FBox tile{float(tx), float(ty), float(tx + 8), float(ty + 8)};
    std::vector<int> stack{0};
    uint64_t bits = 0;
    while (!stack.empty()) {
        const Node& n = nodes_[size_t(stack.back())];
        stack.pop_back();
        if (n.complete) continue;
        const FBox c{std::max(n.bounds.x0, tile.x0), std::max(n.bounds.y0, tile.y0), std::min(n.bounds.x1, tile.x1),
                     std::min(n.bounds.y1, tile.y1)};
        if (c.x1 <= c.x0 || c.y1 <= c.y0) continue;
        if (n.children >= 0) {
            stack.push_back(n.children);
            stack.push_back(n.children + 1);
            continue;
        }
        const IBox r = outward(c);
        for (int y = std::max(r.y0, ty); y < std::min(r.y1, ty + 8); ++y)
            for (int x = std::max(r.x0, tx); x < std::min(r.x1, tx + 8); ++x) bits |= 1ull << ((y - ty) * 8 + (x - tx));
    }
    return fillRectangle(bits);
}

Regions::Regions(int width, int height, float target) : width_(width) {
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

// The regions are independent: each rebuilds on a thread of its own (the reference's tree pool, async_render.h:205).
void Regions::update(const float* errors) {
    std::vector<std::thread> pool;
    pool.reserve(trees_.size());
    for (Tree& t : trees_) pool.emplace_back([&t, errors, this] { t.update(errors, width_); });
    for (std::thread& th : pool) th.join();
}

bool Regions::done() const {
    return std::all_of(trees_.begin(), trees_.end(), [](const Tree& t) { return t.done(); });
}

uint64_t Regions::sampleArea(int tx, int ty) const {
    const int rx = std::min(tx / regionW_, perDim_ - 1), ry = std::min(ty / regionH_, perDim_ - 1);
    return trees_[size_t(ry * perDim_ + rx)].sampleArea(tx, ty);
}

Cellular::Cellular(int width, int height, float target)
    : width_(width), height_(height), tilesX_((width + 7) / 8), tilesY_((height + 7) / 8), target_(target),
      alive_(size_t(tilesX_) * tilesY_, 1), next_(alive_.size(), 0), ratio_(alive_.size(), 0.0f), error_(alive_.size(), 0.0f),
      pixels_(alive_.size(), 0.0), samples_(alive_.size(), 0.0) {}

void Cellular::deadForever(const std::vector<uint64_t>& inImage) {
    dead_.assign(alive_.size(), 0);
    for (size_t t = 0; t < alive_.size() && t < inImage.size(); ++t)
        if (!inImage[t]) dead_[t] = 1, alive_[t] = 0;
}

size_t Cellular::update(const float* errors, const Budget* budget) {
    for (int ty = 0; ty < tilesY_; ++ty)
        for (int tx = 0; tx < tilesX_; ++tx) {
            const size_t t = size_t(ty) * tilesX_ + tx;
            double sum = 0, most = 0, rendered = 0;
            int n = 0;
            bool unknown = false;
            for (int y = ty * 8; y < std::min(height_, ty * 8 + 8); ++y)
                for (int x = tx * 8; x < std::min(width_, tx * 8 + 8); ++x) {
                    const size_t i = size_t(y) * size_t(width_) + size_t(x);
                    if (errors[i] < 0) unknown = true;
                    else sum += errors[i];
                    ++n;
                    if (budget) {
                        const double s = (*budget->samples)[i];
                        most = std::max(most, s);
                        rendered += s > 0 ? 1 : 0;
                    }
                }
            error_[t] = unknown ? -1.0f : nodeError(float(sum), float(n));
            pixels_[t] = rendered;
            samples_[t] = most;
        }
    const float target = budget ? std::max(budgetTarget(*budget), budget->floor) : target_;
    for (size_t t = 0; t < ratio_.size(); ++t) ratio_[t] = error_[t] < 0 ? 3.4e38f : error_[t] / target;
    size_t live = 0;
    constexpr float Survive = 0.5f, Birth = 0.75f;
    constexpr int SurviveNeighbours = 3, BirthNeighbours = 5;
    for (int ty = 0; ty < tilesY_; ++ty)
        for (int tx = 0; tx < tilesX_; ++tx) {
            const size_t t = size_t(ty) * tilesX_ + tx;
            int around = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x = tx + dx, y = ty + dy;
                    if ((dx || dy) && x >= 0 && y >= 0 && x < tilesX_ && y < tilesY_ && alive_[size_t(y) * tilesX_ + x]) ++around;
                }
            const float r = ratio_[t];
            const bool stays = alive_[t] ? (r >= 1.0f || (r >= Survive && around >= SurviveNeighbours))
                                         : (r >= 1.0f || (r >= Birth && around >= BirthNeighbours));
            next_[t] = stays && (dead_.empty() || !dead_[t]) ? 1 : 0;
            live += next_[t] ? 1 : 0;
        }
    alive_.swap(next_);
    return live;
}

// The target whose cost — each tile's rendered pixels times the samples that bring its error down to it, n (error /
// target)^2 less the n it has, at most the most a pixel may take — is the samples left; bisected in log space (the
// cost falls as the target rises). No samples left, or nothing left to refine: no tile lives on its own error.
float Cellular::budgetTarget(const Budget& b) {
    double lo = 3.4e38, hi = 0;
    for (size_t t = 0; t < error_.size(); ++t)
        if ((dead_.empty() || !dead_[t]) && error_[t] > 0 && samples_[t] > 0)
            lo = std::min(lo, double(error_[t])), hi = std::max(hi, double(error_[t]));
    if (b.remaining <= 0 || hi <= 0) return lastTarget_ = 3.4e38f;
    auto cost = [&](double target) {
        double c = 0;
        for (size_t t = 0; t < error_.size(); ++t) {
            if ((!dead_.empty() && dead_[t]) || error_[t] <= 0 || samples_[t] <= 0) continue;
            const double n = samples_[t], r = double(error_[t]) / target;
            c += pixels_[t] * (std::min(std::max(n * r * r, n), double(b.most)) - std::min(n, double(b.most)));
        }
        return c;
    };
    double low = lo * 1e-3, high = hi;
    if (cost(low) <= b.remaining) return lastTarget_ = float(low);
    for (int i = 0; i < 48; ++i) {
        const double middle = sqrt(low * high);
        (cost(middle) > b.remaining ? low : high) = middle;
    }
    return lastTarget_ = float(high);
}

}  // namespace x2::adaptive
