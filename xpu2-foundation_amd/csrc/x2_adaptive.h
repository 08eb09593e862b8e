// x2_adaptive.h — an adaptive frame's host side: MoonRay's pass schedule, the per-pixel error, and the region error
// trees that choose each pass's active pixels (xpu2/interactive/async_adaptive.h:22-144, async_adaptive_core.h,
// async_integrator.h:3102-3125). The device renders the passes and works out the errors; the trees run here.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace x2::adaptive {

struct Pass {
    uint32_t start, end;  // each active pixel goes from start to end samples
};
// One sample, then 4, 6, 8, ... more, the rest in a last pass (async_adaptive.h:27-38).
std::vector<Pass> schedule(uint32_t maxSamples);

// A pixel's error from its sums, odd-index sums and count; -1 until both halves have a sample
// (async_integrator.h:3109-3125). Every product and sum rounds on its own (the build never contracts).
float pixelError(const double total[4], const double odd[4], uint32_t samples);

struct IBox {
    int x0, y0, x1, y1;
};
struct FBox {
    float x0, y0, x1, y1;
};
struct Node {
    FBox bounds;
    int children;  // the first of two, or -1: a leaf
    int complete;  // converged (an inner node: both children have)
};

// One region's tree (AdaptiveRegionTree): rebuilt from the image's errors before each adaptive pass.
class Tree {
public:
    Tree(IBox root, float target);
    void update(const float* errors, int stride);
    bool done() const { return nodes_[0].complete != 0; }
    // The pixels of the 8x8 tile at (tx, ty) still to sample (bit y * 8 + x), filled to their bounding rectangle.
    uint64_t sampleArea(int tx, int ty) const;
    const std::vector<Node>& nodes() const { return nodes_; }
    IBox root() const { return root_; }

private:
    void updateNode(int index);
    IBox root_;
    float target_;
    std::vector<double> sat_;  // the region's summed-area table of errors, (w + 1) x (h + 1), first row/column zero
    std::vector<Node> nodes_;
};

// The image's overlapping regions (up to 4 x 4, whole 8-pixel tiles, widened by 8 pixels), each with its tree; a
// tile reads the tree of the region its corner lies in (async_adaptive.h:99-145, async_adaptive_core.h:182-191).
class Regions {
public:
    Regions(int width, int height, float target);
    void update(const float* errors);
    bool done() const;
    uint64_t sampleArea(int tx, int ty) const;
    size_t count() const { return trees_.size(); }
    const Tree& tree(size_t i) const { return trees_[i]; }
    int perDim() const { return perDim_; }
    int regionWidth() const { return regionW_; }
    int regionHeight() const { return regionH_; }

private:
    int width_, perDim_ = 1, regionW_ = 0, regionH_ = 0;
    std::vector<Tree> trees_;
};

// Cellular generations (async_adaptive.h:147-317, the preview session's default): each adaptive pass a generation of
// the image's 8x8 tiles as cells. A tile's error is MoonRay's node error over it against the target; its own error
// keeps a tile alive at 1 or more, live neighbours keep it alive down to Survive (3 of its 8) and bring a dead one
// back from Birth (5 of 8). With a budget the target is the one the samples still to spend would bring every tile to.
struct Budget {
    const std::vector<uint32_t>* samples;  // each pixel's samples planned so far
    double remaining;                      // the frame's samples still to spend
    uint32_t most;                         // the most samples a pixel may take
    float floor;                           // the error no tile need go below (0: none)
};
class Cellular {
public:
    Cellular(int width, int height, float target);
    void deadForever(const std::vector<uint64_t>& inImage);  // tiles without a pixel in the render region
    size_t update(const float* errors, const Budget* budget);  // the next generation; its live tiles
    bool alive(size_t tile) const { return alive_[tile] != 0; }
    float ratio(size_t tile) const { return ratio_[tile]; }
    const std::vector<char>& aliveTiles() const { return alive_; }
    void setAlive(const std::vector<char>& a) { alive_ = a; }

private:
    float budgetTarget(const Budget& b);
    int width_, height_, tilesX_, tilesY_;
    float target_, lastTarget_ = 0;
    std::vector<char> alive_, next_, dead_;
    std::vector<float> ratio_, error_;
    std::vector<double> pixels_, samples_;
};

}  // namespace x2::adaptive
