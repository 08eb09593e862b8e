// x2_adaptive_core.h — the arithmetic of MoonRay's adaptive error trees shared by the host's trees (x2_adaptive.hip)
// and the device's pass chain (x2_kernel.hip chainTreesKernel), so both build the same trees, and so the same tile
// masks, from the same errors (async_adaptive_core.h:29-164, async_integrator.h:3109-3125). Every product and sum
// rounds on its own (the build never contracts).
//
// Adapted from MoonRay rndr/adaptive/AdaptiveRegionTree.{h,cc} and ActivePixelMask.h, by way of the reference's
// restatement (xpu2/interactive/async_adaptive_core.h): Apache License 2.0, copyright 2023-2024 DreamWorks Animation
// LLC (http://www.apache.org/licenses/LICENSE-2.0). Changes: restated for HIP host and device code with this
// build's names; the node arithmetic is unchanged so that the same errors give the same trees.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "x2_adaptive.h"

namespace x2::adaptive {

__host__ __device__ inline int min2(int a, int b) { return b < a ? b : a; }
__host__ __device__ inline int max2(int a, int b) { return a < b ? b : a; }

constexpr float MinNodeSize = 1.27201964951407f;  // sqrt of the golden ratio (async_adaptive_core.h:29)
constexpr float MaxNodeSize = 16.0f;
constexpr float SplitThreshold = 256.0f;
constexpr float NoSplit = -1.0f;
// MoonRay judges a node's error against a fixed image area, whatever the image (async_adaptive_core.h:53-56).
constexpr float ReferenceArea = (815.0f * 1967.0f) / 4.0f;

__host__ __device__ inline IBox outward(const FBox& b) {
    return {int(floorf(b.x0)), int(floorf(b.y0)), int(ceilf(b.x1)), int(ceilf(b.y1))};
}

__host__ __device__ inline float nodeError(float summed, float area) { return summed / sqrtf(area * ReferenceArea); }

// The offset along a node (a fraction of its length) for a split at `at`, kept MinNodeSize from either end.
__host__ __device__ inline float keepFromEdge(float at, float length) {
    if (length * at >= MinNodeSize) return at;
    if (length >= MinNodeSize * 2) return MinNodeSize / length;
    return NoSplit;
}
__host__ __device__ inline float splitOffset(float at, float length) {
    if (at <= 0.5f) return keepFromEdge(at, length);
    const float mirrored = keepFromEdge(1.0f - at, length);
    return mirrored == NoSplit ? NoSplit : 1.0f - mirrored;
}

// A region's summed-area table: sum(x0..x, y0..y) of errors at [(y - y0) * w1 + (x - x0)], exclusive corners.
struct Table {
    const double* s;
    int x0, y0, w1;
    __host__ __device__ double at(int x, int y) const { return s[size_t(y - y0) * size_t(w1) + size_t(x - x0)]; }
    // The summed error of [ax, bx) x [ay, by).
    __host__ __device__ double box(int ax, int ay, int bx, int by) const { return at(bx, by) - at(ax, by) - at(bx, ay) + at(ax, ay); }
};

// The error a node has accumulated in its first k + 1 columns (rows when it splits across y), as a float.
struct Prefix {
    const Table& t;
    IBox b;
    bool acrossX;
    int n;
    __host__ __device__ float operator()(int k) const {
        return acrossX ? float(t.box(b.x0, b.y0, b.x0 + k + 1, b.y1)) : float(t.box(b.x0, b.y0, b.x1, b.y0 + k + 1));
    }
};

// Where along the node half its error lies: the first column whose prefix passes half the total (bisection),
// interpolated within that column.
__host__ __device__ inline float halfErrorAt(const Prefix& p) {
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
    k = min2(max2(k, 1), p.n - 1);
    const float below = p(k - 1), above = p(k), step = above - below;
    const float within = step > 0 ? (half - below) / step : 0.5f;
    return (float(k) + within) / float(p.n);
}

enum Fate { Open, Converged, Split };

// A node's fate (AdaptiveRegionTree::updateImpl): converged, split at the half-error point of its longer side, or
// left open.
__host__ __device__ inline Fate judge(const Table& t, const IBox& root, const FBox& b, float target, FBox& first, FBox& second) {
    const float w = b.x1 - b.x0, h = b.y1 - b.y0;
    const bool acrossX = w > h;
    const float length = acrossX ? w : h;
    IBox ib = outward(b);
    ib = {max2(ib.x0, root.x0), max2(ib.y0, root.y0), min2(ib.x1, root.x1), min2(ib.y1, root.y1)};
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


// The bounding rectangle of a tile's active bits, filled (ActivePixelMask::fillGaps, async_adaptive_core.h:135-151).
__host__ __device__ inline uint64_t fillRectangle(uint64_t bits) {
    if (!bits) return 0;
    int row0 = 8, row1 = -1, col0 = 8, col1 = -1;
    for (int q = 0; q < 64; ++q)
        if (bits >> q & 1) {
            row0 = min2(row0, q / 8), row1 = max2(row1, q / 8);
            col0 = min2(col0, q % 8), col1 = max2(col1, q % 8);
        }
    uint64_t filled = 0;
    for (int y = row0; y <= row1; ++y)
        for (int x = col0; x <= col1; ++x) filled |= 1ull << (y * 8 + x);
    return filled;
}

// A pixel's error (async_integrator.h:3109-3125): the luma of the difference between the even- and odd-index
// half-images over the square root of the pixel's luma, plus the alpha difference; -1 until both halves have a sample.
__host__ __device__ inline float pixelErrorOf(const double total[4], const double odd[4], uint32_t samples) {
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

}  // namespace x2::adaptive
