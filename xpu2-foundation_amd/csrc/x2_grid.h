// x2_grid.h — the work grid that hands a frame's chunks to the devices, and its cellular rules, host side
// (xpu2/interactive/async_grid.h:38-439).
//
// The image is cut into square cells (64 pixels by default); each cell keeps its unstarted chunks in a grow-only queue
// ordered by sample range and then tile. A claim is one compare-and-swap on the cell's cursor, so every chunk goes to
// exactly one claimant. Cells start owned by one device (the caller's owners, else a checkerboard); workers of both
// devices finish their cell, then their device's reserved cells, then the open ones (most important first), then —
// by a right of passage — the other device's, a neighbour before the best cell anywhere. A uniform frame goes out in
// progressive sample ranges (1, 1, 2, then the chunk size), none handed to a device past the least-sampled range of
// the cells it works (its frontier). Ownership persists across frames and flips towards the device that rendered most
// of a cell.
#pragma once
#include <stdint.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/x2dev.h"

namespace x2::grid {

enum Device : uint32_t { Cpu = X2_DEVICE_CPU, Gpu = X2_DEVICE_GPU };

// A cell's chunks: appended by one coordinator while any number of workers claim. Chunks live in fixed blocks that
// never move, so a claimant copies out of them while the coordinator appends (async_grid.h:41-101).
class CellQueue {
public:
    static constexpr uint32_t BlockChunks = 4096, MaxBlocks = 2048;
    CellQueue();
    ~CellQueue();
    CellQueue(const CellQueue&) = delete;
    CellQueue& operator=(const CellQueue&) = delete;
    bool append(const x2_chunk* chunks, uint32_t n);  // false: the queue is full
    uint32_t published() const { return published_.load(std::memory_order_acquire); }
    uint32_t taken() const { return next_.load(std::memory_order_acquire); }
    uint32_t remaining() const {
        const uint32_t end = published(), at = taken();
        return at >= end ? 0 : end - at;
    }
    // Up to n chunks, none at or past queue position limit.
    uint32_t claim(uint32_t n, x2_chunk* out, uint32_t limit);

private:
    std::unique_ptr<std::atomic<x2_chunk*>[]> blocks_;
    alignas(64) std::atomic<uint32_t> published_{0};
    alignas(64) std::atomic<uint32_t> next_{0};
};

struct Range {
    uint32_t first, count;
};

class WorkGrid {
public:
    // Every chunk of a uniform frame is queued at once; with empty the cells start with nothing (adaptive passes are
    // appended later). viewport: {x0, y0, x1, y1} half open, or null for the whole image. owners: one per cell.
    WorkGrid(uint32_t width, uint32_t height, uint32_t spp, const x2_grid_options& options, const uint8_t* owners,
             uint32_t ownerCount, bool empty, const int32_t* viewport);
    uint32_t cellCount() const { return uint32_t(cells_.size()); }
    uint64_t totalChunks() const { return total_.load(std::memory_order_relaxed); }
    uint32_t spp() const { return spp_; }
    uint32_t samplesPerChunk() const { return options_.samples_per_chunk; }
    int32_t mode() const { return options_.mode; }
    Device owner(uint32_t cell) const { return cells_[cell].owner; }
    uint32_t remaining(uint32_t cell) const { return closed() ? 0u : cells_[cell].queue.remaining(); }
    uint64_t remainingChunks() const;
    bool empty() const;
    bool append(uint32_t cell, const x2_chunk* chunks, uint32_t n);
    uint32_t cellOf(uint32_t tile) const;
    uint32_t tilePixels(uint32_t tile) const;
    uint32_t tileCount() const { return tilesX_ * tilesY_; }

    uint32_t claim(uint32_t cell, uint32_t n, x2_chunk* out, Device device);
    uint32_t claimable(uint32_t cell, Device device) const;
    uint32_t frontier(Device device) const { return frontier_[device].load(std::memory_order_acquire); }
    const std::vector<Range>& ranges() const { return ranges_; }

    void setScores(const float* scores, uint32_t n);
    bool scored() const { return scored_.load(std::memory_order_acquire); }
    bool important(uint32_t cell) const { return cells_[cell].important.load(std::memory_order_relaxed); }
    bool reserved(uint32_t cell) const { return !scored() || important(cell); }
    void setSuitability(const float* p, uint32_t n);

    void setRates(double gpuSamples, double cpuWorkerSamples, uint32_t cpuWorkers, double windowSeconds);
    void gpuClaimed();
    bool passage(Device device, uint32_t cell) const;

    void close() { closed_.store(true, std::memory_order_release); }
    bool closed() const { return closed_.load(std::memory_order_acquire); }
    void rendered(uint32_t cell, Device device, uint64_t samples) {
        cells_[cell].samples[device].fetch_add(samples, std::memory_order_relaxed);
    }
    uint64_t renderedBy(uint32_t cell, Device device) const {
        return cells_[cell].samples[device].load(std::memory_order_relaxed);
    }
    void deferredToCpu(uint32_t tile);

    uint32_t choose(Device device, uint32_t current, uint32_t salt) const;
    uint32_t seed(Device device, uint32_t worker, uint32_t workers) const;

private:
    struct Cell {
        Device owner = Cpu;
        CellQueue queue;
        uint32_t tiles = 0;  // its tiles with pixels in the frame: the chunks of one sample range
        float gpu = 0.5f;    // the GPU's probability of being the cell's efficient owner
        std::atomic<float> score{0.0f};
        std::atomic<bool> important{false};
        std::atomic<uint64_t> samples[2] = {};
    };
    uint32_t limit(uint32_t cell, Device device) const;
    void refreshFrontiers();
    uint64_t reservedChunks(Device device) const;
    template <class F>
    void neighbours(uint32_t cell, F&& f) const;

    uint32_t width_, height_, spp_, cellsX_ = 0, cellsY_ = 0, tilesX_ = 0, tilesY_ = 0;
    int32_t viewport_[4] = {0, 0, 0x7fffffff, 0x7fffffff};
    x2_grid_options options_;
    std::vector<Cell> cells_;
    std::vector<Range> ranges_;
    std::atomic<uint64_t> total_{0};
    std::atomic<bool> closed_{false}, scored_{false};
    std::atomic<uint32_t> frontier_[2] = {};
    std::mutex frontierLock_;  // each refresh reads the queues and publishes both frontiers as one step
    std::atomic<double> gpuRate_{0}, cpuRate_{0}, window_{0.01};
    std::atomic<uint32_t> cpuWorkers_{1};
    std::atomic<int64_t> gpuClaim_{0};
};

// Cell owners from frame to frame (async_grid.h:366-437): flips towards the device that rendered most of a cell over
// the last 8 frames, at most 8 a frame and a tenth of the cells, 16 frames apart for a cell; a cell flipping back
// within 32 frames freezes every owner for 64.
class Ownership {
public:
    bool owners(uint32_t cells, uint8_t* out) const;  // false: not known for this many cells (the checkerboard)
    void finished(const WorkGrid& grid);

private:
    static constexpr uint32_t Window = 8;
    std::vector<uint8_t> owners_;
    std::vector<float> history_;
    std::vector<int64_t> lastFlip_;
    int64_t frame_ = 0, frozenUntil_ = 0;
};

}  // namespace x2::grid
