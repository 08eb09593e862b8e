// x2_grid.h — the work grid that hands a frame's chunks to the devices, and the cell owners across frames, host side.
//
// It follows the reference's grid rules (xpu2/interactive/async_grid.h:38-439), written here from a statement of them
// (the cells and their chunk order, the progressive ranges, the frontier and claim limit, importance, the rights of
// passage, a worker's choice and starting cell, ownership across frames), so that a broker driving it makes the
// reference's choices. The structures are this build's own:
//   - a cell's chunks sit in a vector behind the cell's own lock; its size and claim cursor are atomics that any thread
//     reads without the lock, so choosing a cell never waits and a claim holds the lock only while it copies out;
//   - the frontier is kept by a census: for each device the number of cells with work at each range index, updated
//     as a claim moves a cell's cursor, so a claim costs a few counter updates instead of a pass over every cell.
#pragma once
#include <stdint.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/x2dev.h"

namespace x2::grid {

enum Device : uint32_t { Cpu = X2_DEVICE_CPU, Gpu = X2_DEVICE_GPU };

struct Range {
    uint32_t first, count;
};

// A cell's queued chunks and how far claims have got: chunks [cursor, size) are unclaimed. Appends (one coordinator)
// and claims (any thread) take `lock`; size and cursor change only under it and are read without it.
struct CellWork {
    std::mutex lock;
    std::vector<x2_chunk> chunks;
    std::atomic<uint32_t> size{0}, cursor{0};
    uint32_t unclaimed() const {
        const uint32_t s = size.load(std::memory_order_acquire), c = cursor.load(std::memory_order_acquire);
        return c < s ? s - c : 0u;
    }
};

class WorkGrid {
public:
    // Every chunk of a uniform frame is queued at once; with empty the cells start with nothing (adaptive passes are
    // appended later). viewport: {x0, y0, x1, y1} half open, or null for the whole image. owners: one per cell.
    WorkGrid(uint32_t width, uint32_t height, uint32_t spp, const x2_grid_options& options, const uint8_t* owners,
             uint32_t ownerCount, bool empty, const int32_t* viewport);
    WorkGrid(const WorkGrid&) = delete;
    WorkGrid& operator=(const WorkGrid&) = delete;

    uint32_t cellCount() const { return cellCount_; }
    uint64_t totalChunks() const { return total_.load(std::memory_order_relaxed); }
    uint32_t spp() const { return spp_; }
    uint32_t samplesPerChunk() const { return options_.samples_per_chunk; }
    int32_t mode() const { return options_.mode; }
    Device owner(uint32_t cell) const { return meta_[cell].owner; }
    uint32_t remaining(uint32_t cell) const { return closed() ? 0u : work_[cell].unclaimed(); }
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
    bool important(uint32_t cell) const { return meta_[cell].important.load(std::memory_order_relaxed); }
    bool reserved(uint32_t cell) const { return !scored() || important(cell); }
    void setSuitability(const float* p, uint32_t n);

    void setRates(double gpuSamples, double cpuWorkerSamples, uint32_t cpuWorkers, double windowSeconds);
    void gpuClaimed();
    bool passage(Device device, uint32_t cell) const;

    void close() { closed_.store(true, std::memory_order_release); }
    bool closed() const { return closed_.load(std::memory_order_acquire); }
    void rendered(uint32_t cell, Device device, uint64_t samples) {
        meta_[cell].samples[device].fetch_add(samples, std::memory_order_relaxed);
    }
    uint64_t renderedBy(uint32_t cell, Device device) const {
        return meta_[cell].samples[device].load(std::memory_order_relaxed);
    }
    void deferredToCpu(uint32_t tile);

    uint32_t choose(Device device, uint32_t current, uint32_t salt) const;
    uint32_t seed(Device device, uint32_t worker, uint32_t workers) const;

private:
    struct CellMeta {
        Device owner = Cpu;
        uint32_t tiles = 0;  // tiles with pixels in the frame: a sample range's chunks in this cell
        float gpu = 0.5f;    // the GPU's probability of being the cell's efficient owner (the indexer's)
        std::atomic<float> score{0.0f};
        std::atomic<bool> important{false};
        std::atomic<uint64_t> samples[2] = {};
        // what the census counts for this cell (under censusLock_): its range index, or NotCounted
        uint32_t counted = NotCounted;
        bool countedOpen = false;
    };
    static constexpr uint32_t NotCounted = ~0u;
    uint32_t limit(uint32_t cell, Device device) const;
    uint64_t reservedChunks(Device device) const;
    // The census: recount one cell (after a claim or an append moved its counts), or every cell (after the scores
    // changed which cells are open); then publish both frontiers.
    void recount(uint32_t cell);
    void recountAll();
    void censusAdd(uint32_t cell, int delta);
    void publishFrontiers();

    uint32_t width_, height_, spp_, cellsX_ = 0, cellsY_ = 0, tilesX_ = 0, tilesY_ = 0, cellCount_ = 0;
    int32_t viewport_[4] = {0, 0, 0x7fffffff, 0x7fffffff};
    x2_grid_options options_;
    std::unique_ptr<CellWork[]> work_;
    std::unique_ptr<CellMeta[]> meta_;
    std::vector<Range> ranges_;
    std::atomic<uint64_t> total_{0};
    std::atomic<bool> closed_{false}, scored_{false};
    std::atomic<uint32_t> frontier_[2] = {};
    std::mutex censusLock_;
    std::vector<uint32_t> census_[3];  // [device or 2: any cell][range index 0..ranges]: cells with work there
    std::vector<uint32_t> stale_;      // cells appended to since the census last counted them
    std::atomic<double> gpuRate_{0}, cpuRate_{0}, window_{0.01};
    std::atomic<uint32_t> cpuWorkers_{1};
    std::atomic<int64_t> gpuClaim_{0};
};

// Cell owners from frame to frame: each cell's GPU share of the samples rendered in it over the last 8 frames; a cell
// whose other device rendered most of it (a share above 0.6) changes hands, at most 8 a frame and a tenth of the
// cells, a cell at most once in 16 frames; a cell changing back within 32 frames of its last change freezes every
// owner for 64 frames.
class Ownership {
public:
    bool owners(uint32_t cells, uint8_t* out) const;  // false: not known for this many cells (the checkerboard)
    void finished(const WorkGrid& grid);

private:
    static constexpr uint32_t Frames = 8;
    struct Record {
        uint8_t owner;
        int64_t changed;        // the frame of its last change of hands
        float share[Frames];    // the GPU's share of its samples, frame by frame (a ring)
    };
    std::vector<Record> cells_;
    int64_t frame_ = 0, frozenUntil_ = 0;
};

int gridFail(int code, const std::string& what);  // sets x2dev_last_error's text; returns code

}  // namespace x2::grid

// The grid behind a C handle (x2_grid.cpp), for the broker's frame (x2_broker.cpp).
x2::grid::WorkGrid& gridOf(x2_grid* g);
