// x2_grid.cpp — the work grid and the cell owners across frames (x2_grid.h), and their C entry points (x2_grid_*,
// x2_ownership_*). Host code only. The rules are the reference's (xpu2/interactive/async_grid.h:38-439); each function
// below names the rule it carries out.
#include "x2_grid.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

namespace x2 {
int reportFailure(int code, const std::string& what);  // sets x2dev_last_error's text (x2_lane.hip)
}

namespace x2::grid {
namespace {
constexpr uint32_t TileSize = 8, TilePixels = 64;

int64_t steadyNs() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
// The tie-break between cells of equal merit for one worker: a 32-bit mix of (cell, salt), the reference's
// (async_grid.h:424-430), so that the same worker picks the same cell.
uint32_t mix(uint32_t cell, uint32_t salt) {
    uint32_t h = (cell * 0x9e3779b1u) ^ (salt + 0x7f4a7c15u + (cell << 6) + (cell >> 2));
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    return h ^ (h >> 13);
}
}  // namespace

// The cells, their owners, the sample ranges and each cell's chunks (async_grid.h:131-177): ranges of 1, 1, 2, then
// samples_per_chunk (progressive) or samples_per_chunk throughout, over the frame's share of sample indices; a cell's
// chunks range by range, each range over the cell's tiles that hold a pixel of the viewport, row by row.
WorkGrid::WorkGrid(uint32_t width, uint32_t height, uint32_t spp, const x2_grid_options& options, const uint8_t* owners,
                   uint32_t ownerCount, bool empty, const int32_t* viewport)
    : width_(width), height_(height), spp_(spp), options_(options) {
    if (!width || !height || !spp || !options.cell_size || options.cell_size % TileSize || !options.samples_per_chunk ||
        options.mode < 0 || options.mode > 2)
        throw std::invalid_argument("invalid work grid (sizes, cell_size a multiple of 8, mode 0..2)");
    if (viewport) std::memcpy(viewport_, viewport, sizeof viewport_);
    const uint32_t side = options.cell_size;
    cellsX_ = (width + side - 1) / side;
    cellsY_ = (height + side - 1) / side;
    tilesX_ = (width + TileSize - 1) / TileSize;
    tilesY_ = (height + TileSize - 1) / TileSize;
    cellCount_ = cellsX_ * cellsY_;
    work_.reset(new CellWork[cellCount_]);
    meta_.reset(new CellMeta[cellCount_]);
    const bool given = owners && ownerCount == cellCount_;
    if (!empty) {
        const uint32_t spc = options.samples_per_chunk;
        const uint32_t lo = options.sample_count ? options.sample_first : 0u;
        const uint32_t hi = options.sample_count ? std::min(spp, options.sample_first + options.sample_count) : spp;
        uint32_t at = lo;
        while (at < hi) {
            const uint32_t into = at - lo;
            uint32_t size = spc;
            if (options.progressive) size = std::min(spc, into < 2 ? 1u : into < 4 ? 2u : spc);
            size = std::min(size, hi - at);
            ranges_.push_back({at, size});
            at += size;
        }
    }
    const uint32_t perSide = side / TileSize;
    for (uint32_t c = 0; c < cellCount_; ++c) {
        const uint32_t cx = c % cellsX_, cy = c / cellsX_;
        CellMeta& m = meta_[c];
        if (options.mode == 0) m.owner = Cpu;
        else if (options.mode == 1) m.owner = Gpu;
        else if (given) m.owner = owners[c] ? Gpu : Cpu;
        else m.owner = (cx + cy) % 2 == 0 ? Gpu : Cpu;
        if (empty) continue;
        std::vector<uint32_t> tiles;
        for (uint32_t ty = cy * perSide; ty < std::min(tilesY_, (cy + 1) * perSide); ++ty)
            for (uint32_t tx = cx * perSide; tx < std::min(tilesX_, (cx + 1) * perSide); ++tx)
                if (tilePixels(ty * tilesX_ + tx)) tiles.push_back(ty * tilesX_ + tx);
        CellWork& w = work_[c];
        w.chunks.reserve(tiles.size() * ranges_.size());
        for (const Range& r : ranges_)
            for (const uint32_t t : tiles) w.chunks.push_back({t, r.first, r.count, c});
        m.tiles = ranges_.empty() ? 0u : uint32_t(tiles.size());
        w.size.store(uint32_t(w.chunks.size()), std::memory_order_release);
        total_ += w.chunks.size();
    }
    if (!ranges_.empty())
        for (auto& counts : census_) counts.assign(ranges_.size() + 1, 0u);
    recountAll();
}

uint64_t WorkGrid::remainingChunks() const {
    uint64_t n = 0;
    for (uint32_t c = 0; c < cellCount_; ++c) n += remaining(c);
    return n;
}
bool WorkGrid::empty() const {
    for (uint32_t c = 0; c < cellCount_; ++c)
        if (remaining(c)) return false;
    return true;
}
// The coordinator's appends (async_grid.h:182-186): they leave the frontiers as they are until the next claim or
// score update counts the cell again.
bool WorkGrid::append(uint32_t cell, const x2_chunk* chunks, uint32_t n) {
    CellWork& w = work_[cell];
    {
        std::lock_guard<std::mutex> hold(w.lock);
        if (uint64_t(w.chunks.size()) + n > 0x7fffffffull) return false;
        try {
            w.chunks.insert(w.chunks.end(), chunks, chunks + n);
        } catch (const std::bad_alloc&) {
            return false;
        }
        w.size.store(uint32_t(w.chunks.size()), std::memory_order_release);
    }
    total_ += n;
    if (!ranges_.empty()) {
        std::lock_guard<std::mutex> hold(censusLock_);
        stale_.push_back(cell);
    }
    return true;
}
uint32_t WorkGrid::cellOf(uint32_t tile) const {
    const uint32_t perSide = options_.cell_size / TileSize;
    return ((tile / tilesX_) / perSide) * cellsX_ + (tile % tilesX_) / perSide;
}
uint32_t WorkGrid::tilePixels(uint32_t tile) const {
    const int x = int((tile % tilesX_) * TileSize), y = int((tile / tilesX_) * TileSize);
    const int x0 = std::max(x, int(viewport_[0])), y0 = std::max(y, int(viewport_[1]));
    const int x1 = std::min({x + int(TileSize), int(width_), int(viewport_[2])});
    const int y1 = std::min({y + int(TileSize), int(height_), int(viewport_[3])});
    return x1 > x0 && y1 > y0 ? uint32_t(x1 - x0) * uint32_t(y1 - y0) : 0u;
}

// The claim limit (async_grid.h:372-377): a device reaches the ranges up to its frontier, one further in an important
// cell of a scored grid; no limit without ranges (appended work).
uint32_t WorkGrid::limit(uint32_t cell, Device device) const {
    if (ranges_.empty()) return ~0u;
    const uint64_t ahead = scored() && important(cell) ? 2 : 1;
    const uint64_t upTo = std::min<uint64_t>(uint64_t(frontier(device)) + ahead, ranges_.size());
    return uint32_t(std::min<uint64_t>(upTo * meta_[cell].tiles, ~0u));
}
// A claim (async_grid.h:231-237): up to n chunks from the cell's cursor, below its limit; each chunk goes out once.
uint32_t WorkGrid::claim(uint32_t cell, uint32_t n, x2_chunk* out, Device device) {
    if (closed() || !n) return 0;
    const uint32_t lim = limit(cell, device);
    CellWork& w = work_[cell];
    uint32_t got = 0;
    {
        std::lock_guard<std::mutex> hold(w.lock);
        const uint32_t from = w.cursor.load(std::memory_order_relaxed);
        const uint32_t end = std::min(w.size.load(std::memory_order_relaxed), lim);
        if (from >= end) return 0;
        got = std::min(n, end - from);
        std::memcpy(out, w.chunks.data() + from, size_t(got) * sizeof(x2_chunk));
        w.cursor.store(from + got, std::memory_order_release);
    }
    if (!ranges_.empty()) recount(cell);
    return got;
}
uint32_t WorkGrid::claimable(uint32_t cell, Device device) const {
    if (closed()) return 0;
    const CellWork& w = work_[cell];
    const uint32_t end = std::min(w.size.load(std::memory_order_acquire), limit(cell, device));
    const uint32_t from = w.cursor.load(std::memory_order_acquire);
    return from < end ? end - from : 0u;
}

// The frontiers (async_grid.h:380-395): a device's is the least range index (cursor / tiles) among the cells with
// unclaimed chunks that are open or its own reserved ones, else among all cells with unclaimed chunks, at most the
// number of ranges. The census keeps, per device and for any cell, how many such cells sit at each range index; each
// cell's counted entry is replaced under censusLock_ whenever its counts move, so the published frontiers always come
// from one consistent census.
void WorkGrid::censusAdd(uint32_t cell, int delta) {
    CellMeta& m = meta_[cell];
    if (m.counted == NotCounted) return;
    const uint32_t r = m.counted;
    census_[2][r] += delta;
    for (uint32_t d = 0; d < 2; ++d)
        if (m.countedOpen || uint32_t(m.owner) == d) census_[d][r] += delta;
}
void WorkGrid::recount(uint32_t cell) {
    std::lock_guard<std::mutex> hold(censusLock_);
    auto place = [&](uint32_t c) {
        censusAdd(c, -1);
        CellMeta& m = meta_[c];
        m.counted = NotCounted;
        const CellWork& w = work_[c];
        if (m.tiles && w.unclaimed()) {
            m.counted = std::min<uint32_t>(w.cursor.load(std::memory_order_acquire) / m.tiles, uint32_t(ranges_.size()));
            m.countedOpen = !reserved(c);
            censusAdd(c, +1);
        }
    };
    for (const uint32_t c : stale_) place(c);
    stale_.clear();
    place(cell);
    publishFrontiers();
}
void WorkGrid::recountAll() {
    if (ranges_.empty()) return;
    {
        std::lock_guard<std::mutex> hold(censusLock_);
        for (auto& counts : census_) std::fill(counts.begin(), counts.end(), 0u);
        stale_.clear();
        for (uint32_t c = 0; c < cellCount_; ++c) {
            CellMeta& m = meta_[c];
            m.counted = NotCounted;
            const CellWork& w = work_[c];
            if (m.tiles && w.unclaimed()) {
                m.counted =
                    std::min<uint32_t>(w.cursor.load(std::memory_order_acquire) / m.tiles, uint32_t(ranges_.size()));
                m.countedOpen = !reserved(c);
                censusAdd(c, +1);
            }
        }
        publishFrontiers();
    }
}
void WorkGrid::publishFrontiers() {  // censusLock_ held
    const uint32_t n = uint32_t(ranges_.size());
    auto least = [&](const std::vector<uint32_t>& counts) {
        for (uint32_t r = 0; r <= n; ++r)
            if (counts[r]) return r;
        return ~0u;
    };
    const uint32_t any = least(census_[2]);
    for (uint32_t d = 0; d < 2; ++d) {
        const uint32_t own = least(census_[d]);
        frontier_[d].store(std::min(own != ~0u ? own : any, n), std::memory_order_release);
    }
}

uint64_t WorkGrid::reservedChunks(Device device) const {
    uint64_t n = 0;
    for (uint32_t c = 0; c < cellCount_; ++c)
        if (owner(c) == device && reserved(c)) n += remaining(c);
    return n;
}

// Important parts first (async_grid.h:250-268): the known scores' mean (negative: unknown); a cell above it is
// important and stays reserved to its owner; the others open to both devices.
void WorkGrid::setScores(const float* scores, uint32_t n) {
    if (n != cellCount_) return;
    double sum = 0;
    uint32_t known = 0;
    for (uint32_t c = 0; c < n; ++c)
        if (scores[c] >= 0) {
            sum += scores[c];
            ++known;
        }
    if (!known) return;
    const float mean = float(sum / known);
    for (uint32_t c = 0; c < n; ++c) {
        meta_[c].score.store(std::max(scores[c], 0.0f), std::memory_order_relaxed);
        meta_[c].important.store(scores[c] > mean, std::memory_order_relaxed);
    }
    scored_.store(true, std::memory_order_release);
    recountAll();
}
void WorkGrid::setSuitability(const float* p, uint32_t n) {
    if (n != cellCount_) return;
    for (uint32_t c = 0; c < n; ++c) meta_[c].gpu = p[c];
}
void WorkGrid::setRates(double gpuSamples, double cpuWorkerSamples, uint32_t cpuWorkers, double windowSeconds) {
    gpuRate_.store(gpuSamples, std::memory_order_relaxed);
    cpuRate_.store(cpuWorkerSamples, std::memory_order_relaxed);
    cpuWorkers_.store(cpuWorkers, std::memory_order_relaxed);
    window_.store(windowSeconds, std::memory_order_relaxed);
}
void WorkGrid::gpuClaimed() { gpuClaim_.store(steadyNs(), std::memory_order_relaxed); }

// The rights of passage (async_grid.h:279-296): whether a device may work a cell the other holds reserved. Before both
// rates are known, yes. The CPU may once the GPU has not claimed for max(4 windows, 50 ms), or when a chunk of its own
// takes under half the time the GPU needs for the GPU's reserved chunks; the GPU may in a cell it suits (probability
// at least 0.2) when the CPU lane's reserved chunks would take more than 4 of the GPU's windows.
bool WorkGrid::passage(Device device, uint32_t cell) const {
    const double gpu = gpuRate_.load(std::memory_order_relaxed), cpu = cpuRate_.load(std::memory_order_relaxed);
    if (gpu <= 0 || cpu <= 0) return true;
    const double chunkSamples = double(options_.samples_per_chunk) * TilePixels;
    const double window = window_.load(std::memory_order_relaxed);
    if (device == Cpu) {
        const int64_t since = gpuClaim_.load(std::memory_order_relaxed);
        if (since && double(steadyNs() - since) * 1e-9 > std::max(4 * window, 0.05)) return true;
        const double ownChunk = chunkSamples / cpu, gpuBacklog = double(reservedChunks(Gpu)) * chunkSamples / gpu;
        return ownChunk < 0.5 * gpuBacklog;
    }
    if (meta_[cell].gpu < 0.2f) return false;
    const double cpuLane = cpu * std::max(1u, cpuWorkers_.load(std::memory_order_relaxed));
    return double(reservedChunks(Cpu)) * chunkSamples / cpuLane > 4 * window;
}

// A GPU sample handed to the CPU lane (a deferred sample): its count moves from the GPU's share to the CPU's.
void WorkGrid::deferredToCpu(uint32_t tile) {
    CellMeta& m = meta_[cellOf(tile)];
    uint64_t g = m.samples[Gpu].load(std::memory_order_relaxed);
    while (g && !m.samples[Gpu].compare_exchange_weak(g, g - 1, std::memory_order_relaxed)) {
    }
    m.samples[Cpu].fetch_add(1, std::memory_order_relaxed);
}

// A worker's next cell (async_grid.h:325-359). It stays in its cell while it may claim there (in the other device's
// reserved cell only by a right of passage). Otherwise it looks tier by tier: its device's reserved cells (merit: the
// chunks it may claim), the open cells (merit: score, then chunks), the other device's reserved cells it has a right
// of passage to (merit: chunks); in each tier first among its cell's eight neighbours, then anywhere; the highest
// merit wins, ties to the higher mix(cell, salt).
uint32_t WorkGrid::choose(Device device, uint32_t current, uint32_t salt) const {
    enum : int { Own = 0, Open = 1, Other = 2 };
    auto tierOf = [&](uint32_t c) { return !reserved(c) ? Open : owner(c) == device ? Own : Other; };
    int cpuMay = -1;  // the CPU's right of passage does not depend on the cell: asked at most once a call
    auto mayEnter = [&](uint32_t c) {
        if (device == Gpu) return passage(Gpu, c);
        if (cpuMay < 0) cpuMay = passage(Cpu, c) ? 1 : 0;
        return cpuMay == 1;
    };
    const bool inGrid = current < cellCount_;
    if (inGrid && claimable(current, device) && (tierOf(current) != Other || mayEnter(current))) return current;
    struct Pick {
        uint32_t cell = X2_NO_CELL;
        float merit = 0.0f;
    };
    auto offer = [&](Pick& p, int tier, uint32_t c) {
        if (tierOf(c) != tier) return;
        const uint32_t chunks = claimable(c, device);
        if (!chunks || (tier == Other && !mayEnter(c))) return;
        const float merit = tier == Open ? meta_[c].score.load(std::memory_order_relaxed) + 1e-9f * float(chunks)
                                         : float(chunks);
        const bool better = p.cell == X2_NO_CELL || merit > p.merit || (merit == p.merit && mix(c, salt) > mix(p.cell, salt));
        if (better) p = Pick{c, merit};
    };
    const int cx = inGrid ? int(current % cellsX_) : 0, cy = inGrid ? int(current / cellsX_) : 0;
    for (int tier = Own; tier <= Other; ++tier) {
        Pick p;
        if (inGrid) {
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int nx = cx + dx, ny = cy + dy;
                    if ((dx || dy) && nx >= 0 && ny >= 0 && nx < int(cellsX_) && ny < int(cellsY_))
                        offer(p, tier, uint32_t(ny) * cellsX_ + uint32_t(nx));
                }
            if (p.cell != X2_NO_CELL) return p.cell;
        }
        for (uint32_t c = 0; c < cellCount_; ++c) offer(p, tier, c);
        if (p.cell != X2_NO_CELL) return p.cell;
    }
    return X2_NO_CELL;
}
// A worker's first cell (async_grid.h:362-368): worker w of n starts at the (w * k / n)-th of its device's k cells in
// index order; a device owning none chooses as from nowhere.
uint32_t WorkGrid::seed(Device device, uint32_t worker, uint32_t workers) const {
    uint32_t own = 0;
    for (uint32_t c = 0; c < cellCount_; ++c) own += owner(c) == device ? 1u : 0u;
    if (!own) return choose(device, X2_NO_CELL, worker);
    const uint64_t k = (uint64_t(worker) * own) / std::max(1u, workers);
    for (uint32_t c = 0, seen = 0; c < cellCount_; ++c)
        if (owner(c) == device && seen++ == k) return c;
    return X2_NO_CELL;  // worker >= workers: past the device's cells
}

// Owners across frames (async_grid.h:402-437).
bool Ownership::owners(uint32_t cells, uint8_t* out) const {
    if (cells_.size() != cells) return false;
    for (uint32_t c = 0; c < cells; ++c) out[c] = cells_[c].owner;
    return true;
}
void Ownership::finished(const WorkGrid& grid) {
    const uint32_t n = grid.cellCount();
    if (cells_.size() != n) {  // first frame, or another grid: start from this grid's owners
        cells_.assign(n, Record{});
        for (uint32_t c = 0; c < n; ++c) {
            cells_[c].owner = uint8_t(grid.owner(c));
            cells_[c].changed = -1000000;
        }
    }
    const uint32_t slot = uint32_t(frame_ % Frames);
    for (uint32_t c = 0; c < n; ++c) {
        const double cpu = double(grid.renderedBy(c, Cpu)), gpu = double(grid.renderedBy(c, Gpu));
        cells_[c].share[slot] = cpu + gpu > 0 ? float(gpu / (cpu + gpu)) : 0.5f;
    }
    ++frame_;
    if (frame_ < int64_t(Frames) || frame_ < frozenUntil_) return;
    std::vector<std::pair<float, uint32_t>> candidates;  // (the other device's share, cell)
    for (uint32_t c = 0; c < n; ++c) {
        const Record& r = cells_[c];
        if (frame_ - r.changed < 16) continue;
        float mean = 0.0f;
        for (uint32_t k = 0; k < Frames; ++k) mean += r.share[k];
        mean /= float(Frames);
        const float other = r.owner == Cpu ? mean : 1.0f - mean;
        if (other > 0.6f) candidates.push_back({other, c});
    }
    // The largest share first; equal shares in cell order.
    std::stable_sort(candidates.begin(), candidates.end(),
                     [](const std::pair<float, uint32_t>& a, const std::pair<float, uint32_t>& b) { return a.first > b.first; });
    const size_t cap = std::min<size_t>(8, std::max<size_t>(1, n / 10));
    for (size_t k = 0; k < std::min(cap, candidates.size()); ++k) {
        Record& r = cells_[candidates[k].second];
        if (frame_ - r.changed < 32) {  // changing back soon after a change: the owners thrash, and freeze
            frozenUntil_ = frame_ + 64;
            return;
        }
        r.owner = r.owner == Cpu ? uint8_t(Gpu) : uint8_t(Cpu);
        r.changed = frame_;
    }
}

int gridFail(int code, const std::string& what) { return reportFailure(code, what); }
}  // namespace x2::grid

using x2::grid::Ownership;
using x2::grid::WorkGrid;
struct x2_grid {
    WorkGrid g;
    template <class... A>
    explicit x2_grid(A&&... a) : g(std::forward<A>(a)...) {}
};
struct x2_ownership {
    Ownership o;
};
x2::grid::WorkGrid& gridOf(x2_grid* g) { return g->g; }  // for the broker (x2_broker.cpp)

namespace {
bool validDevice(uint32_t device) { return device == X2_DEVICE_CPU || device == X2_DEVICE_GPU; }
x2::grid::Device dev(uint32_t device) { return device == X2_DEVICE_GPU ? x2::grid::Gpu : x2::grid::Cpu; }
bool validCell(const x2_grid* g, uint32_t cell) { return g && cell < g->g.cellCount(); }
}  // namespace

extern "C" {

int x2_grid_create(uint32_t width, uint32_t height, uint32_t spp, const x2_grid_options* options, const uint8_t* owners,
                   uint32_t owner_count, int empty, const int32_t* viewport, x2_grid** out) {
    if (!out) return x2::grid::gridFail(X2_ERR_ARGUMENT, "null grid out");
    *out = nullptr;
    const x2_grid_options o = options ? *options : x2_grid_options{64, 4, 2, 1, 0, 0};
    try {
        *out = new x2_grid(width, height, spp, o, owners, owner_count, empty != 0, viewport);
    } catch (const std::bad_alloc&) {
        return x2::grid::gridFail(X2_ERR_MEMORY, "work grid allocation failed");
    } catch (const std::exception& e) {
        return x2::grid::gridFail(X2_ERR_ARGUMENT, e.what());
    }
    return 0;
}
void x2_grid_destroy(x2_grid* g) { delete g; }
uint32_t x2_grid_cell_count(const x2_grid* g) { return g ? g->g.cellCount() : 0u; }
uint64_t x2_grid_total_chunks(const x2_grid* g) { return g ? g->g.totalChunks() : 0u; }
uint32_t x2_grid_owner(const x2_grid* g, uint32_t cell) {
    return validCell(g, cell) ? uint32_t(g->g.owner(cell)) : X2_NO_CELL;
}
uint32_t x2_grid_remaining(const x2_grid* g, uint32_t cell) { return validCell(g, cell) ? g->g.remaining(cell) : 0u; }
uint32_t x2_grid_claimable(const x2_grid* g, uint32_t cell, uint32_t device) {
    return validCell(g, cell) && validDevice(device) ? g->g.claimable(cell, dev(device)) : 0u;
}
int x2_grid_empty(const x2_grid* g) { return g ? int(g->g.empty()) : 1; }
uint32_t x2_grid_claim(x2_grid* g, uint32_t cell, uint32_t n, x2_chunk* out, uint32_t device) {
    if (!out || !validCell(g, cell) || !validDevice(device)) return 0u;
    return g->g.claim(cell, n, out, dev(device));
}
int x2_grid_append(x2_grid* g, uint32_t cell, const x2_chunk* chunks, uint32_t n) {
    if ((!chunks && n) || !validCell(g, cell)) return x2::grid::gridFail(X2_ERR_ARGUMENT, "bad append");
    for (uint32_t i = 0; i < n; ++i)
        if (chunks[i].tile >= g->g.tileCount() || g->g.cellOf(chunks[i].tile) != cell)
            return x2::grid::gridFail(X2_ERR_ARGUMENT, "an appended chunk's tile lies outside its cell");
    return g->g.append(cell, chunks, n) ? 0 : x2::grid::gridFail(X2_ERR_MEMORY, "a cell's chunk queue is full");
}
uint32_t x2_grid_choose(const x2_grid* g, uint32_t device, uint32_t current, uint32_t salt) {
    return g && validDevice(device) ? g->g.choose(dev(device), current, salt) : X2_NO_CELL;
}
uint32_t x2_grid_seed(const x2_grid* g, uint32_t device, uint32_t worker, uint32_t workers) {
    return g && validDevice(device) ? g->g.seed(dev(device), worker, workers) : X2_NO_CELL;
}
uint32_t x2_grid_frontier(const x2_grid* g, uint32_t device) {
    return g && validDevice(device) ? g->g.frontier(dev(device)) : 0u;
}
uint32_t x2_grid_ranges(const x2_grid* g, uint32_t* out, uint32_t capacity) {
    if (!g) return 0u;
    const auto& r = g->g.ranges();
    for (uint32_t i = 0; out && i < capacity && i < r.size(); ++i) {
        out[2 * i] = r[i].first;
        out[2 * i + 1] = r[i].count;
    }
    return uint32_t(r.size());
}
int x2_grid_set_scores(x2_grid* g, const float* scores, uint32_t n) {
    if (!g || !scores || n != g->g.cellCount()) return x2::grid::gridFail(X2_ERR_ARGUMENT, "scores: one per cell");
    g->g.setScores(scores, n);
    return 0;
}
int x2_grid_important(const x2_grid* g, uint32_t cell) { return validCell(g, cell) ? int(g->g.important(cell)) : 0; }
int x2_grid_reserved(const x2_grid* g, uint32_t cell) { return validCell(g, cell) ? int(g->g.reserved(cell)) : 0; }
int x2_grid_set_suitability(x2_grid* g, const float* p, uint32_t n) {
    if (!g || !p || n != g->g.cellCount()) return x2::grid::gridFail(X2_ERR_ARGUMENT, "suitability: one per cell");
    g->g.setSuitability(p, n);
    return 0;
}
void x2_grid_set_rates(x2_grid* g, double gpu_samples, double cpu_worker_samples, uint32_t cpu_workers,
                       double window_seconds) {
    if (g) g->g.setRates(gpu_samples, cpu_worker_samples, cpu_workers, window_seconds);
}
void x2_grid_gpu_claimed(x2_grid* g) {
    if (g) g->g.gpuClaimed();
}
int x2_grid_passage(const x2_grid* g, uint32_t device, uint32_t cell) {
    return validCell(g, cell) && validDevice(device) ? int(g->g.passage(dev(device), cell)) : 0;
}
void x2_grid_close(x2_grid* g) {
    if (g) g->g.close();
}
void x2_grid_rendered(x2_grid* g, uint32_t cell, uint32_t device, uint64_t samples) {
    if (validCell(g, cell) && validDevice(device)) g->g.rendered(cell, dev(device), samples);
}
uint64_t x2_grid_rendered_by(const x2_grid* g, uint32_t cell, uint32_t device) {
    return validCell(g, cell) && validDevice(device) ? g->g.renderedBy(cell, dev(device)) : 0u;
}
void x2_grid_deferred_to_cpu(x2_grid* g, uint32_t tile) {
    if (g && tile < g->g.tileCount()) g->g.deferredToCpu(tile);
}
uint32_t x2_grid_cell_of(const x2_grid* g, uint32_t tile) {
    return g && tile < g->g.tileCount() ? g->g.cellOf(tile) : X2_NO_CELL;
}
uint32_t x2_grid_tile_pixels(const x2_grid* g, uint32_t tile) {
    return g && tile < g->g.tileCount() ? g->g.tilePixels(tile) : 0u;
}

int x2_ownership_create(x2_ownership** out) {
    if (!out) return x2::grid::gridFail(X2_ERR_ARGUMENT, "null ownership out");
    *out = new (std::nothrow) x2_ownership();
    return *out ? 0 : x2::grid::gridFail(X2_ERR_MEMORY, "ownership allocation failed");
}
void x2_ownership_destroy(x2_ownership* o) { delete o; }
int x2_ownership_owners(const x2_ownership* o, uint32_t cells, uint8_t* owners) {
    return o && owners ? int(o->o.owners(cells, owners)) : 0;
}
int x2_ownership_finished(x2_ownership* o, const x2_grid* g) {
    if (!o || !g) return x2::grid::gridFail(X2_ERR_ARGUMENT, "null ownership or grid");
    o->o.finished(g->g);
    return 0;
}

}  // extern "C"
