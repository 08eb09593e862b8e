// x2_grid.cpp — the work grid, the cell owners across frames, and the broker's frame that feeds the GPU lane and the
// CPU lane from the grid (async_grid.h:38-439, async_render.h:590-830,1668-1718). Host code: the broker reaches the
// device only through the lane's C-ABI (x2dev_begin / ready / launch / poll / download_film).
#include "x2_grid.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>

namespace x2 {
int reportFailure(int code, const std::string& what);  // sets x2dev_last_error's text (x2_lane.hip)
}

namespace x2::grid {
namespace {
constexpr uint32_t TileSize = 8, TilePixels = 64;

int64_t nowNs() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
// Spreads workers of a device over cells of equal merit: an integer mix of (cell, salt), the same as the reference's
// tie-break so that both pick the same cell (async_grid.h:424-430).
uint32_t spread(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9e3779b1u ^ (b + 0x7f4a7c15u + (a << 6) + (a >> 2));
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    return h;
}
}  // namespace

CellQueue::CellQueue() : blocks_(new std::atomic<x2_chunk*>[MaxBlocks]) {
    for (uint32_t i = 0; i < MaxBlocks; ++i) blocks_[i].store(nullptr, std::memory_order_relaxed);
}
CellQueue::~CellQueue() {
    for (uint32_t i = 0; i < MaxBlocks; ++i) delete[] blocks_[i].load(std::memory_order_relaxed);
}
bool CellQueue::append(const x2_chunk* chunks, uint32_t n) {
    uint32_t at = published_.load(std::memory_order_relaxed);
    if (uint64_t(at) + n > uint64_t(BlockChunks) * MaxBlocks) return false;
    for (uint32_t i = 0; i < n; ++i, ++at) {
        std::atomic<x2_chunk*>& slot = blocks_[at / BlockChunks];
        x2_chunk* block = slot.load(std::memory_order_relaxed);
        if (!block) {
            block = new x2_chunk[BlockChunks];
            slot.store(block, std::memory_order_relaxed);
        }
        block[at % BlockChunks] = chunks[i];
    }
    published_.store(at, std::memory_order_release);  // the chunks before the count
    return true;
}
uint32_t CellQueue::claim(uint32_t n, x2_chunk* out, uint32_t limit) {
    const uint32_t end = std::min(published(), limit);
    uint32_t start = next_.load(std::memory_order_relaxed);
    do {
        if (start >= end) return 0;
    } while (!next_.compare_exchange_weak(start, std::min(end, start + n), std::memory_order_acq_rel,
                                          std::memory_order_relaxed));
    const uint32_t got = std::min(end, start + n) - start;
    for (uint32_t i = 0; i < got; ++i)
        out[i] = blocks_[(start + i) / BlockChunks].load(std::memory_order_relaxed)[(start + i) % BlockChunks];
    return got;
}

WorkGrid::WorkGrid(uint32_t width, uint32_t height, uint32_t spp, const x2_grid_options& options, const uint8_t* owners,
                   uint32_t ownerCount, bool empty, const int32_t* viewport)
    : width_(width), height_(height), spp_(spp), options_(options) {
    if (!width || !height || !spp || !options.cell_size || options.cell_size % TileSize || !options.samples_per_chunk ||
        options.mode < 0 || options.mode > 2)
        throw std::invalid_argument("invalid work grid (sizes, cell_size a multiple of 8, mode 0..2)");
    if (viewport) std::memcpy(viewport_, viewport, sizeof viewport_);
    cellsX_ = (width + options.cell_size - 1) / options.cell_size;
    cellsY_ = (height + options.cell_size - 1) / options.cell_size;
    tilesX_ = (width + TileSize - 1) / TileSize;
    tilesY_ = (height + TileSize - 1) / TileSize;
    const uint32_t per = options.cell_size / TileSize;
    cells_ = std::vector<Cell>(size_t(cellsX_) * cellsY_);
    if (owners && ownerCount != cells_.size()) owners = nullptr;
    if (!empty) {
        // The progressive order's ranges grow 1, 1, 2, then the chunk size (never above it: a chunk holds at most
        // samples_per_chunk indices, which sizes the device's items per chunk).
        const uint32_t from = options.sample_count ? options.sample_first : 0u;
        const uint32_t to = options.sample_count ? std::min(spp, options.sample_first + options.sample_count) : spp;
        const uint32_t spc = options.samples_per_chunk;
        for (uint32_t first = from; first < to;) {
            const uint32_t done = first - from;
            const uint32_t size = !options.progressive ? spc : std::min(spc, done < 2 ? 1u : done < 4 ? 2u : spc);
            ranges_.push_back({first, std::min(size, to - first)});
            first += ranges_.back().count;
        }
    }
    std::vector<x2_chunk> chunks;
    for (uint32_t cy = 0; cy < cellsY_; ++cy)
        for (uint32_t cx = 0; cx < cellsX_; ++cx) {
            const uint32_t index = cy * cellsX_ + cx;
            Cell& cell = cells_[index];
            cell.owner = options.mode == 0 ? Cpu
                         : options.mode == 1 ? Gpu
                         : owners            ? (owners[index] ? Gpu : Cpu)
                                             : ((cx + cy) % 2 == 0 ? Gpu : Cpu);
            if (empty) continue;
            chunks.clear();
            for (const Range& r : ranges_)
                for (uint32_t ty = cy * per; ty < std::min(tilesY_, (cy + 1) * per); ++ty)
                    for (uint32_t tx = cx * per; tx < std::min(tilesX_, (cx + 1) * per); ++tx)
                        if (tilePixels(ty * tilesX_ + tx)) chunks.push_back({ty * tilesX_ + tx, r.first, r.count, index});
            cell.tiles = ranges_.empty() ? 0u : uint32_t(chunks.size() / ranges_.size());
            if (!cell.queue.append(chunks.data(), uint32_t(chunks.size())))
                throw std::length_error("a cell's chunk queue is full");
            total_ += chunks.size();
        }
    refreshFrontiers();
}

uint64_t WorkGrid::remainingChunks() const {
    uint64_t n = 0;
    for (uint32_t i = 0; i < cellCount(); ++i) n += remaining(i);
    return n;
}
bool WorkGrid::empty() const {
    for (uint32_t i = 0; i < cellCount(); ++i)
        if (remaining(i)) return false;
    return true;
}
bool WorkGrid::append(uint32_t cell, const x2_chunk* chunks, uint32_t n) {
    if (!cells_[cell].queue.append(chunks, n)) return false;
    total_ += n;
    return true;
}
uint32_t WorkGrid::cellOf(uint32_t tile) const {
    const uint32_t per = options_.cell_size / TileSize;
    return ((tile / tilesX_) / per) * cellsX_ + (tile % tilesX_) / per;
}
uint32_t WorkGrid::tilePixels(uint32_t tile) const {
    const int x = int((tile % tilesX_) * TileSize), y = int((tile / tilesX_) * TileSize);
    const int x0 = std::max(x, int(viewport_[0])), y0 = std::max(y, int(viewport_[1]));
    const int x1 = std::min({x + int(TileSize), int(width_), int(viewport_[2])});
    const int y1 = std::min({y + int(TileSize), int(height_), int(viewport_[3])});
    return x1 > x0 && y1 > y0 ? uint32_t(x1 - x0) * uint32_t(y1 - y0) : 0u;
}

// A device's claims in a cell reach the ranges up to its frontier, one more in an important cell
// (async_grid.h:372-377).
uint32_t WorkGrid::limit(uint32_t cell, Device device) const {
    if (ranges_.empty()) return ~0u;
    const uint64_t lead = scored() && important(cell) ? 1 : 0;
    const uint64_t upTo = std::min<uint64_t>(uint64_t(frontier(device)) + lead + 1, ranges_.size());
    return uint32_t(std::min<uint64_t>(upTo * cells_[cell].tiles, ~0u));
}
uint32_t WorkGrid::claim(uint32_t cell, uint32_t n, x2_chunk* out, Device device) {
    if (closed()) return 0;
    const uint32_t got = cells_[cell].queue.claim(n, out, limit(cell, device));
    if (got && !ranges_.empty()) refreshFrontiers();
    return got;
}
uint32_t WorkGrid::claimable(uint32_t cell, Device device) const {
    if (closed()) return 0;
    const CellQueue& q = cells_[cell].queue;
    const uint32_t end = std::min(q.published(), limit(cell, device)), at = q.taken();
    return at >= end ? 0 : end - at;
}
// Each device's frontier: the least range still to hand out among its own reserved cells and the open ones (where
// none is left there, among all) (async_grid.h:380-395). Computed and published under one lock, so a refresh from an
// older state never lands after a newer one.
void WorkGrid::refreshFrontiers() {
    if (ranges_.empty()) return;
    std::lock_guard<std::mutex> hold(frontierLock_);
    uint32_t least[3] = {~0u, ~0u, ~0u};
    for (uint32_t i = 0; i < cellCount(); ++i) {
        const Cell& c = cells_[i];
        if (!c.tiles || !c.queue.remaining()) continue;
        const uint32_t at = c.queue.taken() / c.tiles;
        const bool open = !reserved(i);
        for (uint32_t dev = 0; dev < 2; ++dev)
            if (open || uint32_t(c.owner) == dev) least[dev] = std::min(least[dev], at);
        least[2] = std::min(least[2], at);
    }
    for (uint32_t dev = 0; dev < 2; ++dev)
        frontier_[dev].store(std::min<uint32_t>(least[dev] != ~0u ? least[dev] : least[2], uint32_t(ranges_.size())),
                             std::memory_order_release);
}
uint64_t WorkGrid::reservedChunks(Device device) const {
    uint64_t n = 0;
    for (uint32_t i = 0; i < cellCount(); ++i)
        if (owner(i) == device && reserved(i)) n += remaining(i);
    return n;
}

void WorkGrid::setScores(const float* scores, uint32_t n) {
    if (n != cellCount()) return;
    double sum = 0;
    uint32_t known = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (scores[i] >= 0) {
            sum += scores[i];
            ++known;
        }
    if (!known) return;
    const float mean = float(sum / known);
    for (uint32_t i = 0; i < n; ++i) {
        cells_[i].score.store(std::max(scores[i], 0.0f), std::memory_order_relaxed);
        cells_[i].important.store(scores[i] > mean, std::memory_order_relaxed);
    }
    scored_.store(true, std::memory_order_release);
    refreshFrontiers();
}
void WorkGrid::setSuitability(const float* p, uint32_t n) {
    if (n != cellCount()) return;
    for (uint32_t i = 0; i < n; ++i) cells_[i].gpu = p[i];
}
void WorkGrid::setRates(double gpuSamples, double cpuWorkerSamples, uint32_t cpuWorkers, double windowSeconds) {
    gpuRate_.store(gpuSamples, std::memory_order_relaxed);
    cpuRate_.store(cpuWorkerSamples, std::memory_order_relaxed);
    cpuWorkers_.store(cpuWorkers, std::memory_order_relaxed);
    window_.store(windowSeconds, std::memory_order_relaxed);
}
void WorkGrid::gpuClaimed() { gpuClaim_.store(nowNs(), std::memory_order_relaxed); }

// Whether a device may work a cell the other has reserved (async_grid.h:279-296). Before both rates are known, yes.
// The CPU may when the GPU has not claimed for a while (its launches are slow to come back), or when one CPU chunk
// ends well before the GPU would get through its reserved work; the GPU may when the CPU's reserved work outlasts a
// few of its launches, in a cell it is suited to.
bool WorkGrid::passage(Device device, uint32_t cell) const {
    const double gpu = gpuRate_.load(std::memory_order_relaxed), cpu = cpuRate_.load(std::memory_order_relaxed);
    if (gpu <= 0 || cpu <= 0) return true;
    const double chunk = double(options_.samples_per_chunk) * TilePixels;
    const double window = window_.load(std::memory_order_relaxed);
    if (device == Cpu) {
        const int64_t last = gpuClaim_.load(std::memory_order_relaxed);
        if (last && double(nowNs() - last) * 1e-9 > std::max(4 * window, 0.05)) return true;
        return chunk / cpu < 0.5 * double(reservedChunks(Gpu)) * chunk / gpu;
    }
    if (cells_[cell].gpu < 0.2f) return false;
    const uint32_t workers = std::max(1u, cpuWorkers_.load(std::memory_order_relaxed));
    return double(reservedChunks(Cpu)) * chunk / (cpu * workers) > 4 * window;
}

void WorkGrid::deferredToCpu(uint32_t tile) {
    Cell& c = cells_[cellOf(tile)];
    uint64_t gpu = c.samples[Gpu].load(std::memory_order_relaxed);
    while (gpu && !c.samples[Gpu].compare_exchange_weak(gpu, gpu - 1, std::memory_order_relaxed)) {
    }
    c.samples[Cpu].fetch_add(1, std::memory_order_relaxed);
}

template <class F>
void WorkGrid::neighbours(uint32_t cell, F&& f) const {
    const int x = int(cell % cellsX_), y = int(cell / cellsX_);
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int nx = x + dx, ny = y + dy;
            if ((dx || dy) && nx >= 0 && ny >= 0 && nx < int(cellsX_) && ny < int(cellsY_))
                f(uint32_t(ny) * cellsX_ + uint32_t(nx));
        }
}

// A worker's next cell (async_grid.h:325-359): the one it is in while it may claim there; else, tier by tier — its
// device's reserved cells (most unclaimed work first), the open ones (most important first), the other device's
// reserved ones by a right of passage — a neighbour of its cell before the best anywhere.
uint32_t WorkGrid::choose(Device device, uint32_t current, uint32_t salt) const {
    auto tier = [&](uint32_t i) { return !reserved(i) ? 1 : owner(i) == device ? 0 : 2; };
    int cpuPassage = -1;  // the CPU's right of passage is the same in every cell: asked once a call
    auto passes = [&](uint32_t i) {
        if (device == Gpu) return passage(device, i);
        if (cpuPassage < 0) cpuPassage = passage(device, i) ? 1 : 0;
        return cpuPassage == 1;
    };
    const uint32_t n = cellCount();
    if (current < n && claimable(current, device) && (tier(current) != 2 || passes(current))) return current;
    for (int t = 0; t < 3; ++t) {
        uint32_t best = X2_NO_CELL;
        float bestKey = 0;
        auto consider = [&](uint32_t i) {
            if (tier(i) != t) return;
            const uint32_t work = claimable(i, device);
            if (!work || (t == 2 && !passes(i))) return;
            const float key = t == 1 ? cells_[i].score.load(std::memory_order_relaxed) + 1e-9f * float(work) : float(work);
            if (best == X2_NO_CELL || key > bestKey || (key == bestKey && spread(i, salt) > spread(best, salt))) {
                best = i;
                bestKey = key;
            }
        };
        if (current < n) {
            neighbours(current, consider);
            if (best != X2_NO_CELL) return best;
        }
        for (uint32_t i = 0; i < n; ++i) consider(i);
        if (best != X2_NO_CELL) return best;
    }
    return X2_NO_CELL;
}
uint32_t WorkGrid::seed(Device device, uint32_t worker, uint32_t workers) const {
    std::vector<uint32_t> own;
    for (uint32_t i = 0; i < cellCount(); ++i)
        if (owner(i) == device) own.push_back(i);
    if (own.empty()) return choose(device, X2_NO_CELL, worker);
    return own[(size_t(worker) * own.size()) / std::max(1u, workers)];
}

bool Ownership::owners(uint32_t cells, uint8_t* out) const {
    if (owners_.size() != cells) return false;
    std::memcpy(out, owners_.data(), cells);
    return true;
}
void Ownership::finished(const WorkGrid& grid) {
    const uint32_t n = grid.cellCount();
    if (owners_.size() != n) {
        owners_.resize(n);
        history_.assign(size_t(n) * Window, 0.0f);
        lastFlip_.assign(n, -1000000);
        for (uint32_t i = 0; i < n; ++i) owners_[i] = uint8_t(grid.owner(i));
    }
    for (uint32_t i = 0; i < n; ++i) {
        const double cpu = double(grid.renderedBy(i, Cpu)), gpu = double(grid.renderedBy(i, Gpu));
        history_[size_t(i) * Window + frame_ % Window] = cpu + gpu > 0 ? float(gpu / (cpu + gpu)) : 0.5f;
    }
    ++frame_;
    if (frame_ < frozenUntil_ || frame_ < int64_t(Window)) return;
    struct Flip {
        uint32_t cell;
        float share;
    };
    std::vector<Flip> flips;
    for (uint32_t i = 0; i < n; ++i) {
        if (frame_ - lastFlip_[i] < 16) continue;
        float gpuShare = 0;
        for (uint32_t k = 0; k < Window; ++k) gpuShare += history_[size_t(i) * Window + k];
        gpuShare /= Window;
        const float other = owners_[i] == Cpu ? gpuShare : 1 - gpuShare;
        if (other > 0.6f) flips.push_back({i, other});
    }
    // Largest share first; equal shares in cell order (a stable sort, so the flips are the same on every build).
    std::stable_sort(flips.begin(), flips.end(), [](const Flip& a, const Flip& b) { return a.share > b.share; });
    const size_t most = std::min<size_t>(8, std::max<size_t>(1, n / 10));
    for (size_t k = 0; k < flips.size() && k < most; ++k) {
        const uint32_t i = flips[k].cell;
        if (frame_ - lastFlip_[i] < 32) {  // flipping back soon after a flip: thrashing
            frozenUntil_ = frame_ + 64;
            return;
        }
        owners_[i] = owners_[i] == Cpu ? uint8_t(Gpu) : uint8_t(Cpu);
        lastFlip_[i] = frame_;
    }
}

namespace {
int gridFail(int code, const std::string& what) { return reportFailure(code, what); }

// The broker's frame (Renderer::render's GPU feeder, async_render.h:590-830, with the CPU lane's workers,
// async_cpu.h:1584-1673).
struct Broker {
    x2dev* d;
    WorkGrid& grid;
    const x2_cpu_lane* cpu;
    double window;
    uint32_t firstChunks, gpuWorkers;
    x2_grid_frame& out;
    std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now();
    std::atomic<uint64_t> cpuSamples{0}, cpuChunks{0}, cpuStolen{0};
    std::atomic<int> failed{0};
    std::atomic<bool> stop{false};
    double gpuRate = 0;
    uint64_t queued = 0;

    double elapsed() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    }
    // One CPU worker: claims a chunk at a time by the cellular rules and renders it.
    void cpuWorker(uint32_t w) {
        uint32_t cell = grid.seed(Cpu, w, cpu->threads);
        while (!stop.load(std::memory_order_relaxed)) {
            const uint32_t next = grid.choose(Cpu, cell, w);
            if (next == X2_NO_CELL) {
                if (grid.empty()) break;
                std::this_thread::sleep_for(std::chrono::microseconds(50));  // the GPU's reserved work, for now
                continue;
            }
            cell = next;
            x2_chunk c;
            if (!grid.claim(cell, 1, &c, Cpu)) continue;
            if (cpu->render(cpu->context, w, &c) != 0) {
                failed.store(1);
                stop.store(true);
                break;
            }
            const uint64_t samples = uint64_t(grid.tilePixels(c.tile)) * c.count;
            grid.rendered(cell, Cpu, samples);
            cpuSamples += samples;
            ++cpuChunks;
            if (grid.owner(cell) != Cpu) ++cpuStolen;
        }
    }
    // How many chunks the next GPU launch takes (launchSize, async_render.h:1668-1718): about one window of the GPU's
    // measured work; near the end, the share that lets both lanes finish together given what the GPU has queued.
    uint32_t launchSize(uint32_t cpuThreads) const {
        const uint64_t remaining = grid.remainingChunks();
        if (!remaining) return 0;
        const auto& ranges = grid.ranges();
        const uint32_t front = grid.frontier(Gpu);
        const double items =
            TilePixels * double(ranges.empty() || front >= ranges.size() ? grid.samplesPerChunk() : ranges[front].count);
        const uint64_t cap = x2dev_chunk_room(d);
        // Never below two samples a launch thread: fewer leave most of the lanes idle and understate the rate.
        const uint64_t least = uint64_t(std::ceil(2.0 * x2dev_launch_threads(d) / items));
        if (gpuRate <= 0) return uint32_t(std::min<uint64_t>({cap, remaining, std::max<uint64_t>(firstChunks, least)}));
        uint64_t want = std::max<uint64_t>(uint64_t(gpuRate * window / items), least);
        if (cpuThreads) {
            const double t = elapsed();
            const double cpuRate = t > 0.01 ? double(cpuSamples.load()) / t : 0.0;
            // The GPU's share L with (queued + L) / gpuRate == (remaining - L) / cpuRate.
            const double share = (double(remaining) * items * gpuRate - double(queued) * cpuRate) / (gpuRate + cpuRate);
            if (share <= 0) return 0;
            want = std::min<uint64_t>(want, uint64_t(share / items) + 1);
        }
        return uint32_t(std::min<uint64_t>({want, cap, remaining}));
    }
    int collect() {
        x2_completion c;
        for (;;) {
            const int r = x2dev_poll(d, &c);
            if (r < 0) return r;
            if (r == 0) return 0;
            queued -= std::min(queued, c.samples);
            if (c.device_seconds > 0 && c.samples >= x2dev_launch_threads(d)) {
                const double rate = double(c.samples) / c.device_seconds;
                gpuRate = gpuRate > 0 ? 0.7 * gpuRate + 0.3 * rate : rate;
            }
            out.errors += c.errors;
            out.rays_trace += c.rays_trace;
            out.rays_shadow += c.rays_shadow;
            out.nodes_visited += c.nodes_visited;
            out.triangles_tested += c.triangles_tested;
        }
    }
    int run() {
        // Mode 1 is the GPU alone; mode 0 the CPU alone (the GPU launches nothing).
        const bool useCpu = grid.mode() != 1 && cpu && cpu->threads && cpu->render;
        const bool useGpu = grid.mode() != 0;
        const uint32_t cpuThreads = useCpu ? cpu->threads : 0u;
        std::vector<std::thread> workers;
        for (uint32_t w = 0; w < cpuThreads; ++w) workers.emplace_back([this, w] { cpuWorker(w); });
        std::vector<uint32_t> gpuCells(gpuWorkers);
        for (uint32_t w = 0; w < gpuWorkers; ++w) gpuCells[w] = grid.seed(Gpu, w, gpuWorkers);
        std::vector<x2_chunk> claim;
        int status = 0;
        for (;;) {
            if ((status = collect()) < 0) break;
            if (failed.load()) {
                status = gridFail(X2_ERR_STATE, "the CPU lane failed a chunk");
                break;
            }
            {
                const double t = elapsed();
                const double perWorker = cpuThreads && t > 0.05 ? double(cpuSamples.load()) / t / cpuThreads : 0.0;
                grid.setRates(gpuRate, perWorker, std::max(1u, cpuThreads), window);
            }
            bool launched = false;
            while (useGpu && x2dev_ready(d)) {
                const uint32_t want = launchSize(cpuThreads);
                if (!want) {
                    // Paged textures: parked samples run again in a launch of their own once no chunk is left.
                    if (grid.empty() && x2dev_parked_outstanding(d)) {
                        const int r = x2dev_launch(d, nullptr, 0, 0);
                        if (r < 0) {
                            status = r;
                            break;
                        }
                        ++out.launches;
                        launched = true;
                    }
                    break;
                }
                claim.resize(want);
                uint32_t got = 0;
                // The GPU's cellular workers take the launch's shares, each from its cell, moving by the rules. A
                // worker moves up to 8 times (the reference's bound), and further while its share of a long launch
                // spans more cells than that: a progressive range holds one chunk a tile, so a share of a
                // window-sized launch crosses many cells.
                const uint32_t share = (want + gpuWorkers - 1) / gpuWorkers;
                const uint32_t moveCap = 8 + share / 32;
                for (uint32_t w = 0; w < gpuWorkers && got < want; ++w)
                    for (uint32_t need = std::min(share, want - got), moves = 0; need && moves < moveCap; ++moves) {
                        const uint32_t cell = grid.choose(Gpu, gpuCells[w], w);
                        if (cell == X2_NO_CELL) break;
                        gpuCells[w] = cell;
                        const uint32_t n = grid.claim(cell, need, claim.data() + got, Gpu);
                        if (n && grid.owner(cell) != Gpu) out.gpu_stolen += n;
                        got += n;
                        need -= n;
                    }
                if (!got) break;
                grid.gpuClaimed();
                uint64_t samples = 0;
                for (uint32_t i = 0; i < got; ++i) {
                    const uint64_t s = uint64_t(grid.tilePixels(claim[i].tile)) * claim[i].count;
                    grid.rendered(claim[i].cell, Gpu, s);
                    samples += s;
                }
                const int r = x2dev_launch(d, claim.data(), got, samples);
                if (r < 0) {
                    status = r;
                    break;
                }
                out.gpu_chunks += got;
                out.gpu_samples += samples;
                ++out.launches;
                queued += samples;
                launched = true;
            }
            if (status < 0) break;
            if (grid.empty() && !x2dev_in_flight(d) && !x2dev_parked_outstanding(d)) break;
            if (!launched) std::this_thread::sleep_for(std::chrono::microseconds(x2dev_in_flight(d) ? 20 : 100));
        }
        if (status < 0) {
            grid.close();
            stop.store(true);
        }
        for (auto& t : workers) t.join();
        if (status < 0) {
            while (x2dev_in_flight(d)) {
                if (collect() < 0) break;
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            return status;
        }
        if (failed.load()) return gridFail(X2_ERR_STATE, "the CPU lane failed a chunk");
        out.cpu_chunks = cpuChunks.load();
        out.cpu_samples = cpuSamples.load();
        out.cpu_stolen = cpuStolen.load();
        out.gpu_rate = gpuRate;
        int r = x2dev_download_film(d);
        if (r < 0) return r;
        while ((r = x2dev_film_ready(d)) == 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
        if (r < 0) return r;
        out.seconds = elapsed();
        out.cpu_rate = out.seconds > 0 ? double(out.cpu_samples) / out.seconds : 0.0;
        return 0;
    }
};
}  // namespace
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

namespace {
bool validDevice(uint32_t device) { return device == X2_DEVICE_CPU || device == X2_DEVICE_GPU; }
x2::grid::Device dev(uint32_t device) { return device == X2_DEVICE_GPU ? x2::grid::Gpu : x2::grid::Cpu; }
}  // namespace

extern "C" {

int x2_grid_create(uint32_t width, uint32_t height, uint32_t spp, const x2_grid_options* options, const uint8_t* owners,
                   uint32_t owner_count, int empty, const int32_t* viewport, x2_grid** out) {
    if (!out) return x2::grid::gridFail(X2_ERR_ARGUMENT, "null grid out");
    *out = nullptr;
    x2_grid_options o = options ? *options : x2_grid_options{64, 4, 2, 1, 0, 0};
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
    return g && cell < g->g.cellCount() ? uint32_t(g->g.owner(cell)) : X2_NO_CELL;
}
uint32_t x2_grid_remaining(const x2_grid* g, uint32_t cell) {
    return g && cell < g->g.cellCount() ? g->g.remaining(cell) : 0u;
}
uint32_t x2_grid_claimable(const x2_grid* g, uint32_t cell, uint32_t device) {
    return g && cell < g->g.cellCount() && validDevice(device) ? g->g.claimable(cell, dev(device)) : 0u;
}
int x2_grid_empty(const x2_grid* g) { return g ? int(g->g.empty()) : 1; }
uint32_t x2_grid_claim(x2_grid* g, uint32_t cell, uint32_t n, x2_chunk* out, uint32_t device) {
    if (!g || !out || cell >= g->g.cellCount() || !validDevice(device)) return 0u;
    return g->g.claim(cell, n, out, dev(device));
}
int x2_grid_append(x2_grid* g, uint32_t cell, const x2_chunk* chunks, uint32_t n) {
    if (!g || (!chunks && n) || cell >= g->g.cellCount()) return x2::grid::gridFail(X2_ERR_ARGUMENT, "bad append");
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
int x2_grid_important(const x2_grid* g, uint32_t cell) {
    return g && cell < g->g.cellCount() ? int(g->g.important(cell)) : 0;
}
int x2_grid_reserved(const x2_grid* g, uint32_t cell) {
    return g && cell < g->g.cellCount() ? int(g->g.reserved(cell)) : 0;
}
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
    return g && validDevice(device) && cell < g->g.cellCount() ? int(g->g.passage(dev(device), cell)) : 0;
}
void x2_grid_close(x2_grid* g) {
    if (g) g->g.close();
}
void x2_grid_rendered(x2_grid* g, uint32_t cell, uint32_t device, uint64_t samples) {
    if (g && cell < g->g.cellCount() && validDevice(device)) g->g.rendered(cell, dev(device), samples);
}
uint64_t x2_grid_rendered_by(const x2_grid* g, uint32_t cell, uint32_t device) {
    return g && cell < g->g.cellCount() && validDevice(device) ? g->g.renderedBy(cell, dev(device)) : 0u;
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

int x2dev_render_grid(x2dev* d, const x2_config* config, x2_grid* grid, const x2_cpu_lane* cpu,
                      const x2_broker_options* options, x2_grid_frame* out) {
    if (!d || !config || !grid || !out) return x2::grid::gridFail(X2_ERR_ARGUMENT, "null device, config, grid or frame");
    if (cpu && cpu->threads && !cpu->render) return x2::grid::gridFail(X2_ERR_ARGUMENT, "a CPU lane without render");
    if (grid->g.mode() == 0 && !(cpu && cpu->threads))
        return x2::grid::gridFail(X2_ERR_ARGUMENT, "a CPU-only grid needs a CPU lane");
    std::memset(out, 0, sizeof *out);
    int r = x2dev_begin(d, config, grid->g.samplesPerChunk());
    if (r < 0) return r;
    const double window = options && options->window_ms > 0 ? options->window_ms / 1000.0 : 0.010;
    const uint32_t first = options && options->first_chunks ? options->first_chunks : 2048u;
    const uint32_t workers = options && options->gpu_workers ? options->gpu_workers : 32u;
    x2::grid::Broker b{d, grid->g, cpu, window, first, workers, *out};
    return b.run();
}

}  // extern "C"
