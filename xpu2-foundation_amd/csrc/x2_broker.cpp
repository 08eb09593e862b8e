// x2_broker.cpp — the broker's frame over the work grid (x2dev_render_grid): the GPU lane fed through the lane's C-ABI
// (x2dev_begin / ready / launch / poll / download_film) in launches sized to its measured rate, beside a CPU lane
// whose workers claim chunks on their own threads and render them through the caller's callback
// (Renderer::render's GPU feeder and the CPU lane's workers, async_render.h:590-830,1668-1718; async_cpu.h:1584-1673).
// Host code only.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "x2_grid.h"

namespace x2::grid {
namespace {
constexpr uint32_t TilePixels = 64;

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
    std::vector<x2_refine> refines;  // host refines: provisional samples taken, to render again

    // A launch with up to a launch's provisional samples ahead of its chunks; those sent leave the list.
    int launchWith(const x2_chunk* chunks, uint32_t count, uint64_t samples) {
        const uint32_t n = uint32_t(std::min<size_t>(refines.size(), X2_REFINES_PER_LAUNCH));
        const int r = x2dev_launch_refines(d, chunks, count, refines.data(), n, samples);
        if (r >= 0 && n) refines.erase(refines.begin(), refines.begin() + n);
        return r;
    }

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
            if (x2dev_refines_waiting(d)) {  // host refines: sent back in the next launches (async_render.h:718-726)
                const size_t at = refines.size();
                refines.resize(at + x2dev_refines_waiting(d));
                uint32_t got = 0;
                const int e = x2dev_take_refines(d, refines.data() + at, uint32_t(refines.size() - at), &got);
                refines.resize(at + got);
                if (e < 0) return e;
            }
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
            if (out.errors) {  // async_render.h:736: the frame stops feeding and fails
                status = gridFail(X2_ERR_STATE, "GPU sample failure; whole epoch must use fallback (async_render.h:736)");
                break;
            }
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
                    // Paged textures: parked and provisional samples run again in a launch of their own once no
                    // chunk is left.
                    if (grid.empty() && (x2dev_parked_outstanding(d) || !refines.empty())) {
                        const int r = launchWith(nullptr, 0, 0);
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
                const int r = launchWith(claim.data(), got, samples);
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
            if (grid.empty() && !x2dev_in_flight(d) && !x2dev_parked_outstanding(d) && refines.empty() &&
                !x2dev_refines_waiting(d))
                break;
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

extern "C" {

int x2dev_render_grid(x2dev* d, const x2_config* config, x2_grid* grid, const x2_cpu_lane* cpu,
                      const x2_broker_options* options, x2_grid_frame* out) {
    if (!d || !config || !grid || !out) return x2::grid::gridFail(X2_ERR_ARGUMENT, "null device, config, grid or frame");
    if (cpu && cpu->threads && !cpu->render) return x2::grid::gridFail(X2_ERR_ARGUMENT, "a CPU lane without render");
    if (gridOf(grid).mode() == 0 && !(cpu && cpu->threads))
        return x2::grid::gridFail(X2_ERR_ARGUMENT, "a CPU-only grid needs a CPU lane");
    std::memset(out, 0, sizeof *out);
    int r = x2dev_begin(d, config, gridOf(grid).samplesPerChunk());
    if (r < 0) return r;
    const double window = options && options->window_ms > 0 ? options->window_ms / 1000.0 : 0.010;
    const uint32_t first = options && options->first_chunks ? options->first_chunks : 2048u;
    const uint32_t workers = options && options->gpu_workers ? options->gpu_workers : 32u;
    x2::grid::Broker b{d, gridOf(grid), cpu, window, first, workers, *out};
    return b.run();
}

}  // extern "C"
