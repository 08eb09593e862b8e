// A concurrent-claim driver for the work grid (x2_grid.cpp), built with the thread sanitizer by
// tests/test_grid_tsan.py on the CPU. CPU workers claim one chunk at a time on their own threads, following the
// cellular rules (seed, choose, claim), while a GPU claimant takes 97 at a time and a scorer updates the cells'
// importance; in a second grid a coordinator appends chunks while workers claim. Every chunk must be claimed exactly
// once; the sanitizer reports any data race. Prints "ok" and exits 0 when the checks hold.
#include <atomic>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "../../xpu2-foundation_amd/csrc/x2_grid.h"

namespace x2 {
int reportFailure(int code, const std::string&) { return code; }  // the device library's error text, not needed here
}

using namespace x2::grid;

static int check(bool ok, const char* what) {
    if (!ok) std::printf("failed: %s\n", what);
    return ok ? 0 : 1;
}

static int frameOnce() {
    x2_grid_options o{32, 4, 2, 1, 0, 0};
    const uint32_t w = 400, h = 232, spp = 24;
    WorkGrid g(w, h, spp, o, nullptr, 0, false, nullptr);
    const uint32_t workers = 6;
    std::vector<std::vector<x2_chunk>> got(workers + 1);
    std::atomic<bool> done{false};
    std::vector<std::thread> threads;
    for (uint32_t k = 0; k < workers; ++k)
        threads.emplace_back([&, k] {
            uint32_t cell = g.seed(Cpu, k, workers);
            for (;;) {
                cell = g.choose(Cpu, cell, k);
                if (cell == X2_NO_CELL) {
                    if (g.empty()) break;
                    std::this_thread::yield();
                    continue;
                }
                x2_chunk c;
                if (g.claim(cell, 1, &c, Cpu)) got[k].push_back(c);
            }
        });
    std::thread scorer([&] {  // importance changes while the workers claim
        std::vector<float> s(g.cellCount());
        for (int round = 0; round < 50 && !done.load(); ++round) {
            for (uint32_t i = 0; i < s.size(); ++i) s[i] = float((i * 7 + round) % 11);
            g.setScores(s.data(), uint32_t(s.size()));
            g.setRates(1e9, 1e6, workers, 0.01);
            std::this_thread::yield();
        }
    });
    std::vector<uint32_t> cells(8);
    for (uint32_t k = 0; k < 8; ++k) cells[k] = g.seed(Gpu, k, 8);
    std::vector<x2_chunk> buf(97);
    while (!g.empty()) {
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t c = g.choose(Gpu, cells[k], k);
            if (c == X2_NO_CELL) continue;
            cells[k] = c;
            const uint32_t n = g.claim(c, 97, buf.data(), Gpu);
            got[workers].insert(got[workers].end(), buf.begin(), buf.begin() + n);
            g.gpuClaimed();
        }
    }
    done = true;
    for (auto& t : threads) t.join();
    scorer.join();
    std::vector<uint32_t> samples(size_t(w) * h, 0);
    std::vector<uint8_t> seen(size_t(g.tileCount()) * spp, 0);
    uint64_t n = 0;
    int bad = 0;
    const uint32_t tx = (w + 7) / 8;
    for (const auto& list : got)
        for (const x2_chunk& c : list) {
            ++n;
            for (uint32_t s = c.first; s < c.first + c.count; ++s) bad += check(!seen[size_t(c.tile) * spp + s]++, "twice");
            for (uint32_t y = (c.tile / tx) * 8; y < std::min(h, (c.tile / tx) * 8 + 8); ++y)
                for (uint32_t x = (c.tile % tx) * 8; x < std::min(w, (c.tile % tx) * 8 + 8); ++x) samples[y * w + x] += c.count;
        }
    bad += check(n == g.totalChunks(), "every chunk claimed");
    for (uint32_t v : samples) bad += v == spp ? 0 : check(false, "every pixel spp samples");
    return bad;
}

static int appendWhileClaiming() {
    x2_grid_options o{64, 4, 2, 1, 0, 0};
    WorkGrid g(256, 256, 64, o, nullptr, 0, true, nullptr);
    const uint32_t perCell = 64 * 16;  // every tile of a cell, 16 sample ranges of 4
    std::atomic<bool> appended{false};
    std::atomic<uint64_t> claimed{0};
    std::vector<std::thread> threads;
    for (uint32_t k = 0; k < 4; ++k)
        threads.emplace_back([&, k] {
            x2_chunk buf[7];
            uint32_t cell = k;
            for (;;) {
                const bool finished = appended.load();
                cell = g.choose(k % 2 ? Gpu : Cpu, cell, k);
                if (cell == X2_NO_CELL) {
                    if (finished && g.empty()) break;
                    std::this_thread::yield();
                    continue;
                }
                claimed += g.claim(cell, 7, buf, k % 2 ? Gpu : Cpu);
            }
        });
    for (uint32_t r = 0; r < 16; ++r)
        for (uint32_t c = 0; c < g.cellCount(); ++c) {
            std::vector<x2_chunk> add;
            const uint32_t cx = c % 4, cy = c / 4;
            for (uint32_t ty = cy * 8; ty < cy * 8 + 8; ++ty)
                for (uint32_t tx = cx * 8; tx < cx * 8 + 8; ++tx) add.push_back({ty * 32 + tx, r * 4, 4, c});
            g.append(c, add.data(), uint32_t(add.size()));
        }
    appended = true;
    for (auto& t : threads) t.join();
    return check(claimed.load() == uint64_t(perCell) * g.cellCount() && claimed.load() == g.totalChunks(),
                 "appended chunks claimed once");
}

int main() {
    int bad = 0;
    for (int i = 0; i < 3; ++i) bad += frameOnce();
    bad += appendWhileClaiming();
    if (!bad) std::printf("ok\n");
    return bad ? 1 : 0;
}
