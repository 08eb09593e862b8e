// x2_files.cpp — the on-disk formats either side of a frame, host side: the film save a checkpoint render writes and
// a render taken up reads (StagedFilm::save/load, xpu2/interactive/xpu2_session.cc:1347-1475, "X2RS"), the staged
// checkpoint render itself (stagedFrame, xpu2_session.cc:1540-1640: the frame in stages of samples, the film saved
// after each, taken up from its save), and the beauty output as a scanline OpenEXR file (RenderOutputWriter through
// OIIO, aov_write.h:176-255; the layout is the published OpenEXR file format, written here directly).
#include <stdint.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/x2dev.h"
#include "x2_half.h"

namespace x2 {
int reportFailure(int code, const std::string& what);  // sets x2dev_last_error's text (x2_lane.hip)
}

namespace {
using x2::reportFailure;
constexpr uint32_t FilmMagic = 0x53523258u;  // "X2RS"
constexpr uint32_t FilmVersion = 1;

// Writes whole or not at all: into path.part, then renamed over path.
struct PartFile {
    std::string path, part;
    FILE* f = nullptr;
    explicit PartFile(const char* p) : path(p), part(std::string(p) + ".part") { f = std::fopen(part.c_str(), "wb"); }
    ~PartFile() {
        if (f) {
            std::fclose(f);
            std::remove(part.c_str());
        }
    }
    bool put(const void* p, size_t n) {
        if (!n || !p) return n == 0;
        return std::fwrite(p, 1, n, f) == n;
    }
    int finish() {
        const bool ok = std::fflush(f) == 0;
        std::fclose(f);
        f = nullptr;
        if (!ok) {
            std::remove(part.c_str());
            return reportFailure(X2_ERR_FILE, "cannot write " + part);
        }
        if (std::rename(part.c_str(), path.c_str()) != 0) {
            std::remove(part.c_str());
            return reportFailure(X2_ERR_FILE, "cannot name " + path);
        }
        return 0;
    }
};

using x2::toHalf;

void put32(std::vector<uint8_t>& b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back(uint8_t(v >> (8 * i)));
}
void putString(std::vector<uint8_t>& b, const char* s) {
    b.insert(b.end(), s, s + std::strlen(s) + 1);
}
void attribute(std::vector<uint8_t>& b, const char* name, const char* type, const std::vector<uint8_t>& value) {
    putString(b, name);
    putString(b, type);
    put32(b, uint32_t(value.size()));
    b.insert(b.end(), value.begin(), value.end());
}
std::vector<uint8_t> box(int x0, int y0, int x1, int y1) {
    std::vector<uint8_t> v;
    for (int a : {x0, y0, x1, y1}) put32(v, uint32_t(a));
    return v;
}
std::vector<uint8_t> floatBytes(std::initializer_list<float> fs) {
    std::vector<uint8_t> v;
    for (float f : fs) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        put32(v, u);
    }
    return v;
}

// OpenEXR's ZIP block transform: the bytes split into their even and odd halves, then each byte stored as its
// difference from the one before plus 128; deflated with zlib. A block that does not shrink is stored as it is.
std::vector<uint8_t> zipBlock(const std::vector<uint8_t>& raw) {
    const size_t n = raw.size();
    std::vector<uint8_t> t(n);
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) t[(i & 1) ? b++ : a++] = raw[i];
    for (size_t i = n; i-- > 1;) t[i] = uint8_t(int(t[i]) - int(t[i - 1]) + 128);
    uLongf size = compressBound(uLong(n));
    std::vector<uint8_t> out(size);
    if (compress2(out.data(), &size, t.data(), uLong(n), 4) != Z_OK || size >= n) return raw;
    out.resize(size);
    return out;
}
}  // namespace

extern "C" {

int x2_film_save(const char* path, const x2_film_file* film) {
    if (!path || !film || !film->sums || !film->counts || !film->width || !film->height)
        return reportFailure(X2_ERR_ARGUMENT, "null path or film");
    if (film->done > film->spp) return reportFailure(X2_ERR_ARGUMENT, "a film with more samples done than it has");
    PartFile f(path);
    if (!f.f) return reportFailure(X2_ERR_FILE, std::string("cannot write ") + path + ".part");
    const uint64_t pixels = uint64_t(film->width) * film->height;
    const uint32_t head[8] = {FilmMagic, FilmVersion, film->width, film->height, film->spp, film->done, 0u, 0u};
    bool ok = f.put(head, sizeof head);
    auto array = [&](const void* p, uint64_t n, size_t elem) {
        ok = ok && f.put(&n, sizeof n) && f.put(p, size_t(n) * elem);
    };
    array(film->sums, pixels * 4, sizeof(double));
    array(film->counts, pixels, sizeof(uint32_t));
    array(film->odd, film->odd ? pixels * 4 : 0, sizeof(double));
    array(nullptr, 0, sizeof(double));  // AOV accumulators: none on this device
    array(nullptr, 0, sizeof(float));   // closest-filter depths: none
    const uint32_t categories = 0;      // Cryptomatte categories: none
    ok = ok && f.put(&categories, sizeof categories);
    if (!ok) return reportFailure(X2_ERR_FILE, std::string("cannot write ") + path + ".part");
    return f.finish();
}

int x2_film_load(const char* path, uint32_t width, uint32_t height, uint32_t spp, double* sums, uint32_t* counts,
                 double* odd, uint32_t* done, uint32_t* has_odd) {
    if (!path || !sums || !counts || !done) return reportFailure(X2_ERR_ARGUMENT, "null path, film or done");
    FILE* f = std::fopen(path, "rb");
    if (!f) return reportFailure(X2_ERR_FILE, std::string("no saved render at ") + path);
    struct Close {
        FILE* f;
        ~Close() { std::fclose(f); }
    } close{f};
    uint32_t head[8];
    if (std::fread(head, sizeof head, 1, f) != 1 || head[0] != FilmMagic || head[1] != FilmVersion)
        return reportFailure(X2_ERR_FILE, std::string(path) + " is not an XPU2 save");
    if (head[2] != width || head[3] != height || head[4] != spp || head[6] || head[7] || head[5] > head[4])
        return reportFailure(X2_ERR_FILE, std::string(path) + " is of another render (its size, samples or outputs differ)");
    const uint64_t pixels = uint64_t(width) * height;
    // Each array: its length, then its elements; read into a copy first, so a file cut short leaves the film as it was.
    const size_t np = size_t(pixels);
    std::vector<double> s4(np * 4), o4;
    std::vector<uint32_t> c1(np);
    auto length = [&](uint64_t& n) { return std::fread(&n, sizeof n, 1, f) == 1; };
    uint64_t n = 0;
    bool bad = !length(n) || n != pixels * 4 || std::fread(s4.data(), sizeof(double), s4.size(), f) != s4.size();
    bad = bad || !length(n) || n != pixels || std::fread(c1.data(), sizeof(uint32_t), c1.size(), f) != c1.size();
    if (!bad) {
        bad = !length(n) || (n != 0 && n != pixels * 4);
        if (!bad && n) {
            o4.resize(size_t(n));
            bad = std::fread(o4.data(), sizeof(double), o4.size(), f) != o4.size();
        }
    }
    bad = bad || !length(n) || n != 0 || !length(n) || n != 0;  // no AOV accumulators, no closest-filter depths
    uint32_t categories = 1;
    bad = bad || std::fread(&categories, sizeof categories, 1, f) != 1 || categories != 0;
    if (bad) return reportFailure(X2_ERR_FILE, std::string(path) + " is cut short or of another render's outputs");
    std::memcpy(sums, s4.data(), s4.size() * sizeof(double));
    std::memcpy(counts, c1.data(), c1.size() * sizeof(uint32_t));
    if (odd && !o4.empty()) std::memcpy(odd, o4.data(), o4.size() * sizeof(double));
    *done = head[5];
    if (has_odd) *has_odd = o4.empty() ? 0u : 1u;
    return 0;
}

int x2_write_exr(const char* path, const x2_exr_image* im) {
    if (!path || !im || !im->sums || !im->counts || !im->width || !im->height)
        return reportFailure(X2_ERR_ARGUMENT, "null path or image");
    if (im->compression != X2_EXR_NONE && im->compression != X2_EXR_ZIPS && im->compression != X2_EXR_ZIP)
        return reportFailure(X2_ERR_ARGUMENT, "compression: none, zips or zip");
    const int W = int(im->width), H = int(im->height);
    int r[4] = {0, 0, W, H};
    if (im->region[2] > im->region[0] && im->region[3] > im->region[1])
        for (int i = 0; i < 4; ++i) r[i] = im->region[i];
    r[0] = std::max(r[0], 0), r[1] = std::max(r[1], 0), r[2] = std::min(r[2], W), r[3] = std::min(r[3], H);
    if (r[2] <= r[0] || r[3] <= r[1]) return reportFailure(X2_ERR_ARGUMENT, "an empty region");
    // The film's rows run from the bottom; the file's from the top (aov_write.h:238).
    const int x0 = r[0], x1 = r[2] - 1, y0 = H - r[3], y1 = H - 1 - r[1];
    const int lines = y1 - y0 + 1;
    struct Channel {
        const char* name;
        int index;  // 0..3 the beauty's r, g, b, alpha; 4 the weight
    };
    std::vector<Channel> channels;  // the header's order: sorted by name
    channels.push_back({"A", 3});
    channels.push_back({"B", 2});
    channels.push_back({"G", 1});
    channels.push_back({"R", 0});
    if (im->weight) channels.push_back({"weight", 4});
    const bool half = im->half != 0;
    std::vector<uint8_t> h;
    put32(h, 20000630u);  // the OpenEXR magic number
    put32(h, 2u);         // version 2, single-part scanline file, short names
    {
        std::vector<uint8_t> v;
        for (const Channel& c : channels) {
            putString(v, c.name);
            const bool whole = c.index == 4;  // the weight is a count: FLOAT whatever the beauty's format
            put32(v, whole || !half ? 2u : 1u);
            put32(v, 0u);  // pLinear and reserved
            put32(v, 1u);  // x sampling
            put32(v, 1u);  // y sampling
        }
        v.push_back(0);
        attribute(h, "channels", "chlist", v);
    }
    attribute(h, "compression", "compression", {uint8_t(im->compression)});
    attribute(h, "dataWindow", "box2i", box(x0, y0, x1, y1));
    attribute(h, "displayWindow", "box2i", box(0, 0, W - 1, H - 1));
    attribute(h, "lineOrder", "lineOrder", {0});
    attribute(h, "pixelAspectRatio", "float", floatBytes({1.0f}));
    attribute(h, "screenWindowCenter", "v2f", floatBytes({0.0f, 0.0f}));
    attribute(h, "screenWindowWidth", "float", floatBytes({1.0f}));
    h.push_back(0);  // the header's end

    const int block = im->compression == X2_EXR_ZIP ? 16 : 1;
    const int chunks = (lines + block - 1) / block;
    std::vector<std::vector<uint8_t>> data(static_cast<size_t>(chunks));
    std::vector<uint8_t> raw;
    for (int c = 0; c < chunks; ++c) {
        const int first = y0 + c * block, last = std::min(y1, first + block - 1);
        raw.clear();
        for (int y = first; y <= last; ++y) {
            const size_t row = size_t(H - 1 - y) * size_t(W);
            for (const Channel& ch : channels)
                for (int x = x0; x <= x1; ++x) {
                    const size_t p = row + size_t(x);
                    const uint32_t n = im->counts[p];
                    float v = 0.0f;
                    if (ch.index == 4) v = float(n);
                    else if (n) v = float(im->sums[p * 4 + size_t(ch.index)] / n);
                    if (half && ch.index != 4) {
                        const uint16_t b = toHalf(v);
                        raw.push_back(uint8_t(b)), raw.push_back(uint8_t(b >> 8));
                    } else {
                        uint32_t u;
                        std::memcpy(&u, &v, 4);
                        put32(raw, u);
                    }
                }
        }
        std::vector<uint8_t>& out = data[size_t(c)];
        const std::vector<uint8_t> payload = im->compression == X2_EXR_NONE ? raw : zipBlock(raw);
        put32(out, uint32_t(first));
        put32(out, uint32_t(payload.size()));
        out.insert(out.end(), payload.begin(), payload.end());
    }
    PartFile f(path);
    if (!f.f) return reportFailure(X2_ERR_FILE, std::string("cannot write ") + path + ".part");
    bool ok = f.put(h.data(), h.size());
    uint64_t at = h.size() + uint64_t(chunks) * 8;
    for (int c = 0; c < chunks && ok; ++c) {  // the offset table: each chunk's place in the file
        ok = f.put(&at, 8);
        at += data[size_t(c)].size();
    }
    for (int c = 0; c < chunks && ok; ++c) ok = f.put(data[size_t(c)].data(), data[size_t(c)].size());
    if (!ok) return reportFailure(X2_ERR_FILE, std::string("cannot write ") + path + ".part");
    return f.finish();
}

int x2dev_render_staged(x2dev* d, const x2_config* config, uint32_t samples_per_chunk, const x2_staged* s,
                        x2_staged_frame* out) {
    if (!d || !config || !s || !out) return reportFailure(X2_ERR_ARGUMENT, "null device, config, stages or frame");
    std::memset(out, 0, sizeof *out);
    const uint32_t spp = config->spp, spc = samples_per_chunk ? samples_per_chunk : 4u;
    if (!spp) return reportFailure(X2_ERR_ARGUMENT, "a frame of no samples");
    const uint32_t step = s->stage_samples ? s->stage_samples : std::max(1u, (spp + 15) / 16);
    const uint32_t W = config->width, H = config->height, tilesX = (W + 7) / 8, tilesY = (H + 7) / 8;
    const size_t pixels = size_t(W) * H;
    const auto start = std::chrono::steady_clock::now();
    int r = x2dev_begin(d, config, spc);
    if (r < 0) return r;
    uint32_t done = 0;
    if (s->resume && s->path) {
        std::vector<double> sums(pixels * 4);
        std::vector<uint32_t> counts(pixels);
        uint32_t saved = 0;
        if (x2_film_load(s->path, W, H, spp, sums.data(), counts.data(), nullptr, &saved, nullptr) == 0) {
            if ((r = x2dev_seed_film(d, sums.data(), nullptr, counts.data())) < 0) return r;
            done = saved;
            out->resumed_from = done;
            out->resumed = 1;
        }  // else the render begins anew (the reference says why and goes on)
    }
    // Cell by cell (64-pixel cells), a stage's sample ranges in turn, as the frame loop queues them.
    std::vector<x2_chunk> chunks;
    x2_completion c;
    auto collect = [&]() {
        const int q = x2dev_poll(d, &c);
        if (q == 1) out->errors += c.errors, out->device_seconds += c.device_seconds;
        return q;
    };
    while (done < spp) {
        if (s->max_stages && out->stages >= s->max_stages) break;
        const uint32_t first = done, end = std::min(spp, done + step);
        chunks.clear();
        for (uint32_t cy = 0; cy * 8 < tilesY; ++cy)
            for (uint32_t cx = 0; cx * 8 < tilesX; ++cx)
                for (uint32_t f = first; f < end; f += spc)
                    for (uint32_t ty = cy * 8; ty < std::min(tilesY, cy * 8 + 8); ++ty)
                        for (uint32_t tx = cx * 8; tx < std::min(tilesX, cx * 8 + 8); ++tx)
                            chunks.push_back({ty * tilesX + tx, f, std::min(spc, end - f), cy * ((tilesX + 7) / 8) + cx});
        for (size_t at = 0; at < chunks.size();) {
            uint32_t room = 0;
            while (!x2dev_ready(d) || !(room = x2dev_chunk_room(d))) {  // room: the ring, and parked samples' ring
                if ((r = collect()) < 0) return r;
                if (!x2dev_in_flight(d) && !room && x2dev_ready(d) && !(room = x2dev_chunk_room(d)))
                    return reportFailure(X2_ERR_STATE, "no room for a launch with nothing in flight");
            }
            const uint32_t n = uint32_t(std::min<size_t>(room, chunks.size() - at));
            if ((r = x2dev_launch(d, chunks.data() + at, n, 0)) < 0) return r;
            ++out->launches;
            at += n;
        }
        // Paged textures: the stage's parked samples run again until none is left.
        for (uint32_t idle = 0; x2dev_in_flight(d) || x2dev_parked_outstanding(d);) {
            if ((r = collect()) < 0) return r;
            if (x2dev_ready(d) && x2dev_parked_outstanding(d)) {
                if (++idle > 100000) return reportFailure(X2_ERR_STATE, "parked samples do not finish");
                if ((r = x2dev_launch(d, nullptr, 0, 0)) < 0) return r;
                ++out->launches;
            }
        }
        done = end;
        ++out->stages;
        if (s->path) {  // the film so far, saved whole before the next stage
            if ((r = x2dev_download_film(d)) < 0) return r;
            while ((r = x2dev_film_ready(d)) == 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
            if (r < 0) return r;
            const double* sums;
            const uint32_t* counts;
            if ((r = x2dev_film(d, &sums, &counts)) < 0) return r;
            x2_film_file film{W, H, spp, done, sums, counts, nullptr};
            if ((r = x2_film_save(s->path, &film)) < 0) return r;
            ++out->saves;
        }
    }
    if ((r = x2dev_download_film(d)) < 0) return r;
    while ((r = x2dev_film_ready(d)) == 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
    if (r < 0) return r;
    out->done = done;
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return 0;
}

}  // extern "C"
