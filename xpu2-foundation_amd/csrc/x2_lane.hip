// x2_lane.hip — the device's host side behind the C-ABI (include/x2dev.h): the counterpart of the reference's
// GpuLane (xpu2/interactive/async_gpu.h:82) over the HIP runtime. One stream; a mailbox ring of chunks in device
// memory fed from pinned staging; at most two launches in flight, each bracketed by HIP events and followed by a
// readback of its counters; a film of FP64 sums and u32 counts downloaded to pinned host memory on request. The
// broker counterpart (x2dev_render_frame) follows Renderer::render's GPU-lane sequence (async_render.h:318-1081).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "x2_bvh.h"
#include "x2_half.h"
#include "x2_layout.h"
#include "x2_volume.h"
#include "x2_adaptive_core.h"

namespace x2 {
hipError_t launchRender(const RenderParams& p, hipStream_t stream);
hipError_t launchAdjoint(const RenderParams& p, hipStream_t stream);
hipError_t launchWave(const RenderParams& p, const WaveParams& w, uint32_t iterations, hipStream_t stream);
hipError_t launchChainStep(const ChainParams& c, const RenderParams& p, bool uniform, hipStream_t stream);
hipError_t launchErrors(const double* sums, const double* odd, const uint32_t* counts, uint32_t pixels, float* errors,
                        hipStream_t stream);
hipError_t launchResolve(const double* sums, const uint32_t* counts, uint32_t pixels, float* rgb, hipStream_t stream);
hipError_t launchClearState(LaunchState* s, hipStream_t stream);
hipError_t launchPyramid(const float* in, uint32_t w, uint32_t h, float* out, hipStream_t stream);
hipError_t launchPushDown(const float* g, uint32_t uw, uint32_t uh, float* below, uint32_t w, uint32_t h, hipStream_t stream);
hipError_t uploadSobol(const uint32_t* table);
hipError_t launchRngTwin(const x2_rng_query* q, uint32_t n, uint32_t draws, uint32_t* words, float* out,
                         uint64_t* branches, hipStream_t stream);
}  // namespace x2

namespace {
using namespace x2;

thread_local std::string g_error;

int fail(int code, const std::string& what) {
    g_error = what;
    return code;
}
int hipFail(hipError_t e, const char* where) {
    return fail(X2_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define X2_HIP(call)                                   \
    do {                                               \
        const hipError_t e_ = (call);                  \
        if (e_ != hipSuccess) return hipFail(e_, #call); \
    } while (0)

// A simulation setting from the environment (multi_device.h:47-50): its value as a count, 0 when unset or negative.
uint32_t simulated(const char* name) {
    const char* v = std::getenv(name);
    return v && *v ? uint32_t(std::max(0, std::atoi(v))) : 0u;
}

// Sobol direction numbers (sample_stream.h:82-100) from the Joe-Kuo primitive polynomials (new-joe-kuo-6.21201,
// dimensions 2..17: degree s, coefficient bits a, initial m), by the standard recurrence; row 0 is van der Corput.
void sobolDirections(uint32_t out[17 * 32]) {
    static const uint32_t s[16] = {1, 2, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6};
    static const uint32_t a[16] = {0, 1, 1, 2, 1, 4, 2, 4, 7, 11, 13, 14, 1, 13, 16, 19};
    static const uint32_t m0[16][6] = {{1},          {1, 3},        {1, 3, 1},        {1, 1, 1},
                                       {1, 1, 3, 3}, {1, 3, 5, 13}, {1, 1, 5, 5, 17}, {1, 1, 5, 5, 5},
                                       {1, 1, 7, 11, 19}, {1, 1, 5, 1, 1}, {1, 1, 1, 3, 11}, {1, 3, 5, 5, 31},
                                       {1, 3, 3, 9, 7, 49}, {1, 1, 1, 15, 21, 21}, {1, 3, 1, 13, 27, 49},
                                       {1, 1, 1, 15, 7, 5}};
    for (uint32_t k = 0; k < 32; ++k) out[k] = 0x80000000u >> k;
    for (uint32_t d = 1; d < 17; ++d) {
        const uint32_t deg = s[d - 1], coef = a[d - 1];
        uint32_t m[33] = {0};
        for (uint32_t i = 1; i <= deg; ++i) m[i] = m0[d - 1][i - 1];
        for (uint32_t i = deg + 1; i <= 32; ++i) {
            uint32_t v = m[i - deg] ^ (m[i - deg] << deg);
            for (uint32_t j = 1; j < deg; ++j)
                if ((coef >> (deg - 1 - j)) & 1u) v ^= m[i - j] << j;
            m[i] = v;
        }
        for (uint32_t k = 1; k <= 32; ++k) out[d * 32 + k - 1] = m[k] << (32 - k);
    }
}

struct Flight {
    uint64_t launch, first, count, samples;
    uint32_t slot;
    bool wave;  // run in the wavefront form
    uint64_t parkedBegin = 0, parkedCount = 0, items = 0;  // parked samples it ran again; items it could park
    uint64_t refines = 0;  // provisional samples it rendered again
};

// On-demand texture tiles (GpuLane's fine-detail cache, async_gpu.h:2911-3200): every paged texture's level 0 as
// 32x32 tiles of half-float rgb on the host, a pool of them on the device with the scene-wide page table, the tiles the
// kernels asked for, the ring of parked samples and the provisional film.
struct Paging {
    bool on = false, partial = false, hostRefines = false;
    // host refines: the ring the kernels write provisional records to (refineCap; those below refineHead are taken,
    // into refinesOut), and each launch slot's records to render again (device, and their pinned staging)
    x2_refine* refineRing = nullptr;
    x2_refine* refineReadback = nullptr;
    x2_refine* refineIn[Slots] = {};
    x2_refine* refineStaging[Slots] = {};
    uint64_t refineCap = 0, refineHead = 0;
    std::deque<x2_refine> refinesOut;
    uint32_t tiles = 0, poolTiles = 0, missCap = 0, uploadCap = 0;
    uint64_t parkCap = 0;
    std::vector<uint16_t> hostTiles;  // tile t's texels at t * TexTileTexels * 3
    std::vector<uint32_t> slotOf;     // host copy of the page table
    std::vector<uint32_t> tileIn;     // each pool slot's tile (NoTileSlot: free)
    std::vector<uint64_t> loadSeq;    // each pool slot's load, in order (ties of last use go to the oldest)
    std::vector<uint32_t> lastUse;    // each tile's last launch to read it, as read back
    uint32_t used = 0;                // pool slots filled so far
    uint64_t loads = 0;
    std::vector<uint8_t> queued;      // asked for and not yet sent up
    std::vector<uint32_t> pending;
    std::vector<uint16_t*> coarse;    // each texture's always-resident coarse levels (device, half floats)
    uint32_t* tileSlot = nullptr;
    uint32_t* tileAsked = nullptr;
    uint32_t* tileStamp = nullptr;
    uint32_t* stampReadback[Slots] = {};  // the tiles' stamps after each launch (pinned)
    uint32_t* missList = nullptr;
    uint16_t* pool = nullptr;
    PagingState* state = nullptr;
    x2_parked* parked = nullptr;
    float* parkedRgba = nullptr;  // a provisional parked sample's value (rgb + alpha) by ring slot
    double* provSums = nullptr;
    uint32_t* provCounts = nullptr;
    double* hostProvSums = nullptr;
    uint32_t* hostProvCounts = nullptr;
    size_t provPixels = 0;
    PagingState* readback = nullptr;             // one per launch slot (pinned)
    uint32_t* missReadback[Slots] = {};          // missCap entries per launch slot (pinned)
    uint16_t* upload[Slots] = {};                // tile staging per launch slot (pinned)
    uint32_t* tableStaging[Slots] = {};          // the page table per launch slot (pinned)
    uint64_t parkTail = 0, parkIssued = 0, inflightItems = 0;
    uint64_t parkedTotal = 0, provisional = 0, misses = 0, uploads = 0, evictions = 0, overflow = 0;
};


}  // namespace

namespace {
// The allocation-free render's device memory (x2dev_render_chunks_static): the launch's counters and claim counter
// (a LaunchState) at offset 0 and the chunks it renders from StaticChunkOffset, X2_STATIC_MAX_CHUNKS of them
// (1 MiB + 256 B). Each GPU the library runs on holds its own copy until the process ends. In host code `scratch`
// names a host-side stand-in, not device memory: host code uses the address hipGetSymbolAddress gives for the
// current device (staticScratch below), never the array's name.
constexpr size_t StaticChunkOffset = 256;
constexpr size_t StaticScratchBytes = StaticChunkOffset + size_t(X2_STATIC_MAX_CHUNKS) * sizeof(x2_chunk);
static_assert(sizeof(LaunchState) <= StaticChunkOffset, "the launch counters fit before the chunks");
static __device__ __align__(256) unsigned char scratch[StaticScratchBytes];

constexpr int MaxDevices = 64;
std::mutex g_scratchLock;
unsigned char* g_scratchAt[MaxDevices] = {};  // each device's address of `scratch` (null: not looked up yet)

// Looks up `scratch`'s address on `device` once per device, with that device current (called from x2dev_create).
hipError_t staticScratch(int device, unsigned char** out) {
    if (device < 0 || device >= MaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(g_scratchLock);
    if (!g_scratchAt[device]) {
        hipError_t e = hipSetDevice(device);
        void* at = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&at, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!at) return hipErrorInvalidSymbol;
        g_scratchAt[device] = static_cast<unsigned char*>(at);
    }
    *out = g_scratchAt[device];
    return hipSuccess;
}
}  // namespace

namespace x2 {  // for the other host files (x2_batch.hip)
void sobolTable(uint32_t out[17 * 32]) { sobolDirections(out); }

// A rectangle light's corners (+-half width, +-half height) through local_to_render, plus its position, in the order
// and with the float operations of the device's canIlluminate (rows(): (m0 x + m3 y) + m6 z, z = 0): made once here
// instead of at every vertex, the same floats. Zeros for an environment light.
static void lightCorners(const x2_light& l, float out[12]) {
    for (int i = 0; i < 12; ++i) out[i] = 0.0f;
    if (l.type != X2_LIGHT_RECT) return;
    const float hx = l.half_width, hy = l.half_height;
    const float cx[4] = {hx, hx, -hx, -hx}, cy[4] = {hy, -hy, hy, -hy};
    const float* m = l.local_to_render;
    for (int i = 0; i < 4; ++i)
        for (int a = 0; a < 3; ++a) {
            const float v = ((m[a] * cx[i]) + (m[3 + a] * cy[i])) + (m[6 + a] * 0.0f);
            out[3 * i + a] = v + l.position[a];
        }
}
size_t lightBufferBytes(uint32_t count) {
    return std::max<size_t>(1, count) * (sizeof(x2_light) + 12 * sizeof(float));
}
std::vector<char> lightBuffer(const x2_light* lights, uint32_t count) {
    std::vector<char> b(lightBufferBytes(count), 0);
    if (count) std::memcpy(b.data(), lights, count * sizeof(x2_light));
    for (uint32_t i = 0; i < count; ++i)
        lightCorners(lights[i], reinterpret_cast<float*>(b.data() + count * sizeof(x2_light)) + 12 * i);
    return b;
}
int reportFailure(int code, const std::string& what) { return fail(code, what); }
}  // namespace x2

struct x2dev {
    int device = 0;
    x2dev_options opt{};
    hipStream_t stream = nullptr;
    uint32_t cus = 0, launchThreads = 0;
    // scene
    DevNode* nodes = nullptr;
    TriRecord* tris = nullptr;
    x2_material* materials = nullptr;
    x2_light* lights = nullptr;
    uint32_t nodeCount = 0, triCount = 0, lightCount = 0, materialCount = 0, maxStack = 0;
    float* triUV = nullptr;
    uint8_t* triCull = nullptr;  // the single-sided meshes' triangles (leaf order), or null
    std::vector<float*> texels;   // one device buffer per texture: its mip pyramid, levels end to end
    float* gradPyramid = nullptr;  // the adjoint's gradients of every texture's levels 1 and up
    size_t gradPyramidFloats = 0;
    std::vector<TexDesc> texDesc;  // host copy of the descriptors
    TexDesc* textures = nullptr;
    // the medium's bricks (volDesc.table == null: none)
    VolumeDesc volDesc{};
    void* volBuffers[6] = {};  // table, density, temperature, majorant, ramp, emission region pool
    bool prepared = false;
    uint32_t failPixel = 0;  // XPU2_SIMULATE_SAMPLE_FAIL as the frame began (RenderParams::failPixel)
    // film
    double* sums = nullptr;
    uint32_t* counts = nullptr;
    size_t filmPixels = 0;
    double* hostSums = nullptr;
    uint32_t* hostCounts = nullptr;
    hipEvent_t filmDone = nullptr;
    bool filmQueued = false;
    // mailbox ring, staging, per-slot state and readback
    x2_chunk* ring = nullptr;
    x2_chunk* staging[Slots] = {};
    LaunchState* state = nullptr;
    LaunchState* readback = nullptr;
    // the wavefront form: one allocation holding every slot array and queue, and each launch's counters read back
    WaveParams wave{};
    void* waveBlock = nullptr;
    WaveCounts* waveReadback = nullptr;
    double waveVertices = 0;  // closest-hit rays an item took in the last wavefront launches
    // an adaptive frame: the odd-index film, the per-pixel errors and the pass tables (async_gpu.h:1836-1916)
    bool adaptive = false;
    size_t adaptivePixels = 0, adaptiveTiles = 0;
    double* odd = nullptr;
    double* hostOdd = nullptr;
    float* errors = nullptr;
    float* hostErrors = nullptr;
    hipEvent_t errorsDone = nullptr;
    bool errorsQueued = false;
    unsigned long long* passMask[PassSlots] = {};
    uint32_t* passBase[PassSlots] = {};
    unsigned long long* passStaging[PassSlots] = {};
    hipEvent_t passCopied[PassSlots] = {};
    hipEvent_t begun[Slots] = {}, ended[Slots] = {}, done[Slots] = {};
    std::deque<Flight> inFlight;
    uint64_t published = 0, launches = 0;
    // frame
    x2_config cfg{};
    uint32_t itemsPerChunk = 0;
    bool begun_ = false;
    // an adjoint frame's inputs and outputs (device pointers the caller owns), null in a forward frame
    const float* dImage = nullptr;
    float* gradTextures = nullptr;
    float* gradRoughness = nullptr;
    Paging pg;
    unsigned char* staticScratch = nullptr;  // this device's address of `scratch` (x2dev_render_chunks_static)
};

namespace {
void releasePaging(x2dev* d) {
    Paging& g = d->pg;
    (void)hipFree(g.tileSlot);
    (void)hipFree(g.tileAsked);
    (void)hipFree(g.tileStamp);
    (void)hipFree(g.missList);
    (void)hipFree(g.pool);
    (void)hipFree(g.state);
    (void)hipFree(g.parked);
    (void)hipFree(g.parkedRgba);
    (void)hipFree(g.provSums);
    (void)hipFree(g.provCounts);
    for (uint16_t* c : g.coarse) (void)hipFree(c);
    (void)hipFree(g.refineRing);
    if (g.refineReadback) (void)hipHostFree(g.refineReadback);
    for (uint32_t i = 0; i < Slots; ++i) {
        (void)hipFree(g.refineIn[i]);
        if (g.refineStaging[i]) (void)hipHostFree(g.refineStaging[i]);
    }
    if (g.hostProvSums) (void)hipHostFree(g.hostProvSums);
    if (g.hostProvCounts) (void)hipHostFree(g.hostProvCounts);
    if (g.readback) (void)hipHostFree(g.readback);
    for (uint32_t i = 0; i < Slots; ++i) {
        if (g.missReadback[i]) (void)hipHostFree(g.missReadback[i]);
        if (g.stampReadback[i]) (void)hipHostFree(g.stampReadback[i]);
        if (g.upload[i]) (void)hipHostFree(g.upload[i]);
        if (g.tableStaging[i]) (void)hipHostFree(g.tableStaging[i]);
    }
    g = Paging{};
}

void release(x2dev* d) {
    if (!d) return;
    (void)hipSetDevice(d->device);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    releasePaging(d);
    (void)hipFree(d->nodes);
    (void)hipFree(d->tris);
    (void)hipFree(d->materials);
    (void)hipFree(d->lights);
    (void)hipFree(d->triUV);
    (void)hipFree(d->triCull);
    (void)hipFree(d->textures);
    for (float* t : d->texels) (void)hipFree(t);
    (void)hipFree(d->gradPyramid);
    for (void* b : d->volBuffers) (void)hipFree(b);
    (void)hipFree(d->sums);
    (void)hipFree(d->counts);
    (void)hipFree(d->ring);
    (void)hipFree(d->state);
    if (d->hostSums) (void)hipHostFree(d->hostSums);
    if (d->hostCounts) (void)hipHostFree(d->hostCounts);
    if (d->readback) (void)hipHostFree(d->readback);
    (void)hipFree(d->waveBlock);
    if (d->waveReadback) (void)hipHostFree(d->waveReadback);
    for (uint32_t i = 0; i < Slots; ++i) {
        if (d->staging[i]) (void)hipHostFree(d->staging[i]);
        if (d->begun[i]) (void)hipEventDestroy(d->begun[i]);
        if (d->ended[i]) (void)hipEventDestroy(d->ended[i]);
        if (d->done[i]) (void)hipEventDestroy(d->done[i]);
    }
    if (d->filmDone) (void)hipEventDestroy(d->filmDone);
    (void)hipFree(d->odd);
    (void)hipFree(d->errors);
    if (d->hostOdd) (void)hipHostFree(d->hostOdd);
    if (d->hostErrors) (void)hipHostFree(d->hostErrors);
    if (d->errorsDone) (void)hipEventDestroy(d->errorsDone);
    for (uint32_t i = 0; i < PassSlots; ++i) {
        (void)hipFree(d->passMask[i]);
        (void)hipFree(d->passBase[i]);
        if (d->passStaging[i]) (void)hipHostFree(d->passStaging[i]);
        if (d->passCopied[i]) (void)hipEventDestroy(d->passCopied[i]);
    }
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

uint64_t oldestChunk(const x2dev* d) { return d->inFlight.empty() ? d->published : d->inFlight.front().first; }

int validateConfig(const x2dev* d, const x2_config& c) {
    if (!c.width || !c.height || !c.spp) return fail(X2_ERR_ARGUMENT, "width, height and spp must be positive");
    if (uint64_t(c.width) * c.height > (1ull << 31)) return fail(X2_ERR_ARGUMENT, "image too large");
    if (c.bsdf_samples == 0 || c.bsdf_samples > 64 || c.light_samples == 0 || c.light_samples > 64)
        return fail(X2_ERR_ARGUMENT, "bsdf_samples and light_samples must be 1..64 (async_integrator.h:1989)");
    if (c.max_depth > 250 || c.max_diffuse_depth > 250 || c.max_glossy_depth > 250)
        return fail(X2_ERR_ARGUMENT, "depth limits above 250 are not supported");
    if (c.filter != 0) return fail(X2_ERR_ARGUMENT, "only the box pixel filter is implemented");
    if (c.light_count > d->lightCount) return fail(X2_ERR_ARGUMENT, "config.light_count exceeds the scene's lights");
    if (c.direct_only && (c.surface_limit == 0 || c.surface_limit > 8))
        return fail(X2_ERR_ARGUMENT, "the prototype estimator takes surface_limit 1..8 (async_integrator.h:65)");
    if (c.direct_only && d->volDesc.table) return fail(X2_ERR_ARGUMENT, "the prototype estimator renders no media");
    if (d->volDesc.table && (c.max_volume_depth > 16 || c.volume_illumination_samples > 64 ||
                             !(c.inv_volume_quality > 0.0f) || !(c.volume_transmittance_threshold >= 0.0f)))
        return fail(X2_ERR_ARGUMENT, "volume settings budget (cpu_scene.h:3168): depth <= 16, samples <= 64, "
                                     "a positive quality");
    return X2_OK;
}
}  // namespace

namespace {
// A texture's pyramid layout (sizes and float offsets of its levels, end to end): level k + 1 is
// max(1, w >> 1) x max(1, h >> 1), down to 1 x 1.
TexDesc pyramidLayout(uint32_t w, uint32_t h) {
    TexDesc td{};
    unsigned long long at = 0;
    for (;;) {
        TexLevel& l = td.levels[td.count++];
        l.width = w;
        l.height = h;
        l.offset = at;
        at += (unsigned long long)w * h * 3;
        if ((w == 1 && h == 1) || td.count == MipLevels) break;
        w = w > 1 ? w >> 1 : 1u;
        h = h > 1 ? h >> 1 : 1u;
    }
    td.firstResident = td.count;
    return td;
}
// Texture t's levels 1 and up from its level 0, on the device (queued on the lane's stream).
hipError_t buildLevels(x2dev* d, uint32_t t) {
    const TexDesc& td = d->texDesc[t];
    for (uint32_t k = 1; k < td.count; ++k) {
        const TexLevel &in = td.levels[k - 1], &out = td.levels[k];
        const hipError_t e = launchPyramid(d->texels[t] + in.offset, in.width, in.height, d->texels[t] + out.offset, d->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
}  // namespace

extern "C" {

const char* x2dev_last_error(void) { return g_error.c_str(); }

int x2dev_create(const x2dev_options* options, x2dev** out) {
    if (!out) return fail(X2_ERR_ARGUMENT, "null output pointer");
    *out = nullptr;
    x2dev* d = new x2dev;
    d->opt.device = options ? options->device : 0;
    d->opt.max_chunks_per_launch = options && options->max_chunks_per_launch ? options->max_chunks_per_launch : 16384u;
    d->opt.ring_chunks = options && options->ring_chunks ? options->ring_chunks : (1u << 17);
    d->opt.counters = options ? options->counters : 1u;
    d->device = d->opt.device;
    d->opt.wave_slots = options ? options->wave_slots : 0u;
    d->opt.lane = options ? options->lane : 0u;
    if (d->opt.wave_slots % 256 || d->opt.wave_slots > (1u << 24)) {
        delete d;
        return fail(X2_ERR_ARGUMENT, "wave_slots must be a multiple of 256, at most 2^24");
    }
    if (d->opt.ring_chunks < d->opt.max_chunks_per_launch) {
        delete d;
        return fail(X2_ERR_ARGUMENT, "ring_chunks must hold at least one launch");
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || d->device < 0 || d->device >= count) {
        delete d;
        return fail(X2_ERR_HIP, e != hipSuccess ? hipGetErrorString(e) : "no such HIP device");
    }
#define X2_CREATE(call)                 \
    do {                                \
        const hipError_t e_ = (call);   \
        if (e_ != hipSuccess) {         \
            const int r_ = hipFail(e_, #call); \
            release(d);                 \
            return r_;                  \
        }                               \
    } while (0)
    X2_CREATE(hipSetDevice(d->device));
    hipDeviceProp_t prop;
    X2_CREATE(hipGetDeviceProperties(&prop, d->device));
    d->cus = uint32_t(prop.multiProcessorCount);
    d->launchThreads = uint32_t(prop.multiProcessorCount) * uint32_t(prop.maxThreadsPerMultiProcessor);
    X2_CREATE(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    X2_CREATE(hipMalloc(&d->ring, size_t(d->opt.ring_chunks) * sizeof(x2_chunk)));
    X2_CREATE(hipMalloc(&d->state, Slots * sizeof(LaunchState)));
    X2_CREATE(hipHostMalloc(&d->readback, Slots * sizeof(LaunchState), hipHostMallocDefault));
    X2_CREATE(hipHostMalloc(&d->waveReadback, Slots * sizeof(WaveCounts), hipHostMallocDefault));
    for (uint32_t i = 0; i < Slots; ++i) {
        X2_CREATE(hipHostMalloc(&d->staging[i], size_t(d->opt.max_chunks_per_launch) * sizeof(x2_chunk), hipHostMallocDefault));
        X2_CREATE(hipEventCreate(&d->begun[i]));
        X2_CREATE(hipEventCreate(&d->ended[i]));
        X2_CREATE(hipEventCreateWithFlags(&d->done[i], hipEventDisableTiming));
    }
    X2_CREATE(hipEventCreateWithFlags(&d->filmDone, hipEventDisableTiming));
    X2_CREATE(hipEventCreateWithFlags(&d->errorsDone, hipEventDisableTiming));
    for (uint32_t i = 0; i < PassSlots; ++i) X2_CREATE(hipEventCreateWithFlags(&d->passCopied[i], hipEventDisableTiming));
    uint32_t table[17 * 32];
    sobolDirections(table);
    X2_CREATE(uploadSobol(table));
    X2_CREATE(staticScratch(d->device, &d->staticScratch));
#undef X2_CREATE
    *out = d;
    return X2_OK;
}

int x2dev_prepare(x2dev* d, const x2_scene* s) {
    if (!d || !s) return fail(X2_ERR_ARGUMENT, "null device or scene");
    if (d->opt.lane && simulated("XPU2_SIMULATE_CARD_FAIL") == d->opt.lane)  // async_gpu.h:685-686
        return fail(X2_ERR_STATE, "failed (simulated, XPU2_SIMULATE_CARD_FAIL)");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "prepare with launches in flight");
    if ((s->triangle_count && (!s->positions || !s->indices)) || !s->geom_first || !s->geom_material ||
        !s->materials || (s->light_count && !s->lights))
        return fail(X2_ERR_ARGUMENT, "scene arrays missing");
    for (uint32_t i = 0; i < s->light_count; ++i)
        if (s->lights[i].type != X2_LIGHT_ENV && s->lights[i].type != X2_LIGHT_RECT)
            return fail(X2_ERR_ARGUMENT, "only environment and rectangle lights are implemented");
    if (s->texture_count > MaxTextures || (s->texture_count && !s->textures))
        return fail(X2_ERR_ARGUMENT, "at most 16 textures, and their records");
    for (uint32_t i = 0; i < s->texture_count; ++i)
        if (!s->textures[i].texels || !s->textures[i].width || !s->textures[i].height)
            return fail(X2_ERR_ARGUMENT, "a texture without texels");
    for (uint32_t i = 0; i < s->material_count; ++i)
        if (s->materials[i].albedo_map > s->texture_count) return fail(X2_ERR_ARGUMENT, "albedo_map names no texture");
    if (s->volume_count > 1 || (s->volume_count && !s->volumes))
        return fail(X2_ERR_ARGUMENT, "at most one volume, and its record");
    if (s->volume_count && s->texture_count) return fail(X2_ERR_ARGUMENT, "a scene with a medium reads no textures");
    if (d->opt.wave_slots && s->texture_count)
        return fail(X2_ERR_ARGUMENT, "the wavefront form renders untextured scenes (the one-kernel form carries differentials)");
    for (uint32_t i = 0; i < s->texture_count; ++i)
        if (s->textures[i].width > 32768u || s->textures[i].height > 32768u)
            return fail(X2_ERR_ARGUMENT, "texture sides are at most 32768 (16 mip levels)");
    VolumeHost vol;
    if (s->volume_count) {
        std::string verr;
        if (!packVolume(s->volumes[0], vol, verr)) return fail(X2_ERR_ARGUMENT, verr);
    }
    Bvh8Result bvh;
    std::string err;
    if (!buildBvh8(*s, bvh, err)) return fail(X2_ERR_ARGUMENT, err);
    if (bvh.maxStack > uint32_t(TraversalStack))
        return fail(X2_ERR_ARGUMENT, "BVH8 too deep for the traversal stack (" + std::to_string(bvh.maxStack) + ")");
    X2_HIP(hipSetDevice(d->device));
    (void)hipFree(d->nodes);
    (void)hipFree(d->tris);
    (void)hipFree(d->materials);
    (void)hipFree(d->lights);
    (void)hipFree(d->triUV);
    (void)hipFree(d->triCull);
    (void)hipFree(d->textures);
    for (float* t : d->texels) (void)hipFree(t);
    (void)hipFree(d->gradPyramid);
    d->gradPyramid = nullptr;
    d->gradPyramidFloats = 0;
    releasePaging(d);  // a new scene's textures are resident until paged again
    for (void*& b : d->volBuffers) {
        (void)hipFree(b);
        b = nullptr;
    }
    d->volDesc = VolumeDesc{};
    d->texels.clear();
    d->texDesc.clear();
    d->triUV = nullptr;
    d->triCull = nullptr;
    d->textures = nullptr;
    d->nodes = nullptr;
    d->tris = nullptr;
    d->materials = nullptr;
    d->lights = nullptr;
    d->prepared = false;
    const size_t nb = bvh.deviceNodeBytes(), tb = std::max<size_t>(1, triBufferBytes(bvh.tris.size()));
    X2_HIP(hipMalloc(&d->nodes, nb));
    X2_HIP(hipMalloc(&d->tris, tb));
    X2_HIP(hipMalloc(&d->materials, std::max<size_t>(1, s->material_count) * sizeof(x2_material)));
    X2_HIP(hipMalloc(&d->lights, lightBufferBytes(s->light_count)));
    X2_HIP(hipMemcpy(d->nodes, bvh.deviceNodes(), nb, hipMemcpyHostToDevice));
    if (!bvh.tris.empty()) X2_HIP(hipMemcpy(d->tris, bvh.tris.data(), bvh.tris.size() * sizeof(TriRecord), hipMemcpyHostToDevice));
    if (!bvh.tris.empty()) {
        const std::vector<float> v = triVertices(bvh.tris);
        X2_HIP(hipMemcpy(reinterpret_cast<char*>(d->tris) + triVertsOffset(bvh.tris.size()), v.data(),
                         v.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    X2_HIP(hipMemcpy(d->materials, s->materials, s->material_count * sizeof(x2_material), hipMemcpyHostToDevice));
    {
        const std::vector<char> lb = lightBuffer(s->lights, s->light_count);
        X2_HIP(hipMemcpy(d->lights, lb.data(), lb.size(), hipMemcpyHostToDevice));
    }
    if (!bvh.triUV.empty()) {
        X2_HIP(hipMalloc(&d->triUV, bvh.triUV.size() * sizeof(float)));
        X2_HIP(hipMemcpy(d->triUV, bvh.triUV.data(), bvh.triUV.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (!bvh.triCull.empty()) {
        X2_HIP(hipMalloc(&d->triCull, bvh.triCull.size()));
        X2_HIP(hipMemcpy(d->triCull, bvh.triCull.data(), bvh.triCull.size(), hipMemcpyHostToDevice));
    }
    unsigned long long gradOffset = 0;
    for (uint32_t i = 0; i < s->texture_count; ++i) {
        const x2_texture& t = s->textures[i];
        TexDesc td = pyramidLayout(t.width, t.height);
        const size_t floats = td.levels[td.count - 1].offset + 3;
        float* buf = nullptr;
        X2_HIP(hipMalloc(&buf, floats * sizeof(float)));
        d->texels.push_back(buf);
        X2_HIP(hipMemcpy(buf, t.texels, size_t(t.width) * t.height * 3 * sizeof(float), hipMemcpyHostToDevice));
        td.texels = buf;
        td.gradOffset = gradOffset;
        gradOffset += size_t(t.width) * t.height * 3;
        // its levels 1 and up in the pyramid gradient, at the same spacing as in its pyramid
        td.gradLevels = td.count > 1 ? (long long)d->gradPyramidFloats - (long long)td.levels[1].offset : 0;
        d->gradPyramidFloats += floats - size_t(t.width) * t.height * 3;
        d->texDesc.push_back(td);
        X2_HIP(buildLevels(d, i));
    }
    X2_HIP(hipStreamSynchronize(d->stream));
    {  // the adjoint's LDS levels: the coarsest of every texture first, while they fit
        for (TexDesc& td : d->texDesc)
            for (uint32_t k = 0; k < td.count; ++k) td.levels[k].lds = NoLds;
        uint32_t used = 0;
        for (uint32_t k = 0; k < MipLevels; ++k)  // k levels up from each texture's 1 x 1
            for (TexDesc& td : d->texDesc) {
                if (k >= td.count) continue;
                TexLevel& l = td.levels[td.count - 1 - k];
                const uint32_t n = l.width * l.height;
                if (n <= AdjLdsTexels - used) {
                    l.lds = used;
                    used += n;
                }
            }
    }
    if (!d->texDesc.empty()) {
        X2_HIP(hipMalloc(&d->textures, d->texDesc.size() * sizeof(TexDesc)));
        X2_HIP(hipMemcpy(d->textures, d->texDesc.data(), d->texDesc.size() * sizeof(TexDesc), hipMemcpyHostToDevice));
    }
    if (s->volume_count) {
        VolumeDesc v = vol.desc;
        auto upload = [&](int slot, const void* src, size_t bytes) -> hipError_t {
            hipError_t e = hipMalloc(&d->volBuffers[slot], std::max<size_t>(bytes, 4));
            if (e == hipSuccess && bytes) e = hipMemcpy(d->volBuffers[slot], src, bytes, hipMemcpyHostToDevice);
            return e;
        };
        X2_HIP(upload(0, vol.table.data(), vol.table.size() * sizeof(uint32_t)));
        X2_HIP(upload(1, vol.density.data(), vol.density.size() * sizeof(uint16_t)));
        X2_HIP(upload(3, vol.majorant.data(), vol.majorant.size() * sizeof(float)));
        if (!vol.temperature.empty() || !vol.ramp.empty()) {
            X2_HIP(upload(2, vol.temperature.data(), vol.temperature.size() * sizeof(uint16_t)));
            X2_HIP(upload(4, vol.ramp.data(), vol.ramp.size() * sizeof(float)));
        }
        v.table = static_cast<const uint32_t*>(d->volBuffers[0]);
        v.density = static_cast<const uint16_t*>(d->volBuffers[1]);
        v.temperature = s->volumes[0].temperature ? static_cast<const uint16_t*>(d->volBuffers[2]) : nullptr;
        v.majorant = static_cast<const float*>(d->volBuffers[3]);
        v.ramp = static_cast<const float*>(d->volBuffers[4]);
        v.region.pool = nullptr;
        if (!vol.emissionPool.empty()) {
            X2_HIP(upload(5, vol.emissionPool.data(), vol.emissionPool.size() * sizeof(float)));
            v.region.pool = static_cast<const float*>(d->volBuffers[5]);
        }
        d->volDesc = v;
    }
    d->nodeCount = uint32_t(bvh.nodes.size());
    d->triCount = uint32_t(bvh.tris.size());
    d->lightCount = s->light_count;
    d->materialCount = s->material_count;
    d->maxStack = bvh.maxStack;
    d->prepared = true;
    return X2_OK;
}

namespace {
// The buffers of an adaptive frame at this image size: the odd film and errors (device and pinned host copies) and
// the four pass slots' tables.
int ensureAdaptive(x2dev* d, size_t pixels, size_t tiles) {
    if (d->adaptivePixels == pixels && d->adaptiveTiles == tiles) return X2_OK;
    (void)hipFree(d->odd);
    (void)hipFree(d->errors);
    if (d->hostOdd) (void)hipHostFree(d->hostOdd);
    if (d->hostErrors) (void)hipHostFree(d->hostErrors);
    d->odd = nullptr;
    d->errors = nullptr;
    d->hostOdd = nullptr;
    d->hostErrors = nullptr;
    d->adaptivePixels = d->adaptiveTiles = 0;
    for (uint32_t i = 0; i < PassSlots; ++i) {
        (void)hipFree(d->passMask[i]);
        (void)hipFree(d->passBase[i]);
        if (d->passStaging[i]) (void)hipHostFree(d->passStaging[i]);
        d->passMask[i] = nullptr;
        d->passBase[i] = nullptr;
        d->passStaging[i] = nullptr;
    }
    X2_HIP(hipMalloc(&d->odd, pixels * 4 * sizeof(double)));
    X2_HIP(hipMalloc(&d->errors, pixels * sizeof(float)));
    X2_HIP(hipHostMalloc(&d->hostOdd, pixels * 4 * sizeof(double), hipHostMallocDefault));
    X2_HIP(hipHostMalloc(&d->hostErrors, pixels * sizeof(float), hipHostMallocDefault));
    for (uint32_t i = 0; i < PassSlots; ++i) {
        X2_HIP(hipMalloc(&d->passMask[i], tiles * sizeof(unsigned long long)));
        X2_HIP(hipMalloc(&d->passBase[i], pixels * sizeof(uint32_t)));
        X2_HIP(hipHostMalloc(&d->passStaging[i], tiles * sizeof(unsigned long long), hipHostMallocDefault));
        X2_HIP(hipMemsetAsync(d->passMask[i], 0, tiles * sizeof(unsigned long long), d->stream));
    }
    d->adaptivePixels = pixels;
    d->adaptiveTiles = tiles;
    return X2_OK;
}
int beginFrame(x2dev* d, const x2_config* c, uint32_t samples_per_chunk, bool adaptive);
}  // namespace

int x2dev_begin(x2dev* d, const x2_config* c, uint32_t samples_per_chunk) {
    return beginFrame(d, c, samples_per_chunk, false);
}
int x2dev_begin_adaptive(x2dev* d, const x2_config* c, uint32_t samples_per_chunk) {
    return beginFrame(d, c, samples_per_chunk, true);
}

namespace {
// A frame's paging state: the parked ring (three launches of the cap), the provisional film, the counters cleared.
int beginPaging(x2dev* d, size_t pixels) {
    Paging& g = d->pg;
    const uint64_t cap = std::max<uint64_t>(uint64_t(3) * d->opt.max_chunks_per_launch * d->itemsPerChunk, 1u << 20);
    if (cap > g.parkCap) {
        (void)hipFree(g.parked);
        (void)hipFree(g.parkedRgba);
        g.parked = nullptr;
        g.parkedRgba = nullptr;
        g.parkCap = 0;
        X2_HIP(hipMalloc(&g.parked, cap * sizeof(x2_parked)));
        X2_HIP(hipMalloc(&g.parkedRgba, cap * 4 * sizeof(float)));
        g.parkCap = cap;
    }
    if (g.hostRefines && cap > g.refineCap) {  // as many provisional records as the parked ring holds samples
        (void)hipFree(g.refineRing);
        if (g.refineReadback) (void)hipHostFree(g.refineReadback);
        g.refineRing = nullptr;
        g.refineReadback = nullptr;
        g.refineCap = 0;
        X2_HIP(hipMalloc(&g.refineRing, cap * sizeof(x2_refine)));
        X2_HIP(hipHostMalloc(&g.refineReadback, cap * sizeof(x2_refine), hipHostMallocDefault));
        g.refineCap = cap;
    }
    g.refineHead = 0;
    g.refinesOut.clear();
    if (pixels != g.provPixels) {
        (void)hipFree(g.provSums);
        (void)hipFree(g.provCounts);
        if (g.hostProvSums) (void)hipHostFree(g.hostProvSums);
        if (g.hostProvCounts) (void)hipHostFree(g.hostProvCounts);
        g.provSums = nullptr;
        g.provCounts = nullptr;
        g.hostProvSums = nullptr;
        g.hostProvCounts = nullptr;
        g.provPixels = 0;
        X2_HIP(hipMalloc(&g.provSums, pixels * 4 * sizeof(double)));
        X2_HIP(hipMalloc(&g.provCounts, pixels * sizeof(uint32_t)));
        X2_HIP(hipHostMalloc(&g.hostProvSums, pixels * 4 * sizeof(double), hipHostMallocDefault));
        X2_HIP(hipHostMalloc(&g.hostProvCounts, pixels * sizeof(uint32_t), hipHostMallocDefault));
        g.provPixels = pixels;
    }
    X2_HIP(hipMemsetAsync(g.provSums, 0, pixels * 4 * sizeof(double), d->stream));
    X2_HIP(hipMemsetAsync(g.provCounts, 0, pixels * sizeof(uint32_t), d->stream));
    X2_HIP(hipMemsetAsync(g.state, 0, sizeof(PagingState), d->stream));
    g.parkTail = g.parkIssued = g.inflightItems = 0;
    g.parkedTotal = g.provisional = g.overflow = 0;
    return X2_OK;
}
int beginFrame(x2dev* d, const x2_config* c, uint32_t samples_per_chunk, bool adaptive) {
    if (!d || !c) return fail(X2_ERR_ARGUMENT, "null device or config");
    if (!d->prepared) return fail(X2_ERR_STATE, "begin before prepare");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "GPU frame overlap (async_gpu.h:1374)");
    if (samples_per_chunk == 0 || samples_per_chunk > 1024) return fail(X2_ERR_ARGUMENT, "samples_per_chunk must be 1..1024");
    const int v = validateConfig(d, *c);
    if (v) return v;
    X2_HIP(hipSetDevice(d->device));
    const size_t pixels = size_t(c->width) * c->height;
    if (pixels != d->filmPixels) {
        (void)hipFree(d->sums);
        (void)hipFree(d->counts);
        if (d->hostSums) (void)hipHostFree(d->hostSums);
        if (d->hostCounts) (void)hipHostFree(d->hostCounts);
        d->sums = nullptr;
        d->counts = nullptr;
        d->hostSums = nullptr;
        d->hostCounts = nullptr;
        d->filmPixels = 0;
        X2_HIP(hipMalloc(&d->sums, pixels * 4 * sizeof(double)));
        X2_HIP(hipMalloc(&d->counts, pixels * sizeof(uint32_t)));
        X2_HIP(hipHostMalloc(&d->hostSums, pixels * 4 * sizeof(double), hipHostMallocDefault));
        X2_HIP(hipHostMalloc(&d->hostCounts, pixels * sizeof(uint32_t), hipHostMallocDefault));
        d->filmPixels = pixels;
    }
    X2_HIP(hipMemsetAsync(d->sums, 0, pixels * 4 * sizeof(double), d->stream));
    X2_HIP(hipMemsetAsync(d->counts, 0, pixels * sizeof(uint32_t), d->stream));
    d->cfg = *c;
    d->failPixel = simulated("XPU2_SIMULATE_SAMPLE_FAIL");
    d->dImage = nullptr;
    d->gradTextures = nullptr;
    d->gradRoughness = nullptr;
    d->itemsPerChunk = samples_per_chunk * TilePixels;
    d->published = 0;
    d->launches = 0;
    d->filmQueued = false;
    d->errorsQueued = false;
    d->adaptive = adaptive;
    if (d->pg.on) {
        if (c->direct_only) return fail(X2_ERR_STATE, "paged textures: the prototype estimator reads resident textures");
        const int r = beginPaging(d, pixels);
        if (r) return r;
    }
    if (adaptive) {
        if (d->volDesc.table || c->direct_only) return fail(X2_ERR_ARGUMENT, "adaptive frames render the path tracer");
        const int r = ensureAdaptive(d, pixels, size_t((c->width + 7) / 8) * ((c->height + 7) / 8));
        if (r) return r;
        X2_HIP(hipMemsetAsync(d->odd, 0, pixels * 4 * sizeof(double), d->stream));
    }
    d->begun_ = true;
    return X2_OK;
}
}  // namespace

int x2dev_ready(const x2dev* d) { return d && d->begun_ && d->inFlight.size() < 2 ? 1 : 0; }
uint32_t x2dev_in_flight(const x2dev* d) { return d ? uint32_t(d->inFlight.size()) : 0u; }
uint32_t x2dev_max_chunks_per_launch(const x2dev* d) { return d ? d->opt.max_chunks_per_launch : 0u; }
uint32_t x2dev_launch_threads(const x2dev* d) { return d ? d->launchThreads : 0u; }
uint32_t x2dev_chunk_room(const x2dev* d) {
    if (!d) return 0u;
    const uint64_t used = d->published - oldestChunk(d);
    uint64_t room = used >= d->opt.ring_chunks ? 0 : d->opt.ring_chunks - used;
    if (d->pg.on && d->itemsPerChunk) {  // what the launch may park, on top of what is parked and may park in flight
        const Paging& g = d->pg;
        const uint64_t bound = g.parkTail + g.inflightItems + (g.parkTail - g.parkIssued);
        const uint64_t limit = g.parkIssued + g.parkCap;
        room = std::min<uint64_t>(room, bound >= limit ? 0 : (limit - bound) / d->itemsPerChunk);
        if (g.hostRefines)  // what the launch may write as provisional records
            room = std::min<uint64_t>(room, g.inflightItems >= g.refineCap ? 0 : (g.refineCap - g.inflightItems) / d->itemsPerChunk);
    }
    return uint32_t(std::min<uint64_t>(room, d->opt.max_chunks_per_launch));
}
uint64_t x2dev_parked_outstanding(const x2dev* d) { return d && d->pg.on ? d->pg.parkTail - d->pg.parkIssued : 0u; }

namespace {
// What every launch of the frame reads (its items and, for the chain, its pass set by the caller).
RenderParams renderParams(const x2dev* d, uint32_t slot) {
    RenderParams p{};
    p.cfg = d->cfg;
    p.nodes = d->nodes;
    p.tris = d->tris;
    p.triVerts = reinterpret_cast<const float*>(reinterpret_cast<const char*>(d->tris) + triVertsOffset(d->triCount));
    p.triUV = d->triUV;
    p.triCull = d->triCull;
    p.textures = d->textures;
    p.materials = d->materials;
    p.lights = d->lights;
    p.lightCorners = reinterpret_cast<const float*>(reinterpret_cast<const char*>(d->lights) + d->lightCount * sizeof(x2_light));
    p.dImage = d->dImage;
    p.gradTextures = d->gradTextures;
    p.gradRoughness = d->gradRoughness;
    p.gradPyramid = d->gradPyramid;
    p.textureCount = uint32_t(d->texDesc.size());
    p.ring = d->ring;
    p.sums = d->sums;
    p.counts = d->counts;
    p.state = d->state + slot;
    p.itemsPerChunk = d->itemsPerChunk;
    p.ringCapacity = d->opt.ring_chunks;
    p.counters = d->opt.counters;
    p.adaptive = d->adaptive ? 1u : 0u;
    p.failPixel = d->failPixel;
    if (d->pg.on) {
        const Paging& g = d->pg;
        p.tileSlot = g.tileSlot;
        p.tilePool = g.pool;
        p.tileAsked = g.tileAsked;
        p.missList = g.missList;
        p.paging = g.state;
        p.parked = g.parked;
        p.parkedRgba = g.parkedRgba;
        p.provSums = g.provSums;
        p.provCounts = g.provCounts;
        p.parkCap = uint32_t(std::min<uint64_t>(g.parkCap, 0xffffffffu));
        p.missCap = g.missCap;
        p.partial = g.partial ? 1u : 0u;
        p.tileStamp = g.tileStamp;
        p.hostRefines = g.hostRefines ? 1u : 0u;
        p.refineRing = g.refineRing;
        p.refineCap = uint32_t(std::min<uint64_t>(g.refineCap, 0xffffffffu));
    }
    p.vol = d->volDesc;
    p.odd = d->adaptive ? d->odd : nullptr;
    for (uint32_t i = 0; i < PassSlots; ++i) {
        p.passMask[i] = d->adaptive ? d->passMask[i] : nullptr;
        p.passBase[i] = d->adaptive ? d->passBase[i] : nullptr;
    }
    return p;
}
// The surface estimator of a forward frame runs in the wavefront form when the build has path slots.
bool waveForm(const x2dev* d) { return d->opt.wave_slots && !d->dImage && !d->volDesc.table && !d->cfg.direct_only; }
// Allocates the slot arrays and queues for `perVertex` shadow rays a vertex (once; again if a frame needs more).
int ensureWave(x2dev* d, uint32_t perVertex) {
    if (d->waveBlock && d->wave.perVertex >= perVertex) return X2_OK;
    X2_HIP(hipStreamSynchronize(d->stream));
    (void)hipFree(d->waveBlock);
    d->waveBlock = nullptr;
    const size_t n = d->opt.wave_slots, e = n * perVertex;
    constexpr uint32_t partialBlocks = 8192;  // trace-pass workgroups with their own counters (at most)
    const size_t sizes[] = {
        n * 4,  n * 8,  n * 4,  n * 32, size_t(MaxPendingRecords) * n * PendingBytes, n * PendingBytes, n * 4, n * 4,
        n * 32, n * 32, n * 4,  n * 4,  e * 32, e * 32, e * 16, e * 16, e * 4, e * 4, e * 4, e * 4,
        sizeof(WaveCounts), size_t(partialBlocks) * 4 * 8};
    constexpr size_t parts = sizeof(sizes) / sizeof(sizes[0]);
    size_t offsets[parts], total = 0;
    for (size_t i = 0; i < parts; ++i) {
        offsets[i] = total;
        total += (sizes[i] + 255) & ~size_t(255);
    }
    X2_HIP(hipMalloc(&d->waveBlock, total));
    char* b = static_cast<char*>(d->waveBlock);
    WaveParams& w = d->wave;
    w.slots = uint32_t(n);
    w.perVertex = perVertex;
    w.flags = reinterpret_cast<uint32_t*>(b + offsets[0]);
    w.sample = reinterpret_cast<uint32_t*>(b + offsets[1]);
    w.top = reinterpret_cast<uint32_t*>(b + offsets[2]);
    w.accum = reinterpret_cast<float*>(b + offsets[3]);
    w.stack = b + offsets[4];
    w.current = b + offsets[5];
    w.hitTri = reinterpret_cast<uint32_t*>(b + offsets[6]);
    w.shadowCount = reinterpret_cast<uint32_t*>(b + offsets[7]);
    for (int q = 0; q < 2; ++q) {
        w.rays[q] = reinterpret_cast<float*>(b + offsets[8 + q]);
        w.rayList[q] = reinterpret_cast<uint32_t*>(b + offsets[10 + q]);
        w.shadowRays[q] = reinterpret_cast<float*>(b + offsets[12 + q]);
        w.shadowValue[q] = reinterpret_cast<float*>(b + offsets[14 + q]);
        w.shadowList[q] = reinterpret_cast<uint32_t*>(b + offsets[16 + q]);
        w.shadowVis[q] = reinterpret_cast<uint32_t*>(b + offsets[18 + q]);
    }
    w.counts = reinterpret_cast<WaveCounts*>(b + offsets[20]);
    w.partials = reinterpret_cast<unsigned long long*>(b + offsets[21]);
    w.partialBlocks = partialBlocks;
    w.pad = 0;
    X2_HIP(hipMemsetAsync(w.flags, 0, n * 4, d->stream));  // every slot empty
    return X2_OK;
}
}  // namespace

namespace {
// GpuLane::commitTiles (async_gpu.h:3164-3200): the tiles asked for go up into the pool, the oldest asked first, at
// most uploadCap a launch; a full pool gives up its oldest load. Then the page table, and every tile's asked-for flag
// cleared (a tile evicted later is asked for again; one asked for by a launch still in flight may be asked twice,
// which the host's queue keeps once). Stream order keeps every copy after the launches before and ahead of the next.
int commitTiles(x2dev* d, uint32_t slot) {
    Paging& g = d->pg;
    if (g.pending.empty()) return X2_OK;
    const size_t tileBytes = size_t(TexTileTexels) * 3 * sizeof(uint16_t);
    const size_t n = std::min<size_t>({g.pending.size(), size_t(g.uploadCap), size_t(g.poolTiles)});
    // The slots the commit fills: free ones first, then those of the least recently read tiles (the oldest load
    // among equals) — a parked sample's tiles are read again each time it runs, so they stay, and the tiles of
    // finished samples go first.
    std::vector<uint32_t> slots;
    const uint32_t resident = g.used;  // the slots holding a tile before this commit: the only candidates to give up
    while (slots.size() < n && g.used < g.poolTiles) slots.push_back(g.used++);
    if (slots.size() < n) {
        std::vector<uint32_t> order(resident);
        for (uint32_t i = 0; i < resident; ++i) order[i] = i;
        const size_t k = n - slots.size();
        auto older = [&](uint32_t a, uint32_t b) {
            const uint32_t ua = g.lastUse[g.tileIn[a]], ub = g.lastUse[g.tileIn[b]];
            return ua != ub ? ua < ub : g.loadSeq[a] < g.loadSeq[b];
        };
        std::nth_element(order.begin(), order.begin() + std::ptrdiff_t(k - 1), order.end(), older);
        slots.insert(slots.end(), order.begin(), order.begin() + std::ptrdiff_t(k));
    }
    for (size_t i = 0; i < n; ++i) {
        const uint32_t tile = g.pending[i];
        const uint32_t at = slots[i];
        if (g.tileIn[at] != NoTileSlot) {
            g.slotOf[g.tileIn[at]] = NoTileSlot;
            ++g.evictions;
        }
        g.loadSeq[at] = ++g.loads;
        // the launch this commit precedes reads it (its parked samples asked for it): a later commit, before that
        // launch's stamps are read back, must not take it for a stale one
        g.lastUse[tile] = std::max(g.lastUse[tile], uint32_t(d->launches + 1));
        g.tileIn[at] = tile;
        g.slotOf[tile] = at;
        g.queued[tile] = 0;
        std::memcpy(g.upload[slot] + i * TexTileTexels * 3, g.hostTiles.data() + size_t(tile) * TexTileTexels * 3,
                    tileBytes);
        X2_HIP(hipMemcpyAsync(g.pool + size_t(at) * TexTileTexels * 3, g.upload[slot] + i * TexTileTexels * 3,
                              tileBytes, hipMemcpyHostToDevice, d->stream));
        ++g.uploads;
    }
    g.pending.erase(g.pending.begin(), g.pending.begin() + std::ptrdiff_t(n));
    std::memcpy(g.tableStaging[slot], g.slotOf.data(), size_t(g.tiles) * sizeof(uint32_t));
    X2_HIP(hipMemcpyAsync(g.tileSlot, g.tableStaging[slot], size_t(g.tiles) * sizeof(uint32_t), hipMemcpyHostToDevice,
                          d->stream));
    X2_HIP(hipMemsetAsync(g.tileAsked, 0, size_t(g.tiles) * sizeof(uint32_t), d->stream));
    return X2_OK;
}
}  // namespace

int x2dev_page_textures(x2dev* d, const x2_paging* o) {
    if (!d || !o) return fail(X2_ERR_ARGUMENT, "null device or paging");
    if (!d->prepared) return fail(X2_ERR_STATE, "paging before prepare");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "paging with launches in flight");
    if (d->pg.on) return fail(X2_ERR_STATE, "the scene's textures are paged already");
    if (d->texDesc.empty()) return fail(X2_ERR_STATE, "the scene has no texture to page");
    if (d->volDesc.table) return fail(X2_ERR_STATE, "paged textures: the medium kernel reads resident textures");
    if (d->opt.wave_slots) return fail(X2_ERR_STATE, "paged textures: the wavefront form reads resident textures");
    if (o->pool_tiles == 0 || o->pool_tiles > (1u << 22)) return fail(X2_ERR_ARGUMENT, "pool_tiles must be 1..2^22");
    if (o->host_refines && !o->partial) return fail(X2_ERR_ARGUMENT, "host_refines needs partial shading");
    const uint32_t coarseSide = o->coarse_side ? o->coarse_side : 64u;
    X2_HIP(hipSetDevice(d->device));
    Paging g;
    g.on = true;
    g.partial = o->partial != 0;
    g.hostRefines = o->host_refines != 0;
    g.poolTiles = o->pool_tiles;
    g.missCap = 65536;
    g.uploadCap = std::min<uint32_t>(o->pool_tiles, 1024u);
    std::vector<TexDesc> desc = d->texDesc;
    std::vector<float> levels;
    for (size_t t = 0; t < desc.size(); ++t) {
        TexDesc& td = desc[t];
        levels.resize(td.levels[td.count - 1].offset + 3);
        X2_HIP(hipMemcpy(levels.data(), d->texels[t], levels.size() * sizeof(float), hipMemcpyDeviceToHost));
        // Every level in half floats (a float file's levels are held so, texture_core.h:28-31): the levels wider or
        // taller than coarseSide as 32 x 32 tiles on demand, the coarser ones always resident (async_gpu_device.h:
        // 49-51). A partial tile's texels past the level's edge are zero (never read: fetches wrap inside it).
        uint32_t first = 0;
        while (first + 1 < td.count && std::max(td.levels[first].width, td.levels[first].height) > coarseSide) ++first;
        td.firstResident = first;
        for (uint32_t k = 0; k < first; ++k) {
            TexLevel& l = td.levels[k];
            const uint32_t w = l.width, h = l.height, tx = (w + TexTile - 1) / TexTile, ty = (h + TexTile - 1) / TexTile;
            l.tileBase = g.tiles;
            l.tilesX = tx;
            g.tiles += tx * ty;
            const size_t base = g.hostTiles.size();
            g.hostTiles.resize(base + size_t(tx) * ty * TexTileTexels * 3, 0);
            const float* src = levels.data() + l.offset;
            for (uint32_t y = 0; y < h; ++y)
                for (uint32_t x = 0; x < w; ++x) {
                    const size_t tile = size_t(y / TexTile) * tx + x / TexTile;
                    uint16_t* dst = g.hostTiles.data() + base + (tile * TexTileTexels + (y % TexTile) * TexTile + x % TexTile) * 3;
                    for (int c = 0; c < 3; ++c) dst[c] = x2::toHalf(src[(size_t(y) * w + x) * 3 + c]);
                }
        }
        const unsigned long long from = td.levels[first].offset;
        std::vector<uint16_t> coarse(levels.size() - from);
        for (size_t i = 0; i < coarse.size(); ++i) coarse[i] = x2::toHalf(levels[from + i]);
        for (uint32_t k = first; k < td.count; ++k) td.levels[k].offset -= from;  // now into the coarse halves
        uint16_t* dc = nullptr;
        X2_HIP(hipMalloc(&dc, coarse.size() * sizeof(uint16_t)));
        g.coarse.push_back(dc);
        X2_HIP(hipMemcpy(dc, coarse.data(), coarse.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        td.coarse = dc;
        td.paged = 1;
        td.texels = nullptr;
    }
    if (!g.tiles) {
        for (uint16_t* c : g.coarse) (void)hipFree(c);
        return fail(X2_ERR_ARGUMENT, "coarse_side leaves no level to page (every level is no larger)");
    }
    g.slotOf.assign(g.tiles, NoTileSlot);
    g.tileIn.assign(g.poolTiles, NoTileSlot);
    g.loadSeq.assign(g.poolTiles, 0);
    g.lastUse.assign(g.tiles, 0);
    g.queued.assign(g.tiles, 0);
    const size_t tileBytes = size_t(TexTileTexels) * 3 * sizeof(uint16_t);
    X2_HIP(hipMalloc(&g.tileSlot, size_t(g.tiles) * sizeof(uint32_t)));
    X2_HIP(hipMalloc(&g.tileAsked, size_t(g.tiles) * sizeof(uint32_t)));
    X2_HIP(hipMalloc(&g.tileStamp, size_t(g.tiles) * sizeof(uint32_t)));
    X2_HIP(hipMemset(g.tileStamp, 0, size_t(g.tiles) * sizeof(uint32_t)));
    X2_HIP(hipMalloc(&g.missList, size_t(g.missCap) * sizeof(uint32_t)));
    X2_HIP(hipMalloc(&g.pool, size_t(g.poolTiles) * tileBytes));
    X2_HIP(hipMalloc(&g.state, sizeof(PagingState)));
    X2_HIP(hipMemcpy(g.tileSlot, g.slotOf.data(), size_t(g.tiles) * sizeof(uint32_t), hipMemcpyHostToDevice));
    X2_HIP(hipMemset(g.tileAsked, 0, size_t(g.tiles) * sizeof(uint32_t)));
    X2_HIP(hipMemset(g.state, 0, sizeof(PagingState)));
    X2_HIP(hipHostMalloc(&g.readback, Slots * sizeof(PagingState), hipHostMallocDefault));
    for (uint32_t i = 0; i < Slots; ++i) {
        X2_HIP(hipHostMalloc(&g.missReadback[i], size_t(g.missCap) * sizeof(uint32_t), hipHostMallocDefault));
        X2_HIP(hipHostMalloc(&g.stampReadback[i], size_t(g.tiles) * sizeof(uint32_t), hipHostMallocDefault));
        X2_HIP(hipHostMalloc(&g.upload[i], size_t(g.uploadCap) * tileBytes, hipHostMallocDefault));
        X2_HIP(hipHostMalloc(&g.tableStaging[i], size_t(g.tiles) * sizeof(uint32_t), hipHostMallocDefault));
        if (g.hostRefines) {
            X2_HIP(hipMalloc(&g.refineIn[i], size_t(X2_REFINES_PER_LAUNCH) * sizeof(x2_refine)));
            X2_HIP(hipHostMalloc(&g.refineStaging[i], size_t(X2_REFINES_PER_LAUNCH) * sizeof(x2_refine), hipHostMallocDefault));
        }
    }
    // The descriptors now read the tiles; level 0's resident copies are let go.
    X2_HIP(hipMemcpy(d->textures, desc.data(), desc.size() * sizeof(TexDesc), hipMemcpyHostToDevice));
    for (float*& t : d->texels) {
        (void)hipFree(t);
        t = nullptr;
    }
    d->texDesc = desc;
    releasePaging(d);
    d->pg = std::move(g);
    d->begun_ = false;  // a frame begun before paging reads the old descriptors: begin again
    return X2_OK;
}

int x2dev_paging_stats(const x2dev* d, x2_paging_stats* out) {
    if (!d || !out) return fail(X2_ERR_ARGUMENT, "null device or stats");
    std::memset(out, 0, sizeof *out);
    const Paging& g = d->pg;
    if (!g.on) return X2_OK;
    out->parked_outstanding = g.parkTail - g.parkIssued;
    out->parked_total = g.parkedTotal;
    out->provisional = g.provisional;
    out->misses = g.misses;
    out->uploads = g.uploads;
    out->evictions = g.evictions;
    out->overflow = g.overflow;
    out->tiles = g.tiles;
    out->pool_tiles = g.poolTiles;
    uint32_t resident = 0;
    for (uint32_t t : g.tileIn) resident += t != NoTileSlot;
    out->resident = resident;
    return X2_OK;
}

int x2dev_provisional_film(const x2dev* d, const double** sums, const uint32_t** counts) {
    if (!d || !sums || !counts) return fail(X2_ERR_ARGUMENT, "null argument");
    if (!d->pg.on || !d->pg.hostProvSums) return fail(X2_ERR_STATE, "no provisional film: the textures are not paged");
    *sums = d->pg.hostProvSums;
    *counts = d->pg.hostProvCounts;
    return X2_OK;
}

int x2dev_launch(x2dev* d, const x2_chunk* chunks, uint32_t count, uint64_t samples) {
    return x2dev_launch_refines(d, chunks, count, nullptr, 0, samples);
}
int x2dev_launch_refines(x2dev* d, const x2_chunk* chunks, uint32_t count, const x2_refine* refines,
                         uint32_t refine_count, uint64_t samples) {
    if (!d) return fail(X2_ERR_ARGUMENT, "null device");
    if (!d->begun_) return fail(X2_ERR_STATE, "launch before begin");
    if (refine_count) {
        if (!refines) return fail(X2_ERR_ARGUMENT, "null refines");
        if (!d->pg.on || !d->pg.hostRefines) return fail(X2_ERR_STATE, "provisional samples to render again, but no host refines");
        if (refine_count > X2_REFINES_PER_LAUNCH)
            return fail(X2_ERR_ARGUMENT, "Too many provisional samples for one launch (async_gpu.h:1436)");
        const uint32_t pixels = d->cfg.width * d->cfg.height;
        for (uint32_t i = 0; i < refine_count; ++i)
            if (refines[i].pixel >= pixels) return fail(X2_ERR_ARGUMENT, "a provisional sample's pixel is outside the image");
    }
    const uint64_t parkedCount = d->pg.on ? d->pg.parkTail - d->pg.parkIssued : 0;
    if (!count && !parkedCount && !refine_count) return 0;  // nothing to run: no chunks, parked or provisional samples
    if (count && !chunks) return fail(X2_ERR_ARGUMENT, "null chunks");
    if (d->inFlight.size() >= 2 || count > d->opt.max_chunks_per_launch)
        return fail(X2_ERR_BACKPRESSURE, "GPU launch backpressure (async_gpu.h:1435)");
    if (d->published + count - oldestChunk(d) > d->opt.ring_chunks) return fail(X2_ERR_BACKPRESSURE, "mailbox ring full");
    const uint32_t spc = d->itemsPerChunk / TilePixels;
    const uint32_t tiles = ((d->cfg.width + 7) / 8) * ((d->cfg.height + 7) / 8);
    for (uint32_t i = 0; i < count; ++i)
        if (chunks[i].count > spc || chunks[i].tile >= tiles || chunks[i].first > 0xffffffffu - chunks[i].count)
            return fail(X2_ERR_ARGUMENT, "chunk out of range (tile, or count above samples_per_chunk)");
    X2_HIP(hipSetDevice(d->device));
    const uint32_t slot = uint32_t(d->launches % Slots);
    const bool wave = waveForm(d);
    uint64_t parkedBegin = 0;
    const uint64_t items = refine_count + parkedCount + uint64_t(count) * d->itemsPerChunk;
    if (d->pg.on) {
        Paging& g = d->pg;
        if (wave) return fail(X2_ERR_STATE, "paged textures: the wavefront form reads resident textures");
        if (parkedCount > 0xffffffffu) return fail(X2_ERR_BACKPRESSURE, "too many parked samples for one launch");
        if (g.parkTail + g.inflightItems + items > g.parkIssued + g.parkCap)
            return fail(X2_ERR_BACKPRESSURE, "parked-sample ring would overflow (async_gpu.h:1447)");
        if (g.hostRefines && g.inflightItems + items > g.refineCap)
            return fail(X2_ERR_BACKPRESSURE, "Provisional-sample ring would overflow (async_gpu.h:1444)");
        const int r = commitTiles(d, slot);  // the tiles asked for so far, before the samples that wait for them
        if (r) return r;
        parkedBegin = g.parkIssued;
        g.parkIssued += parkedCount;
        g.inflightItems += items;
    }
    if (count) std::memcpy(d->staging[slot], chunks, size_t(count) * sizeof(x2_chunk));
    for (uint32_t done = 0; done < count;) {  // the ring wraps
        const uint64_t at = (d->published + done) % d->opt.ring_chunks;
        const uint32_t n = uint32_t(std::min<uint64_t>(count - done, d->opt.ring_chunks - at));
        X2_HIP(hipMemcpyAsync(d->ring + at, d->staging[slot] + done, size_t(n) * sizeof(x2_chunk), hipMemcpyHostToDevice,
                              d->stream));
        done += n;
    }
    if (refine_count) {
        std::memcpy(d->pg.refineStaging[slot], refines, size_t(refine_count) * sizeof(x2_refine));
        X2_HIP(hipMemcpyAsync(d->pg.refineIn[slot], d->pg.refineStaging[slot], size_t(refine_count) * sizeof(x2_refine),
                              hipMemcpyHostToDevice, d->stream));
    }
    X2_HIP(launchClearState(d->state + slot, d->stream));
    RenderParams p = renderParams(d, slot);
    p.itemBegin = d->published * d->itemsPerChunk;
    p.itemEnd = (d->published + count) * d->itemsPerChunk;
    p.parkedBegin = parkedBegin;
    p.parkedCount = uint32_t(parkedCount);
    p.refineIn = d->pg.refineIn[slot];
    p.refineInCount = refine_count;
    p.refineHead = d->pg.refineHead;
    p.stamp = uint32_t(d->launches + 1);
    if (wave) {
        const uint32_t perVertex = d->cfg.light_count * std::max(1u, d->cfg.light_samples) +
                                   2u * std::max(1u, d->cfg.bsdf_samples);
        const int r = ensureWave(d, perVertex);
        if (r) return r;
    }
    X2_HIP(hipEventRecord(d->begun[slot], d->stream));
    if (wave) {
        // Iterations: each busy slot shades one vertex an iteration, plus one to finish a sample; the drain takes
        // whatever is left, so the estimate only sets how much of the launch runs in the wavefront form.
        const double perItem = (d->waveVertices > 0 ? d->waveVertices : 2.0) + 1.0;
        const double want = double(count) * d->itemsPerChunk * perItem / double(d->wave.slots);
        const uint32_t iterations = uint32_t(std::min(4096.0, std::ceil(want))) + 8u;
        X2_HIP(launchWave(p, d->wave, iterations, d->stream));
    } else {
        X2_HIP(d->dImage ? launchAdjoint(p, d->stream) : launchRender(p, d->stream));
    }
    X2_HIP(hipEventRecord(d->ended[slot], d->stream));
    X2_HIP(hipMemcpyAsync(d->readback + slot, d->state + slot, sizeof(LaunchState), hipMemcpyDeviceToHost, d->stream));
    if (wave)
        X2_HIP(hipMemcpyAsync(d->waveReadback + slot, d->wave.counts, sizeof(WaveCounts), hipMemcpyDeviceToHost,
                              d->stream));
    if (d->pg.on) {  // the launch's paging counters and miss list, then its counters cleared for the next launch
        Paging& g = d->pg;
        X2_HIP(hipMemcpyAsync(g.readback + slot, g.state, sizeof(PagingState), hipMemcpyDeviceToHost, d->stream));
        X2_HIP(hipMemcpyAsync(g.missReadback[slot], g.missList, size_t(g.missCap) * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, d->stream));
        X2_HIP(hipMemcpyAsync(g.stampReadback[slot], g.tileStamp, size_t(g.tiles) * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, d->stream));
        X2_HIP(hipMemsetAsync(reinterpret_cast<char*>(g.state) + offsetof(PagingState, missCount), 0,
                              sizeof(PagingState) - offsetof(PagingState, missCount), d->stream));
    }
    X2_HIP(hipEventRecord(d->done[slot], d->stream));
    Flight flight{d->launches, d->published, count, samples, slot, wave};
    flight.parkedBegin = parkedBegin;
    flight.parkedCount = parkedCount;
    flight.items = items;
    flight.refines = refine_count;
    d->inFlight.push_back(flight);
    d->published += count;
    ++d->launches;
    return 1;
}

// The allocation-free render (no counterpart in the reference): x2dev_launch's render with the chunks copied into
// the static `scratch` array instead of the lane's ring, its counters there instead of the lane's launch slots, and
// the caller's film instead of the lane's. No hipMalloc, hipFree or other allocating call: the kernels take this
// device's address of `scratch` as their ring and state arguments. Calls that overlap on one GPU share the array and
// race; each call clears the counters before its kernel runs.
int x2dev_render_chunks_static(x2dev* d, const x2_config* config, uint32_t samples_per_chunk, const x2_chunk* chunks,
                               uint32_t count, double* sums, uint32_t* counts, x2_completion* out) {
    if (!d || !config || (count && !chunks) || !sums || !counts) return fail(X2_ERR_ARGUMENT, "null argument");
    if (count > X2_STATIC_MAX_CHUNKS)
        return fail(X2_ERR_ARGUMENT, "x2dev_render_chunks_static takes at most X2_STATIC_MAX_CHUNKS (65536) chunks");
    if (samples_per_chunk == 0 || samples_per_chunk > 1024) return fail(X2_ERR_ARGUMENT, "samples_per_chunk must be 1..1024");
    if (!d->prepared) return fail(X2_ERR_STATE, "render before prepare");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "x2dev_render_chunks_static with launches in flight");
    if (d->pg.on) return fail(X2_ERR_STATE, "x2dev_render_chunks_static reads resident textures only");
    if (!d->staticScratch) return fail(X2_ERR_STATE, "the static scratch array was not looked up");
    const int v = validateConfig(d, *config);
    if (v) return v;
    const uint32_t tiles = ((config->width + 7) / 8) * ((config->height + 7) / 8);
    for (uint32_t i = 0; i < count; ++i)
        if (chunks[i].count > samples_per_chunk || chunks[i].tile >= tiles || chunks[i].first > 0xffffffffu - chunks[i].count)
            return fail(X2_ERR_ARGUMENT, "chunk out of range (tile, or count above samples_per_chunk)");
    if (out) *out = x2_completion{};
    if (!count) return X2_OK;
    X2_HIP(hipSetDevice(d->device));
    LaunchState* state = reinterpret_cast<LaunchState*>(d->staticScratch);
    x2_chunk* ring = reinterpret_cast<x2_chunk*>(d->staticScratch + StaticChunkOffset);
    X2_HIP(hipMemcpyAsync(ring, chunks, size_t(count) * sizeof(x2_chunk), hipMemcpyHostToDevice, d->stream));
    X2_HIP(launchClearState(state, d->stream));  // a fill kernel: the last call's counters and claims are gone
    RenderParams p = renderParams(d, 0);
    p.cfg = *config;
    p.failPixel = simulated("XPU2_SIMULATE_SAMPLE_FAIL");
    p.dImage = nullptr;
    p.gradTextures = nullptr;
    p.gradRoughness = nullptr;
    p.ring = ring;
    p.ringCapacity = X2_STATIC_MAX_CHUNKS;
    p.sums = sums;
    p.counts = counts;
    p.state = state;
    p.itemsPerChunk = samples_per_chunk * TilePixels;
    p.itemBegin = 0;
    p.itemEnd = uint64_t(count) * p.itemsPerChunk;
    p.adaptive = 0;
    p.odd = nullptr;
    for (uint32_t i = 0; i < PassSlots; ++i) {
        p.passMask[i] = nullptr;
        p.passBase[i] = nullptr;
    }
    // the lane's slot-0 events and readback are free: no launch of the lane is in flight
    X2_HIP(hipEventRecord(d->begun[0], d->stream));
    X2_HIP(launchRender(p, d->stream));
    X2_HIP(hipEventRecord(d->ended[0], d->stream));
    X2_HIP(hipMemcpyAsync(d->readback, state, sizeof(LaunchState), hipMemcpyDeviceToHost, d->stream));
    X2_HIP(hipStreamSynchronize(d->stream));
    if (out) {
        const LaunchState& s = d->readback[0];
        float ms = 0.0f;
        X2_HIP(hipEventElapsedTime(&ms, d->begun[0], d->ended[0]));
        out->chunks = count;
        out->items = p.itemEnd;
        uint64_t samples = 0;
        for (uint32_t i = 0; i < count; ++i) {  // in-image pixels of the chunk's tile, times its sample indices
            const uint32_t tilesX = (config->width + 7) / 8, tx = chunks[i].tile % tilesX, ty = chunks[i].tile / tilesX;
            const uint64_t w = std::min(8u, config->width - tx * 8), h = std::min(8u, config->height - ty * 8);
            samples += w * h * chunks[i].count;
        }
        out->samples = samples;
        out->errors = s.errors;
        out->device_seconds = double(ms) * 1e-3;
        out->rays_trace = s.traces;
        out->rays_shadow = s.shadows;
        out->nodes_visited = s.nodes;
        out->triangles_tested = s.triangles;
        out->voxel_taps = s.voxels;
    }
    return X2_OK;
}

int x2dev_poll(x2dev* d, x2_completion* out) {
    if (!d || !out) return fail(X2_ERR_ARGUMENT, "null device or completion");
    if (d->inFlight.empty()) return 0;
    const Flight f = d->inFlight.front();
    const hipError_t q = hipEventQuery(d->done[f.slot]);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) return hipFail(q, "hipEventQuery");
    d->inFlight.pop_front();
    const LaunchState& s = d->readback[f.slot];
    uint32_t lost = 0;
    if (d->pg.on) {  // the samples it parked, the tiles it asked for
        Paging& g = d->pg;
        const PagingState& rb = g.readback[f.slot];
        const uint64_t tail = std::min<uint64_t>(rb.parkTail, f.parkedBegin + g.parkCap);
        if (tail > g.parkTail) {
            g.parkedTotal += tail - g.parkTail;
            g.parkTail = tail;
        }
        g.inflightItems -= std::min(g.inflightItems, f.items);
        g.provisional += rb.provisional;
        g.overflow += rb.parkOverflow + rb.refineOverflow;
        lost = rb.parkOverflow + rb.refineOverflow;
        if (g.hostRefines) {  // the provisional records it wrote, to the host (async_gpu.h:1520-1525)
            const uint64_t tail = std::min<uint64_t>(rb.refineTail, g.refineHead + g.refineCap);
            while (g.refineHead < tail) {
                const uint64_t at = g.refineHead % g.refineCap, n = std::min<uint64_t>(tail - g.refineHead, g.refineCap - at);
                X2_HIP(hipMemcpy(g.refineReadback, g.refineRing + at, size_t(n) * sizeof(x2_refine), hipMemcpyDeviceToHost));
                g.refinesOut.insert(g.refinesOut.end(), g.refineReadback, g.refineReadback + n);
                g.refineHead += n;
            }
            g.refineHead = std::max<uint64_t>(g.refineHead, rb.refineTail);  // records past the ring's room were counted lost
        }
        for (uint32_t t = 0; t < g.tiles; ++t) g.lastUse[t] = std::max(g.lastUse[t], g.stampReadback[f.slot][t]);
        const uint32_t n = std::min(rb.missCount, g.missCap);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t tile = g.missReadback[f.slot][i];
            if (tile >= g.tiles || g.queued[tile] || g.slotOf[tile] != NoTileSlot) continue;
            g.queued[tile] = 1;
            g.pending.push_back(tile);
            ++g.misses;
        }
    }
    if (f.wave) {
        const double v = double(d->waveReadback[f.slot].traced) / double(f.count * uint64_t(d->itemsPerChunk));
        d->waveVertices = d->waveVertices > 0 ? 0.5 * (d->waveVertices + v) : v;
    }
    float ms = 0;
    X2_HIP(hipEventElapsedTime(&ms, d->begun[f.slot], d->ended[f.slot]));
    out->chunks = f.count;
    out->items = f.count * uint64_t(d->itemsPerChunk);
    out->samples = f.samples;
    out->parked_run = f.parkedCount;
    out->refines_run = f.refines;
    out->errors = s.errors + lost;  // a sample the parked ring could not hold is lost: a failure
    out->pad = 0;
    out->device_seconds = ms / 1000.0;
    out->rays_trace = s.traces;
    out->rays_shadow = s.shadows;
    out->nodes_visited = s.nodes;
    out->triangles_tested = s.triangles;
    out->voxel_taps = s.voxels;
    return 1;
}

int x2dev_take_refines(x2dev* d, x2_refine* out, uint32_t capacity, uint32_t* taken) {
    if (!d || !taken || (capacity && !out)) return fail(X2_ERR_ARGUMENT, "null device, output or count");
    auto& q = d->pg.refinesOut;
    const uint32_t n = uint32_t(std::min<size_t>(capacity, q.size()));
    std::copy(q.begin(), q.begin() + n, out);
    q.erase(q.begin(), q.begin() + n);
    *taken = n;
    return X2_OK;
}
uint64_t x2dev_refines_waiting(const x2dev* d) { return d ? d->pg.refinesOut.size() : 0u; }

int x2dev_take_deferred(x2dev* d, x2_parked* out, uint32_t capacity, uint32_t* taken) {
    (void)out;
    (void)capacity;
    if (!d || !taken) return fail(X2_ERR_ARGUMENT, "null device or count");
    *taken = 0;  // this device renders every sample it claims: nothing is deferred to the CPU lane
    return X2_OK;
}

int x2dev_download_film(x2dev* d) {
    if (!d) return fail(X2_ERR_ARGUMENT, "null device");
    if (!d->begun_) return fail(X2_ERR_STATE, "download before begin");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "Film download before the last launch was collected");
    X2_HIP(hipSetDevice(d->device));
    X2_HIP(hipMemcpyAsync(d->hostSums, d->sums, d->filmPixels * 4 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    X2_HIP(hipMemcpyAsync(d->hostCounts, d->counts, d->filmPixels * sizeof(uint32_t), hipMemcpyDeviceToHost, d->stream));
    if (d->adaptive)
        X2_HIP(hipMemcpyAsync(d->hostOdd, d->odd, d->filmPixels * 4 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    if (d->pg.on && d->pg.provPixels == d->filmPixels) {
        X2_HIP(hipMemcpyAsync(d->pg.hostProvSums, d->pg.provSums, d->filmPixels * 4 * sizeof(double),
                              hipMemcpyDeviceToHost, d->stream));
        X2_HIP(hipMemcpyAsync(d->pg.hostProvCounts, d->pg.provCounts, d->filmPixels * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, d->stream));
    }
    X2_HIP(hipEventRecord(d->filmDone, d->stream));
    d->filmQueued = true;
    return X2_OK;
}

int x2dev_film_ready(x2dev* d) {
    if (!d || !d->filmQueued) return 0;
    const hipError_t q = hipEventQuery(d->filmDone);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) return hipFail(q, "hipEventQuery");
    return 1;
}

int x2dev_film(const x2dev* d, const double** sums, const uint32_t** counts) {
    if (!d || !sums || !counts) return fail(X2_ERR_ARGUMENT, "null argument");
    *sums = d->hostSums;
    *counts = d->hostCounts;
    return X2_OK;
}

int x2dev_seed_film(x2dev* d, const double* sums, const double* odd, const uint32_t* counts) {
    if (!d || !sums || !counts) return fail(X2_ERR_ARGUMENT, "null device, sums or counts");
    if (!d->begun_) return fail(X2_ERR_STATE, "seed before begin");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "seed with launches in flight");
    X2_HIP(hipSetDevice(d->device));
    const size_t pixels = d->filmPixels;
    X2_HIP(hipMemcpyAsync(d->sums, sums, pixels * 4 * sizeof(double), hipMemcpyHostToDevice, d->stream));
    X2_HIP(hipMemcpyAsync(d->counts, counts, pixels * sizeof(uint32_t), hipMemcpyHostToDevice, d->stream));
    if (odd && d->adaptive)
        X2_HIP(hipMemcpyAsync(d->odd, odd, pixels * 4 * sizeof(double), hipMemcpyHostToDevice, d->stream));
    X2_HIP(hipStreamSynchronize(d->stream));  // the caller's buffers may go once this returns
    return X2_OK;
}

int x2dev_odd(const x2dev* d, const double** odd) {
    if (!d || !odd) return fail(X2_ERR_ARGUMENT, "null argument");
    if (!d->adaptive) return fail(X2_ERR_STATE, "the odd-index film is kept in adaptive frames only");
    *odd = d->hostOdd;
    return X2_OK;
}

int x2dev_upload_pass(x2dev* d, uint32_t slot, const uint64_t* masks, uint32_t uniform_base) {
    if (!d || !masks) return fail(X2_ERR_ARGUMENT, "null device or masks");
    if (!d->begun_ || !d->adaptive) return fail(X2_ERR_STATE, "pass tables before an adaptive begin");
    if (slot >= PassSlots) return fail(X2_ERR_ARGUMENT, "pass slot out of range (4 slots)");
    X2_HIP(hipSetDevice(d->device));
    X2_HIP(hipEventSynchronize(d->passCopied[slot]));  // the slot's last upload has left its staging
    std::memcpy(d->passStaging[slot], masks, d->adaptiveTiles * sizeof(unsigned long long));
    X2_HIP(hipMemcpyAsync(d->passMask[slot], d->passStaging[slot], d->adaptiveTiles * sizeof(unsigned long long),
                          hipMemcpyHostToDevice, d->stream));
    if (uniform_base != X2_NO_SLOT)
        X2_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(d->passBase[slot]), int(uniform_base),
                                 d->adaptivePixels, d->stream));
    else  // a later pass starts from the counts every earlier pass has left
        X2_HIP(hipMemcpyAsync(d->passBase[slot], d->counts, d->adaptivePixels * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, d->stream));
    X2_HIP(hipEventRecord(d->passCopied[slot], d->stream));
    return X2_OK;
}

int x2dev_request_errors(x2dev* d) {
    if (!d) return fail(X2_ERR_ARGUMENT, "null device");
    if (!d->begun_ || !d->adaptive) return fail(X2_ERR_STATE, "errors before an adaptive begin");
    X2_HIP(hipSetDevice(d->device));
    X2_HIP(launchErrors(d->sums, d->odd, d->counts, uint32_t(d->adaptivePixels), d->errors, d->stream));
    X2_HIP(hipMemcpyAsync(d->hostErrors, d->errors, d->adaptivePixels * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    X2_HIP(hipEventRecord(d->errorsDone, d->stream));
    d->errorsQueued = true;
    return X2_OK;
}

int x2dev_errors_ready(x2dev* d) {
    if (!d || !d->errorsQueued) return 0;
    const hipError_t q = hipEventQuery(d->errorsDone);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) return hipFail(q, "hipEventQuery");
    return 1;
}

int x2dev_errors(const x2dev* d, const float** errors) {
    if (!d || !errors) return fail(X2_ERR_ARGUMENT, "null argument");
    if (!d->adaptive) return fail(X2_ERR_STATE, "errors are kept in adaptive frames only");
    *errors = d->hostErrors;
    return X2_OK;
}

int x2dev_device_film(const x2dev* d, void** sums, void** counts) {
    if (!d || !sums || !counts) return fail(X2_ERR_ARGUMENT, "null argument");
    *sums = d->sums;
    *counts = d->counts;
    return X2_OK;
}

void x2dev_destroy(x2dev* d) { release(d); }

int x2dev_set_texture(x2dev* d, uint32_t texture, const float* texels, int texels_on_device) {
    if (!d || !texels) return fail(X2_ERR_ARGUMENT, "null device or texels");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "texture update with launches in flight");
    if (texture >= d->texDesc.size()) return fail(X2_ERR_ARGUMENT, "no such texture");
    if (d->pg.on) return fail(X2_ERR_STATE, "a paged texture's texels are the tiles x2dev_page_textures made");
    const TexDesc& t = d->texDesc[texture];
    X2_HIP(hipSetDevice(d->device));
    X2_HIP(hipMemcpyAsync(d->texels[texture], texels, size_t(t.levels[0].width) * t.levels[0].height * 3 * sizeof(float),
                          texels_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, d->stream));
    X2_HIP(buildLevels(d, texture));  // the pyramid made again from the new level 0
    X2_HIP(hipStreamSynchronize(d->stream));
    return X2_OK;
}

int x2dev_rng_twin(x2dev* d, const x2_rng_query* queries, uint32_t n, uint32_t draws, uint32_t* words,
                   float* stream_out, uint64_t* branches) {
    if (!d || !queries || !words || !branches || (draws && !stream_out)) return fail(X2_ERR_ARGUMENT, "null argument");
    if (n == 0 || n > (1u << 20) || draws > 256) return fail(X2_ERR_ARGUMENT, "n must be 1..2^20, draws <= 256");
    X2_HIP(hipSetDevice(d->device));
    void *dq = nullptr, *dw = nullptr, *ds = nullptr, *db = nullptr;
    auto done = [&](int r) {
        (void)hipFree(dq);
        (void)hipFree(dw);
        (void)hipFree(ds);
        (void)hipFree(db);
        return r;
    };
    hipError_t e = hipMalloc(&dq, size_t(n) * sizeof(x2_rng_query));
    if (e == hipSuccess) e = hipMalloc(&dw, size_t(n) * 4);
    if (e == hipSuccess) e = hipMalloc(&ds, std::max<size_t>(4, size_t(n) * draws * 4));
    if (e == hipSuccess) e = hipMalloc(&db, size_t(n) * 8);
    if (e == hipSuccess) e = hipMemcpy(dq, queries, size_t(n) * sizeof(x2_rng_query), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launchRngTwin(static_cast<const x2_rng_query*>(dq), n, draws, static_cast<uint32_t*>(dw),
                          static_cast<float*>(ds), static_cast<uint64_t*>(db), d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e == hipSuccess) e = hipMemcpy(words, dw, size_t(n) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && draws) e = hipMemcpy(stream_out, ds, size_t(n) * draws * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(branches, db, size_t(n) * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(hipFail(e, "x2dev_rng_twin"));
    return done(X2_OK);
}

int x2dev_set_material(x2dev* d, uint32_t material, const x2_material* m) {
    if (!d || !m) return fail(X2_ERR_ARGUMENT, "null device or material");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "material update with launches in flight");
    if (material >= d->materialCount) return fail(X2_ERR_ARGUMENT, "no such material");
    if (m->albedo_map > d->texDesc.size()) return fail(X2_ERR_ARGUMENT, "albedo_map names no texture");
    X2_HIP(hipSetDevice(d->device));
    X2_HIP(hipMemcpyAsync(d->materials + material, m, sizeof(x2_material), hipMemcpyHostToDevice, d->stream));
    X2_HIP(hipStreamSynchronize(d->stream));
    return X2_OK;
}

int x2dev_resolve_image(x2dev* d, float* device_rgb) {
    if (!d || !device_rgb) return fail(X2_ERR_ARGUMENT, "null device or image");
    if (!d->begun_ || !d->filmPixels) return fail(X2_ERR_STATE, "no film to resolve");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "resolve with launches in flight");
    X2_HIP(hipSetDevice(d->device));
    X2_HIP(launchResolve(d->sums, d->counts, uint32_t(d->filmPixels), device_rgb, d->stream));
    X2_HIP(hipStreamSynchronize(d->stream));
    return X2_OK;
}

namespace {
// The broker's GPU-lane sequence over one uniform frame's chunks (async_render.h:318-1081 with the GPU lane alone):
// chunks in the work grid's order (async_grid.h:129-175: sample ranges 1, 1, 2, then samples_per_chunk; each range
// over the 64-pixel cells' tiles), launches of about windowMs at the measured rate (launchSize,
// async_render.h:1674-1718: firstChunks first, never under two items a resident thread), at most two in flight,
// completions polled until the last.
int runChunks(x2dev* d, const std::vector<x2_chunk>& chunks, const std::vector<uint32_t>* chunkSamples,
              x2_frame_stats& st);

int runFrame(x2dev* d, const x2_config* c, uint32_t spc, x2_frame_stats& st) {
    std::vector<x2_chunk> chunks;
    {
        const uint32_t tilesX = (c->width + 7) / 8, tilesY = (c->height + 7) / 8;
        const uint32_t cellsX = (c->width + 63) / 64, cellsY = (c->height + 63) / 64;
        std::vector<std::pair<uint32_t, uint32_t>> ranges;
        for (uint32_t first = 0; first < c->spp;) {
            const uint32_t size = first < 2 ? 1u : first < 4 ? 2u : spc;
            ranges.push_back({first, std::min(size, c->spp - first)});
            first += ranges.back().second;
        }
        // Cell by cell, each cell's sample ranges in order (the work grid's cell queues, async_grid.h:160-172): the
        // lanes in flight work a few cells at a time, whose BVH and triangles the Infinity Cache then holds (1.035x
        // over sweeping the whole image range by range on the headline frame).
        for (uint32_t cy = 0; cy < cellsY; ++cy)
            for (uint32_t cx = 0; cx < cellsX; ++cx) {
                for (const auto& rg : ranges)
                    for (uint32_t ty = cy * 8; ty < std::min(tilesY, (cy + 1) * 8); ++ty)
                        for (uint32_t tx = cx * 8; tx < std::min(tilesX, (cx + 1) * 8); ++tx)
                            chunks.push_back({ty * tilesX + tx, rg.first, rg.second, cy * cellsX + cx});
            }
    }
    return runChunks(d, chunks, nullptr, st);
}

// Launches the chunks in order, at most two launches in flight, sized by the measured rate, and collects every
// completion (chunkSamples: each chunk's samples, else every in-image pixel of its tile). A completion that counted
// an error ends the feed: nothing more is launched, the launches in flight are collected, and the caller fails the
// frame (async_render.h:679, 736). With host refines the provisional samples the launches make go back in the next.
int runChunks(x2dev* d, const std::vector<x2_chunk>& chunks, const std::vector<uint32_t>* chunkSamples,
              x2_frame_stats& st) {
    const x2_config* c = &d->cfg;
    // A launch's device time once the rate is known: long enough that max_chunks_per_launch bounds it at the
    // headline's rate (a launch's last samples leave lanes idle, so fewer, larger launches waste less).
    const double windowMs = 100.0;
    const uint64_t firstChunks = 2048;
    const uint64_t least = (uint64_t(d->launchThreads) * 2 + d->itemsPerChunk - 1) / d->itemsPerChunk;
    double rate = 0;  // items per device second, from completed launches
    size_t next = 0;
    int r = 0;
    // Host refines: the provisional samples the collected launches made, sent back in the next launches (the tiles
    // they found missing are sent up first, async_render.h:718-726).
    std::vector<x2_refine> refines;
    auto takeRefines = [&]() -> int {
        const uint64_t n = x2dev_refines_waiting(d);
        if (!n) return X2_OK;
        const size_t at = refines.size();
        refines.resize(at + n);
        uint32_t got = 0;
        const int e = x2dev_take_refines(d, refines.data() + at, uint32_t(n), &got);
        refines.resize(at + got);
        return e;
    };
    // Launch with up to a launch's refines and `want` chunks from `next`; the refines sent are dropped from the list.
    auto launchWith = [&](uint64_t want, uint64_t samples) -> int {
        const uint32_t n = uint32_t(std::min<size_t>(refines.size(), X2_REFINES_PER_LAUNCH));
        const int e = x2dev_launch_refines(d, want ? chunks.data() + next : nullptr, uint32_t(want), refines.data(), n, samples);
        if (e >= 0 && n) refines.erase(refines.begin(), refines.begin() + n);
        return e;
    };
    // Paged textures: parked samples run again (in the next launches, or in launches of them alone at the end) until
    // none is left; a stretch of launches of parked samples alone that ends none means the pool cannot hold what they
    // need (async_render.h:816-818).
    uint64_t fewest = ~0ull, idle = 0, rounds = 0;
    auto outstanding = [&]() { return x2dev_parked_outstanding(d) + refines.size() + x2dev_refines_waiting(d); };
    while (!d->inFlight.empty() || (!st.errors && (next < chunks.size() || outstanding()))) {
        x2_completion done;
        while ((r = x2dev_poll(d, &done)) == 1) {
            ++st.launches;
            st.chunks += done.chunks;
            st.samples += done.samples;
            st.errors += done.errors;
            st.device_seconds += done.device_seconds;
            st.rays_trace += done.rays_trace;
            st.rays_shadow += done.rays_shadow;
            st.nodes_visited += done.nodes_visited;
            st.triangles_tested += done.triangles_tested;
            st.voxel_taps += done.voxel_taps;
            if (done.device_seconds > 0) {
                const double now = double(done.items) / done.device_seconds;
                rate = rate ? 0.5 * rate + 0.5 * now : now;
            }
        }
        if (r < 0) return r;
        if ((r = takeRefines()) < 0) return r;
        if (st.errors) {  // the launches in flight finish; none follows them
            if (!d->inFlight.empty()) {
                const hipError_t e = hipEventSynchronize(d->done[d->inFlight.front().slot]);
                if (e != hipSuccess) return hipFail(e, "hipEventSynchronize");
            }
            continue;
        }
        if (next >= chunks.size() && x2dev_ready(d) && outstanding()) {
            const uint64_t left = outstanding() + d->pg.inflightItems;
            if (left < fewest) {
                fewest = left;
                idle = 0;
            } else {
                ++idle;
            }
            // A pool smaller than what a sample's path reads at once parks it again and again: the frame fails (as
            // the reference's does, async_render.h:816-818) rather than run on.
            if (idle > 256 || ++rounds > 4096)
                return fail(X2_ERR_STATE, "the tile pool cannot hold the parked samples' working set");
            if ((r = launchWith(0, 0)) < 0) return r;
        }
        while (x2dev_ready(d) && next < chunks.size()) {
            const uint64_t remaining = chunks.size() - next;
            uint64_t want = rate ? std::max<uint64_t>(uint64_t(rate * windowMs / 1000.0 / d->itemsPerChunk), least)
                                 : std::max<uint64_t>(firstChunks, least);
            const uint64_t refineChunks =
                (std::min<size_t>(refines.size(), X2_REFINES_PER_LAUNCH) + d->itemsPerChunk - 1) / d->itemsPerChunk;
            const uint64_t room = x2dev_chunk_room(d);
            want = std::min<uint64_t>({want, room > refineChunks ? room - refineChunks : 0, remaining});
            if (!want) break;  // the parked ring is full: the launches in flight make room
            uint64_t samples = 0;
            const uint32_t tilesX = (c->width + 7) / 8;
            for (uint64_t i = 0; i < want; ++i) {
                if (chunkSamples) {
                    samples += (*chunkSamples)[next + i];
                    continue;
                }
                const x2_chunk& ch = chunks[next + i];
                for (uint32_t p = 0; p < TilePixels; ++p) {
                    const uint32_t x = (ch.tile % tilesX) * 8 + p % 8, y = (ch.tile / tilesX) * 8 + p / 8;
                    if (x < c->width && y < c->height) samples += ch.count;
                }
            }
            r = launchWith(want, samples);
            if (r < 0) return r;
            next += want;
        }
        if (!d->inFlight.empty()) {
            const hipError_t e = hipEventSynchronize(d->done[d->inFlight.front().slot]);
            if (e != hipSuccess) return hipFail(e, "hipEventSynchronize");
        }
    }
    return X2_OK;
}
}  // namespace

// One uniform frame: runFrame, then the film downloaded and every pixel's count audited against spp
// (async_render.h:1035-1038).
int x2dev_render_frame(x2dev* d, const x2_config* c, uint32_t spc, x2_frame_stats* stats) {
    using clock = std::chrono::steady_clock;
    const auto start = clock::now();
    int r = x2dev_begin(d, c, spc);
    if (r) return r;
    x2_frame_stats st{};
    r = runFrame(d, c, spc, st);
    if (r) return r;
    if (st.errors) {  // before the audit, as the reference's frame fails (async_render.h:955-958)
        st.wall_seconds = std::chrono::duration<double>(clock::now() - start).count();
        if (stats) *stats = st;
        return fail(X2_ERR_STATE, "GPU sample failure; whole epoch must use fallback (async_render.h:736)");
    }
    r = x2dev_download_film(d);
    if (r) return r;
    X2_HIP(hipEventSynchronize(d->filmDone));
    for (uint32_t y = 0; y < c->height; ++y)
        for (uint32_t x = 0; x < c->width; ++x) {
            if (int(x) < c->viewport[0] || int(y) < c->viewport[1] || int(x) >= c->viewport[2] || int(y) >= c->viewport[3])
                continue;
            if (d->hostCounts[size_t(y) * c->width + x] != c->spp) ++st.audit_failures;
        }
    st.wall_seconds = std::chrono::duration<double>(clock::now() - start).count();
    if (stats) *stats = st;
    if (st.audit_failures) return fail(X2_ERR_STATE, "Sample audit failure (async_render.h:1035)");
    return X2_OK;
}

namespace {
// The device pass chain over a frame (GpuLane::beginChain / launchChain, async_gpu.h:1960-2085): the regions' bounds,
// node room (MoonRay's maxNodes) and summed-area tables, the tiles in cell order; then every step queued at once —
// the uniform stage as one step, each later pass as one — and one wait. `planned` receives each pixel's samples.
int runChain(x2dev* d, const adaptive::Regions& regions, const std::vector<adaptive::Pass>& passes,
             const std::vector<uint64_t>& inImage, uint32_t minSamples, float target, std::vector<uint32_t>& planned,
             x2_adaptive& a, x2_frame_stats& st) {
    const uint32_t width = d->cfg.width, height = d->cfg.height, tilesX = (width + 7) / 8, tilesY = (height + 7) / 8;
    const uint32_t tiles = tilesX * tilesY, cellsX = (width + 63) / 64, n = uint32_t(regions.count());
    const size_t pixels = size_t(width) * height;
    std::vector<ChainRegion> bounds(n);
    uint64_t nodes = 0, sat = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const adaptive::IBox b = regions.tree(i).root();
        const int w = b.x1 - b.x0, h = b.y1 - b.y0;
        const float cells = (float(w) / adaptive::MinNodeSize) * (float(h) / adaptive::MinNodeSize);
        const uint32_t room = uint32_t(std::max(1, 2 * int(ceilf(cells)) - 1));  // AdaptiveRegionTree::maxNodes
        bounds[i] = {b.x0, b.y0, b.x1, b.y1, uint32_t(nodes), room, sat};
        nodes += room;
        sat += uint64_t(w + 1) * uint64_t(h + 1);
    }
    std::vector<uint32_t> order;
    for (uint32_t cy = 0; cy < (tilesY + 7) / 8; ++cy)
        for (uint32_t cx = 0; cx < cellsX; ++cx)
            for (uint32_t ty = cy * 8; ty < std::min(tilesY, (cy + 1) * 8); ++ty)
                for (uint32_t tx = cx * 8; tx < std::min(tilesX, (cx + 1) * 8); ++tx) order.push_back(ty * tilesX + tx);
    // One allocation for the chain's tables (freed when the frame returns).
    const size_t sizes[] = {n * sizeof(ChainRegion), nodes * sizeof(adaptive::Node), sat * sizeof(double), n * 4,
                            tiles * 8, tiles * 8, tiles * 4, tiles * 4, tiles * 8, pixels * 4, pixels * 4,
                            sizeof(ChainState)};
    size_t offsets[12], total = 0;
    for (int i = 0; i < 12; ++i) {
        offsets[i] = total;
        total += (sizes[i] + 255) & ~size_t(255);
    }
    char* block = nullptr;
    X2_HIP(hipMalloc(&block, total));
    struct Free {
        char* p;
        ~Free() { (void)hipFree(p); }
    } release{block};
    ChainParams c{};
    c.sums = d->sums;
    c.odd = d->odd;
    c.counts = d->counts;
    c.errors = d->errors;
    c.regions = reinterpret_cast<const ChainRegion*>(block + offsets[0]);
    c.nodes = block + offsets[1];
    c.sat = reinterpret_cast<double*>(block + offsets[2]);
    c.regionDone = reinterpret_cast<uint32_t*>(block + offsets[3]);
    c.tileMask = reinterpret_cast<unsigned long long*>(block + offsets[4]);
    c.inImage = reinterpret_cast<const unsigned long long*>(block + offsets[5]);
    c.tileOrder = reinterpret_cast<const uint32_t*>(block + offsets[6]);
    c.activeTiles = reinterpret_cast<uint32_t*>(block + offsets[7]);
    c.activeMasks = reinterpret_cast<unsigned long long*>(block + offsets[8]);
    c.planned = reinterpret_cast<uint32_t*>(block + offsets[9]);
    c.base = reinterpret_cast<uint32_t*>(block + offsets[10]);
    c.state = reinterpret_cast<ChainState*>(block + offsets[11]);
    c.width = width;
    c.height = height;
    c.tiles = tiles;
    c.tilesX = tilesX;
    c.regionCount = n;
    c.perDim = uint32_t(regions.perDim());
    c.regionW = uint32_t(regions.regionWidth());
    c.regionH = uint32_t(regions.regionHeight());
    c.target = target;
    X2_HIP(hipMemcpyAsync(block + offsets[0], bounds.data(), sizes[0], hipMemcpyHostToDevice, d->stream));
    X2_HIP(hipMemcpyAsync(block + offsets[5], inImage.data(), sizes[5], hipMemcpyHostToDevice, d->stream));
    X2_HIP(hipMemcpyAsync(block + offsets[6], order.data(), sizes[6], hipMemcpyHostToDevice, d->stream));
    X2_HIP(hipMemsetAsync(block + offsets[3], 0, sizes[3], d->stream));
    X2_HIP(hipMemsetAsync(block + offsets[9], 0, sizes[9], d->stream));
    X2_HIP(hipMemsetAsync(block + offsets[11], 0, sizes[11], d->stream));
    X2_HIP(launchClearState(d->state, d->stream));
    RenderParams p = renderParams(d, 0);
    p.chain = c.state;
    p.chainTiles = c.activeTiles;
    p.chainMasks = c.activeMasks;
    p.chainBase = c.base;
    uint32_t uniformEnd = 0;
    while (uniformEnd < passes.size() && passes[uniformEnd].start < minSamples) ++uniformEnd;
    X2_HIP(hipEventRecord(d->begun[0], d->stream));
    if (uniformEnd) {
        p.chainSize = uniformEnd < passes.size() ? passes[uniformEnd].start : passes.back().end;
        X2_HIP(launchChainStep(c, p, true, d->stream));
    }
    for (uint32_t j = uniformEnd; j < passes.size(); ++j) {
        p.chainSize = passes[j].end - passes[j].start;
        X2_HIP(launchChainStep(c, p, false, d->stream));
    }
    X2_HIP(hipEventRecord(d->ended[0], d->stream));
    ChainState cs{};
    X2_HIP(hipMemcpyAsync(d->readback, d->state, sizeof(LaunchState), hipMemcpyDeviceToHost, d->stream));
    X2_HIP(hipMemcpyAsync(&cs, c.state, sizeof(ChainState), hipMemcpyDeviceToHost, d->stream));
    X2_HIP(hipMemcpyAsync(planned.data(), c.planned, pixels * 4, hipMemcpyDeviceToHost, d->stream));
    X2_HIP(hipStreamSynchronize(d->stream));
    float ms = 0;
    X2_HIP(hipEventElapsedTime(&ms, d->begun[0], d->ended[0]));
    const LaunchState& ls = d->readback[0];
    st.launches += 1;
    st.errors += ls.errors;
    st.device_seconds += ms / 1000.0;
    st.rays_trace += ls.traces;
    st.rays_shadow += ls.shadows;
    st.nodes_visited += ls.nodes;
    st.triangles_tested += ls.triangles;
    st.voxel_taps += ls.voxels;
    for (uint32_t v : planned) st.samples += v;
    // The passes rendered, counted as the reference counts a chain's (async_render.h:1069).
    a.passes = uniformEnd + (cs.passes ? cs.passes - (uniformEnd ? 1u : 0u) : 0u);
    a.tree_updates = cs.treeUpdates;
    return X2_OK;
}
}  // namespace

// One adaptive frame (the broker's adaptive passes, async_render.h:1219-1440, with the host's region trees,
// async_adaptive.h): MoonRay's schedule of passes; the passes below min_samples sample every pixel; before each later
// pass the device works out every pixel's error (only the errors cross to the host), the trees rebuild from them,
// and the pass samples the pixels their open leaves cover, each from its own count. A pass with no samples ends the
// frame. The film is downloaded and every pixel's count audited against the samples its passes planned.
int x2dev_render_adaptive(x2dev* d, const x2_config* c, uint32_t spc, x2_adaptive* a, x2_frame_stats* stats) {
    using clock = std::chrono::steady_clock;
    const auto start = clock::now();
    if (!d || !c || !a) return fail(X2_ERR_ARGUMENT, "null device, config or adaptive settings");
    if (a->max_samples == 0 || a->max_samples > 65535 || !(a->target_error > 0.0f))
        return fail(X2_ERR_ARGUMENT, "max_samples must be 1..65535 and target_error positive");
    if (a->cellular && a->device_chain) return fail(X2_ERR_ARGUMENT, "cellular generations run on the host, not the chain");
    if (a->device_chain && d->pg.on) return fail(X2_ERR_STATE, "the device pass chain reads resident textures");
    x2_config cfg = *c;
    cfg.spp = a->max_samples;
    int r = x2dev_begin_adaptive(d, &cfg, spc);
    if (r) return r;
    const uint32_t width = cfg.width, height = cfg.height, tilesX = (width + 7) / 8, tilesY = (height + 7) / 8;
    const uint32_t cellsX = (width + 63) / 64, tiles = tilesX * tilesY;
    const std::vector<adaptive::Pass> passes = adaptive::schedule(a->max_samples);
    if (passes.size() > MaxPasses) return fail(X2_ERR_ARGUMENT, "adaptive pass budget (160 passes)");
    adaptive::Regions regions(int(width), int(height), a->target_error);
    adaptive::Cellular cells(int(width), int(height), a->target_error);
    std::vector<uint64_t> inImage(tiles, 0), masks(tiles);
    for (uint32_t t = 0; t < tiles; ++t)
        for (uint32_t q = 0; q < TilePixels; ++q) {
            const uint32_t x = (t % tilesX) * 8 + q % 8, y = (t / tilesX) * 8 + q / 8;
            if (x < width && y < height && int(x) >= cfg.viewport[0] && int(y) >= cfg.viewport[1] &&
                int(x) < cfg.viewport[2] && int(y) < cfg.viewport[3])
                inImage[t] |= 1ull << q;
        }
    if (a->cellular) cells.deadForever(inImage);
    std::vector<uint32_t> planned(size_t(width) * height, 0);
    double spent = 0;  // the samples of the passes generated (a frame by sampling generations)
    x2_frame_stats st{};
    a->passes = 0;
    a->tree_updates = 0;
    if (a->device_chain) {
        r = runChain(d, regions, passes, inImage, a->min_samples, a->target_error, planned, *a, st);
        if (r) return r;
    }
    for (uint32_t j = 0; j < passes.size() && !a->device_chain; ++j) {
        const bool uniform = passes[j].start < a->min_samples;
        const uint32_t size = passes[j].end - passes[j].start;
        const bool budgeted = a->cellular && a->sample_budget;
        if (!uniform) {  // every pass before this one has completed: the trees (or cells) see all of them
            if (budgeted && spent >= double(a->sample_budget)) break;  // the budget is spent
            r = x2dev_request_errors(d);
            if (r) return r;
            X2_HIP(hipEventSynchronize(d->errorsDone));
            ++a->tree_updates;
            if (a->cellular) {
                const adaptive::Budget budget{&planned, std::max(0.0, double(a->sample_budget) - spent), a->max_samples,
                                              a->noise_floor};
                if (!cells.update(d->hostErrors, budgeted ? &budget : nullptr)) break;  // no tile lives: converged
            } else {
                regions.update(d->hostErrors);
            }
        }
        // A frame by sampling generations: a pass no larger than the samples left buys, the tiles furthest above the
        // target first (ties by tile), async_render.h:1372-1394.
        std::vector<char> dropped;
        if (!uniform && budgeted) {
            std::vector<std::pair<float, uint32_t>> live;
            double wanted = 0;
            for (uint32_t t = 0; t < tiles; ++t)
                if (inImage[t] && cells.alive(t)) {
                    live.push_back({cells.ratio(t), t});
                    wanted += double(__builtin_popcountll(inImage[t])) * size;
                }
            const double left = std::max(0.0, double(a->sample_budget) - spent);
            if (wanted > left) {
                std::stable_sort(live.begin(), live.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
                dropped.assign(tiles, 1);
                double taken = 0;
                for (const auto& [ratio, t] : live) {
                    const double cost = double(__builtin_popcountll(inImage[t])) * size;
                    if (taken + cost > left) break;
                    taken += cost;
                    dropped[t] = 0;
                }
            }
        }
        uint64_t samples = 0;
        std::vector<x2_chunk> chunks;
        std::vector<uint32_t> chunkSamples;
        for (uint32_t cy = 0; cy < (tilesY + 7) / 8; ++cy)  // the frame loop's cell order
            for (uint32_t cx = 0; cx < cellsX; ++cx)
                for (uint32_t ty = cy * 8; ty < std::min(tilesY, (cy + 1) * 8); ++ty)
                    for (uint32_t tx = cx * 8; tx < std::min(tilesX, (cx + 1) * 8); ++tx) {
                        const uint32_t t = ty * tilesX + tx;
                        uint64_t active = inImage[t];
                        if (!uniform && a->cellular)
                            active &= cells.alive(t) && (dropped.empty() || !dropped[t]) ? ~0ull : 0ull;
                        else if (!uniform)
                            active &= regions.sampleArea(int(tx * 8), int(ty * 8));
                        masks[t] = active;
                        if (!active) continue;
                        const uint32_t n = uint32_t(__builtin_popcountll(active));
                        for (uint32_t q = 0; q < TilePixels; ++q)
                            if (active >> q & 1) planned[size_t(ty * 8 + q / 8) * width + tx * 8 + q % 8] += size;
                        samples += uint64_t(n) * size;
                        for (uint32_t off = 0; off < size; off += spc) {
                            const uint32_t count = std::min(spc, size - off);
                            chunks.push_back({t, (j << PassShift) | off, count, cy * cellsX + cx});
                            chunkSamples.push_back(n * count);
                        }
                    }
        if (!samples) break;  // the trees have converged: nothing changes after this
        spent += double(samples);
        r = x2dev_upload_pass(d, j % PassSlots, masks.data(), uniform ? passes[j].start : X2_NO_SLOT);
        if (r) return r;
        r = runChunks(d, chunks, &chunkSamples, st);
        if (r) return r;
        ++a->passes;
        if (st.errors) break;
    }
    if (st.errors) {  // before the audit (async_render.h:955-958)
        st.wall_seconds = std::chrono::duration<double>(clock::now() - start).count();
        if (stats) *stats = st;
        return fail(X2_ERR_STATE, "GPU sample failure; whole epoch must use fallback (async_render.h:736)");
    }
    r = x2dev_download_film(d);
    if (r) return r;
    X2_HIP(hipEventSynchronize(d->filmDone));
    for (size_t i = 0; i < planned.size(); ++i)
        if (d->hostCounts[i] != planned[i]) ++st.audit_failures;
    st.wall_seconds = std::chrono::duration<double>(clock::now() - start).count();
    if (stats) *stats = st;
    if (st.audit_failures) return fail(X2_ERR_STATE, "Sample audit failure (async_render.h:1035)");
    return X2_OK;
}

int x2_adaptive_masks(uint32_t width, uint32_t height, float target_error, const float* errors, uint64_t* masks,
                      uint32_t* regions_done) {
    if (!errors || !masks || !width || !height) return fail(X2_ERR_ARGUMENT, "null buffer or empty image");
    if (!(target_error > 0.0f)) return fail(X2_ERR_ARGUMENT, "target_error must be positive");
    adaptive::Regions regions(int(width), int(height), target_error);
    regions.update(errors);
    const uint32_t tilesX = (width + 7) / 8, tilesY = (height + 7) / 8;
    for (uint32_t t = 0; t < tilesX * tilesY; ++t) masks[t] = regions.sampleArea(int(t % tilesX) * 8, int(t / tilesX) * 8);
    if (regions_done) *regions_done = regions.done() ? 1u : 0u;
    return X2_OK;
}

int x2_cellular_step(uint32_t width, uint32_t height, float target_error, const uint8_t* alive_in,
                     const uint64_t* in_image, const float* errors, const uint32_t* planned, double samples_left,
                     uint32_t max_samples, float noise_floor, uint8_t* alive_out, float* ratios) {
    if (!alive_in || !in_image || !errors || !alive_out || !width || !height) return fail(X2_ERR_ARGUMENT, "null buffer");
    const size_t tiles = size_t((width + 7) / 8) * ((height + 7) / 8);
    adaptive::Cellular cells(int(width), int(height), target_error);
    cells.deadForever(std::vector<uint64_t>(in_image, in_image + tiles));
    std::vector<char> alive(tiles);
    for (size_t t = 0; t < tiles; ++t) alive[t] = alive_in[t] && in_image[t] ? 1 : 0;  // dead tiles stay dead
    cells.setAlive(alive);
    std::vector<uint32_t> plan;
    if (planned) plan.assign(planned, planned + size_t(width) * height);
    const adaptive::Budget budget{&plan, samples_left, max_samples, noise_floor};
    const size_t live = cells.update(errors, planned ? &budget : nullptr);
    for (size_t t = 0; t < tiles; ++t) {
        alive_out[t] = cells.alive(t) ? 1 : 0;
        if (ratios) ratios[t] = cells.ratio(t);
    }
    return int(live);
}

int x2_pixel_errors(const double* sums, const double* odd, const uint32_t* counts, uint32_t pixels, float* errors) {
    if (!sums || !odd || !counts || !errors) return fail(X2_ERR_ARGUMENT, "null buffer");
    for (uint32_t i = 0; i < pixels; ++i) errors[i] = adaptive::pixelError(sums + size_t(i) * 4, odd + size_t(i) * 4, counts[i]);
    return X2_OK;
}

// The adjoint of one uniform frame: the same chunks as runFrame, each sample replayed by the adjoint kernel. The
// film is left as it was (the forward frame's), so x2dev_resolve_image still reads the forward image.
int x2dev_adjoint_frame(x2dev* d, const x2_config* c, uint32_t spc, const float* d_image, float* grad_textures,
                        float* grad_roughness, x2_frame_stats* stats) {
    using clock = std::chrono::steady_clock;
    const auto start = clock::now();
    if (d && d->pg.on) return fail(X2_ERR_STATE, "the adjoint reads resident textures (the textures are paged)");
    if (!d || !c || !d_image || !grad_roughness || (!d->texDesc.empty() && !grad_textures))
        return fail(X2_ERR_ARGUMENT, "null device, config, image adjoint or gradient buffer");
    if (!d->prepared) return fail(X2_ERR_STATE, "adjoint before prepare");
    if (!d->inFlight.empty()) return fail(X2_ERR_STATE, "adjoint with launches in flight");
    if (spc == 0 || spc > 1024) return fail(X2_ERR_ARGUMENT, "samples_per_chunk must be 1..1024");
    int r = validateConfig(d, *c);
    if (r) return r;
    if (c->max_depth > 32) return fail(X2_ERR_ARGUMENT, "the adjoint supports depth limits up to 32");
    if (d->volDesc.table) return fail(X2_ERR_ARGUMENT, "the adjoint does not differentiate through a medium");
    if (c->direct_only) return fail(X2_ERR_ARGUMENT, "the adjoint replays the path tracer, not the prototype estimator");
    X2_HIP(hipSetDevice(d->device));
    d->cfg = *c;
    d->failPixel = simulated("XPU2_SIMULATE_SAMPLE_FAIL");
    d->itemsPerChunk = spc * TilePixels;
    d->published = 0;
    d->launches = 0;
    d->begun_ = true;
    d->dImage = d_image;
    d->gradTextures = grad_textures;
    d->gradRoughness = grad_roughness;
    // The levels above 0 gather their texels' gradients apart; after the frame each level's go down to the level
    // below, a quarter to each texel that made it, ending in the caller's level-0 gradient.
    if (d->gradPyramidFloats && !d->gradPyramid) X2_HIP(hipMalloc(&d->gradPyramid, d->gradPyramidFloats * sizeof(float)));
    if (d->gradPyramidFloats)
        X2_HIP(hipMemsetAsync(d->gradPyramid, 0, d->gradPyramidFloats * sizeof(float), d->stream));
    x2_frame_stats st{};
    r = runFrame(d, c, spc, st);
    for (size_t t = 0; !r && t < d->texDesc.size(); ++t) {
        const TexDesc& td = d->texDesc[t];
        for (uint32_t k = td.count - 1; k >= 1; --k) {
            const TexLevel &up = td.levels[k], &in = td.levels[k - 1];
            float* below = k == 1 ? grad_textures + td.gradOffset : d->gradPyramid + (td.gradLevels + (long long)in.offset);
            const hipError_t e = launchPushDown(d->gradPyramid + (td.gradLevels + (long long)up.offset), up.width,
                                                up.height, below, in.width, in.height, d->stream);
            if (e != hipSuccess) r = hipFail(e, "launchPushDown");
        }
    }
    if (!r) {
        const hipError_t e = hipStreamSynchronize(d->stream);
        if (e != hipSuccess) r = hipFail(e, "hipStreamSynchronize");
    }
    d->dImage = nullptr;
    d->gradTextures = nullptr;
    d->gradRoughness = nullptr;
    if (r) return r;
    st.wall_seconds = std::chrono::duration<double>(clock::now() - start).count();
    if (stats) *stats = st;
    if (st.errors) return fail(X2_ERR_STATE, "GPU sample failure in the adjoint");
    return X2_OK;
}

}  // extern "C"
