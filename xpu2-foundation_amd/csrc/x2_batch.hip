// x2_batch.hip — batched render jobs on one device (no counterpart in the reference, which renders one frame at a
// time). Each job is its own scene (BVH8 built on the host at add time), settings and film. A run renders every
// job once: the jobs' items are laid end to end in a device queue, shortest job first (paths x log2 triangles), and
// one persistent kernel claims them (runBatch in x2_kernel.hip). The run — film clears, a start stamp, the kernel —
// is captured once in a hipGraph and replayed; the comparison mode issues one launch per job instead. The kernel's
// shape is chosen by a small autotuner that times each candidate on a one-sample pass over every job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "x2_bvh.h"
#include "x2_host.h"
#include "x2_layout.h"

namespace {
using namespace x2;

int fail(int code, const std::string& what) { return reportFailure(code, what); }
int hipFail(hipError_t e, const char* where) { return fail(X2_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e)); }
#define X2_HIP(call)                                     \
    do {                                                 \
        const hipError_t e_ = (call);                    \
        if (e_ != hipSuccess) return hipFail(e_, #call); \
    } while (0)

struct JobHost {
    x2_config cfg{};
    RenderParams params{};
    void* buffers[7] = {};  // nodes, tris, materials, lights, single-sided flags, sums, counts
    uint64_t pixels = 0, items = 0, itemsOneSample = 0, paths = 0;
    uint32_t triangles = 0;
    std::vector<double> hostSums;
    std::vector<uint32_t> hostCounts;
};
}  // namespace

struct x2batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double clockHz = 1e8;
    std::vector<JobHost> jobs;
    // the queue's device tables, rebuilt when a job is added
    bool dirty = true;
    RenderParams* dJobs = nullptr;
    uint32_t* dOrder = nullptr;
    unsigned long long *dFirst = nullptr, *dFirstOne = nullptr, *dHeads = nullptr, *dDone = nullptr,
                       *dFinish = nullptr, *dStamp = nullptr;
    BatchState* dState = nullptr;
    FillRegion* dRegions = nullptr;
    uint32_t regionCount = 0;
    std::vector<unsigned long long> first, firstOne;
    x2_batch_tuning tuning{256, 64, 3, 0, 0.0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {
void freeTables(x2batch* b) {
    for (void* p : {(void*)b->dJobs, (void*)b->dOrder, (void*)b->dFirst, (void*)b->dFirstOne, (void*)b->dHeads,
                    (void*)b->dDone, (void*)b->dFinish, (void*)b->dStamp, (void*)b->dState, (void*)b->dRegions})
        (void)hipFree(p);
    b->dJobs = nullptr;
    b->dOrder = nullptr;
    b->dFirst = b->dFirstOne = b->dHeads = b->dDone = b->dFinish = b->dStamp = nullptr;
    b->dState = nullptr;
    b->dRegions = nullptr;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}
template <class T>
hipError_t upload(T*& dst, const std::vector<T>& src) {
    hipError_t e = hipMalloc(&dst, std::max<size_t>(1, src.size()) * sizeof(T));
    if (e == hipSuccess && !src.empty()) e = hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}
// The queue: shortest job first by paths x log2(2 + triangles); item prefixes for the full run and for the tuner's
// one-sample pass; the regions a run clears (films, counters).
int buildTables(x2batch* b) {
    freeTables(b);
    const uint32_t n = uint32_t(b->jobs.size());
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        const JobHost &a = b->jobs[x], &c = b->jobs[y];
        return double(a.paths) * std::log2(2.0 + a.triangles) < double(c.paths) * std::log2(2.0 + c.triangles);
    });
    b->first.assign(n + 1, 0);
    b->firstOne.assign(n + 1, 0);
    for (uint32_t p = 0; p < n; ++p) {
        b->first[p + 1] = b->first[p] + b->jobs[order[p]].items;
        b->firstOne[p + 1] = b->firstOne[p] + b->jobs[order[p]].itemsOneSample;
    }
    std::vector<RenderParams> params(n);
    for (uint32_t j = 0; j < n; ++j) params[j] = b->jobs[j].params;
    X2_HIP(upload(b->dJobs, params));
    X2_HIP(upload(b->dOrder, order));
    X2_HIP(upload(b->dFirst, b->first));
    X2_HIP(upload(b->dFirstOne, b->firstOne));
    X2_HIP(hipMalloc(&b->dHeads, (n + 1) * sizeof(unsigned long long)));
    X2_HIP(hipMalloc(&b->dDone, std::max(1u, n) * sizeof(unsigned long long)));
    X2_HIP(hipMalloc(&b->dFinish, std::max(1u, n) * sizeof(unsigned long long)));
    X2_HIP(hipMalloc(&b->dStamp, sizeof(unsigned long long)));
    X2_HIP(hipMalloc(&b->dState, sizeof(BatchState)));
    std::vector<FillRegion> regions;
    for (JobHost& j : b->jobs) {
        regions.push_back(FillRegion{static_cast<uint32_t*>(j.buffers[5]), j.pixels * 8});  // 4 doubles a pixel
        regions.push_back(FillRegion{static_cast<uint32_t*>(j.buffers[6]), j.pixels});
    }
    regions.push_back(FillRegion{reinterpret_cast<uint32_t*>(b->dHeads), (n + 1) * 2ull});
    regions.push_back(FillRegion{reinterpret_cast<uint32_t*>(b->dDone), std::max(1u, n) * 2ull});
    regions.push_back(FillRegion{reinterpret_cast<uint32_t*>(b->dFinish), std::max(1u, n) * 2ull});
    regions.push_back(FillRegion{reinterpret_cast<uint32_t*>(b->dState), sizeof(BatchState) / 4});
    X2_HIP(upload(b->dRegions, regions));
    b->regionCount = uint32_t(regions.size());
    b->dirty = false;
    return X2_OK;
}
BatchParams paramsFor(const x2batch* b, bool oneSample, uint32_t posBegin, uint32_t posEnd, uint32_t headSlot) {
    BatchParams p{};
    p.jobs = b->dJobs;
    p.order = b->dOrder;
    p.first = oneSample ? b->dFirstOne : b->dFirst;
    p.head = b->dHeads + headSlot;
    p.jobDone = b->dDone;
    p.jobFinish = b->dFinish;
    p.state = b->dState;
    p.posBegin = posBegin;
    p.posEnd = posEnd;
    p.claim = b->tuning.claim_items;
    return p;
}
bool validTuning(const x2_batch_tuning& t) {
    return (t.block_threads == 64 || t.block_threads == 128 || t.block_threads == 256) && t.claim_items >= 64 &&
           t.claim_items % 64 == 0 && t.claim_items <= 65536 && (t.waves_per_simd == 2 || t.waves_per_simd == 3);
}
// The kernel over queue positions [posBegin, posEnd) with the batch's shape (blocks: the persistent grid for it,
// asked of the runtime before any capture begins).
int residentBlocks(const x2batch* b, int& blocks) {
    X2_HIP(batchResident(int(b->tuning.waves_per_simd), b->tuning.block_threads, blocks));
    return X2_OK;
}
int launchRange(x2batch* b, int blocks, bool oneSample, uint32_t posBegin, uint32_t posEnd, uint32_t headSlot) {
    X2_HIP(launchBatch(paramsFor(b, oneSample, posBegin, posEnd, headSlot), int(b->tuning.waves_per_simd),
                       b->tuning.block_threads, blocks, b->stream));
    return X2_OK;
}
}  // namespace

extern "C" {

int x2batch_create(int device, x2batch** out) {
    if (!out) return fail(X2_ERR_ARGUMENT, "null output pointer");
    *out = nullptr;
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || device < 0 || device >= count)
        return fail(X2_ERR_HIP, e != hipSuccess ? hipGetErrorString(e) : "no such HIP device");
    x2batch* b = new x2batch;
    b->device = device;
    int rateKhz = 0;
    uint32_t table[17 * 32];
    sobolTable(table);
    hipError_t r = hipSetDevice(device);
    if (r == hipSuccess) r = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (r == hipSuccess) r = hipEventCreate(&b->t0);
    if (r == hipSuccess) r = hipEventCreate(&b->t1);
    if (r == hipSuccess) r = hipDeviceGetAttribute(&rateKhz, hipDeviceAttributeWallClockRate, device);
    if (r == hipSuccess) r = uploadSobol(table);
    if (r != hipSuccess) {
        x2batch_destroy(b);
        return hipFail(r, "x2batch_create");
    }
    b->clockHz = rateKhz > 0 ? double(rateKhz) * 1e3 : 1e8;
    *out = b;
    return X2_OK;
}

int x2batch_add_job(x2batch* b, const x2_scene* s, const x2_config* c, uint32_t* job) {
    if (!b || !s || !c) return fail(X2_ERR_ARGUMENT, "null batch, scene or config");
    if ((s->triangle_count && (!s->positions || !s->indices)) || !s->geom_first || !s->geom_material ||
        !s->materials || (s->light_count && !s->lights))
        return fail(X2_ERR_ARGUMENT, "scene arrays missing");
    if (s->texture_count || s->volume_count) return fail(X2_ERR_ARGUMENT, "a batched job's scene has no textures or media");
    for (uint32_t i = 0; i < s->light_count; ++i)
        if (s->lights[i].type != X2_LIGHT_ENV && s->lights[i].type != X2_LIGHT_RECT)
            return fail(X2_ERR_ARGUMENT, "only environment and rectangle lights are implemented");
    if (!c->width || !c->height || !c->spp || uint64_t(c->width) * c->height > (1ull << 28))
        return fail(X2_ERR_ARGUMENT, "width, height and spp must be positive, at most 2^28 pixels");
    if (c->bsdf_samples == 0 || c->bsdf_samples > 64 || c->light_samples == 0 || c->light_samples > 64 ||
        c->max_depth > 250 || c->max_diffuse_depth > 250 || c->max_glossy_depth > 250 || c->filter != 0 ||
        c->light_count > s->light_count)
        return fail(X2_ERR_ARGUMENT, "job settings out of range (samples 1..64, depth <= 250, box filter, lights)");
    for (uint32_t i = 0; i < s->material_count; ++i)
        if (s->materials[i].albedo_map) return fail(X2_ERR_ARGUMENT, "a batched job's materials have no maps");
    Bvh8Result bvh;
    std::string err;
    if (!buildBvh8(*s, bvh, err)) return fail(X2_ERR_ARGUMENT, err);
    if (bvh.maxStack > uint32_t(TraversalStack)) return fail(X2_ERR_ARGUMENT, "BVH8 too deep for the traversal stack");
    X2_HIP(hipSetDevice(b->device));
    JobHost j;
    j.cfg = *c;
    j.pixels = uint64_t(c->width) * c->height;
    const uint64_t tiles = uint64_t((c->width + 7) / 8) * ((c->height + 7) / 8);
    j.items = tiles * 64 * c->spp;
    j.itemsOneSample = tiles * 64;
    j.paths = j.pixels * c->spp;
    j.triangles = s->triangle_count;
    const size_t nodeBytes = bvh.deviceNodeBytes(), triBytes = std::max<size_t>(1, triBufferBytes(bvh.tris.size()));
    auto cleanup = [&](hipError_t e, const char* where) {
        for (void* p : j.buffers) (void)hipFree(p);
        return hipFail(e, where);
    };
    hipError_t e = hipSuccess;
    const size_t sizes[7] = {nodeBytes, triBytes, std::max<size_t>(1, s->material_count) * sizeof(x2_material),
                             lightBufferBytes(s->light_count), bvh.triCull.size(),
                             size_t(j.pixels) * 4 * sizeof(double),
                             size_t(j.pixels) * sizeof(uint32_t)};
    for (int i = 0; i < 7 && e == hipSuccess; ++i)
        if (sizes[i]) e = hipMalloc(&j.buffers[i], sizes[i]);
    if (e == hipSuccess) e = hipMemcpy(j.buffers[0], bvh.deviceNodes(), nodeBytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && !bvh.tris.empty())
        e = hipMemcpy(j.buffers[1], bvh.tris.data(), bvh.tris.size() * sizeof(TriRecord), hipMemcpyHostToDevice);
    if (e == hipSuccess && !bvh.tris.empty()) {
        const std::vector<float> v = triVertices(bvh.tris);
        e = hipMemcpy(static_cast<char*>(j.buffers[1]) + triVertsOffset(bvh.tris.size()), v.data(), v.size() * sizeof(float),
                      hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(j.buffers[2], s->materials, s->material_count * sizeof(x2_material), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const std::vector<char> lb = lightBuffer(s->lights, s->light_count);
        e = hipMemcpy(j.buffers[3], lb.data(), lb.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !bvh.triCull.empty())
        e = hipMemcpy(j.buffers[4], bvh.triCull.data(), bvh.triCull.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(e, "x2batch_add_job upload");
    RenderParams& p = j.params;
    p.cfg = *c;
    p.nodes = static_cast<const DevNode*>(j.buffers[0]);
    p.tris = static_cast<const TriRecord*>(j.buffers[1]);
    p.triVerts = reinterpret_cast<const float*>(static_cast<const char*>(j.buffers[1]) + triVertsOffset(bvh.tris.size()));
    p.materials = static_cast<const x2_material*>(j.buffers[2]);
    p.lights = static_cast<const x2_light*>(j.buffers[3]);
    p.lightCorners = reinterpret_cast<const float*>(static_cast<const char*>(j.buffers[3]) + s->light_count * sizeof(x2_light));
    p.triCull = static_cast<const uint8_t*>(j.buffers[4]);
    p.sums = static_cast<double*>(j.buffers[5]);
    p.counts = static_cast<uint32_t*>(j.buffers[6]);
    if (job) *job = uint32_t(b->jobs.size());
    b->jobs.push_back(std::move(j));
    b->dirty = true;
    return X2_OK;
}

int x2batch_set_tuning(x2batch* b, const x2_batch_tuning* t) {
    if (!b || !t) return fail(X2_ERR_ARGUMENT, "null batch or tuning");
    if (!validTuning(*t)) return fail(X2_ERR_ARGUMENT, "tuning: 64/128/256 threads, claims a multiple of 64, 2 or 3 waves");
    b->tuning = *t;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return X2_OK;
}

int x2batch_tune(x2batch* b, x2_batch_tuning* best) {
    if (!b || b->jobs.empty()) return fail(X2_ERR_STATE, "tune needs at least one job");
    X2_HIP(hipSetDevice(b->device));
    if (b->dirty) {
        const int r = buildTables(b);
        if (r) return r;
    }
    const uint32_t n = uint32_t(b->jobs.size());
    x2_batch_tuning chosen{};
    chosen.seconds = 1e30;
    for (uint32_t waves : {3u, 2u})
        for (uint32_t threads : {256u, 128u, 64u})
            for (uint32_t claim : {64u, 256u, 1024u}) {
                b->tuning = x2_batch_tuning{threads, claim, waves, 0, 0.0};
                int blocks = 0;
                if (const int r = residentBlocks(b, blocks)) return r;
                double bestRun = 1e30;
                for (int rep = 0; rep < 2; ++rep) {
                    X2_HIP(launchFill(b->dRegions, b->regionCount, b->stream));
                    X2_HIP(hipEventRecord(b->t0, b->stream));
                    const int r = launchRange(b, blocks, true, 0, n, 0);
                    if (r) return r;
                    X2_HIP(hipEventRecord(b->t1, b->stream));
                    X2_HIP(hipEventSynchronize(b->t1));
                    float ms = 0;
                    X2_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
                    bestRun = std::min(bestRun, double(ms) * 1e-3);
                }
                if (bestRun < chosen.seconds) {
                    chosen = b->tuning;
                    chosen.seconds = bestRun;
                }
            }
    const int r = x2batch_set_tuning(b, &chosen);
    if (r) return r;
    if (best) *best = chosen;
    return X2_OK;
}

int x2batch_run(x2batch* b, int per_job_launches, double* job_seconds, x2_batch_stats* stats) {
    if (!b || b->jobs.empty()) return fail(X2_ERR_STATE, "run needs at least one job");
    using clock = std::chrono::steady_clock;
    const auto start = clock::now();
    X2_HIP(hipSetDevice(b->device));
    if (b->dirty) {
        const int r = buildTables(b);
        if (r) return r;
    }
    const uint32_t n = uint32_t(b->jobs.size());
    uint32_t launches = 0;
    int blocks = 0;
    if (const int r = residentBlocks(b, blocks)) return r;
    if (!per_job_launches) {
        if (!b->exec) {  // capture once: clears (a fill kernel, not memsets), the start stamp, the persistent kernel
            X2_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = launchFill(b->dRegions, b->regionCount, b->stream);
            if (e == hipSuccess) e = launchStamp(b->dStamp, b->stream);
            const int r = e == hipSuccess ? launchRange(b, blocks, false, 0, n, 0) : X2_OK;
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess) return hipFail(e, "batch capture");
            if (r) return r;
            X2_HIP(ee);
            b->graph = g;
            X2_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        X2_HIP(hipEventRecord(b->t0, b->stream));
        X2_HIP(hipGraphLaunch(b->exec, b->stream));
        X2_HIP(hipEventRecord(b->t1, b->stream));
        launches = 3;
    } else {
        X2_HIP(launchFill(b->dRegions, b->regionCount, b->stream));
        X2_HIP(hipEventRecord(b->t0, b->stream));
        X2_HIP(launchStamp(b->dStamp, b->stream));
        for (uint32_t p = 0; p < n; ++p) {
            const int r = launchRange(b, blocks, false, p, p + 1, p + 1);
            if (r) return r;
        }
        X2_HIP(hipEventRecord(b->t1, b->stream));
        launches = n + 2;
    }
    X2_HIP(hipEventSynchronize(b->t1));
    float ms = 0;
    X2_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
    std::vector<unsigned long long> finish(n);
    unsigned long long stamp = 0;
    BatchState st{};
    X2_HIP(hipMemcpy(finish.data(), b->dFinish, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    X2_HIP(hipMemcpy(&stamp, b->dStamp, sizeof(stamp), hipMemcpyDeviceToHost));
    X2_HIP(hipMemcpy(&st, b->dState, sizeof(st), hipMemcpyDeviceToHost));
    if (job_seconds)
        for (uint32_t j = 0; j < n; ++j) job_seconds[j] = finish[j] >= stamp ? double(finish[j] - stamp) / b->clockHz : -1.0;
    x2_batch_stats s{};
    s.wall_seconds = std::chrono::duration<double>(clock::now() - start).count();
    s.device_seconds = double(ms) * 1e-3;
    for (const JobHost& j : b->jobs) s.paths += j.paths;
    s.rays_trace = st.traces;
    s.rays_shadow = st.shadows;
    s.jobs = n;
    s.errors = st.errors;
    s.launches = launches;
    s.graph = per_job_launches ? 0u : 1u;
    s.tuning = b->tuning;
    if (stats) *stats = s;
    for (uint32_t j = 0; j < n; ++j)
        if (finish[j] < stamp) return fail(X2_ERR_STATE, "a job did not finish (its completion stamp is missing)");
    if (st.errors) return fail(X2_ERR_STATE, "GPU sample failure in a batched job");
    return X2_OK;
}

int x2batch_job_film(x2batch* b, uint32_t job, const double** sums, const uint32_t** counts) {
    if (!b || !sums || !counts || job >= b->jobs.size()) return fail(X2_ERR_ARGUMENT, "no such job");
    JobHost& j = b->jobs[job];
    X2_HIP(hipSetDevice(b->device));
    j.hostSums.resize(size_t(j.pixels) * 4);
    j.hostCounts.resize(size_t(j.pixels));
    X2_HIP(hipMemcpy(j.hostSums.data(), j.buffers[5], j.hostSums.size() * sizeof(double), hipMemcpyDeviceToHost));
    X2_HIP(hipMemcpy(j.hostCounts.data(), j.buffers[6], j.hostCounts.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    *sums = j.hostSums.data();
    *counts = j.hostCounts.data();
    return X2_OK;
}

uint32_t x2batch_job_count(const x2batch* b) { return b ? uint32_t(b->jobs.size()) : 0u; }

void x2batch_destroy(x2batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    freeTables(b);
    for (JobHost& j : b->jobs)
        for (void* p : j.buffers) (void)hipFree(p);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
