// x2_layout.h — the device's data layout in HBM, shared by the host side (BVH8 builder, lane) and the kernels.
//
//   BVH8 node     128 B (Node8Q, one cache line): the node's origin, a power-of-two step per axis, the 8 children's
//                 bounds as 8-bit steps above the origin, 8 child references, 32 B pad. A reference is a node index
//                 (bit 31 clear), a leaf (bit 31 set: first triangle << 3 | count - 1, 1..8 triangles from first,
//                 first < 2^28), or EmptyChild, whose box no ray enters. The builder makes float nodes (Node8) first.
//   triangle      48 B: three float4, the vertices in xyz and (geom, prim, material) as the w words, in leaf order;
//                 the same vertices packed at 36 B a triangle after the records (the leaf tests' array).
//   film          FP64 sums[pixel][4] (rgb, alpha) + u32 counts[pixel] (async_gpu_abi.h:182-183 sums/counts).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/x2dev.h"

namespace x2 {

constexpr uint32_t EmptyChild = 0xffffffffu;
constexpr uint32_t LeafBit = 0x80000000u;
constexpr uint32_t MaxLeafTriangles = 8;
constexpr uint32_t TilePixels = 64;
constexpr uint32_t Slots = 4;  // launches whose per-launch counters may be live at once (async_gpu_abi.h:14)
constexpr int TraversalStack = 128;  // per-lane traversal stack entries; prepare() checks the BVH fits

// The builder's node: float bounds, quantized to Node8Q for the device.
struct alignas(256) Node8 {
    float lox[8], hix[8], loy[8], hiy[8], loz[8], hiz[8];
    uint32_t child[8];
    uint32_t pad[8];
};
static_assert(sizeof(Node8) == 256, "BVH8 node is 256 B");

// The device's node in one 128-B cache line: the children's bounds as 8-bit steps above the node's origin,
// each axis's step a power of two (its float exponent byte), lows rounded down and highs up by the builder, which
// checks each decoded bound in the device's arithmetic (origin + q * step, two roundings) so that every child's box
// contains its true box; the traversal then finds the same closest hits (Ylitie, Karras and Laine 2017).
struct alignas(128) Node8Q {
    float ox, oy, oz;
    uint8_t ex, ey, ez, pad0;
    uint8_t q[6][8];  // lox, hix, loy, hiy, loz, hiz
    uint32_t child[8];
    uint32_t pad[8];
};
static_assert(sizeof(Node8Q) == 128, "a quantized BVH8 node is one 128-B line");
using DevNode = Node8Q;
static_assert(sizeof(DevNode) == X2_NODE_BYTES, "the node the device reads");

struct alignas(16) TriRecord {
    float ax, ay, az;
    uint32_t geom;
    float bx, by, bz;
    uint32_t prim;
    float cx, cy, cz;
    uint32_t material;
};
static_assert(sizeof(TriRecord) == X2_TRIANGLE_BYTES, "triangle record is 48 B");
// The leaf tests read a triangle's vertices from a packed array of 36 B a triangle placed after the records in the
// same allocation (at a 256-B boundary); the ids are read from the record only for an exact tie in distance.
inline size_t triVertsOffset(size_t triangles) { return (triangles * sizeof(TriRecord) + 255) / 256 * 256; }
inline size_t triBufferBytes(size_t triangles) {
    return triVertsOffset(triangles) + triangles * 9 * sizeof(float);
}

// Per-launch state: the claim counter and the totals the kernel adds up (one atomic per wave).
struct alignas(64) LaunchState {
    unsigned long long traces, shadows, nodes, triangles;
    unsigned long long claimed;  // items claimed by the launch's persistent lanes
    unsigned long long voxels;   // the medium's trilinear tap sets (8 FP16 voxels of one grid each)
    uint32_t errors, pad[3];
};
static_assert(sizeof(LaunchState) == 64, "one line of counters a launch");

// A texture on the device: its texels (RGB f32, row-major) and where its gradient starts in the adjoint's buffer.
// A texture's mip pyramid: level k + 1 is max(1, w >> 1) x max(1, h >> 1), each texel the mean of the 2 x 2 below
// it, down to 1 x 1 (at most MipLevels levels). Resident, every level is rgb floats end to end in `texels` (level k
// at levels[k].offset floats). Paged (texture_core.h:37-60, the GPU's fine-detail cache, async_gpu.h:2911-3200),
// every level is half floats: the levels finer than firstResident as 32 x 32 tiles in the device's tile pool (level
// k's tile (x >> 5, y >> 5) is the scene-wide tile levels[k].tileBase + (y >> 5) * levels[k].tilesX + (x >> 5)), the
// coarser ones always resident in `coarse` (level k at levels[k].offset halves).
constexpr uint32_t MipLevels = 16;
// The adjoint gathers the gradients of small levels (at most AdjLdsTexels texels over all textures, the coarsest
// first: the levels diffuse paths read, where every lane's atomics would meet on a few texels) in each workgroup's
// LDS and adds them to the global gradients once, at the workgroup's end. lds: the level's first texel there, or
// NoLds.
constexpr uint32_t AdjLdsTexels = 2048, NoLds = 0xffffffffu;
struct TexLevel {
    uint32_t width, height, tileBase, tilesX;
    unsigned long long offset;
    uint32_t lds, ldsPad;
};
struct TexDesc {
    const float* texels;
    const uint16_t* coarse;
    uint32_t count, paged, firstResident, pad;
    unsigned long long gradOffset;  // level 0's texels in the caller's texture gradient (floats)
    long long gradLevels;  // level k's gradient (k >= 1) is at gradLevels + levels[k].offset in the device's pyramid
                           // gradient (floats)
    TexLevel levels[MipLevels];
};
constexpr uint32_t TexTile = 32, TexTileTexels = TexTile * TexTile;
constexpr uint32_t NoTileSlot = 0xffffffffu;
// The paging state the kernels write and the host reads back after each launch: samples parked for a missing tile
// (a ring the next launches run again), the tiles asked for, provisional samples (partial shading).
struct PagingState {
    unsigned long long parkTail;    // parked samples ever written (ring position of the next)
    unsigned long long refineTail;  // provisional records ever written for the host (host refines)
    uint32_t missCount;             // tiles this launch asked for (entries of the miss list)
    uint32_t parkOverflow;          // samples that could not be parked (the ring was full): failures
    uint32_t provisional;
    uint32_t refineOverflow;        // provisional records the ring could not hold: failures
};
constexpr uint32_t MaxTextures = 16;
constexpr uint32_t RoughnessSlots = 8;  // materials whose roughness gradient a lane accumulates in registers

// The medium on the device (x2_volume as bricks): the brick table (pool index or ~0u), the FP16 pools (512 halves a
// brick, z, y, x within), each brick's majorant density, the emission ramp, and the derived sizes (host-computed in
// float so that every reader rounds alike). table == null: no medium.
// The glowing medium as a light source (an emissive region over a dense emission distribution: MoonRay's
// EmissiveRegion and DenseEmissionDistribution, volume_emission.h:43-60): a box of res cells over the medium's box, a
// cell's value the largest emission luminance within one voxel of it, sampled through the marginals of a 3D
// histogram (x within each (z, y) row, then y within each z, then z), each 1D distribution a record in the pool:
// funcInt, n, n values, n + 1 cdf entries. distToWorld maps the cells' coordinates (0..res per axis) to render space
// and worldToDist back (row-vector affine, 12 floats: the rows x, y, z, then the translation); invUnitVolume is one
// over a cell's volume. pool == null: the medium does not light the scene.
struct EmissionRegion {
    int res[3];
    uint32_t visibility;  // the ray kinds it lights (lobeRayMask bits; 0x100: volume scattering)
    float distToWorld[12], worldToDist[12];
    float invUnitVolume;
    uint32_t conditionalX, marginalY, marginalZ;  // pool offsets of each set's first record
    const float* pool;
};
constexpr uint32_t EmissionCell = 4;  // voxels a region cell spans per axis

struct VolumeDesc {
    float bmin[3], bmax[3];
    uint32_t res[3], nb[3];
    float scaleVox[3], brickExtent[3];
    float minFeature, sigmaTScale, g, emissionScale, rampMax;
    float albedo[3];
    uint32_t rampSize, pad;
    const uint32_t* table;
    const uint16_t* density;
    const uint16_t* temperature;  // null: no emission
    const float* majorant;
    const float* ramp;
    EmissionRegion region;  // region.pool == null: none
};

// Everything a render launch reads, by value.
struct RenderParams {
    x2_config cfg;
    const DevNode* nodes;
    const TriRecord* tris;
    const float* triVerts;  // the records' vertices alone, 9 floats a triangle, for the leaf tests
    const float* triUV;  // 6 floats per triangle record (leaf order), or null: (0,0), (1,0), (1,1)
    const uint8_t* triCull;  // 1 per triangle record (leaf order) of a single-sided mesh, or null: none is
    const TexDesc* textures;
    const x2_material* materials;
    const x2_light* lights;
    const float* lightCorners;  // 12 floats a light, beside the lights: a rectangle's corners (lightCorners, host)
    // the adjoint's image adjoint (dLoss / d mean rgb per pixel) and gradient outputs (adjoint launches only)
    const float* dImage;
    float* gradTextures;
    float* gradRoughness;
    float* gradPyramid;  // the texture gradients of levels 1 and up (TexDesc::gradLevels), pushed down after a frame
    const x2_chunk* ring;
    double* sums;
    uint32_t* counts;
    LaunchState* state;
    unsigned long long itemBegin, itemEnd;
    uint32_t itemsPerChunk, ringCapacity;
    uint32_t counters;  // nonzero: count rays, nodes, triangles
    uint32_t adaptive;  // nonzero: an adaptive frame (the chunks carry pass and offset; odd-index sums kept)
    VolumeDesc vol;
    // An adaptive frame's film of odd-index sums and its pass tables (async_gpu.cu:337-345, :440-446): a chunk's
    // first = pass << PassShift | offset; its tile's active pixels (passMask) take index passBase[pixel] + offset + s.
    double* odd;
    const unsigned long long* passMask[4];
    const uint32_t* passBase[4];
    // The device pass chain's render (async_gpu.cu:448-466, Feed phase 2): items over the chain's active tiles, each
    // chainSize samples of its active pixels from chainBase; null: chunks from the ring.
    struct ChainState* chain;
    const uint32_t* chainTiles;
    const unsigned long long* chainMasks;
    const uint32_t* chainBase;
    uint32_t chainSize, chainPad;
    // On-demand texture tiles: the scene-wide page table (a tile's pool slot, or NoTileSlot), the pool
    // (TexTileTexels * 3 halves a slot), each tile's asked-for flag and the launch's miss list; the parked ring
    // (parkCap records) whose [parkedBegin, parkedBegin + parkedCount) this launch runs before its chunks; with
    // partial shading the provisional film, which holds the samples parked after a coarse render (each parked
    // record's value, parkedRgba[slot * 4..], leaves it when the sample runs again in full). Null tileSlot: no
    // texture pages.
    const uint32_t* tileSlot;
    const uint16_t* tilePool;
    uint32_t* tileAsked;
    uint32_t* missList;
    PagingState* paging;
    x2_parked* parked;
    float* parkedRgba;
    double* provSums;
    uint32_t* provCounts;
    unsigned long long parkedBegin;
    uint32_t parkedCount, parkCap, missCap, partial;
    uint32_t* tileStamp;  // each tile's last launch to read it (the pool evicts the least recently used)
    // Host refines: the launch renders refineInCount provisional samples again ahead of its parked samples and
    // chunks; the provisional samples it makes go to the refine ring (refineCap records; those below refineHead are
    // taken) instead of parking.
    x2_refine* refineRing;
    const x2_refine* refineIn;
    unsigned long long refineHead;
    uint32_t refineCap, refineInCount, hostRefines, refinePad;
    uint32_t stamp;
    // Failure injection: nonzero k makes every sample of pixel k - 1 fail as it starts, counted in the launch's
    // errors like any sample the estimator cannot run (async_gpu.cu:363), so the frame fails as the reference's does.
    uint32_t failPixel;
    uint32_t textureCount, textureCountPad;
};
constexpr uint32_t PassSlots = 4, PassShift = 16, MaxPasses = 160;  // async_integrator.h:3083

// The device pass chain (x2_kernel.hip chain*Kernel; async_adaptive.cu): a frame's adaptive passes queued at once,
// each pass's trees, masks and active tiles made on the device from the errors of every pass before it.
struct ChainState {
    unsigned long long claimed;  // the current pass's items claimed (a line of its own: every claim's atomic)
    uint32_t padClaimed[30];
    uint32_t stop;               // the trees converged or a pass came out empty: every later step does nothing
    uint32_t activeTiles;        // the current pass's active tiles
    uint32_t passes, treeUpdates, pad[28];
};
struct ChainRegion {
    int x0, y0, x1, y1;
    uint32_t nodeOffset, nodeCapacity;
    unsigned long long satOffset;
};
struct ChainParams {
    const double* sums;
    const double* odd;
    const uint32_t* counts;
    float* errors;
    const ChainRegion* regions;
    void* nodes;  // adaptive::Node
    double* sat;
    uint32_t* regionDone;
    unsigned long long* tileMask;
    const unsigned long long* inImage;
    const uint32_t* tileOrder;  // the tiles in the work grid's cell order
    uint32_t* activeTiles;
    unsigned long long* activeMasks;
    uint32_t* planned;
    uint32_t* base;
    ChainState* state;
    uint32_t width, height, tiles, tilesX, regionCount, perDim, regionW, regionH;
    float target;
    uint32_t pad;
};

// The wavefront form of the surface estimator (x2_kernel.hip waveShadeKernel / waveTraceKernel / waveDrainKernel):
// `slots` path slots, each holding one sample between its vertices, in SoA arrays over the slots; and two sets of ray
// queues (ping-pong by iteration parity). A shading pass writes a slot's closest-hit ray at the slot's position and
// its shadow rays (at most `perVertex`, each with the contribution it adds when unoccluded) at slot * perVertex + j;
// each workgroup then appends its queued positions to the compacted lists with one atomic.
//   pending records   72 B, level-major: level i of slot s at stack[i * slots + s] (MaxPending levels)
//   ray               two float4: origin + near, direction + far
struct WaveCounts {
    uint32_t rays[2], shadows[2];  // per parity: the compacted lists' lengths
    unsigned long long traced;     // closest-hit rays traced in the launch (the host's iteration estimate)
};
struct WaveParams {
    uint32_t slots, perVertex;
    uint32_t* flags;         // [slots] WaveLive | WaveTraced | WaveShadows
    uint32_t* sample;        // [slots * 2] pixel, sample index
    uint32_t* top;           // [slots] pending records
    float* accum;            // [slots * 8] radiance rgb, alpha, camera near, camera far, 2 unused
    void* stack;             // [MaxPending * slots] pending records
    void* current;           // [slots] the pending record whose ray is being traced
    uint32_t* hitTri;        // [slots] its closest hit (a triangle record, or ~0u)
    uint32_t* shadowCount;   // [slots] shadow rays of the slot's last vertex
    float* rays[2];          // [slots * 8] by slot
    uint32_t* rayList[2];    // [slots] queued slots, compacted
    float* shadowRays[2];    // [slots * perVertex * 8] by slot and entry
    float* shadowValue[2];   // [slots * perVertex * 4] contribution rgb, unused
    uint32_t* shadowList[2]; // [slots * perVertex] queued entries, compacted
    uint32_t* shadowVis[2];  // [slots * perVertex] 1 unoccluded
    WaveCounts* counts;
    unsigned long long* partials;  // [partialBlocks * 4] the trace passes' counters by workgroup
    uint32_t partialBlocks, pad;
};
constexpr uint32_t WaveLive = 1, WaveTraced = 2, WaveShadows = 4;
constexpr uint32_t MaxPendingRecords = 32, PendingBytes = 72;  // a sample's pending stack (x2_kernel.hip Pending)

// The batched queue's launch record and totals (x2_kernel.hip runBatch, x2_batch.hip).
struct BatchState {
    unsigned long long traces, shadows;
    uint32_t errors, pad;
};
struct BatchParams {
    const RenderParams* jobs;
    const uint32_t* order;             // queue position -> job
    const unsigned long long* first;   // [positions + 1] item prefix over queue positions
    unsigned long long* head;          // the queue's claim counter
    unsigned long long* jobDone;       // per job: items finished
    unsigned long long* jobFinish;     // per job: wall clock when its last group finished
    BatchState* state;
    uint32_t posBegin, posEnd, claim, pad;
};
// A region the batch clears before a run: `count` 32-bit words at `words`.
struct FillRegion {
    uint32_t* words;
    unsigned long long count;
};
}  // namespace x2
