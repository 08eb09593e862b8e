/* x2dev.h — C-ABI of the MI355X (gfx950) path-tracing device for XPU2's shared integrator.
 *
 * This header is the drop-in boundary: the entry points below replace the reference's GPU lane
 * (the C++ class `GpuLane`, xpu2/interactive/async_gpu.h:82) that the broker (`Renderer::render`,
 * xpu2/interactive/async_render.h:318) drives from its feeder thread. Every entry point returns a status code
 * instead of throwing; no C++ exception and no torch type crosses it. The caller owns every array it passes in;
 * arrays are copied before the call returns. The device owns its film and the host copies of it.
 *
 * Record layouts (x2_chunk, the film) are bit-identical to the reference's so a broker written against the
 * reference's types can hand them over unchanged.
 */
#ifndef X2DEV_H
#define X2DEV_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (negative: failure; x2dev_last_error() says what failed). */
#define X2_OK 0
#define X2_ERR_ARGUMENT (-1)   /* a null pointer, a size out of range, a malformed record */
#define X2_ERR_HIP (-2)        /* the HIP runtime reported an error */
#define X2_ERR_STATE (-3)      /* a call out of order (launch before begin, download with launches in flight ...) */
#define X2_ERR_BACKPRESSURE (-4) /* launch() while not ready(), or a launch larger than the mailbox ring allows */
#define X2_ERR_MEMORY (-5)     /* device or pinned allocation failed */
#define X2_ERR_FILE (-6)       /* a file could not be written or read, or is not the save asked for */

/* The unit of work: one 8x8 tile of the image and a range of sample indices (async_integrator.h:3057 Chunk). */
typedef struct x2_chunk {
    uint32_t tile;  /* tile index, row-major over the image's 8x8 tiles */
    uint32_t first; /* first sample index */
    uint32_t count; /* number of sample indices */
    uint32_t cell;  /* owning cell of the work grid, for per-cell accounting */
} x2_chunk;

/* A sample handed back to the host (async_gpu_abi.h:24 Parked): re-rendered elsewhere from its identity. */
typedef struct x2_parked {
    uint32_t pixel, index, tries, pass;
} x2_parked;

/* A provisional sample (Refine, async_integrator.h:3065-3073): one the device finished reading a missing texel's
 * parent a level up (partial shading). Its value stays in the provisional film until the sample is rendered again
 * in full, which takes it out. tile: the first missing tile's scene-wide id + 1 (0: not known). 32 bytes. */
typedef struct x2_refine {
    uint32_t pixel, index, pass, tile;
    float rgba[4];
} x2_refine;

/* Light types and sidedness (light_core.h:19-20). Only the environment and rectangle are implemented. */
#define X2_LIGHT_ENV 0u
#define X2_LIGHT_RECT 1u
#define X2_SIDE_REGULAR 0u
#define X2_SIDE_REVERSE 1u
#define X2_SIDE_TWO 2u

/* One light in render space, the subset of `lights::Light` (light_core.h:59-117) that a rectangle or a constant
 * environment reads. Linear parts are rows (a local vector v maps to v.x*row0 + v.y*row1 + v.z*row2). */
typedef struct x2_light {
    uint32_t type, sidedness, hemispherical, visibility_mask;
    float radiance[3];
    float inv_area;            /* rect: 1 / render-space area; env: 1 / (4 pi) (or 1 / (2 pi) hemispherical) */
    float position[3];
    float plane_offset;        /* -dot(direction, position) */
    float direction[3];        /* local +Z, unit: the side a one-sided light lights */
    float scale;               /* local to render length scale */
    float local_to_render[9];  /* rows, with scale */
    float half_width;          /* local units */
    float render_to_local[9];  /* rows, its inverse */
    float half_height;         /* local units */
    float render_to_local_rot[9]; /* rotation only */
    float pad;
} x2_light;

/* A material: a Lambert lobe and a GGX lobe (no Fresnel, no energy compensation), plus self-emission.
 * Lambert albedo = albedo * diffuse_weight; GGX lobe scale = specular_weight (white). A weight of zero omits the
 * lobe. Lobes are recorded in that order (Lambert first), which fixes their branch addresses. */
typedef struct x2_material {
    float albedo[3];
    float roughness;       /* MoonRay's input roughness; the lobe uses max(0.001, r^2) (shade_lobes.h:536) */
    float diffuse_weight;
    float specular_weight;
    float emission[3];
    uint32_t albedo_map;   /* 0: none; k: the scene's texture k-1 multiplies the albedo */
} x2_material;

/* An RGB FP32 texture, row-major texels (y * width + x) * 3, sides at most 32768. The device makes its mip pyramid
 * (level k + 1 max(1, w >> 1) x max(1, h >> 1), each texel the mean of the 2 x 2 below it, down to 1 x 1) and reads it
 * as the path tracer's map lookups do (texture_core.h:161-206; shade_mapnodes.h:344-347, :420-470): anisotropic on
 * camera paths, trilinear closest texels on glossy ones, one level's closest texel on diffuse ones, at the footprint
 * the ray's differentials give (shade_material.h:388-558), periodic, t flipped. */
typedef struct x2_texture {
    uint32_t width, height;
    const float* texels;
} x2_texture;

/* A heterogeneous participating medium filling a render-space box (the build's own volume model: the reference binds
 * volumes to closed meshes and reads NanoVDB grids, async_integrator.h:415-1291). Two grids of res[0] x res[1] x res[2]
 * IEEE half values, x fastest, each res a multiple of 8: density and temperature (kelvin; null: no emission). The
 * device keeps them as 8^3 bricks, only the bricks holding a nonzero density. A point's values are trilinear in the
 * voxel centres; extinction = density * sigma_t_scale (grey), scattering = albedo * extinction, Henyey-Greenstein
 * phase g, emission = emission_scale * density * ramp(temperature), ramp being ramp_size rgb values over
 * [0, ramp_max_kelvin] (a blackbody colour ramp), linearly interpolated. */
typedef struct x2_volume {
    float bounds_min[3], bounds_max[3];
    uint32_t res[3];
    float sigma_t_scale;
    float albedo[3];
    float g;
    float emission_scale;
    float ramp_max_kelvin;
    uint32_t ramp_size, pad;
    const uint16_t* density;
    const uint16_t* temperature;
    const float* ramp;             /* ramp_size x 3 */
} x2_volume;

/* A triangle scene. Triangles are ordered mesh by mesh; mesh g owns triangles [geom_first[g], geom_first[g+1]),
 * and a triangle's primitive id is its index less geom_first of its mesh. Meshes are two-sided unless geom_flags
 * marks them single-sided (X2_GEOM_SINGLE_SIDED, MoonRay's side_type single, cpu_scene.h:110-111,729): then a hit
 * whose ray runs along the triangle's (b - a) x (c - a), dot >= 0, is no hit, for every ray (the reference's
 * backFaceCullingFilter, async_cpu.h:122-133, and OPTIX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES,
 * async_gpu_device.h:333,527). A material whose
 * albedo_map is k > 0 multiplies its albedo by texture k-1 at the hit's (s, t), interpolated from per-vertex uvs
 * (shade_material.h:121-160; a triangle without uvs uses (0,0), (1,0), (1,1)). */
typedef struct x2_scene {
    const float* positions;        /* vertex_count x 3 */
    const uint32_t* indices;       /* triangle_count x 3 */
    const uint32_t* geom_first;    /* geom_count + 1 */
    const uint32_t* geom_material; /* geom_count */
    const x2_material* materials;  /* material_count */
    const x2_light* lights;        /* light_count */
    uint32_t vertex_count, triangle_count, geom_count, material_count, light_count, pad;
    const float* uvs;              /* vertex_count x 2, or null */
    const x2_texture* textures;    /* texture_count */
    uint32_t texture_count, pad2;
    const x2_volume* volumes;      /* volume_count (0 or 1) */
    uint32_t volume_count, pad3;
    const uint32_t* geom_flags;    /* geom_count, or null: every mesh two-sided */
} x2_scene;
#define X2_GEOM_SINGLE_SIDED 1u

/* The integrator's settings: the used subset of `mixed::Config` (gpu_jobs.h:10-44), plus rr_min_depth (an
 * extension: Russian roulette applies from that non-mirror depth on; 1 is the reference's behaviour). */
typedef struct x2_config {
    uint32_t width, height, spp, seed;
    uint32_t sampling;          /* sampling policy bits (sample_stream.h:58-144): 1 camera, 2 first BSDF, 4 first light */
    uint32_t filter;            /* 0: box (the only filter implemented) */
    float filter_width;
    float near_clip, far_clip;
    float raster_to_camera[16]; /* m[row][col], row-vector convention */
    float camera_to_render[16];
    uint32_t light_count;
    uint32_t bsdf_samples, light_samples;
    uint32_t max_depth, max_diffuse_depth, max_glossy_depth, max_mirror_depth;
    uint32_t clamp_depth;
    float clamp_value;
    float rr_threshold;
    uint32_t rr_min_depth;
    uint32_t shadow_reuse;      /* bounce-ray reuse: ray visibility bits whose bounces stand for shadow rays */
    int32_t viewport[4];        /* min x, min y, max x, max y, half open */
    uint32_t pad[2];
    /* Volumes (mixed::Config, gpu_jobs.h:19-28; read only when the scene has a volume). */
    uint32_t max_volume_depth, volume_illumination_samples;
    float volume_transmittance_threshold;
    float inv_volume_quality;   /* march step = voxel size * inv_volume_quality * max(depth - 1, 1) */
    float volume_attenuation_factor, volume_contribution_factor, volume_phase_attenuation_factor;
    uint32_t pad_volume;
    /* The estimator: 0 the path tracer (integratePaths); nonzero the prototype (integrateDirect,
     * async_integrator.h:61-126, mixed::Config::paths == 0): the camera hit's emission and surface_limit (1..8)
     * cosine-sampled Lambert bounces, the last looking for a constant sky of radiance sky. */
    uint32_t direct_only, surface_limit;
    float sky[3];
    uint32_t pad_direct;
} x2_config;

/* What the host learns when a launch has finished (async_gpu.h:65 Completion), plus traversal counters. */
typedef struct x2_completion {
    uint64_t chunks, items, samples;
    uint64_t parked_run;  /* parked samples the launch ran again (Completion::parkedRun) */
    uint64_t refines_run; /* provisional samples it rendered again (Completion::refinesRun) */
    uint32_t errors, pad;
    double device_seconds;     /* the launch's own duration, from HIP events around it */
    uint64_t rays_trace;       /* closest-hit rays traced */
    uint64_t rays_shadow;      /* any-hit shadow rays traced */
    uint64_t nodes_visited;    /* BVH8 nodes fetched (X2_NODE_BYTES = 128 B each) */
    uint64_t triangles_tested; /* triangles tested (their packed vertices, X2_LEAF_TRIANGLE_BYTES = 36 B each) */
    uint64_t voxel_taps;       /* the medium's trilinear tap sets (8 FP16 voxels of one grid, X2_VOXEL_TAP_BYTES) */
} x2_completion;

/* Totals over a frame's launches (x2dev_render_frame, x2dev_adjoint_frame). */
typedef struct x2_frame_stats {
    uint64_t launches, chunks, samples;
    uint32_t errors, audit_failures;
    double device_seconds, wall_seconds;
    uint64_t rays_trace, rays_shadow, nodes_visited, triangles_tested;
    uint64_t voxel_taps;
} x2_frame_stats;

typedef struct x2dev_options {
    int32_t device;                 /* HIP device ordinal */
    uint32_t max_chunks_per_launch; /* 16384 in the reference (async_gpu.h:43) */
    uint32_t ring_chunks;           /* mailbox ring capacity, in chunks (2^17 in the reference) */
    uint32_t counters;              /* nonzero: count rays, nodes and triangles per launch */
    uint32_t wave_slots;            /* 0: the surface estimator runs as one kernel with path regeneration (the
                                       default); else the wavefront form with this many path slots (a multiple of
                                       256): shading and traversal in separate kernels. The same samples either way.
                                       No reference counterpart (the reference runs one OptiX kernel). */
    uint32_t lane;                  /* the lane's number (GpuOptions::lane, async_gpu.h:685): lane k >= 1 fails its
                                       prepare when XPU2_SIMULATE_CARD_FAIL=k, as a card that fails as it takes a
                                       scene (multi_device.h:42-46); lane 0 never does */
} x2dev_options;

typedef struct x2dev x2dev;

/* Replaces GpuLane(ptx, ..., GpuOptions) (async_gpu.h:95-175): stream, mailbox ring, pinned staging, events. */
int x2dev_create(const x2dev_options* options, x2dev** out);
/* Replaces GpuLane::prepare(const cpu::PlanView&) (async_gpu.h:677): builds the BVH8 on the host and uploads the
 * scene. */
int x2dev_prepare(x2dev* d, const x2_scene* scene);
/* Replaces GpuLane::begin(samplesPerChunk, adaptive, partial) (async_gpu.h:1372-1422): zeroes film and state;
 * items per chunk = samples_per_chunk * 64. */
int x2dev_begin(x2dev* d, const x2_config* config, uint32_t samples_per_chunk);
/* Replaces GpuLane::ready() (async_gpu.h:1425): 1 when another launch may be queued (at most two in flight). */
int x2dev_ready(const x2dev* d);
/* Replaces GpuLane::launch(chunks, count, ..., samples) (async_gpu.h:1434-1500): copies the chunks into the ring and
 * queues the render. Returns 1 when queued, 0 when there was nothing to run, negative on failure. */
int x2dev_launch(x2dev* d, const x2_chunk* chunks, uint32_t count, uint64_t samples);
/* Replaces GpuLane::launch(chunks, count, refines, refineCount, samples) (async_gpu.h:1434-1500): the same, with
 * refine_count provisional samples (at most X2_REFINES_PER_LAUNCH, from x2dev_take_refines) rendered again first. */
#define X2_REFINES_PER_LAUNCH (1u << 17)
int x2dev_launch_refines(x2dev* d, const x2_chunk* chunks, uint32_t count, const x2_refine* refines,
                         uint32_t refine_count, uint64_t samples);
/* Replaces GpuLane::takeRefines (async_gpu.h:1713-1716): the provisional samples the collected launches made (paged
 * textures with partial shading and host_refines), at most capacity of them, oldest first. */
int x2dev_take_refines(x2dev* d, x2_refine* out, uint32_t capacity, uint32_t* taken);
uint64_t x2dev_refines_waiting(const x2dev* d);
/* Replaces GpuLane::poll(Completion&) (async_gpu.h:1503-1564): never waits. 1: the oldest launch finished and
 * *out holds it; 0: nothing finished; negative on failure. */
int x2dev_poll(x2dev* d, x2_completion* out);
/* Replaces GpuLane::takeDeferred (async_gpu.h:297-300): samples handed to the CPU lane (none in this device). */
int x2dev_take_deferred(x2dev* d, x2_parked* out, uint32_t capacity, uint32_t* taken);
/* Replaces GpuLane::downloadFilm() (async_gpu.h:1567-1581): queues the film's copy to pinned host memory. */
int x2dev_download_film(x2dev* d);
/* Replaces GpuLane::filmReady() (async_gpu.h:1583-1590): 1 once the download has landed. */
int x2dev_film_ready(x2dev* d);
/* Replaces GpuLane::sums()/counts() (async_gpu.h:1705-1707): sums[pixel*4 + c] (rgb, alpha) in FP64 and the
 * per-pixel sample counts, valid until the next download. */
int x2dev_film(const x2dev* d, const double** sums, const uint32_t** counts);
/* Replaces GpuLane::inFlight(), maxChunksPerLaunch() and launchThreads() (async_gpu.h:1426-1428,1725-1731). */
uint32_t x2dev_in_flight(const x2dev* d);
uint32_t x2dev_max_chunks_per_launch(const x2dev* d);
uint32_t x2dev_launch_threads(const x2dev* d);
/* Device pointers of the film (sums, counts) for callers that keep the film on the card (torch interop). */
int x2dev_device_film(const x2dev* d, void** sums, void** counts);
/* Frees everything the device holds. */
void x2dev_destroy(x2dev* d);
/* The last failure's text on the calling thread ("" when none). */
const char* x2dev_last_error(void);

/* Parameter updates between frames, for optimisation loops: a texture's texels (texels_on_device != 0: a device
 * pointer, copied on the device's stream) and a material record. No counterpart in the reference (its scene is
 * rebuilt by the plan); they stand beside prepare. */
int x2dev_set_texture(x2dev* d, uint32_t texture, const float* texels, int texels_on_device);
int x2dev_set_material(x2dev* d, uint32_t material, const x2_material* m);

/* The film's mean rgb per pixel (sums / counts; 0 where a pixel has no sample) into a caller-owned device buffer of
 * width * height * 3 floats, on the device's stream, waited for before return. */
int x2dev_resolve_image(x2dev* d, float* device_rgb);

/* The path-replay adjoint of one uniform frame: every sample of the frame is rendered again from its identity (the
 * same random numbers as the forward frame), and the image adjoint d_image (device, width * height * 3 floats:
 * dLoss / d(mean rgb of the pixel)) is carried back through it to the gradients of the textures' texels and the
 * materials' roughness (device buffers the caller owns and zeroes: grad_textures, the scene's textures' texels
 * concatenated in order; grad_roughness, material_count floats). Sampling, MIS weights and Russian roulette are
 * held fixed (detached), so the result is an unbiased estimate of the gradient of the expected image's loss.
 * No counterpart in the reference (it has no differentiable code). */
int x2dev_adjoint_frame(x2dev* d, const x2_config* config, uint32_t samples_per_chunk, const float* d_image,
                        float* grad_textures, float* grad_roughness, x2_frame_stats* stats);

/* Diagnostics (no counterpart in the reference): the device's sample-identity RNG evaluated on the device, for a
 * twin test against a host restatement. For query i: words[i] = samplingWord(identity, dimension, policy)
 * (sample_stream.h:127-144), stream_out[i * draws ...] = the first `draws` values of paths::Stream(identity,
 * bounce) (async_integrator.h:151-176), branches[i] = childBranch(branch, a, b) (async_integrator.h:281-286).
 * Host arrays, copied in and out; n <= 2^20, draws <= 256. */
typedef struct x2_rng_query {
    uint32_t pixel, dimension, policy, bounce;
    uint64_t index, stream, branch;
    uint32_t a, b;
} x2_rng_query;
int x2dev_rng_twin(x2dev* d, const x2_rng_query* queries, uint32_t n, uint32_t draws, uint32_t* words,
                   float* stream_out, uint64_t* branches);

/* Algorithmic byte sizes the roofline counts: one BVH8 node (one 128-B cache line: 8-bit child bounds above the
 * node's origin), a triangle a leaf test reads (its three vertices from the packed array, 36 B) and the whole record
 * (48 B, read once a hit). */
#define X2_NODE_BYTES 128u
#define X2_TRIANGLE_BYTES 48u
#define X2_LEAF_TRIANGLE_BYTES 36u
#define X2_FILM_BYTES_PER_PATH 72u
#define X2_VOXEL_TAP_BYTES 16u /* a trilinear tap set: 8 FP16 voxels */

/* The broker counterpart (async_render.h:318 Renderer::render, GPU lane only): one uniform frame, chunks made as
 * the work grid makes them (async_grid.h:129-175, progressive ranges 1,1,2,spc over 64-px cells), launches sized
 * to the measured rate with at most two in flight, completions polled, the film downloaded and every pixel's count
 * audited against spp (async_render.h:1035). stats (may be null) receives totals over the frame's launches. */
int x2dev_render_frame(x2dev* d, const x2_config* config, uint32_t samples_per_chunk, x2_frame_stats* stats);

/* ---- Adaptive frames (MoonRay's adaptive sampling as the reference restates it: async_adaptive.h,
 * async_adaptive_core.h, async_integrator.h:3077-3125). A pass is a range of sample counts; a chunk of a pass
 * carries first = pass << X2_PASS_SHIFT | offset, and each active pixel of its tile takes sample indices from its
 * own base, so a pixel's indices stay contiguous. The film also keeps the odd-index sums, from which the device works
 * out every pixel's error for the host's error trees. ---- */
#define X2_PASS_SHIFT 16u
#define X2_PASS_SLOTS 4u
#define X2_NO_SLOT 0xffffffffu /* x2dev_upload_pass: no uniform base, the pass starts from the device's counts */

typedef struct x2_adaptive {
    uint32_t min_samples; /* passes that start below this sample every pixel (MoonRay's min samples) */
    uint32_t max_samples; /* the most samples a pixel takes: the schedule's end */
    float target_error;   /* a region tree's node converges below this error */
    uint32_t passes;      /* out: passes rendered */
    uint32_t tree_updates; /* out: times the trees were rebuilt */
    uint32_t device_chain; /* nonzero: the device pass chain (async_adaptive.cu) builds the trees and masks and the
                              frame's passes are queued at once; zero: the host's trees between passes */
    uint32_t cellular;      /* nonzero: cellular generations of tiles decide each pass (adaptive::Cellular,
                               async_adaptive.h:147-317) in place of the region trees (host side only) */
    uint32_t sample_budget; /* cellular: a frame by sampling generations spends this many samples (0: none) */
    float noise_floor;      /* with a budget: the error no tile need go below (0: none) */
    uint32_t pad;
} x2_adaptive;

/* GpuLane::begin(spc, adaptive = true) (async_gpu.h:1372-1422): begin, with the odd-index film cleared and chunks
 * read through the pass tables. */
int x2dev_begin_adaptive(x2dev* d, const x2_config* config, uint32_t samples_per_chunk);
/* GpuLane::uploadPass (async_gpu.h:1906-1916): pass slot `slot`'s tile masks (one u64 a tile, bit y*8+x) and sample
 * bases: every pixel uniform_base, or X2_NO_SLOT for the device's counts (every earlier pass completed). Queued on
 * the lane's stream ahead of any launch that reads it. */
int x2dev_upload_pass(x2dev* d, uint32_t slot, const uint64_t* masks, uint32_t uniform_base);
/* GpuLane::requestErrors / errorsReady / errors (async_gpu.h:1836-1899; the mode-2 launch, async_gpu.cu:364-369):
 * every pixel's error behind the launches already queued, copied to a pinned host buffer. */
int x2dev_request_errors(x2dev* d);
int x2dev_errors_ready(x2dev* d);
int x2dev_errors(const x2dev* d, const float** errors);
/* GpuLane::seedFilm (async_gpu.h:1594-1609): a frame taken up from a saved film starts with its sums, odd-index sums
 * (may be null outside adaptive frames) and counts, so the samples still to come, and an adaptive frame's bases and
 * errors, go on from them. After begin / begin_adaptive, ahead of any launch; the host buffers are copied before it
 * returns. */
int x2dev_seed_film(x2dev* d, const double* sums, const double* odd, const uint32_t* counts);
/* GpuLane::odd() (async_gpu.h:1706): the odd-index sums, downloaded with the film in an adaptive frame. */
int x2dev_odd(const x2dev* d, const double** odd);
/* One adaptive frame through the broker's sequence (async_render.h:1219-1440): the host's error trees between passes,
 * or with a->device_chain the device pass chain (GpuLane::beginChain / launchChain, async_gpu.h:1960-2085), the
 * uniform stage one step and each later pass one. config->spp is ignored (a->max_samples bounds every pixel); counts
 * are audited against the planned samples. */
int x2dev_render_adaptive(x2dev* d, const x2_config* config, uint32_t samples_per_chunk, x2_adaptive* a,
                          x2_frame_stats* stats);
/* The host's tree step on its own (Regions::update + sampleArea, async_adaptive.h:99-145): from every pixel's error
 * (width x height), each tile's active pixels; regions_done: nonzero when every region's tree converged. Needs no
 * device. */
int x2_adaptive_masks(uint32_t width, uint32_t height, float target_error, const float* errors, uint64_t* masks,
                      uint32_t* regions_done);
/* One generation of the cellular rules on the host (Cellular::update, async_adaptive.h:183-230): from every tile's
 * state alive_in (1 live), the image's errors, the tiles' in-image masks and, with sample_budget > 0, each pixel's
 * planned samples and the samples left, the next generation in alive_out and each tile's error over the target in
 * ratios; returns the live tiles (or a negative status). Needs no device. */
int x2_cellular_step(uint32_t width, uint32_t height, float target_error, const uint8_t* alive_in,
                     const uint64_t* in_image, const float* errors, const uint32_t* planned, double samples_left,
                     uint32_t max_samples, float noise_floor, uint8_t* alive_out, float* ratios);
/* pixelError (async_integrator.h:3109-3125) on the host over a film: errors[i] from sums[i*4..], odd[i*4..],
 * counts[i]. Needs no device. */
int x2_pixel_errors(const double* sums, const double* odd, const uint32_t* counts, uint32_t pixels, float* errors);

/* ---- On-demand texture tiles and partial shading (GpuLane's fine-detail cache, async_gpu.h:2911-3200, and its
 * parked and provisional samples, async_gpu.cu:274-295,350-357; texture_core.h's tiles). After prepare, every level of
 * every texture's pyramid is held in half floats (a float texture's levels are held so, texture_core.h:28-31): the
 * levels wider or taller than coarse_side (0: 64) are kept by the host as 32x32 tiles of half-float rgb, of which the
 * device holds pool_tiles at a time (a full pool gives up its least recently read tile); the coarser levels are
 * always resident. A sample whose texture lookup finds a tile missing asks for every tile its footprint reads (at
 * the levels the footprint picks) and is parked: it ends, and runs again in a later launch (ahead of its chunks, or in
 * a launch of parked samples alone, x2dev_launch with no chunks) once x2dev_launch has sent up the tiles asked for.
 * With partial, each missing texel is read from its parent on the nearest level up that is resident instead
 * (async_gpu_device.h:49-58); the sample finishes into the provisional film and is parked to run again in full. The
 * frame calls
 * (x2dev_render_frame, x2dev_render_adaptive with host trees, x2dev_render_grid, x2dev_render_staged) run parked
 * samples until none is left, so their films are the resident frame's. The surface estimator's forward frames in the
 * one-kernel form only: the wavefront form, the adjoint, the prototype estimator, the device pass chain and media
 * refuse paged textures; x2dev_set_texture refuses a paged one. No counterpart for the pool size in the reference,
 * which sizes its cache from the card's memory. */
typedef struct x2_paging {
    uint32_t pool_tiles, coarse_side, partial;
    uint32_t host_refines; /* with partial: provisional samples go to the host as x2_refine records (the reference's
                              flow, async_gpu.cu:274-295), to be rendered again through x2dev_launch_refines,
                              instead of parking on the device */
} x2_paging;
typedef struct x2_paging_stats {
    uint64_t parked_outstanding; /* parked, not yet run again (GpuLane::parkedOutstanding, async_gpu.h:1427) */
    uint64_t parked_total, provisional;   /* this frame's (since begin) */
    uint64_t misses, uploads, evictions, overflow; /* since x2dev_page_textures */
    uint32_t tiles, pool_tiles, resident, pad;
} x2_paging_stats;
int x2dev_page_textures(x2dev* d, const x2_paging* paging);
int x2dev_paging_stats(const x2dev* d, x2_paging_stats* out);
uint64_t x2dev_parked_outstanding(const x2dev* d);
/* The provisional film (partial shading; downloaded with the film): the samples parked after a coarse render, each
 * one's latest rgb + alpha and a count of them. A sample run again in full leaves it (async_gpu.cu:282-295), so after
 * each launch it counts exactly the parked samples outstanding and a finished frame's is empty. */
int x2dev_provisional_film(const x2dev* d, const double** sums, const uint32_t** counts);

/* ---- Batched jobs. No counterpart in the reference, which renders one frame at a time and decided against CUDA
 * graphs (ASYNC.md:176-177): many independent render jobs, each its own scene, settings and film, run in one
 * persistent kernel that claims work from a device queue (the jobs' items end to end, shortest job first by
 * paths x log2(triangles)); the whole batch (film clears, a start stamp, the kernel) is captured once in a hipGraph
 * and replayed. The kernel's shape (workgroup threads, items a wave claims, the occupancy target) is chosen by
 * x2batch_tune, which times each candidate on a one-sample pass over every job. */
typedef struct x2batch x2batch;
typedef struct x2_batch_tuning {
    uint32_t block_threads;   /* 64, 128 or 256 */
    uint32_t claim_items;     /* items a wave claims at once: a multiple of 64 */
    uint32_t waves_per_simd;  /* occupancy target the kernel was built for: 2 or 3 */
    uint32_t pad;
    double seconds;           /* the tuning pass's device time with this shape (x2batch_tune) */
} x2_batch_tuning;
typedef struct x2_batch_stats {
    double wall_seconds, device_seconds;  /* host clock around the run; HIP events around its device work */
    uint64_t paths, rays_trace, rays_shadow;
    uint32_t jobs, errors, launches, graph; /* launches: kernel launches issued; graph: 1 when replayed from one */
    x2_batch_tuning tuning;
} x2_batch_stats;
int x2batch_create(int device, x2batch** out);
/* Builds the job's BVH8, uploads its scene and allocates its film; *job receives its index (submission order). */
int x2batch_add_job(x2batch* b, const x2_scene* scene, const x2_config* config, uint32_t* job);
/* Times the candidate shapes; the fastest becomes the batch's shape and is written to *best (may be null). */
int x2batch_tune(x2batch* b, x2_batch_tuning* best);
int x2batch_set_tuning(x2batch* b, const x2_batch_tuning* tuning);
/* Renders every job once (every pixel, every sample index). per_job_launches = 0: the persistent kernel over the
 * whole queue, from the captured graph; 1: one launch per job, in the same order (the comparison). job_seconds
 * (job_count doubles, may be null): each job's completion time from the run's start, by the device clock. */
int x2batch_run(x2batch* b, int per_job_launches, double* job_seconds, x2_batch_stats* stats);
/* The job's film (downloaded into host memory the batch owns, valid until the next run). */
int x2batch_job_film(x2batch* b, uint32_t job, const double** sums, const uint32_t** counts);
uint32_t x2batch_job_count(const x2batch* b);
void x2batch_destroy(x2batch* b);

/* ---- The work grid and the broker (async_grid.h:38-439, async_render.h:590-830,1668-1718). A frame's chunks wait in
 * the grid's cells; the GPU lane's broker and the CPU lane's workers claim them by the cellular rules, each chunk
 * exactly once, so a frame can be rendered by the GPU alone or by both devices together. Host code only: every call
 * here but x2dev_render_grid runs without a device. Cells and devices are indices; X2_NO_CELL: none. */
#define X2_DEVICE_CPU 0u
#define X2_DEVICE_GPU 1u
#define X2_NO_CELL 0xffffffffu
/* GridOptions (async_grid.h:103-113). */
typedef struct x2_grid_options {
    uint32_t cell_size;         /* pixels, a multiple of 8 (64) */
    uint32_t samples_per_chunk; /* sample indices a chunk holds at most (4) */
    int32_t mode;               /* 0: CPU only, 1: GPU only, 2: both (2) */
    uint32_t progressive;       /* nonzero: ranges of 1, 1, 2, then samples_per_chunk, handed out by the frontier */
    uint32_t sample_first;      /* a share of the frame: its sample indices [sample_first, sample_first + */
    uint32_t sample_count;      /* sample_count) (0: all of them) */
} x2_grid_options;
typedef struct x2_grid x2_grid;
/* Grid::Grid (async_grid.h:131-177): owners (owner_count == the cell count, else ignored) X2_DEVICE_* per cell, null for
 * the checkerboard; empty != 0: the cells start empty (chunks appended later); viewport {x0, y0, x1, y1} half open, or
 * null. The defaults above apply to a null options. */
int x2_grid_create(uint32_t width, uint32_t height, uint32_t spp, const x2_grid_options* options, const uint8_t* owners,
                   uint32_t owner_count, int empty, const int32_t* viewport, x2_grid** out);
void x2_grid_destroy(x2_grid* g);
uint32_t x2_grid_cell_count(const x2_grid* g);
uint64_t x2_grid_total_chunks(const x2_grid* g);
uint32_t x2_grid_owner(const x2_grid* g, uint32_t cell);
/* A cell's unclaimed chunks (0 once closed) and those a device may claim now (its frontier allowing). */
uint32_t x2_grid_remaining(const x2_grid* g, uint32_t cell);
uint32_t x2_grid_claimable(const x2_grid* g, uint32_t cell, uint32_t device);
int x2_grid_empty(const x2_grid* g);
/* Grid::claim (async_grid.h:231-237): up to n chunks of the cell into out; returns how many. */
uint32_t x2_grid_claim(x2_grid* g, uint32_t cell, uint32_t n, x2_chunk* out, uint32_t device);
/* Grid::append (:182-186), the coordinator only: more chunks for a cell. */
int x2_grid_append(x2_grid* g, uint32_t cell, const x2_chunk* chunks, uint32_t n);
/* Grid::choose (:325-359): the cell a worker of the device works next from `current` (X2_NO_CELL: none yet), or
 * X2_NO_CELL when it may take nothing now; salt spreads workers over equal cells. Grid::seed (:362-368): the w-th of
 * n workers' starting cell. */
uint32_t x2_grid_choose(const x2_grid* g, uint32_t device, uint32_t current, uint32_t salt);
uint32_t x2_grid_seed(const x2_grid* g, uint32_t device, uint32_t worker, uint32_t workers);
/* The progressive order's frontier of a device, and the frame's sample ranges (first, count pairs into out, up to
 * capacity; returns how many there are). */
uint32_t x2_grid_frontier(const x2_grid* g, uint32_t device);
uint32_t x2_grid_ranges(const x2_grid* g, uint32_t* out, uint32_t capacity);
/* Important parts first (:250-268): each cell's score (below zero: unknown); cells above the mean are important. */
int x2_grid_set_scores(x2_grid* g, const float* scores, uint32_t n);
int x2_grid_important(const x2_grid* g, uint32_t cell);
int x2_grid_reserved(const x2_grid* g, uint32_t cell);
/* The scene indexer's per-cell probability that the GPU is the efficient owner (device_suit.h). */
int x2_grid_set_suitability(x2_grid* g, const float* p, uint32_t n);
/* The rights of passage (:279-308): the GPU's samples a second, one CPU worker's, the CPU workers, a GPU launch's
 * target seconds; the time of the GPU's last claim; whether a device may work the other's reserved cell now. */
void x2_grid_set_rates(x2_grid* g, double gpu_samples, double cpu_worker_samples, uint32_t cpu_workers,
                       double window_seconds);
void x2_grid_gpu_claimed(x2_grid* g);
int x2_grid_passage(const x2_grid* g, uint32_t device, uint32_t cell);
/* A drain: nothing more is handed out and the grid counts as empty. */
void x2_grid_close(x2_grid* g);
/* Samples rendered in a cell by a device (for the ownership flips), and a GPU sample re-rendered by the CPU. */
void x2_grid_rendered(x2_grid* g, uint32_t cell, uint32_t device, uint64_t samples);
uint64_t x2_grid_rendered_by(const x2_grid* g, uint32_t cell, uint32_t device);
void x2_grid_deferred_to_cpu(x2_grid* g, uint32_t tile);
uint32_t x2_grid_cell_of(const x2_grid* g, uint32_t tile);
uint32_t x2_grid_tile_pixels(const x2_grid* g, uint32_t tile);

/* Ownership (async_grid.h:366-437): owners persisting across frames. x2_ownership_owners fills cells owners and
 * returns 1 once known for that many cells, 0 for the checkerboard; x2_ownership_finished records a finished frame. */
typedef struct x2_ownership x2_ownership;
int x2_ownership_create(x2_ownership** out);
void x2_ownership_destroy(x2_ownership* o);
int x2_ownership_owners(const x2_ownership* o, uint32_t cells, uint8_t* owners);
int x2_ownership_finished(x2_ownership* o, const x2_grid* g);

/* The CPU lane the broker drives (CpuLane, async_cpu.h:1584-1673): `threads` workers, each calling render for the
 * chunks it claims, on its own thread; render adds the chunk's samples to the caller's own film and returns 0, or
 * nonzero to fail the frame. */
typedef struct x2_cpu_lane {
    uint32_t threads;
    uint32_t pad;
    void* context;
    int (*render)(void* context, uint32_t worker, const x2_chunk* chunk);
} x2_cpu_lane;
/* RenderOptions (async_render.h:43-69): a GPU launch's target duration, the first launches' size before the GPU's
 * rate is known, and the GPU's cellular workers (each a share of a launch). Zero fields take the defaults 10 ms, 2048,
 * 32. */
typedef struct x2_broker_options {
    double window_ms;
    uint32_t first_chunks;
    uint32_t gpu_workers;
} x2_broker_options;
typedef struct x2_grid_frame {
    uint64_t gpu_chunks, cpu_chunks;     /* chunks each device rendered */
    uint64_t gpu_samples, cpu_samples;   /* pixel samples each device rendered */
    uint64_t gpu_stolen, cpu_stolen;     /* chunks taken from cells the other device owns */
    uint64_t launches, errors;           /* GPU launches; device sample errors */
    double seconds, gpu_rate, cpu_rate;  /* the frame's wall time; the GPU's samples a second (device time); the CPU
                                            lane's, over the frame */
    uint64_t rays_trace, rays_shadow, nodes_visited, triangles_tested;
} x2_grid_frame;
/* The broker's frame (Renderer::render, async_render.h:590-830): begins the device (x2dev_begin with the grid's
 * samples_per_chunk), starts the CPU lane's workers (none for a null cpu or mode 1), and feeds the GPU from the grid
 * (launchSize, async_render.h:1668-1718: launches of about window_ms at its measured rate, the last ones sized so both
 * lanes finish together) until every chunk is rendered; then downloads the GPU's film (x2dev_film). The grid's chunks
 * are consumed: a grid renders one frame. */
int x2dev_render_grid(x2dev* d, const x2_config* config, x2_grid* grid, const x2_cpu_lane* cpu,
                      const x2_broker_options* options, x2_grid_frame* out);
/* ---- The on-disk formats either side of a frame. Host code (x2dev_render_staged excepted). */
/* A film as a checkpoint render saves it (StagedFilm, xpu2/interactive/xpu2_session.cc:1338-1420): sums[pixel*4 + c]
 * FP64, counts[pixel], odd (the odd-index sums, or null), `done` of `spp` samples a pixel rendered. */
typedef struct x2_film_file {
    uint32_t width, height, spp, done;
    const double* sums;
    const uint32_t* counts;
    const double* odd;
} x2_film_file;
/* StagedFilm::save (xpu2_session.cc:1416-1452): the "X2RS" version 1 file (magic, version, width, height, spp, done,
 * AOV channels 0, closest filters 0; then the arrays sums, counts, odd, AOVs, depths, each its u64 length and its
 * elements; then 0 Cryptomatte categories), written to path.part and then renamed over path. */
int x2_film_save(const char* path, const x2_film_file* film);
/* StagedFilm::load (xpu2_session.cc:1454-1513): the save at path into caller-owned width*height films (odd may be
 * null; *has_odd says whether the save held one); X2_ERR_FILE, the film untouched, when there is no save, it is not
 * one, it is of another render (size, spp, outputs) or it is cut short. */
int x2_film_load(const char* path, uint32_t width, uint32_t height, uint32_t spp, double* sums, uint32_t* counts,
                 double* odd, uint32_t* done, uint32_t* has_odd);

/* The beauty output as a scanline OpenEXR file (RenderOutputWriter through OIIO, aov_write.h:176-255): channels A, B,
 * G, R (each pixel's sums / counts; 0 where it has no sample) as half or float and, with weight, "weight" (the count,
 * float); the display window the whole image, the data window the region (the film's pixels [x0, x1) x [y0, y1),
 * the whole image when empty); the film's rows run from the bottom, the file's from the top. Compression: */
#define X2_EXR_NONE 0u
#define X2_EXR_ZIPS 2u /* zlib, one scanline a block */
#define X2_EXR_ZIP 3u  /* zlib, 16 scanlines a block (OIIO's default for EXR) */
typedef struct x2_exr_image {
    uint32_t width, height;
    const double* sums;
    const uint32_t* counts;
    int32_t region[4];
    uint32_t half, compression, weight, pad;
} x2_exr_image;
int x2_write_exr(const char* path, const x2_exr_image* image);

/* A checkpoint render (stagedFrame, xpu2_session.cc:1540-1640): the uniform frame in stages of stage_samples sample
 * indices a pixel (0: a sixteenth of spp, at least 1), every pixel alike; after each stage the film is saved to path
 * (null: not saved); with resume, a save at path of the same render is taken up (x2dev_seed_film) and the frame goes
 * on from its samples (none, or another render's: the frame begins anew). max_stages (0: all) stops after that many
 * stages, as a render drained part way. The device's film (x2dev_film) holds the frame so far at return. */
typedef struct x2_staged {
    const char* path;
    uint32_t stage_samples, resume, max_stages, pad;
} x2_staged;
typedef struct x2_staged_frame {
    uint32_t stages, saves, launches, errors;
    uint32_t resumed, resumed_from, done, pad;
    double seconds, device_seconds;
} x2_staged_frame;
int x2dev_render_staged(x2dev* d, const x2_config* config, uint32_t samples_per_chunk, const x2_staged* stages,
                        x2_staged_frame* out);

/* GpuLane::chunkRoom (async_gpu.h:1725-1729): the chunks the next launch may take (the launch cap, and the mailbox
 * ring's free room). */
uint32_t x2dev_chunk_room(const x2dev* d);

/* The render that allocates nothing. No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). x2dev_launch's render
 * of `count` chunks (a host array, samples_per_chunk sample indices a chunk at most) with the prepared scene and
 * `config`, waited for before it returns, each sample's radiance added to the caller's device film (sums: width *
 * height * 4 FP64, counts: width * height u32; the caller zeroes them for a new frame). The chunks and the launch's
 * counters go to a static device array of the library's, 1 MiB + 256 B a GPU, held until the process ends, so at most
 * X2_STATIC_MAX_CHUNKS chunks a call (more: X2_ERR_ARGUMENT); no hipMalloc, hipFree or allocating library call is
 * made. Every call on one GPU shares that array: calls that overlap on one GPU race. Needs a prepared lane with no
 * launch in flight and resident textures; out (may be null) receives the completion. */
#define X2_STATIC_MAX_CHUNKS 65536u
int x2dev_render_chunks_static(x2dev* d, const x2_config* config, uint32_t samples_per_chunk, const x2_chunk* chunks,
                               uint32_t count, double* sums, uint32_t* counts, x2_completion* out);

#ifdef __cplusplus
}
#endif
#endif
