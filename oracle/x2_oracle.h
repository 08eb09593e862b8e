/* x2_oracle.h — TEST INFRASTRUCTURE. The CPU oracle: a plain C++ restatement of the reference's shared
 * integrator subset (xpu2/interactive/async_integrator.h:1986-3042 and the headers it calls), used only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg to check and time the HIP device. Nothing in the product
 * links, loads or calls this library.
 *
 * Inputs use the product's boundary records (include/x2dev.h) so both sides read the same bytes. */
#ifndef X2_ORACLE_H
#define X2_ORACLE_H
#include <stdint.h>
#include "../include/x2dev.h"
#ifdef __cplusplus
extern "C" {
#endif

/* RNG (sample_stream.h). */
void oracle_philox(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);
void oracle_block(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t group, uint32_t out[4]);
uint32_t oracle_word(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t dimension);
float oracle_unit(uint32_t bits);
uint32_t oracle_sobol(uint32_t index, uint32_t dimension);
void oracle_directions(uint32_t out[17 * 32]);
uint32_t oracle_owen_scramble(uint32_t x, uint32_t seed);
uint32_t oracle_sampling_word(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t dimension,
                              uint32_t policy);
uint64_t oracle_child_branch(uint64_t parent, uint32_t a, uint32_t b);
/* n consecutive draws of a vertex stream (async_integrator.h:151-176 paths::Stream). */
void oracle_stream(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t bounce, uint32_t n,
                   float* out);

/* Lobes (shade_lobes_eval.h). kind 0 Lambert (albedo 1), 1 GGX (scale 1). Returns f (with cosine) and pdf. */
void oracle_lobe_eval(uint32_t kind, const float N[3], const float Ng[3], const float wo[3], const float wi[3],
                      float roughness, float f[3], float* pdf);
void oracle_lobe_sample(uint32_t kind, const float N[3], const float Ng[3], const float wo[3], float roughness,
                        float r1, float r2, float wi[3], float f[3], float* pdf);
void oracle_sincos(float angle, float* s, float* c);

/* Lights (light_core.h). n may be null (no culling). Returns 1 when the sample/intersection exists. */
int oracle_light_sample(const x2_light* l, const float p[3], const float* n, const float r[3], float wi[3],
                        float* distance, float Li[3], float* pdf);
int oracle_light_intersect(const x2_light* l, const float p[3], const float wi[3], float Li[3], float* pdf,
                           float* distance);
int oracle_light_can_illuminate(const x2_light* l, const float p[3], const float* n);

/* The watertight triangle test. Returns 1 with *t on a hit within [near, far]. */
int oracle_triangle(const float o[3], const float d[3], float near_t, float far_t, const float a[3],
                    const float b[3], const float c[3], float* t);

/* The pinhole camera's ray through raster (x, y) (camera_core.h:217-250). */
int oracle_camera_ray(const x2_config* c, float x, float y, float origin[3], float dir[3], float* near_t, float* far_t);

/* FP64 hit reconstruction (resume_core.h:17 geometryHit): out = t, point[3], normal[3], origin[3], below[3]. */
int oracle_geometry_hit(const float o[3], const float d[3], const float a[3], const float b[3], const float c[3],
                        float out[13]);

/* A scene: copied, with the oracle's own binary SAH BVH built over it. */
typedef struct oracle_scene oracle_scene;
oracle_scene* oracle_scene_create(const x2_scene* scene);
void oracle_scene_destroy(oracle_scene* s);
/* Closest hit (geom, prim, t; geom = ~0u on a miss); brute != 0 tests every triangle instead of the BVH. */
void oracle_closest(const oracle_scene* s, const float o[3], const float d[3], float near_t, float far_t, int brute,
                    uint32_t* geom, uint32_t* prim, float* t);
int oracle_occluded(const oracle_scene* s, const float o[3], const float d[3], float near_t, float far_t, int brute);

/* One sample (async_integrator.h:1986 integratePaths). rgba receives radiance and alpha; counters[0] += traces,
 * counters[1] += shadow rays. Returns the SampleStatus (0 Done, 2 Error, 3 Skip). */
/* Test hook: the stratified draws of sample (pixel, index)'s first non-mirror vertex, 5 floats a draw (kind: 0 light,
 * 1 lobe; which; sample number; r0; r1) into out (capacity draws); *count receives how many were drawn. */
int oracle_first_hit_draws(const oracle_scene* o, const x2_config* c, uint32_t pixel, uint32_t index, float* out,
                           uint32_t capacity, uint32_t* count);
int oracle_sample(const oracle_scene* s, const x2_config* c, uint32_t pixel, uint32_t index, double rgba[4],
                  uint64_t counters[2]);
/* A whole uniform frame: every pixel's samples [0, spp) accumulated in sample-index order into FP64 sums (rgba) and
 * counts, over `threads` worker threads claiming 8x8 tiles. counters: traces, shadow rays, errors. */
int oracle_render(const oracle_scene* s, const x2_config* c, uint32_t threads, double* sums, uint32_t* counts,
                  uint64_t counters[3]);
/* The same over chunks (tile, first, count), each sample of each chunk once: the film of a partial frame. */
int oracle_render_chunks(const oracle_scene* s, const x2_config* c, const x2_chunk* chunks, uint32_t n,
                         uint32_t threads, double* sums, uint32_t* counts, uint64_t counters[3]);
int oracle_render_pixels(const oracle_scene* s, const x2_config* c, const uint32_t* first, const uint32_t* count,
                         uint32_t threads, double* sums, double* odd, uint32_t* counts, uint64_t counters[3]);

/* The path-replay adjoint of a uniform frame (the device's algorithm, FP64 gradients): d_image is dLoss / d(mean
 * rgb) per pixel; grad_textures (the scene's textures' texels concatenated) and grad_roughness (per material) are
 * added to. counters: traces, shadow rays, errors. */
int oracle_adjoint(const oracle_scene* s, const x2_config* c, uint32_t threads, const float* d_image,
                   double* grad_textures, double* grad_roughness, uint64_t counters[3]);
/* Parameter updates (the device's x2dev_set_texture / x2dev_set_material). */
void oracle_set_texture(oracle_scene* s, uint32_t texture, const float* texels);
/* Every level of a texture rounded to half floats (as paged tiles hold them), or kept in floats (0). */
void oracle_set_texture_half(oracle_scene* s, uint32_t texture, int half);
/* Every fetch at a level finer than `from` reads its ancestor at level `from` (a paged frame's first launch with
 * partial shading and no tile resident); from >= the level count: off. */
void oracle_set_texture_coarse(oracle_scene* s, uint32_t texture, uint32_t from);
void oracle_set_material(oracle_scene* s, uint32_t material, const x2_material* m);
// Test hook: the GGX lobe's value (not its sampling, pdf or MIS weights) uses roughness r for this material;
// r < 0 restores the material's own.
void oracle_set_eval_roughness(oracle_scene* s, uint32_t material, float r);
/* d ln f / d r of the GGX lobe, and one bilinear texture lookup at (u, v). */
float oracle_ggx_roughness_log_derivative(float r, const float N[3], const float wo[3], const float wi[3]);
int oracle_texture_lookup(const x2_texture* t, float u, float v, float rgb[3]);
/* The path tracer's map lookup over the texture's pyramid (mode 0 anisotropic, 1 trilinear closest texels, 2 one
 * level's closest texel) at footprint f = (s, t, ds/dx, dt/dx, ds/dy, dt/dy); half: levels rounded to half floats.
 * 0: an invalid value. The pyramid itself: its levels' texels end to end, sizes[2k], sizes[2k+1] level k's width and
 * height; returns the level count. */
int oracle_map_lookup(const x2_texture* t, int half, uint32_t mode, const float f[6], float rgb[3]);
/* The same with the cross terms swapped (1, the path tracer's maps) or in order (0, the prototype's binding). */
int oracle_map_lookup_swap(const x2_texture* t, int half, uint32_t mode, int swap, const float f[6], float rgb[3]);
uint32_t oracle_pyramid(const x2_texture* t, int half, float* out, uint64_t capacity, uint32_t* sizes);
/* Ray differentials, each as the estimator uses it (the reference's functions are named where they are defined). */
void oracle_hit_tex_derivatives(const float a[3], const float b[3], const float c[3], const float uv[6],
                                 const float org[3], const float dir[3], const float dx[3], const float dy[3],
                                 uint32_t spp, float out[4]);
int oracle_transfer(const float diff[12], int valid, const float org[3], const float dir[3], const float Ng[3],
                    const float P[3], const float dPds[3], const float dPdt[3], float out[10]);
int oracle_triangle_partials(const float x0[3], const float x1[3], const float x2[3], const float uv[6], float dPds[3],
                             float dPdt[3]);
void oracle_hemisphere_differentials(float r1, float r2, float out[6]);
int oracle_scatter_differentials(uint32_t kind, const float N[3], float roughness, const float incoming[12],
                                 const float dP[6], int diffValid, const float P[3], const float wo[3], const float wi[3],
                                 float r1, float r2, const float childOrigin[3], float primaryScale, float out[12]);
/* The prototype estimator's Lambert sample (surface::lambert): 1 with (wi, f, pdf), 0 for an invalid input. */
int oracle_surface_lambert(const float albedo[3], const float N[3], float u, float v, float wi[3], float f[3],
                           float* pdf);
/* The medium, piece by piece (for the tests): density at p; one ratio-tracking transmittance estimate along a ray
 * with its stream at (pixel, index, the config's seed, branch); the Henyey-Greenstein phase; one march (returns its
 * recorded steps; out: transmittance, emitted rgb, first step start, last step end, weight total, reached). */
float oracle_volume_density(const oracle_scene* s, const float p[3]);
float oracle_volume_transmittance(const oracle_scene* s, const x2_config* c, const float org[3], const float dir[3],
                                  float tnear, float tfar, uint32_t pixel, uint64_t index, uint64_t branch);
float oracle_phase_eval(float g, const float wo[3], const float wi[3]);
void oracle_phase_sample(float g, const float wo[3], float u1, float u2, float wi[3]);
int oracle_volume_march(const oracle_scene* s, const x2_config* c, const float org[3], const float dir[3], float tnear,
                        float tfar, uint32_t depth, uint64_t branch, float out[8]);

#ifdef __cplusplus
}
#endif
#endif
