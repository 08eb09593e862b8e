// ref_driver.cpp — TEST INFRASTRUCTURE. A thin C entry layer over the reference's own component headers, compiled
// in place from the reference tree (the Makefile's `ref` target, output oracle/_ref/libref.so). The tests call it to
// check the oracle's restatement against the reference's code on the same inputs. It wraps, it does not restate:
// every computation below is the reference's (sample_stream.h, surface_core.h, camera_core.h, triangle_hit.h,
// resume_core.h, light_core.h, shade_lobes_eval.h, texture_core.h). The shared integrator (async_integrator.h) is not buildable here
// (it includes NanoVDB, which the image lacks), so it is not wrapped.
#include <cstdint>
#include <cstring>

#include "camera_core.h"
#include "light_core.h"
#include "resume_core.h"
#include "sample_stream.h"
#include "shade_lobes_eval.h"
#include "surface_core.h"
#include "texture_core.h"
#include "triangle_hit.h"
#include <vector>

#include "../include/x2dev.h"

namespace ii = xpu2::interactive;
using ii::surface::V3;

namespace {
V3 v3(const float* p) { return V3{p[0], p[1], p[2]}; }
void put(float* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
ii::lights::Light toLight(const x2_light& x) {
    ii::lights::Light l{};
    l.type = x.type;
    l.sidedness = x.sidedness;
    l.index = 0;
    l.rayTerminator = 0;
    l.radiance = v3(x.radiance);
    l.position = v3(x.position);
    l.direction = v3(x.direction);
    for (int i = 0; i < 3; ++i) {
        l.local2Render[i] = v3(x.local_to_render + 3 * i);
        l.render2Local[i] = v3(x.render_to_local + 3 * i);
        l.render2LocalRot[i] = v3(x.render_to_local_rot + 3 * i);
    }
    l.scale = x.scale;
    l.invArea = x.inv_area;
    l.planeOffset = x.plane_offset;
    l.halfWidth = x.half_width;
    l.halfHeight = x.half_height;
    l.image = nullptr;
    l.hemispherical = x.hemispherical;
    l.fallback = V3{1, 1, 1};
    l.spread = 1.0f;
    l.visibilityMask = x.visibility_mask;
    l.presenceShadows = 0;
    l.hasPortal = 0;
    return l;
}
ii::shade::Bsdf g_bsdf;  // the reference's Bsdf is large; one per call on a single-threaded test driver
}  // namespace

extern "C" {
void ref_philox(const uint32_t c[4], const uint32_t k[2], uint32_t out[4]) {
    const auto w = ii::sample::philox({{c[0], c[1], c[2], c[3]}}, {k[0], k[1]});
    std::memcpy(out, w.v, 16);
}
void ref_block(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t group, uint32_t out[4]) {
    const auto w = ii::sample::block({pixel, index, stream, branch}, group);
    std::memcpy(out, w.v, 16);
}
uint32_t ref_sampling_word(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t dimension,
                           uint32_t policy) {
    return ii::sample::samplingWord({pixel, index, stream, branch}, dimension, policy);
}
uint32_t ref_sobol(uint32_t index, uint32_t dimension) { return ii::sample::sobol(index, dimension); }
uint32_t ref_owen_scramble(uint32_t x, uint32_t seed) { return ii::sample::owenScramble(x, seed); }
float ref_unit(uint32_t bits) { return ii::sample::unit(bits); }
void ref_directions(uint32_t out[17 * 32]) { std::memcpy(out, ii::sample::directions, sizeof(ii::sample::directions)); }
uint32_t ref_frame_seed(uint32_t frame, uint32_t eye, int lock) { return ii::surface::frameSeed(frame, eye, lock != 0); }
void ref_sincos(float angle, float* s, float* c) { ii::surface::sincosSample(angle, *s, *c); }

// The camera ray of a config's pinhole camera at raster (x, y) (camera_core.h:651 cameraSample).
int ref_camera_ray(const x2_config* cfg, float x, float y, float origin[3], float dir[3], float* nearT, float* farT) {
    ii::surface::Camera c{};
    std::memcpy(c.rasterToCamera.m, cfg->raster_to_camera, 64);
    std::memcpy(c.cameraToRender.m, cfg->camera_to_render, 64);
    c.nearClip = cfg->near_clip;
    c.farClip = cfg->far_clip;
    ii::surface::Ray r;
    if (ii::surface::cameraSample(c, x, y, 0.0f, 0.0f, 0.0f, r) != ii::surface::CameraResult::Made) return 0;
    put(origin, r.origin);
    put(dir, r.direction);
    *nearT = r.nearT;
    *farT = r.farT;
    return 1;
}
int ref_box_filter(float width, float r, float* out) { return ii::surface::filter(ii::surface::Filter::Box, width, r, *out) ? 1 : 0; }

// FP64 hit reconstruction (resume_core.h:17 geometryHit): out = t, point[3], normal[3], origin[3], below[3].
int ref_geometry_hit(const float o[3], const float d[3], const float a[3], const float b[3], const float c[3],
                     float out[13]) {
    ii::resume::Hit h;
    ii::surface::Ray ray{v3(o), v3(d), {0, 0, 0}, {0, 0, 0}, 0.0f, 3.4e38f};
    if (!ii::resume::geometryHit(h, ray, v3(a), v3(b), v3(c))) return 0;
    out[0] = h.t;
    put(out + 1, h.point);
    put(out + 4, h.normal);
    put(out + 7, h.origin);
    put(out + 10, h.below);
    return 1;
}

int ref_light_sample(const x2_light* lx, const float p[3], const float* n, const float r[3], float wi[3],
                     float* distance, float Li[3], float* pdf) {
    const auto l = toLight(*lx);
    V3 nn = n ? v3(n) : V3{0, 0, 0}, w;
    ii::lights::Hit h{};
    if (!ii::lights::sample(l, v3(p), n ? &nn : nullptr, r, w, h)) return 0;
    put(Li, ii::lights::eval(l, v3(p), w, h, *pdf));
    put(wi, w);
    *distance = h.distance;
    return 1;
}
int ref_light_intersect(const x2_light* lx, const float p[3], const float wi[3], float Li[3], float* pdf,
                        float* distance) {
    const auto l = toLight(*lx);
    ii::lights::Hit h{};
    if (!ii::lights::intersect(l, v3(p), nullptr, v3(wi), 3.4028234663852886e+38f, h)) return 0;
    put(Li, ii::lights::eval(l, v3(p), v3(wi), h, *pdf));
    *distance = h.distance;
    return 1;
}
int ref_light_can_illuminate(const x2_light* lx, const float p[3], const float* n) {
    const auto l = toLight(*lx);
    V3 nn = n ? v3(n) : V3{0, 0, 0};
    return ii::lights::canIlluminate(l, v3(p), n ? &nn : nullptr, 0.0f) ? 1 : 0;
}

// A one-lobe Bsdf: kind 0 Lambert (albedo 1, lambertLobe), 1 GGX (cookTorranceLobe KindGGX, no compensation).
static uint32_t makeLobe(uint32_t kind, const float N[3], float roughness) {
    ii::shade::reset(g_bsdf);
    uint32_t rec;
    if (kind == 0) rec = ii::shade::lambertLobe(g_bsdf, v3(N), V3{1, 1, 1}, true);
    else rec = ii::shade::cookTorranceLobe(g_bsdf, ii::shade::KindGGX, v3(N), roughness, V3{0, 0, 0}, V3{0, 0, 0});
    ii::shade::addLobe(g_bsdf, rec);
    return rec;
}
void ref_lobe_eval(uint32_t kind, const float N[3], const float Ng[3], const float wo[3], const float wi[3],
                   float roughness, float f[3], float* pdf) {
    const uint32_t rec = makeLobe(kind, N, roughness);
    const ii::shade::Slice s{v3(Ng), v3(wo), ii::shade::AllLobeTypes, ii::shade::StfOff, true, true};
    put(f, ii::shade::lobeEval(g_bsdf, rec, s, v3(wi), *pdf));
}
void ref_lobe_sample(uint32_t kind, const float N[3], const float Ng[3], const float wo[3], float roughness, float r1,
                     float r2, float wi[3], float f[3], float* pdf) {
    const uint32_t rec = makeLobe(kind, N, roughness);
    const ii::shade::Slice s{v3(Ng), v3(wo), ii::shade::AllLobeTypes, ii::shade::StfOff, true, true};
    V3 w;
    put(f, ii::shade::lobeSample(g_bsdf, rec, s, r1, r2, w, *pdf));
    put(wi, w);
}
// The prototype estimator's Lambert sample (surface::lambert, surface_core.h:158-165).
int ref_surface_lambert(const float albedo[3], const float N[3], float u, float v, float wi[3], float f[3], float* pdf) {
    ii::surface::Lambert m{v3(albedo), V3{0, 0, 0}};
    ii::surface::Sample s{};
    if (!ii::surface::lambert(m, v3(N), u, v, s)) return 0;
    put(wi, s.wi);
    put(f, s.f);
    *pdf = s.pdf;
    return 1;
}
// The reference's bilinear filter (tex::bilinear on level 0, periodic, float texels) at the lookup's (s, t) =
// (u, 1 - v) with unit scale and no offset; the flat row-major texels are laid into its 32 x 32 tile pages first.
int ref_bilinear(uint32_t width, uint32_t height, const float* texels, float u, float v, float out[3]) {
    namespace tx = ii::tex;
    const uint32_t nx = (width + tx::Tile - 1) / tx::Tile, ny = (height + tx::Tile - 1) / tx::Tile;
    std::vector<std::vector<V3>> tiles(size_t(nx) * ny);
    std::vector<const V3*> pages(size_t(nx) * ny);
    for (uint32_t ty = 0; ty < ny; ++ty)
        for (uint32_t tx0 = 0; tx0 < nx; ++tx0) {
            const uint32_t stride = std::min<uint32_t>(tx::Tile, width - tx0 * tx::Tile);
            const uint32_t rows = std::min<uint32_t>(tx::Tile, height - ty * tx::Tile);
            std::vector<V3>& t = tiles[size_t(ty) * nx + tx0];
            t.resize(size_t(stride) * rows);
            for (uint32_t y = 0; y < rows; ++y)
                for (uint32_t x = 0; x < stride; ++x)
                    t[size_t(y) * stride + x] = v3(texels + (size_t(ty * tx::Tile + y) * width + tx0 * tx::Tile + x) * 3);
            pages[size_t(ty) * nx + tx0] = t.data();
        }
    tx::Image image;
    image.count = 1;
    image.levels[0].width = width;
    image.levels[0].height = height;
    image.levels[0].pages = pages.data();
    tx::NoSink sink;
    V3 r;
    const auto status = tx::bilinear(image, 0, u, 1 - v, true, r, sink);
    if (status != tx::Ready) return 0;
    put(out, r);
    return 1;
}
}
