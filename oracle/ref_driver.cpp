// ref_driver.cpp — TEST INFRASTRUCTURE. A thin C entry layer over the reference's own component headers, compiled
// in place from the reference tree (the Makefile's `ref` target, output oracle/_ref/libref.so). The tests call it to
// check the oracle's restatement against the reference's code on the same inputs. It wraps, it does not restate:
// every computation below is the reference's (sample_stream.h, surface_core.h, camera_core.h, triangle_hit.h,
// resume_core.h, light_core.h, shade_lobes_eval.h, shade_material.h, texture_core.h). The shared integrator (async_integrator.h) is not buildable here
// (it includes NanoVDB, which the image lacks), so it is not wrapped.
#include <cstdint>
#include <cstring>

#include "camera_core.h"
#include "light_core.h"
#include "resume_core.h"
#include "sample_stream.h"
#include "shade_lobes_eval.h"
#include "shade_material.h"
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

// tex::lookup (texture_core.h:179-206) over the given levels (rgb floats, format 0 pages), with ImageMap's binding
// (unit scale, no offset or bias, periodic, no gamma) and the mode and swapCross given.
int ref_map_lookup(uint32_t count, const uint32_t* sizes, const float* texels, uint32_t mode, uint32_t swap,
                   const float f[6], float out[3]) {
    namespace tx = ii::tex;
    std::vector<std::vector<std::vector<V3>>> tiles(count);
    std::vector<std::vector<const V3*>> pages(count);
    tx::Image image;
    image.count = count;
    const float* at = texels;
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t width = sizes[2 * k], height = sizes[2 * k + 1];
        const uint32_t nx = (width + tx::Tile - 1) / tx::Tile, ny = (height + tx::Tile - 1) / tx::Tile;
        tiles[k].resize(size_t(nx) * ny);
        pages[k].resize(size_t(nx) * ny);
        for (uint32_t ty = 0; ty < ny; ++ty)
            for (uint32_t tx0 = 0; tx0 < nx; ++tx0) {
                const uint32_t stride = std::min<uint32_t>(tx::Tile, width - tx0 * tx::Tile);
                const uint32_t rows = std::min<uint32_t>(tx::Tile, height - ty * tx::Tile);
                std::vector<V3>& t = tiles[k][size_t(ty) * nx + tx0];
                t.resize(size_t(stride) * rows);
                for (uint32_t y = 0; y < rows; ++y)
                    for (uint32_t x = 0; x < stride; ++x)
                        t[size_t(y) * stride + x] = v3(at + (size_t(ty * tx::Tile + y) * width + tx0 * tx::Tile + x) * 3);
                pages[k][size_t(ty) * nx + tx0] = t.data();
            }
        image.levels[k].width = width;
        image.levels[k].height = height;
        image.levels[k].pages = pages[k].data();
        at += size_t(width) * height * 3;
    }
    tx::Binding b;
    b.image = &image;
    b.sx = b.sy = 1;
    b.ox = b.oy = 0;
    b.bias = 0;
    b.wrap = 1;
    b.gamma = 0;
    b.mode = mode;
    b.swapCross = swap;
    V3 r;
    if (tx::lookup(b, tx::Footprint{f[0], f[1], f[2], f[3], f[4], f[5]}, r) != tx::Ready) return 0;
    put(out, r);
    return 1;
}
// tex::footprint (texture_core.h:208-221), the spatial form: (u + v, v) and its derivatives from the ray's dx, dy.
int ref_footprint(const float a[3], const float b[3], const float c[3], const float org[3], const float dir[3],
                  const float dx[3], const float dy[3], float u, float v, uint32_t spp, float out[6]) {
    ii::surface::Ray ray{};
    ray.origin = v3(org);
    ray.direction = v3(dir);
    ray.dx = v3(dx);
    ray.dy = v3(dy);
    ii::tex::Footprint f{};
    if (!ii::tex::footprint(v3(a), v3(b), v3(c), ray, u, v, spp, f, true)) return 0;
    out[0] = f.u; out[1] = f.v; out[2] = f.dudx; out[3] = f.dvdx; out[4] = f.dudy; out[5] = f.dvdy;
    return 1;
}
// shade::hitTexDerivatives (shade_material.h:392-432) with the given corners' texture coordinates (uv: 6 floats).
void ref_hit_tex_derivatives(const float a[3], const float b[3], const float c[3], const float uv[6], const float org[3],
                             const float dir[3], const float dx[3], const float dy[3], uint32_t spp, float out[4]) {
    ii::shade::MeshAttributes m{};
    m.uvs = uv;
    m.uvRate = ii::shade::RateVertex;
    ii::surface::Ray ray{};
    ray.origin = v3(org);
    ray.direction = v3(dir);
    ray.dx = v3(dx);
    ray.dy = v3(dy);
    ii::shade::State st{};
    ii::shade::hitTexDerivatives(m, 0, 0, 1, 2, v3(a), v3(b), v3(c), ray, spp, st);
    out[0] = st.dsdx; out[1] = st.dtdx; out[2] = st.dsdy; out[3] = st.dtdy;
}
// shade::transferDifferentials and differentialTexDerivatives (shade_material.h:436-495): diff = dOx, dOy, dDx, dDy
// (12 floats); out = dPdx, dPdy, dsdx, dtdx, dsdy, dtdy; returns diffValid.
int ref_transfer(const float diff[12], int valid, const float org[3], const float dir[3], const float Ng[3],
                 const float P[3], const float dPds[3], const float dPdt[3], float out[10]) {
    ii::shade::RayDiff d;
    d.dOx = v3(diff); d.dOy = v3(diff + 3); d.dDx = v3(diff + 6); d.dDy = v3(diff + 9);
    d.valid = valid ? 1u : 0u;
    ii::shade::State st{};
    st.Ng = v3(Ng);
    st.P = v3(P);
    st.dPds = v3(dPds);
    st.dPdt = v3(dPdt);
    ii::shade::transferDifferentials(d, v3(org), v3(dir), st);
    ii::shade::differentialTexDerivatives(st);
    put(out, st.dPdx);
    put(out + 3, st.dPdy);
    out[6] = st.dsdx; out[7] = st.dtdx; out[8] = st.dsdy; out[9] = st.dtdy;
    return st.diffValid ? 1 : 0;
}
// shade::trianglePartials (shade_material.h:103-115).
int ref_triangle_partials(const float x0[3], const float x1[3], const float x2[3], const float uv[6], float dPds[3],
                          float dPdt[3]) {
    V3 s, t;
    if (!ii::shade::trianglePartials(v3(x0), v3(x1), v3(x2), uv, uv + 2, uv + 4, s, t)) return 0;
    put(dPds, s);
    put(dPdt, t);
    return 1;
}
// shade::localHemisphereCosineDifferentials (shade_math.h:279-288).
void ref_hemisphere_differentials(float r1, float r2, float out[6]) {
    V3 a, b;
    ii::shade::localHemisphereCosineDifferentials(r1, r2, a, b);
    put(out, a);
    put(out + 3, b);
}
// shade::scatterDifferentials (shade_material.h:691-705) for a one-lobe Bsdf (makeLobe's); the hit's state: its
// incoming differentials (12 floats), dPdx and dPdy, diffValid, P and wo. out: dOx, dOy, dDx, dDy; returns valid.
int ref_scatter_differentials(uint32_t kind, const float N[3], float roughness, const float incoming[12], const float dP[6],
                              int diffValid, const float P[3], const float wo[3], const float wi[3], float r1, float r2,
                              const float childOrigin[3], float primaryScale, float out[12]) {
    const uint32_t rec = makeLobe(kind, N, roughness);
    ii::shade::State st{};
    st.incoming.dOx = v3(incoming); st.incoming.dOy = v3(incoming + 3);
    st.incoming.dDx = v3(incoming + 6); st.incoming.dDy = v3(incoming + 9);
    st.incoming.valid = 1;
    st.dPdx = v3(dP);
    st.dPdy = v3(dP + 3);
    st.diffValid = diffValid != 0;
    st.P = v3(P);
    st.wo = v3(wo);
    const ii::shade::RayDiff d = ii::shade::scatterDifferentials(g_bsdf, rec, st, v3(wi), r1, r2, v3(childOrigin), primaryScale);
    put(out, d.dOx); put(out + 3, d.dOy); put(out + 6, d.dDx); put(out + 9, d.dDy);
    return int(d.valid);
}
}
