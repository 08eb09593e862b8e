// x2_kernel.hip — the render megakernel for gfx950: one lane = one sample (pixel, index), a wave64 = one 8x8 tile
// at one sample index (the reference's item order, async_gpu.cu:429-436), so the wave's camera rays are coherent.
//
// A sample runs the reference's estimator whole (async_integrator.h:1986-3042, the Lambert/GGX/rect/env subset):
// camera ray, closest-hit trace through the BVH8, hit reconstruction in FP64, next-event estimation over every
// light with per-lobe MIS, a BSDF sample per lobe with light MIS and bounce-ray reuse, threshold Russian roulette,
// and the pending-ray stack with thinning. Its radiance goes to the FP64 film by atomics with an exact count
// (async_gpu.cu:331-336).
//
// Built with -ffp-contract=off and correctly rounded division and square root, so every float expression of the
// estimator rounds as written, as the host oracle's does. Only the box tests use fused multiply-adds: they are
// conservative (Ize 2013) and decide what is tested, never the result.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cfloat>

#include "x2_adaptive_core.h"
#include "x2_layout.h"
#include "x2_rng.h"

namespace x2 {

__constant__ uint32_t kSobolDirections[17 * 32];

// ---------------------------------------------------------------------------------------------------------------
// Vectors in the reference's operation order (surface_core.h:10-18, shade_math.h:15-26, 65-69, 113-120).
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
__device__ __forceinline__ V3 operator*(float b, V3 a) { return {a.x * b, a.y * b, a.z * b}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float b) { return {a.x / b, a.y / b, a.z / b}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ bool finite1(float a) { return a == a && a <= 3.402823466e38f && a >= -3.402823466e38f; }
__device__ __forceinline__ bool finite3(V3 a) { return finite1(a.x) && finite1(a.y) && finite1(a.z); }
__device__ __forceinline__ bool normalizeTo(V3 a, V3& b) {
    const float l = dot(a, a);
    if (!finite1(l) || l <= 0) return false;
    b = a * (1.0f / sqrtf(l));
    return true;
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float luminance(V3 c) { return 0.299f * c.x + 0.587f * c.y + 0.114f * c.z; }
__device__ __forceinline__ bool isExactlyZero(V3 c) { return c.x == 0 && c.y == 0 && c.z == 0; }
__device__ __forceinline__ bool isEqual0(float x) { return fabsf(x - 0.0f) <= fmaxf(fabsf(x), 1.0f) * 1e-6f; }
__device__ __forceinline__ bool isBlack(V3 c) { return isEqual0(c.x) && isEqual0(c.y) && isEqual0(c.z); }
__device__ __forceinline__ V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

constexpr float Pi = 3.14159265358979323846f, TwoPi = 6.28318530717958647692f;
constexpr float OneOverPi = 0.318309886183790671538f, Epsilon = 1e-6f;
constexpr float MaxFloat = 3.402823466e+38f;

// The shared sine/cosine polynomial (surface_core.h:116-126, ISPC 1.31's coefficients) for every sin/cos on the
// path, so the device and the host round alike.
__device__ __forceinline__ void sincosPoly(float angle, float& sn, float& cs) {
    const float kf = floorf(angle * .6366197466850281f), x = angle - kf * 1.5707963705062866f, x2 = x * x;
    const int k = int(kf) & 3;
    float s = -2.5029327943570934e-8f;
    s = s * x2 + .000002760012648650445f;
    s = s * x2 - .00019842604524455965f;
    s = s * x2 + .008333347737789154f;
    s = s * x2 - .1666666716337204f;
    s = (s * x2 + 1) * x;
    float c = -.0000002596301840185333f;
    c = c * x2 + .000024756234779488295f;
    c = c * x2 - .001388833043165505f;
    c = c * x2 + .04166664183139801f;
    c = c * x2 - .5f;
    c = c * x2 + 1;
    sn = k == 0 ? s : k == 1 ? c : k == 2 ? -s : -c;
    cs = k == 0 ? c : k == 1 ? -s : k == 2 ? -c : s;
}

// ---------------------------------------------------------------------------------------------------------------
// Camera (camera_core.h:53-67, :217-250 without depth of field; surface_core.h:51-52 box filter).
struct Ray {
    V3 origin, direction;
    float nearT, farT;
};
__device__ __forceinline__ V3 vectorM(const float* m, V3 a) {
    return {(a.x * m[0] + a.y * m[4]) + a.z * m[8], (a.x * m[1] + a.y * m[5]) + a.z * m[9],
            (a.x * m[2] + a.y * m[6]) + a.z * m[10]};
}
__device__ __forceinline__ bool cameraRay(const x2_config& c, float x, float y, Ray& out) {
    const float* m = c.raster_to_camera;
    const V3 p{x, y, -1.0f};
    const V3 v = vectorM(m, p) + V3{m[12], m[13], m[14]};
    const float w = ((p.x * m[3] + p.y * m[7]) + p.z * m[11]) + m[15];
    const V3 dir = w == 0.0f ? V3{0, 0, 0} : V3{v.x / w, v.y / w, v.z / w};
    V3 n;
    if (!normalizeTo(dir, n) || n.z >= 0) return false;
    const float* q = c.camera_to_render;
    Ray r;
    r.origin = vectorM(q, V3{0, 0, 0}) + V3{q[12], q[13], q[14]};
    r.nearT = (-c.near_clip - 0.0f) / n.z;
    r.farT = (-c.far_clip - 0.0f) / n.z;
    if (!normalizeTo(vectorM(q, dir), r.direction) || !finite3(r.origin) || !finite1(r.nearT) || !finite1(r.farT))
        return false;
    out = r;
    return true;
}

// ---------------------------------------------------------------------------------------------------------------
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013). The permutation and shear are per ray; the edge
// functions fall back to double only when they round to zero.
struct WoopRay {
    V3 org;
    int kx, ky, kz;
    float sx, sy, sz;
    float nearT, farT;
};
__device__ __forceinline__ float comp(V3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; }
__device__ __forceinline__ WoopRay woopSetup(V3 org, V3 dir, float nearT, float farT) {
    WoopRay w;
    w.org = org;
    const float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
    w.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    w.kx = w.kz == 2 ? 0 : w.kz + 1;
    w.ky = w.kx == 2 ? 0 : w.kx + 1;
    if (comp(dir, w.kz) < 0.0f) {
        const int t = w.kx;
        w.kx = w.ky;
        w.ky = t;
    }
    const float dz = comp(dir, w.kz);
    w.sx = comp(dir, w.kx) / dz;
    w.sy = comp(dir, w.ky) / dz;
    w.sz = 1.0f / dz;
    w.nearT = nearT;
    w.farT = farT;
    return w;
}
__device__ __forceinline__ float edgeFunction(float a, float b, float c, float d) {
    const float e = a * b - c * d;
    if (e != 0.0f) return e;
    return float(double(a) * double(b) - double(c) * double(d));
}
__device__ __forceinline__ bool woopHit(const WoopRay& w, V3 a, V3 b, V3 c, float& t) {
    const V3 A = a - w.org, B = b - w.org, C = c - w.org;
    const float Az = comp(A, w.kz), Bz = comp(B, w.kz), Cz = comp(C, w.kz);
    const float Ax = comp(A, w.kx) - w.sx * Az, Ay = comp(A, w.ky) - w.sy * Az;
    const float Bx = comp(B, w.kx) - w.sx * Bz, By = comp(B, w.ky) - w.sy * Bz;
    const float Cx = comp(C, w.kx) - w.sx * Cz, Cy = comp(C, w.ky) - w.sy * Cz;
    const float U = edgeFunction(Cx, By, Cy, Bx), V = edgeFunction(Ax, Cy, Ay, Cx), W = edgeFunction(Bx, Ay, By, Ax);
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = (U * (w.sz * Az) + V * (w.sz * Bz)) + W * (w.sz * Cz);
    t = T / det;
    return t >= w.nearT && t <= w.farT;
}

// ---------------------------------------------------------------------------------------------------------------
// BVH8 traversal. Closest hit keeps the least (t, geom, prim) key (triangle_hit.h:26 order), so the result does not
// depend on the order nodes are visited in. Boxes are culled conservatively: the far distance is scaled by
// 1 + 2 gamma(3) (Ize 2013) and compared against the best distance with a relative slack of 2^-12.
constexpr float FarScale = 1.0f + 2.0f * (3.0f * 0x1p-24f) / (1.0f - 3.0f * 0x1p-24f);
static_assert(FarScale - 1.0f == 3.0f * 0x1p-23f, "the far scale is 1 + 3 ulp of 1");
constexpr float Slack = 1.0f + 0x1p-12f;

// A lane's traversal stack in private memory: the closest-hit walk's entries are (child ref, entry distance) pairs,
// the any-hit walk's a child ref alone (one word a push).
struct TravStack {
    alignas(8) uint32_t spill[2 * TraversalStack];  // words
    __device__ __forceinline__ void put(int i, uint2 e) {
        spill[2 * i] = e.x;
        spill[2 * i + 1] = e.y;
    }
    __device__ __forceinline__ uint2 get(int i) const { return make_uint2(spill[2 * i], spill[2 * i + 1]); }
    __device__ __forceinline__ void put1(int i, uint32_t e) { spill[i] = e; }
    __device__ __forceinline__ uint32_t get1(int i) const { return spill[i]; }
};

struct Counters {
    uint32_t traces = 0, shadows = 0, nodes = 0, triangles = 0;
};

struct BoxRay {
    float ix, iy, iz, ox, oy, oz;  // 1/d and -o/d
    bool sx, sy, sz;               // d >= 0 per axis: which plane is near
};
__device__ __forceinline__ BoxRay boxRay(V3 o, V3 d) {
    BoxRay b;
    const float dx = d.x != 0.0f ? d.x : copysignf(1e-30f, d.x), dy = d.y != 0.0f ? d.y : copysignf(1e-30f, d.y),
                dz = d.z != 0.0f ? d.z : copysignf(1e-30f, d.z);
    b.ix = 1.0f / dx;
    b.iy = 1.0f / dy;
    b.iz = 1.0f / dz;
    b.ox = -o.x * b.ix;
    b.oy = -o.y * b.iy;
    b.oz = -o.z * b.iz;
    b.sx = dx >= 0.0f;
    b.sy = dy >= 0.0f;
    b.sz = dz >= 0.0f;
    return b;
}

__device__ __forceinline__ void cswap(unsigned long long& a, unsigned long long& b) {
    const unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Tests a node's 8 children (a node in one 128-B line, Node8Q: six 16-B loads); returns the entry distance (>= 0)
// packed over the child reference, ~0 for a miss. Each child's bound is decoded as origin + q * step (two roundings,
// as the builder checked it), then the slab test.
__device__ __forceinline__ void nodeChildren(const Node8Q& n, const BoxRay& r, float lo, float hi,
                                             unsigned long long key[8]) {
    const uint4* q = reinterpret_cast<const uint4*>(&n);
    const uint4 h = q[0], b0 = q[1], b1 = q[2], b2 = q[3], c0 = q[4], c1 = q[5];
    const float ox = __uint_as_float(h.x), oy = __uint_as_float(h.y), oz = __uint_as_float(h.z);
    const float sx = __uint_as_float((h.w & 0xffu) << 23), sy = __uint_as_float(((h.w >> 8) & 0xffu) << 23),
                sz = __uint_as_float(((h.w >> 16) & 0xffu) << 23);
    const uint32_t ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    auto at = [](uint32_t a, uint32_t b, int i) {  // byte i of the eight in (a, b)
        return float(((i < 4 ? a : b) >> (8 * (i & 3))) & 0xffu);
    };
    // the near and far planes' bytes per axis, chosen once for the node by the ray's direction (lows near when the
    // direction is positive)
    const uint32_t nx0 = r.sx ? b0.x : b0.z, nx1 = r.sx ? b0.y : b0.w, fx0 = r.sx ? b0.z : b0.x, fx1 = r.sx ? b0.w : b0.y;
    const uint32_t ny0 = r.sy ? b1.x : b1.z, ny1 = r.sy ? b1.y : b1.w, fy0 = r.sy ? b1.z : b1.x, fy1 = r.sy ? b1.w : b1.y;
    const uint32_t nz0 = r.sz ? b2.x : b2.z, nz1 = r.sz ? b2.y : b2.w, fz0 = r.sz ? b2.z : b2.x, fz1 = r.sz ? b2.w : b2.y;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        // q * step is exact (an 8-bit integer times a power of two), so the fused form rounds once, as the builder's
        // origin + q * step does: the same bound in one instruction
        const float nx = __builtin_fmaf(__builtin_fmaf(at(nx0, nx1, i), sx, ox), r.ix, r.ox);
        const float fx = __builtin_fmaf(__builtin_fmaf(at(fx0, fx1, i), sx, ox), r.ix, r.ox);
        const float ny = __builtin_fmaf(__builtin_fmaf(at(ny0, ny1, i), sy, oy), r.iy, r.oy);
        const float fy = __builtin_fmaf(__builtin_fmaf(at(fy0, fy1, i), sy, oy), r.iy, r.oy);
        const float nz = __builtin_fmaf(__builtin_fmaf(at(nz0, nz1, i), sz, oz), r.iz, r.oz);
        const float fz = __builtin_fmaf(__builtin_fmaf(at(fz0, fz1, i), sz, oz), r.iz, r.oz);
        const float tn = fmaxf(fmaxf(nx, ny), nz);
        float tf = fminf(fminf(fx, fy), fz);
        // tf (1 + 2 gamma(3)) for tf > 0 in one rounding, as tf * FarScale (FarScale - 1 = 3 * 2^-23 exactly);
        // towards zero otherwise
        tf = __builtin_fmaf(fabsf(tf), FarScale - 1.0f, tf);
        const bool hit = ch[i] != EmptyChild && tn <= tf && tn <= hi && tf >= lo;
        key[i] = hit ? ((unsigned long long)__float_as_uint(fmaxf(tn, 0.0f)) << 32) | ch[i] : ~0ull;
    }
}

__device__ __forceinline__ bool keyLess(float t, uint32_t g, uint32_t p, float bt, uint32_t bg, uint32_t bp) {
    return t < bt || (t == bt && (g < bg || (g == bg && p < bp)));
}

// Traversal order is "speculative while-while" (Aila and Laine 2009): a lane descends inner nodes, postponing the
// first leaf it reaches, until every lane of the wave still descending holds a leaf; then the wave tests leaves
// together. Node and leaf work stay apart, so a wave is not split between the two on every step. The closest hit
// is the least key whatever the order, so the result is the same as a depth-first walk's.
constexpr uint32_t Sentinel = 0x7fffffffu;  // an empty stack's pop: neither a node index nor a leaf
// traceWave's wave leaves its inner nodes for the postponed leaves once at most this many of its searching lanes
// still hold none (rather than none): the wave then does not wait on a few lanes' long descents, and the leaves'
// closer hits cull sooner (the 1080p icospheres frame: 820 Mpaths/s at 4 or 8, 814 at 16, 803 at 1, 780 at 0).
constexpr int LeafWait = 8;


// Back-face culling of single-sided meshes (backFaceCullingFilter, async_cpu.h:122-133; the GPU lane's
// OPTIX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES, async_gpu_device.h:333,527): a hit whose ray runs along the triangle's
// (b - a) x (c - a) is no hit. triCull is null unless the scene has a single-sided mesh.
__device__ __forceinline__ bool culled(const RenderParams& P, uint32_t tri, V3 a, V3 b, V3 c, V3 dir) {
    return P.triCull && P.triCull[tri] && dot(dir, cross(b - a, c - a)) >= 0.0f;
}

// Closest hit among triangles with t in [near, far]; returns the triangle record index (~0u: none).
__device__ uint32_t closestHit(const RenderParams& P, V3 org, V3 dir, float nearT, float farT, float& tOut,
                               Counters& k, TravStack& st) {
    const WoopRay w = woopSetup(org, dir, nearT, farT);
    const BoxRay br = boxRay(org, dir);
    const float lo = nearT > 0 ? nearT / Slack : nearT * Slack;
    uint32_t best = ~0u;
    float bestT = farT;
    int sp = 0;
    bool cached = false;  // the stack's top entry held in registers (traceWave's form)
    uint2 top = make_uint2(0u, 0u);
    // The nearest stacked entry still worth visiting, or Sentinel.
    auto pop = [&]() -> uint32_t {
        while (sp > 0) {
            --sp;
            const float bound = best == ~0u ? farT : bestT;
            const uint2 e = cached ? top : st.get(sp);
            cached = false;
            if (__uint_as_float(e.y) <= (bound > 0 ? bound * Slack : bound / Slack)) return e.x;
        }
        return Sentinel;
    };
    auto visitNode = [&](uint32_t node) -> uint32_t {  // pushes the far children, returns the nearest (or a pop)
        ++k.nodes;
        const float hi = best == ~0u ? (farT > 0 ? farT * Slack : farT / Slack) : (bestT > 0 ? bestT * Slack : bestT / Slack);
        unsigned long long key[8];
        nodeChildren(P.nodes[node], br, lo, hi, key);
        // The nearest child next, the others stacked in child order: sorting all eight (a 19-comparator network on
        // the 64-bit keys) culled no more than it cost (583 against 578 Mpaths/s on the 1080p icospheres frame).
        unsigned long long m = key[0];
#pragma unroll
        for (int i = 1; i < 8; ++i) m = key[i] < m ? key[i] : m;
#pragma unroll
        for (int i = 7; i >= 0; --i)
            if (key[i] != ~0ull && key[i] != m) {
                if (cached) st.put(sp - 1, top);
                top = make_uint2(uint32_t(key[i]), uint32_t(key[i] >> 32));
                cached = true;
                ++sp;
            }
        return m != ~0ull ? uint32_t(m) : pop();
    };
    auto visitLeaf = [&](uint32_t leafRef) {
        const uint32_t first = (leafRef & 0x7fffffffu) >> 3, count = (leafRef & 7u) + 1;
        for (uint32_t i = 0; i < count; ++i) {
            ++k.triangles;
            float t;
            const float* v = P.triVerts + size_t(first + i) * 9;
            const V3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
            if (!woopHit(w, a, b, c, t) || culled(P, first + i, a, b, c, dir)) continue;
            if (best == ~0u || t < bestT) {
                best = first + i;
                bestT = t;
            } else if (t == bestT) {  // a tie in distance: the least (geom, prim) wins, from the records
                // One 64-bit comparison and a select: hipcc (ROCm 7.2) has compiled the two-level comparison's
                // nested branches with its taken side emptied, keeping the old best on an equal geom and a lesser
                // prim (renderTexKernel, once culled() was added above).
                const TriRecord &r = P.tris[first + i], &q = P.tris[best];
                const unsigned long long rk = (unsigned long long)r.geom << 32 | r.prim,
                                         qk = (unsigned long long)q.geom << 32 | q.prim;
                best = rk < qk ? first + i : best;
            }
        }
    };
    uint32_t ref = 0, leaf = 0;  // leaf: a postponed leaf reference (0: none)
    for (;;) {
        while (ref < Sentinel) {  // an inner node
            ref = visitNode(ref);
            if ((ref & LeafBit) && leaf == 0) {
                leaf = ref;
                ref = pop();
            }
            if (!__any(leaf == 0)) break;
        }
        while (leaf || (ref & LeafBit)) {
            if (!leaf) {
                leaf = ref;
                ref = pop();
            }
            visitLeaf(leaf);
            leaf = 0;
        }
        if (ref == Sentinel) break;
    }
    tOut = bestT;
    return best;
}

// Any hit with t in [near, far] (a shadow ray; TERMINATE_ON_FIRST_HIT, async_gpu_device.h:521-533), children
// unsorted, the same while-while order.
__device__ bool anyHit(const RenderParams& P, V3 org, V3 dir, float nearT, float farT, Counters& k, TravStack& st) {
    const WoopRay w = woopSetup(org, dir, nearT, farT);
    const BoxRay br = boxRay(org, dir);
    const float lo = nearT > 0 ? nearT / Slack : nearT * Slack, hi = farT > 0 ? farT * Slack : farT / Slack;
    int sp = 0;
    bool cached = false;  // the stack's top entry held in a register (traceWave's form)
    uint32_t top = 0;
    auto pop = [&]() -> uint32_t {
        if (sp == 0) return Sentinel;
        --sp;
        const uint32_t e = cached ? top : st.get1(sp);
        cached = false;
        return e;
    };
    // Unsorted: sorting the keys measured slower here.
    auto visitNode = [&](uint32_t node) -> uint32_t {
        ++k.nodes;
        unsigned long long key[8];
        nodeChildren(P.nodes[node], br, lo, hi, key);
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (key[i] != ~0ull) {
                if (cached) st.put1(sp - 1, top);
                top = uint32_t(key[i]);
                cached = true;
                ++sp;
            }
        return pop();
    };
    auto visitLeaf = [&](uint32_t leafRef) -> bool {
        const uint32_t first = (leafRef & 0x7fffffffu) >> 3, count = (leafRef & 7u) + 1;
        for (uint32_t i = 0; i < count; ++i) {
            ++k.triangles;
            float t;
            const float* v = P.triVerts + size_t(first + i) * 9;
            const V3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
            if (woopHit(w, a, b, c, t) && !culled(P, first + i, a, b, c, dir)) return true;
        }
        return false;
    };
    uint32_t ref = 0, leaf = 0;
    for (;;) {
        while (ref < Sentinel) {
            ref = visitNode(ref);
            if ((ref & LeafBit) && leaf == 0) {
                leaf = ref;
                ref = pop();
            }
            if (!__any(leaf == 0)) break;
        }
        while (leaf || (ref & LeafBit)) {
            if (!leaf) {
                leaf = ref;
                ref = pop();
            }
            if (visitLeaf(leaf)) return true;
            leaf = 0;
        }
        if (ref == Sentinel) return false;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Hit reconstruction (resume_core.h:17-47, triangle_hit.h:6-21): FP64 t, u, v from the float vertices and the ray,
// the point rounded once, conservative offset origins above and below the surface.
struct Hit {
    uint32_t geom, tri;
    float t, u, v;
    V3 normal, origin, below, point;
};
__device__ bool geometryHit(Hit& h, const Ray& ray, V3 a, V3 b, V3 c) {
    const V3 e1 = b + (a * -1.0f), e2 = c + (a * -1.0f);
    if (!normalizeTo(cross(e1, e2), h.normal)) return false;
    if (dot(h.normal, ray.direction) > 0) h.normal = h.normal * -1.0f;
    const double ax = a.x, ay = a.y, az = a.z;
    const double ex = double(b.x) - ax, ey = double(b.y) - ay, ez = double(b.z) - az;
    const double fx = double(c.x) - ax, fy = double(c.y) - ay, fz = double(c.z) - az;
    const double dx = ray.direction.x, dy = ray.direction.y, dz = ray.direction.z;
    const double px = dy * fz - dz * fy, py = dz * fx - dx * fz, pz = dx * fy - dy * fx;
    const double determinant = (ex * px + ey * py) + ez * pz;
    if (determinant == 0) return false;
    const double tx = double(ray.origin.x) - ax, ty = double(ray.origin.y) - ay, tz = double(ray.origin.z) - az;
    const double qx = ty * ez - tz * ey, qy = tz * ex - tx * ez, qz = tx * ey - ty * ex;
    const double pu = ((tx * px + ty * py) + tz * pz) / determinant;
    const double pv = ((dx * qx + dy * qy) + dz * qz) / determinant;
    const double pt = ((fx * qx + fy * qy) + fz * qz) / determinant;
    if (!finite1(float(pt)) || !finite1(float(pu)) || !finite1(float(pv))) return false;
    h.t = float(pt);
    h.u = float(pu);
    h.v = float(pv);
    const V3 p{float(ax + ((double(b.x) - ax) * pu + (double(c.x) - ax) * pv)),
               float(ay + ((double(b.y) - ay) * pu + (double(c.y) - ay) * pv)),
               float(az + ((double(b.z) - az) * pu + (double(c.z) - az) * pv))};
    h.point = p;
    const V3 error{fabsf(a.x) + fabsf(e1.x) + fabsf(e2.x), fabsf(a.y) + fabsf(e1.y) + fabsf(e2.y),
                   fabsf(a.z) + fabsf(e1.z) + fabsf(e2.z)};
    const float extent = fmaxf(fmaxf(fabsf(e1.x), fabsf(e1.y)),
                               fmaxf(fabsf(e1.z), fmaxf(fabsf(e2.x), fmaxf(fabsf(e2.y), fabsf(e2.z)))));
    const float d = 16 * 0x1p-23f *
                    (fabsf(h.normal.x) * error.x + fabsf(h.normal.y) * error.y + fabsf(h.normal.z) * error.z + extent);
    h.origin = p + h.normal * d;
    if (h.normal.x != 0) h.origin.x = nextafterf(h.origin.x, h.normal.x > 0 ? MaxFloat : -MaxFloat);
    if (h.normal.y != 0) h.origin.y = nextafterf(h.origin.y, h.normal.y > 0 ? MaxFloat : -MaxFloat);
    if (h.normal.z != 0) h.origin.z = nextafterf(h.origin.z, h.normal.z > 0 ? MaxFloat : -MaxFloat);
    h.below = p + h.normal * -d;
    if (h.normal.x != 0) h.below.x = nextafterf(h.below.x, h.normal.x > 0 ? -MaxFloat : MaxFloat);
    if (h.normal.y != 0) h.below.y = nextafterf(h.below.y, h.normal.y > 0 ? -MaxFloat : MaxFloat);
    if (h.normal.z != 0) h.below.z = nextafterf(h.below.z, h.normal.z > 0 ? -MaxFloat : MaxFloat);
    return finite3(h.origin) && finite3(h.below);
}

// ---------------------------------------------------------------------------------------------------------------
// Lights (light_core.h): rectangle and constant environment.
constexpr float InfiniteDistance = 3.4028234663852886e+38f - 1e32f;
constexpr float EnvDistance = InfiniteDistance * 0.9f;
constexpr float HitEpsilonEnd = 1.0f - 1e-3f;
struct LightHit {
    V3 N;
    float distance;
};
__device__ __forceinline__ V3 rows(const float* m, V3 v) {
    return ((ld3(m) * v.x) + (ld3(m + 3) * v.y)) + (ld3(m + 6) * v.z);
}
__device__ __forceinline__ V3 polar(float cosTheta, float sinTheta, float phi) {
    float s, c;
    sincosPoly(phi, s, c);
    return {sinTheta * c, sinTheta * s, cosTheta};
}
__device__ __forceinline__ float planeDistance(const x2_light& l, V3 p) {
    float d = dot(ld3(l.direction), p) + l.plane_offset;
    if (l.sidedness == X2_SIDE_REVERSE) d = -d;
    else if (l.sidedness == X2_SIDE_TWO) d = fabsf(d);
    return d;
}
__device__ __forceinline__ float areaToSolidAngleScale(V3 wi, V3 n, float dist) {
    const float denom = -dot(wi, n);
    return denom < 0.00001f ? 0.0f : (dist * dist) / denom;
}
__device__ bool lightSample(const x2_light& l, V3 p, bool cull, V3 n, const float r[3], V3& wi, LightHit& h) {
    if (l.type == X2_LIGHT_ENV) {
        if (l.hemispherical) {
            const float cosTheta = r[0] > 1.0f ? 1.0f : r[0];
            wi = rows(l.local_to_render, polar(cosTheta, sqrtf(1.0f - cosTheta * cosTheta), r[1] * TwoPi));
        } else {
            const float cosTheta = 1.0f - 2.0f * r[0];
            wi = polar(cosTheta, sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta)), r[1] * TwoPi);
        }
        if (cull && dot(n, wi) < Epsilon) return false;
        h.N = wi * -1.0f;
        h.distance = EnvDistance;
        return true;
    }
    if (l.type != X2_LIGHT_RECT) return false;
    const float distance = planeDistance(l, p);
    if (l.sidedness != X2_SIDE_TWO && distance < Epsilon) return false;
    const V3 local{(0.5f - r[0]) * (2 * l.half_width), (0.5f - r[1]) * (2 * l.half_height), 0.0f};
    wi = (rows(l.local_to_render, local) + ld3(l.position)) - p;
    h.distance = sqrtf(dot(wi, wi));
    if (h.distance < Epsilon) return false;
    wi = wi * (1.0f / h.distance);
    if (cull && dot(n, wi) < Epsilon) return false;
    const V3 dir = ld3(l.direction);
    h.N = dot(dir, wi) < 0.0f ? dir : dir * -1.0f;
    return true;
}
__device__ __forceinline__ V3 lightEval(const x2_light& l, V3 wi, const LightHit& h, float& pdf) {
    const V3 radiance = ld3(l.radiance) * V3{1, 1, 1};
    pdf = l.inv_area;
    if (l.type != X2_LIGHT_ENV) pdf *= areaToSolidAngleScale(wi, h.N, h.distance);
    return radiance;
}
__device__ bool lightIntersect(const x2_light& l, V3 p, V3 wi, float maxDistance, LightHit& h) {
    if (l.type == X2_LIGHT_ENV) {
        if (l.hemispherical && dot(wi, ld3(l.direction)) < 0.0f) return false;
        if (EnvDistance > maxDistance) return false;
        h.N = wi * -1.0f;
        h.distance = EnvDistance;
        return true;
    }
    if (l.type != X2_LIGHT_RECT) return false;
    const V3 localWi = rows(l.render_to_local_rot, wi), localP = rows(l.render_to_local, p - ld3(l.position));
    if ((localP.z < 0.0f) == (localWi.z < 0.0f)) return false;
    if (localP.z < 0.0f ? l.sidedness == X2_SIDE_REGULAR : l.sidedness == X2_SIDE_REVERSE) return false;
    const float localDistance = -localP.z / localWi.z, renderDistance = l.scale * localDistance;
    if (renderDistance > maxDistance) return false;
    const V3 hit = localP + localWi * localDistance;
    if (fabsf(hit.x) >= l.half_width || fabsf(hit.y) >= l.half_height) return false;
    const V3 dir = ld3(l.direction);
    h.N = localWi.z < 0.0f ? dir : dir * -1.0f;
    h.distance = renderDistance;
    return true;
}
// corners: the light's four corners, rows(local_to_render, (+-half_width, +-half_height, 0)) + position, made on the
// host with the same float operations (lightCorners, beside the lights).
__device__ bool canIlluminate(const x2_light& l, const float* corners, V3 p, bool cull, V3 n) {
    const float threshold = Epsilon - 0.0f;
    if (l.type == X2_LIGHT_ENV) return true;
    if (l.sidedness != X2_SIDE_TWO && planeDistance(l, p) < threshold) return false;
    if (!cull) return true;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (dot(n, ld3(corners + 3 * i) - p) > threshold) return true;
    return false;
}
__device__ __forceinline__ float powerHeuristic(float f, float g) { return (f * f) / (f * f + g * g); }
// Light i of a loop over the lights, at which every active lane of the wave is: one index for the wave, and the
// lights (never written by a kernel) read as constant memory, so that its fields are scalar loads. As a plain global
// read every field was a vector load, each waited on in turn.
__device__ __forceinline__ const x2_light& waveLight(const x2_light* lights, uint32_t i) {
    const auto c = (const __attribute__((address_space(4))) x2_light*)lights;
    return *(const x2_light*)(c + __builtin_amdgcn_readfirstlane(i));
}

// ---------------------------------------------------------------------------------------------------------------
// Lobes: Lambert and GGX (shade_lobes.h:481-556; shade_lobes_eval.h:11, 93-146, 609-668; shade_math.h:135-262).
constexpr uint32_t Reflection = 1u, Transmission = 2u, Diffuse = 4u, Glossy = 16u, Mirror = 64u;
constexpr uint32_t AllSurfaceSides = Reflection | Transmission, AllLobes = Diffuse | Glossy | Mirror;
constexpr uint32_t AllLobeTypes = AllSurfaceSides | AllLobes;
__device__ __forceinline__ bool matchesFlags(uint32_t type, uint32_t flags) {
    return (type & AllSurfaceSides & flags) && (type & AllLobes & flags);
}
__device__ __forceinline__ uint32_t lobeRayMask(uint32_t t) {
    uint32_t mask = 0;
    if (t & 1u) mask |= t & 0x54u;
    if (t & 2u) mask |= (t & 0x54u) << 1;
    return mask;
}
struct Frame {
    V3 x, y, z;
};
__device__ __forceinline__ Frame frameOf(V3 N) {
    const float sign = copysignf(1.0f, N.z), a = -1.0f / (sign + N.z), b = N.x * N.y * a;
    return {{1.0f + sign * N.x * N.x * a, sign * b, -sign * N.x}, {b, sign + N.y * N.y * a, -N.y}, N};
}
__device__ __forceinline__ V3 localToGlobal(const Frame& f, V3 v) { return (f.x * v.x + f.y * v.y) + f.z * v.z; }
__device__ __forceinline__ V3 localSpherical(float cosTheta, float sinTheta, float phi) {
    float s, c;
    sincosPoly(phi, s, c);
    return {sinTheta * c, sinTheta * s, cosTheta};
}
// A hit's Bsdf: lobe 0 Lambert (albedo * diffuse weight), lobe 1 GGX (scale = specular weight), either omitted at
// weight 0. Both share the shading frame.
struct Bsdf {
    uint32_t count, type;
    uint32_t kind[2], ltype[2];
    V3 albedo, scale;  // Lambert's albedo; GGX's scale
    float roughness;   // GGX: max(0.001, r^2)
    Frame frame;
    V3 emission;
};
__device__ __forceinline__ V3 evalLambert(const Bsdf& b, V3 wi, float& pdf) {
    const V3 N = b.frame.z;
    const float cosWi = fmaxf(dot(N, wi), 0.0f);
    pdf = cosWi * OneOverPi;
    return (((1.0f * b.albedo) * V3{1, 1, 1}) * OneOverPi) * cosWi;
}
__device__ __forceinline__ float widenRoughness(float roughness, float cosNO) {
    return (1.2f - 0.2f * sqrtf(cosNO)) * roughness;
}
__device__ V3 evalGGX(const Bsdf& b, V3 wo, V3 wi, float& pdf) {
    const V3 N = b.frame.z;
    const float cosNO = dot(N, wo), cosNI = dot(N, wi), roughness = b.roughness;
    pdf = 0.0f;
    if (cosNO <= 0.0f || cosNI <= 0.0f) return {0, 0, 0};
    V3 m = wo + wi;
    {
        const float l = length(m);
        if (!(l > 0.001f)) return {0, 0, 0};
        m = m * (1.0f / l);
    }
    const float cosMI = dot(m, wi);
    if (cosMI <= 0.0f) return {0, 0, 0};
    const float cosThetaM = dot(m, N);
    if (cosThetaM <= 0.0f) return {0, 0, 0};
    float alpha2 = roughness * roughness;
    const float cosThetaM2 = cosThetaM * cosThetaM, piCosThetaM4 = Pi * cosThetaM2 * cosThetaM2;
    const float tanThetaM2 = (1.0f - cosThetaM2) / cosThetaM2;
    float tmp = alpha2 + tanThetaM2, D4 = 0.25f * alpha2 / (piCosThetaM4 * tmp * tmp);
    const float invCosNO = 1.0f / cosNO, invCosNI = 1.0f / cosNI;
    const float G1o = 2.0f / (1.0f + sqrtf(1.0f + alpha2 * (1.0f - cosNO * cosNO) * (invCosNO * invCosNO)));
    const float G1i = 2.0f / (1.0f + sqrtf(1.0f + alpha2 * (1.0f - cosNI * cosNI) * (invCosNI * invCosNI)));
    const V3 fr = b.scale * G1o * G1i * D4 * invCosNO * 1.0f;
    const float w2 = 0.0f;
    const float alpha = widenRoughness(roughness, cosNO);
    alpha2 = alpha * alpha;
    tmp = alpha2 + tanThetaM2;
    D4 = 0.25f * alpha2 / (piCosThetaM4 * tmp * tmp);
    const float p2 = 0.0f, p1 = D4 * cosThetaM / cosMI;
    pdf = (1.0f - w2) * p1 + w2 * p2;
    return 1.0f * fr;
}
__device__ __forceinline__ V3 lobeEval(const Bsdf& b, uint32_t q, V3 wo, V3 wi, float& pdf) {
    return b.kind[q] == 0 ? evalLambert(b, wi, pdf) : evalGGX(b, wo, wi, pdf);
}
__device__ V3 lobeSample(const Bsdf& b, uint32_t q, V3 wo, float r1, float r2, V3& wi, float& pdf) {
    if (b.kind[q] == 0) {
        const float cosTheta = r1 > 1.0f ? 1.0f : sqrtf(r1), sinTheta = sqrtf(1.0f - r1);
        wi = localToGlobal(b.frame, localSpherical(cosTheta, sinTheta, r2 * TwoPi));
        return evalLambert(b, wi, pdf);
    }
    const float cosNO = dot(b.frame.z, wo);
    if (cosNO <= 0.0f) {
        wi = {0, 0, 0};
        pdf = 0.0f;
        return {0, 0, 0};
    }
    const float w2 = 0.0f, u1 = (r1 - w2) / (1.0f - w2);
    const float alpha = widenRoughness(b.roughness, cosNO), alpha2 = alpha * alpha;
    const float tanThetaM2 = alpha2 * u1 / (1.0f - u1), cosThetaM = 1.0f / sqrtf(1.0f + tanThetaM2);
    const float sinThetaM = cosThetaM * sqrtf(tanThetaM2);
    const V3 m = localToGlobal(b.frame, localSpherical(cosThetaM, sinThetaM, 2.0f * Pi * r2));
    const float NdotWo = dot(m, wo);
    wi = -1.0f * wo + 2.0f * NdotWo * m;
    return evalGGX(b, wo, wi, pdf);
}

// ---------------------------------------------------------------------------------------------------------------
// The estimator (async_integrator.h:1986-3042, the subset above). Pending rays live in the lane's private memory.
// ---------------------------------------------------------------------------------------------------------------
// Textures: mip pyramids read by the path tracer's map lookups (texture_core.h:86-206, which adapt OpenImageIO
// 3.1.17's texturesys.cpp, Apache-2.0, copyright Contributors to the OpenImageIO project): the anisotropic filter
// on camera paths, trilinear closest texels on glossy ones, one level's closest texel on diffuse ones
// (shade_mapnodes.h:344-347), as MoonRay's ImageMap binds a map (shade_mapnodes.h:420-470: unit scale, no offset or
// bias, periodic, the t flip, the footprint's cross terms swapped as XPU's vector shading passes them,
// async_gpu_device.h:71). theta's sine and cosine come from the shared polynomial and hypot(A - C, B) is the square
// root of the FP64 sum of squares, on both devices. A texture's levels are FP32 rgb end to end (resident), or, paged,
// 32 x 32 tiles of half floats in the tile pool for the finer levels and an always-resident array of half floats for
// the coarse ones.
struct Footprint {
    float u, v, dudx, dvdx, dudy, dvdy;
};
__device__ __forceinline__ uint32_t wrapTexel(int x, uint32_t n) {
    const int m = x % int(n);
    return uint32_t(m < 0 ? m + int(n) : m);
}
// OpenImageIO's fast_atan2 and fast_exp (fmath.h), as texture_core.h:94-97 gives them.
__device__ __forceinline__ float oiioAtan2(float y, float x) {
    const float a = fabsf(x), b = fabsf(y);
    const float k = b == 0 ? 0.f : (a == b ? 1.f : (b > a ? a / b : b / a));
    const float s = 1.f - (1.f - k), t = s * s;
    float r = s * (0.430165678f * t + 1.f) / ((0.0579354987f * t + 0.763007998f) * t + 1.f);
    if (b > a) r = 1.5707963267948966f - r;
    if (copysignf(1.f, x) < 0) r = Pi - r;
    return copysignf(r, y);
}
__device__ __forceinline__ float oiioExp(float a) {
    float x = fminf(fmaxf(a * 1.4426950408889634f, -126.f), 126.f);
    const int m = int(x);
    x -= float(m);
    x = 1.f - (1.f - x);
    float r = 1.33336498402e-3f;
    r = x * r + 9.810352697968e-3f;
    r = x * r + 5.551834031939e-2f;
    r = x * r + 0.2401793301105f;
    r = x * r + 0.693144857883f;
    r = x * r + 1.f;
    return ldexpf(r, m);
}
// The filter's plan (texture_core.h:161-206 up to the reads): the levels and their weights, and for the
// anisotropic filter the taps along the major axis (n 0: one closest texel a level).
struct TexPlan {
    float s, t;
    int levels[2];
    float lw[2];
    int n;
    float sm, tm;
    float pos[7], weight[7];
};
__device__ bool planLookup(const TexDesc& tex, uint32_t mode, const Footprint& f, bool swap, TexPlan& p) {
    const uint32_t count = tex.count;
    if (!count || count > MipLevels) return false;
    float s = f.u, t = 1 - f.v;
    float dx = f.dudx, tx = -f.dvdx, dy = f.dudy, ty = -f.dvdy;
    if (swap) {
        const float c = tx;
        tx = dy;
        dy = c;
    }
    if (!finite1(s) || !finite1(t) || !finite1(dx) || !finite1(tx) || !finite1(dy) || !finite1(ty)) return false;
    if (fabsf(s) > 1024) s = s - floorf(s);
    if (fabsf(t) > 1024) t = t - floorf(t);
    p.s = s;
    p.t = t;
    p.levels[0] = p.levels[1] = -1;
    p.lw[0] = p.lw[1] = 0;
    p.n = 0;
    p.sm = p.tm = 0;
    float minor, major = 0, aspect = 1;
    if (mode != 0) {
        const float eps = 1.0e-8f, eps2 = eps * eps;
        const float dxlen2 = dx * dx + tx * tx, dylen2 = dy * dy + ty * ty;
        if (dxlen2 < eps2) {
            if (dylen2 < eps2) {
                dx = eps;
                dy = 0;
                tx = 0;
                ty = eps;
            } else {
                const float sc = eps / sqrtf(dylen2);
                dx = ty * sc;
                tx = -dy * sc;
            }
        } else if (dylen2 < eps2) {
            const float sc = eps / sqrtf(dxlen2);
            dy = -tx * sc;
            ty = dx * sc;
        }
        minor = fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fmaxf(fabsf(tx), fabsf(ty)));  // the filter width
    } else {
        const double A = double(tx * tx + ty * ty), B = -2.0 * double(dx * tx + dy * ty), C = double(dx * dx + dy * dy);
        const double root = sqrt((A - C) * (A - C) + B * B);
        major = fminf(sqrtf(fmaxf(float((A + C + root) * .5), 0.f)), 1000.f);
        minor = fminf(sqrtf(fmaxf(float((A + C - root) * .5), 0.f)), 1000.f);
        const float theta = oiioAtan2(float(B), float(A - C)) * .5f + Pi * .5f;
        aspect = minor > 0 ? fminf(fmaxf(major / minor, 1.f), 1.e6f) : (major > 0 ? 1.e6f : 1.f);
        if (aspect > 4) {
            aspect = 4;
            major = .5f * (major + minor * 4);
            minor = major / 4;
        }
        float sn, cs;
        sincosPoly(theta, sn, cs);
        p.sm = cs;  // scaled by the axis length below
        p.tm = sn;
    }
    for (uint32_t m = 0; m < count; ++m) {
        const TexLevel& l = tex.levels[m];
        const float w = minor * float(l.width < l.height ? l.width : l.height);
        if (w <= 1) {
            p.levels[0] = int(m) - 1;
            p.levels[1] = int(m);
            p.lw[1] = fminf(fmaxf(2 * w - 1, 0.f), 1.f);
            p.lw[0] = 1 - p.lw[1];
            break;
        }
    }
    if (p.levels[1] < 0) {
        p.levels[0] = p.levels[1] = int(count) - 1;
        p.lw[0] = 1;
        p.lw[1] = 0;
    } else if (p.levels[0] < 0) {
        p.levels[0] = p.levels[1] = 0;
        p.lw[0] = 1;
        p.lw[1] = 0;
        if (mode == 0) {
            const TexLevel& l = tex.levels[0];
            const uint32_t r = l.width > l.height ? l.width : l.height;
            if (minor * float(r) < .5f) aspect = fminf(fmaxf(major * float(r) * 2, 1.f), 4.f);
        }
    } else if (mode == 2) {
        p.levels[0] = p.levels[1];
        p.lw[0] = 1;
        p.lw[1] = 0;
    }
    if (mode != 0) return true;
    int n = int(2 * aspect - 1);
    if (n < 1) n = 1;
    if (n > 7) return false;
    const float LL = 2 * (major - minor);
    p.sm = p.sm * LL * .5f;
    p.tm = p.tm * LL * .5f;
    p.n = n;
    if (n == 1) {
        p.weight[0] = 1;
        p.pos[0] = 0;
    } else if (n == 2) {
        p.weight[0] = p.weight[1] = .5f;
        p.pos[0] = -.5f;
        p.pos[1] = .5f;
    } else {
        float sum = 0;
        const float scale = major / LL;
        for (int i = 0; i < n; ++i) {
            p.pos[i] = 2 * (float(i) + .5f) * (1.f / float(n)) - 1;
            const float x = p.pos[i] * scale;
            p.weight[i] = oiioExp(-2 * x * x);
            sum += p.weight[i];
        }
        for (int i = 0; i < n; ++i) p.weight[i] /= sum;
    }
    return true;
}
enum : int { TapOk = 0, TapInvalid = 1, TapMiss = 2, TapCoarse = 3 };
// A resident texture's texel (level k at levels[k].offset floats).
struct ResidentTexels {
    const TexDesc& t;
    __device__ __forceinline__ int fetch(uint32_t level, uint32_t x, uint32_t y, V3& out) const {
        const TexLevel& l = t.levels[level];
        const float* p = t.texels + l.offset + (size_t(y) * l.width + x) * 3;
        out = V3{p[0], p[1], p[2]};
        return TapOk;
    }
};
// A paged texture's texel: the coarse levels from the always-resident halves; a finer level's through the page table
// (half-float rgb in its tile's pool slot). A missing tile is asked for once (its flag set, its id on the launch's
// miss list; with the list full the flag is cleared, for a later lookup to ask) and the fetch answers TapMiss, or,
// with partial shading, the texel's parent on the nearest level up that is here, TapCoarse (async_gpu_device.h:49-58:
// the sample is then provisional). Every fetch of a lookup runs, so a sample asks for all its tiles at once.
struct PagedTexels {
    const RenderParams& P;
    const TexDesc& t;
    uint32_t* firstMiss;  // the first tile found missing, + 1 (null: not kept)
    __device__ __forceinline__ bool here(uint32_t level, uint32_t x, uint32_t y, V3& out) const {
        const TexLevel& l = t.levels[level];
        if (level >= t.firstResident) {
            const uint16_t* h = t.coarse + l.offset + (size_t(y) * l.width + x) * 3;
            out = V3{__half2float(__ushort_as_half(h[0])), __half2float(__ushort_as_half(h[1])),
                     __half2float(__ushort_as_half(h[2]))};
            return true;
        }
        const uint32_t tile = l.tileBase + (y >> 5) * l.tilesX + (x >> 5);
        const uint32_t slot = P.tileSlot[tile];
        if (slot == NoTileSlot) return false;
        if (P.tileStamp[tile] != P.stamp) P.tileStamp[tile] = P.stamp;  // read this launch: recently used
        const uint16_t* h = P.tilePool + (size_t(slot) * TexTileTexels + (y & (TexTile - 1)) * TexTile + (x & (TexTile - 1))) * 3;
        out = V3{__half2float(__ushort_as_half(h[0])), __half2float(__ushort_as_half(h[1])),
                 __half2float(__ushort_as_half(h[2]))};
        return true;
    }
    __device__ int fetch(uint32_t level, uint32_t x, uint32_t y, V3& out) const {
        if (here(level, x, y, out)) return TapOk;
        const TexLevel& l = t.levels[level];
        const uint32_t tile = l.tileBase + (y >> 5) * l.tilesX + (x >> 5);
        if (firstMiss && !*firstMiss) *firstMiss = tile + 1;
        if (atomicExch(P.tileAsked + tile, 1u) == 0u) {
            const uint32_t at = atomicAdd(&P.paging->missCount, 1u);
            if (at < P.missCap) P.missList[at] = tile;
            else atomicExch(P.tileAsked + tile, 0u);
        }
        if (!P.partial) return TapMiss;
        for (uint32_t k = level + 1; k < t.count; ++k) {
            x >>= 1;
            y >>= 1;
            const TexLevel& u = t.levels[k];
            if (here(k, x < u.width ? x : u.width - 1, y < u.height ? y : u.height - 1, out)) return TapCoarse;
        }
        return TapMiss;
    }
};
// Folds a fetch's status into the lookup's: an invalid value wins, then a miss, then a provisional read.
__device__ __forceinline__ int worse(int a, int b) {
    if (a == TapInvalid || b == TapInvalid) return TapInvalid;
    if (a == TapMiss || b == TapMiss) return TapMiss;
    return a == TapCoarse || b == TapCoarse ? TapCoarse : TapOk;
}
// texture_core.h:144-153 at one level: the four texels about (u, v), each fetch wrapped.
template <class Texels>
__device__ int bilinearAt(const Texels& T, const TexLevel& l, uint32_t level, float u, float v, V3& out) {
    float s = u * float(l.width) - .5f, t = v * float(l.height) - .5f;
    if (!finite1(s) || !finite1(t) || fabsf(s) > 1.e8f || fabsf(t) > 1.e8f) return TapInvalid;
    const int x = int(floorf(s)), y = int(floorf(t));
    s -= float(x);
    t -= float(y);
    const uint32_t x0 = wrapTexel(x, l.width), x1 = wrapTexel(x + 1, l.width);
    const uint32_t y0 = wrapTexel(y, l.height), y1 = wrapTexel(y + 1, l.height);
    V3 a, b, c, d;
    int st = T.fetch(level, x0, y0, a);
    st = worse(st, T.fetch(level, x1, y0, b));
    st = worse(st, T.fetch(level, x0, y1, c));
    st = worse(st, T.fetch(level, x1, y1, d));
    if (st == TapInvalid || st == TapMiss) return st;
    out = ((a * (1 - s)) + (b * s)) * (1 - t) + ((c * (1 - s)) + (d * s)) * t;
    return finite3(out) ? st : TapInvalid;
}
// texture_core.h:155-160: the texel nearest (u, v); an exact half-texel tie takes the lower one.
__device__ __forceinline__ bool closestAt(const TexLevel& l, float u, float v, uint32_t& x, uint32_t& y) {
    const float s = u * float(l.width) - .5f, t = v * float(l.height) - .5f;
    if (!finite1(s) || !finite1(t) || fabsf(s) > 1.e8f || fabsf(t) > 1.e8f) return false;
    int ix = int(floorf(s)), iy = int(floorf(t));
    if (s - float(ix) > .5f) ++ix;
    if (t - float(iy) > .5f) ++iy;
    x = wrapTexel(ix, l.width);
    y = wrapTexel(iy, l.height);
    return true;
}
// The map's value under the plan (texture_core.h:171-177, :199-206).
template <class Texels>
__device__ int mapRead(const Texels& T, const TexDesc& tex, const TexPlan& p, V3& out) {
    V3 result{0, 0, 0};
    int status = TapOk;
    for (int k = 0; k < 2; ++k) {
        if (p.lw[k] == 0) continue;
        const uint32_t level = uint32_t(p.levels[k]);
        const TexLevel& l = tex.levels[level];
        if (p.n == 0) {
            uint32_t x, y;
            if (!closestAt(l, p.s, p.t, x, y)) return TapInvalid;
            V3 value;
            const int st = T.fetch(level, x, y, value);
            if (st == TapMiss) {
                status = TapMiss;
                continue;
            }
            if (!finite3(value)) return TapInvalid;
            status = worse(status, st);
            result = result + value * p.lw[k];
            continue;
        }
        V3 acc{0, 0, 0};
        for (int i = 0; i < p.n; ++i) {
            V3 value;
            const int st = bilinearAt(T, l, level, p.s + p.pos[i] * p.sm, p.t + p.pos[i] * p.tm, value);
            if (st == TapInvalid) return TapInvalid;
            status = worse(status, st);
            if (st == TapMiss) continue;
            acc = acc + value * p.weight[i];
        }
        result = result + acc * p.lw[k];
    }
    if (status == TapMiss) return TapMiss;
    if (!finite3(result)) return TapInvalid;
    out = result;
    return status;
}
// A lookup of a texture, resident or paged.
__device__ int mapLookup(const RenderParams& P, const TexDesc& tex, uint32_t mode, const Footprint& f, bool swap, V3& out,
                         uint32_t* firstMiss = nullptr) {
    TexPlan p;
    if (!planLookup(tex, mode, f, swap, p)) return TapInvalid;
    if (tex.paged) return mapRead(PagedTexels{P, tex, firstMiss}, tex, p, out);
    return mapRead(ResidentTexels{tex}, tex, p, out);
}

// ---------------------------------------------------------------------------------------------------------------
// Ray differentials (shade_state.h:30-33; shade_material.h:388-558, :691-705; shade_lobes_eval.h:725-741;
// shade_math.h:185-207, :279-288), for Lambert and GGX lobes on flat triangles (dN/dx = dN/dy = 0). A pending ray
// carries its differentials in scenes with textures; the camera's carries its auxiliary directions instead.
// Invalid differentials are marked by a NaN in dOx.x (48 bytes a record, three aligned 16-byte words).
struct alignas(16) RayDiff {
    V3 dOx, dOy, dDx, dDy;
};
__device__ __forceinline__ bool diffValid(const RayDiff& d) { return d.dOx.x == d.dOx.x; }
__device__ __forceinline__ float notANumber() { return __int_as_float(0x7fc00000); }
struct HitDiff {
    float dsdx, dtdx, dsdy, dtdy;
    V3 dPdx, dPdy;
    bool valid;
    V3 inDx, inDy;  // the incoming direction differentials (the GGX lobe reflects them)
};
__device__ __forceinline__ float differenceOfProducts(float a, float b, float c, float d) {  // Kahan's
    const float cd = c * d;
    const float err = fmaf(-c, d, cd);
    const float dop = fmaf(a, b, -cd);
    return dop + err;
}
__device__ __forceinline__ V3 normalized(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ bool trianglePartials(V3 x0, V3 x1, V3 x2, const float st[6], V3& dPds, V3& dPdt) {
    const float A[4] = {st[2] - st[0], st[3] - st[1], st[4] - st[0], st[5] - st[1]};
    const V3 b0 = x1 - x0, b1 = x2 - x0;
    const float det = differenceOfProducts(A[0], A[3], A[1], A[2]);
    if (!(fabsf(det) >= 1.17549435e-38f) || !(fabsf(det) <= 3.402823466e38f)) return false;
    const float inv = 1.0f / det;
    dPds = V3{differenceOfProducts(b0.x, A[3], b1.x, A[1]), differenceOfProducts(b0.y, A[3], b1.y, A[1]),
              differenceOfProducts(b0.z, A[3], b1.z, A[1])} * inv;
    dPdt = V3{differenceOfProducts(b1.x, A[0], b0.x, A[2]), differenceOfProducts(b1.y, A[0], b0.y, A[2]),
              differenceOfProducts(b1.z, A[0], b0.z, A[2])} * inv;
    return finite3(dPds) && finite3(dPdt);
}
// shade_material.h:392-432: a camera ray's auxiliary rays meet the triangle's plane; their barycentric offsets map
// through the texture coordinates (the pinhole's offsets are zero).
__device__ void hitTexDerivatives(V3 a, V3 b, V3 c, const float st[6], const Ray& ray, V3 camDx, V3 camDy, uint32_t spp,
                                  HitDiff& h) {
    h.dsdx = h.dsdy = h.dtdx = h.dtdy = 0.0f;
    if (isExactlyZero(camDx) || isExactlyZero(camDy) || !spp) return;
    const V3 e1 = b - a, e2 = c - a, n = cross(e1, e2);
    uint32_t axis = fabsf(n.x) > fabsf(n.y) ? 0 : 1;
    if (fabsf(n.z) > (axis == 0 ? fabsf(n.x) : fabsf(n.y))) axis = 2;
    auto pa = [&](V3 v) { return axis == 0 ? v.y : axis == 1 ? v.z : v.x; };
    auto pb = [&](V3 v) { return axis == 0 ? v.z : axis == 1 ? v.x : v.y; };
    const float det = pa(e1) * pb(e2) - pb(e1) * pa(e2), den = dot(n, ray.direction);
    if (!(fabsf(det) > 0) || !(fabsf(den) > 0)) return;
    const float plane = dot(n, a - ray.origin);
    const V3 primary = ray.direction * (plane / den);
    const float scale = 1.f / fminf(sqrtf(float(spp)), 2.f);
    float ds[2], dt[2];
    for (int i = 0; i < 2; ++i) {
        V3 d = i ? camDy : camDx;
        if (scale < 1) d = normalized(ray.direction + (d - ray.direction) * scale);
        const float dd = dot(n, d);
        if (!(fabsf(dd) > 0)) return;
        const V3 o = V3{0, 0, 0} * scale;
        const V3 delta = o + d * (dot(n, a - ray.origin - o) / dd) - primary;
        const float x = (pa(delta) * pb(e2) - pb(delta) * pa(e2)) / det, y = (pa(e1) * pb(delta) - pb(e1) * pa(delta)) / det;
        ds[i] = (st[2] - st[0]) * x + (st[4] - st[0]) * y;
        dt[i] = (st[3] - st[1]) * x + (st[5] - st[1]) * y;
    }
    if (!(finite1(ds[0]) && finite1(ds[1]) && finite1(dt[0]) && finite1(dt[1]))) return;
    h.dsdx = ds[0];
    h.dtdx = dt[0];
    h.dsdy = ds[1];
    h.dtdy = dt[1];
}
// shade_material.h:436-451.
__device__ __forceinline__ void transferDifferentials(const RayDiff& d, V3 origin, V3 dir, V3 Ng, V3 Pt, HitDiff& h) {
    h.dPdx = h.dPdy = V3{0, 0, 0};
    h.valid = false;
    if (!diffValid(d)) return;
    const V3 oX = origin + d.dOx, oY = origin + d.dOy, dX = dir + d.dDx, dY = dir + d.dDy;
    const float denomX = dot(Ng, dX), denomY = dot(Ng, dY);
    if (denomX == 0.0f || denomY == 0.0f) return;
    const float planeD = -dot(Ng, Pt);
    const float distX = -(dot(oX, Ng) + planeD) / denomX, distY = -(dot(oY, Ng) + planeD) / denomY;
    h.dPdx = (oX + dX * distX) - Pt;
    h.dPdy = (oY + dY * distY) - Pt;
    h.valid = finite3(h.dPdx) && finite3(h.dPdy);
    if (!h.valid) h.dPdx = h.dPdy = V3{0, 0, 0};
}
// shade_material.h:454-495; without differentials, the coarsest level (1024).
__device__ void differentialTexDerivatives(V3 dPds, V3 dPdt, V3 Ng, HitDiff& h) {
    h.dsdx = h.dtdx = h.dsdy = h.dtdy = 0.0f;
    if (!h.valid) {
        h.dsdx = 1024.0f;
        h.dtdy = 1024.0f;
        return;
    }
    const float e = 1e-5f;
    if (fabsf(dot(dPds, dPdt)) <= e) {
        const float du = dot(dPds, dPds), dv = dot(dPdt, dPdt);
        if (du > e && dv > e) {
            const float iu = 1.0f / dot(dPds, dPds), iv = 1.0f / dot(dPdt, dPdt);
            h.dsdx = dot(h.dPdx, dPds) * iu;
            h.dtdx = dot(h.dPdx, dPdt) * iv;
            h.dsdy = dot(h.dPdy, dPds) * iu;
            h.dtdy = dot(h.dPdy, dPdt) * iv;
        }
        return;
    }
    const V3 n2 = Ng * Ng;
    float A[4], bx[2], by[2];
    if (n2.x > n2.y && n2.x > n2.z) {
        A[0] = dPds.y; A[1] = dPdt.y; A[2] = dPds.z; A[3] = dPdt.z;
        bx[0] = h.dPdx.y; bx[1] = h.dPdx.z; by[0] = h.dPdy.y; by[1] = h.dPdy.z;
    } else if (n2.y > n2.z) {
        A[0] = dPds.x; A[1] = dPdt.x; A[2] = dPds.z; A[3] = dPdt.z;
        bx[0] = h.dPdx.x; bx[1] = h.dPdx.z; by[0] = h.dPdy.x; by[1] = h.dPdy.z;
    } else {
        A[0] = dPds.x; A[1] = dPdt.x; A[2] = dPds.y; A[3] = dPdt.y;
        bx[0] = h.dPdx.x; bx[1] = h.dPdx.y; by[0] = h.dPdy.x; by[1] = h.dPdy.y;
    }
    const float det = A[0] * A[3] - A[1] * A[2];
    if (fabsf(det) <= e) return;
    const float detA = differenceOfProducts(A[0], A[3], A[1], A[2]);
    if (!(fabsf(detA) >= 1.17549435e-38f) || !(fabsf(detA) <= 3.402823466e38f)) return;
    const float inv = 1.0f / detA;
    h.dsdx = differenceOfProducts(bx[0], A[3], bx[1], A[1]) * inv;
    h.dtdx = differenceOfProducts(bx[1], A[0], bx[0], A[2]) * inv;
    h.dsdy = differenceOfProducts(by[0], A[3], by[1], A[1]) * inv;
    h.dtdy = differenceOfProducts(by[1], A[0], by[0], A[2]) * inv;
}
// shade_material.h:538-558 for a mesh without subdivision derivatives: a camera ray's texture derivatives by
// hitTexDerivatives and its auxiliary rays, scaled for the samples per pixel, as the incoming differentials; any
// other ray's transferred and turned into texture derivatives. in: the pending ray's record (the camera's holds its
// auxiliary directions in dDx and dDy).
// mapped: the hit's material reads a map (else its texture derivatives are not needed, only dP/dx and dP/dy for the
// rays it scatters).
__device__ void hitDifferentials(V3 a, V3 b, V3 c, const float st[6], V3 N, V3 Ng, V3 Pt, const Ray& ray, bool camera,
                                 const RayDiff& in, uint32_t spp, bool mapped, HitDiff& h) {
    if (camera && !isExactlyZero(in.dDx) && !isExactlyZero(in.dDy) && spp) {
        if (mapped) hitTexDerivatives(a, b, c, st, ray, in.dDx, in.dDy, spp, h);
        const float scale = 1.f / fminf(sqrtf(float(spp)), 2.f);
        RayDiff d;
        d.dOx = V3{0, 0, 0} * scale;
        d.dOy = V3{0, 0, 0} * scale;
        d.dDx = normalized(ray.direction + (in.dDx - ray.direction) * scale) - ray.direction;
        d.dDy = normalized(ray.direction + (in.dDy - ray.direction) * scale) - ray.direction;
        h.inDx = d.dDx;
        h.inDy = d.dDy;
        transferDifferentials(d, ray.origin, ray.direction, Ng, Pt, h);
        return;
    }
    h.inDx = in.dDx;
    h.inDy = in.dDy;
    transferDifferentials(in, ray.origin, ray.direction, Ng, Pt, h);
    if (!mapped) return;
    V3 dPds, dPdt;
    if (!trianglePartials(a, b, c, st, dPds, dPdt)) {
        const Frame f = frameOf(N);
        dPds = f.x;
        dPdt = f.y;
    }
    differentialTexDerivatives(dPds, dPdt, Ng, h);
}
// A scattered ray's differentials: Lambert's from the cosine sample's (clamped 85 degrees from the normal, scaled by
// the primary scale as the lobe ignores the incoming ones), GGX's the incoming ones reflected about the half vector
// and widened by 1 + 7 alpha (shade_lobes.h:479, :548). Invalid where the hit's are.
__device__ RayDiff scatterDifferentials(bool lambert, const Frame& frame, float alpha, const HitDiff& h, V3 Pt, V3 wo,
                                        V3 wi, float r1, float r2, V3 childOrigin, float primaryScale) {
    RayDiff d;
    d.dOx = d.dOy = d.dDx = d.dDy = V3{0, 0, 0};
    d.dOx.x = notANumber();
    if (!h.valid) return d;
    V3 dDdx = h.inDx, dDdy = h.inDy;
    if (lambert) {
        const float c85 = 0.08715580403804779f;  // cosf(85 * (pi / 180)) in binary32
        const float minR1 = c85 * c85, maxR1 = 1.0f - minR1;
        r1 = fminf(fmaxf(r1, minR1), maxR1);
        const float cosTheta = r1 > 1.0f ? 1.0f : sqrtf(r1), sinTheta = sqrtf(1.0f - r1);
        float sinPhi, cosPhi;
        sincosPoly(r2 * TwoPi, sinPhi, cosPhi);
        const float scale = 1.0f / 8.0f, halfScale = 0.5f * scale, twoPiScale = TwoPi * scale;
        dDdx = {-halfScale * cosPhi / sinTheta, -halfScale * sinPhi / sinTheta, halfScale / cosTheta};
        dDdy = {-twoPiScale * sinPhi * sinTheta, twoPiScale * cosPhi * sinTheta, 0.0f};
        const float l1 = length(dDdx), l2 = length(dDdy), l = sqrtf(l1 * l2);
        dDdx = dDdx * (l / l1);
        dDdy = dDdy * (l / l2);
        dDdx = localToGlobal(frame, dDdx);
        dDdy = localToGlobal(frame, dDdy);
    } else {
        V3 H = wo + wi;
        const float hl = length(H);
        if (hl > 0.001f) H = H * (1.0f / hl);
        else H = frame.z;
        const V3 dN{0, 0, 0};
        const float Cx = dot(dDdx, H) - dot(wo, dN), Cy = dot(dDdy, H) - dot(wo, dN), dDotN = -dot(wo, H);
        dDdx = dDdx - (2.0f * (dDotN * dN + Cx * H));
        dDdy = dDdy - (2.0f * (dDotN * dN + Cy * H));
        const float factor = 1.0f + alpha * 7.0f;
        dDdx = dDdx * factor;
        dDdy = dDdy * factor;
    }
    const float scale = lambert ? primaryScale : 1.0f;
    const V3 shift = Pt - childOrigin;
    d.dOx = h.dPdx * scale + shift;
    d.dOy = h.dPdy * scale + shift;
    d.dDx = normalized(wi + dDdx * scale) - wi;
    d.dDy = normalized(wi + dDdy * scale) - wi;
    if (!(finite3(d.dOx) && finite3(d.dOy) && finite3(d.dDx) && finite3(d.dDy))) d.dOx.x = notANumber();
    return d;
}
// The camera ray's auxiliary directions through the next pixel in x and in y (camera_core.h:222-224, :245-246).
__device__ bool cameraDifferentials(const x2_config& c, float x, float y, V3& dx, V3& dy) {
    const float* m = c.raster_to_camera;
    auto dirAt = [&](V3 p) {
        const V3 v = vectorM(m, p) + V3{m[12], m[13], m[14]};
        const float w = ((p.x * m[3] + p.y * m[7]) + p.z * m[11]) + m[15];
        return w == 0.0f ? V3{0, 0, 0} : V3{v.x / w, v.y / w, v.z / w};
    };
    const V3 dirX = dirAt(V3{x + 1, y, -1.0f}), dirY = dirAt(V3{x, y + 1, -1.0f});
    return normalizeTo(vectorM(c.camera_to_render, dirX), dx) && normalizeTo(vectorM(c.camera_to_render, dirY), dy);
}

// d ln f / d r of the GGX lobe (shade_lobes_eval.h:93-121) at (wo, wi) for MoonRay's input roughness r (the lobe's
// alpha = r^2, alpha^2 = A = r^4; 0 where alpha is clamped to 0.001): D4 ~ A / (A + tan^2 theta_m)^2 and Smith's
// G1 = 2 / (1 + sqrt(1 + A tan^2 theta)) for wo and wi; scale and 1 / cos(theta_o) do not depend on r.
__device__ float ggxRoughnessLogDerivative(float r, V3 N, V3 wo, V3 wi) {
    const float alpha = r * r;
    if (!(alpha > 0.001f)) return 0.0f;
    const float A = alpha * alpha;
    const float co = dot(N, wo), ci = dot(N, wi);
    if (co <= 0.0f || ci <= 0.0f) return 0.0f;
    V3 m = wo + wi;
    const float l = length(m);
    if (!(l > 0.001f)) return 0.0f;
    m = m * (1.0f / l);
    const float cm = dot(m, N);
    if (cm <= 0.0f) return 0.0f;
    const float tm = (1.0f - cm * cm) / (cm * cm), to = (1.0f - co * co) / (co * co), ti = (1.0f - ci * ci) / (ci * ci);
    const float so = sqrtf(1.0f + A * to), si = sqrtf(1.0f + A * ti);
    const float dA = 1.0f / A - 2.0f / (A + tm) - to / (2.0f * so * (1.0f + so)) - ti / (2.0f * si * (1.0f + si));
    return dA * (4.0f * r * r * r);
}

// The adjoint's bookkeeping. A gradient sink is what a contribution's derivative goes to: the texels of a textured
// Lambert lobe (its lookup: texture, mode, footprint, and the value read) or a material's roughness (with the GGX
// log-derivative at the lobe's direction).
enum : uint32_t { SinkNone = 0, SinkTexture = 1, SinkRoughness = 2 };
struct Sink {
    uint32_t kind, material;
    float g;
    uint32_t tex, mode;
    Footprint f;
    V3 color;
};
// An open subtree: the pending ray popped at stack position pos, the running radiance when it was popped, and the
// sink of the lobe that made it. Its subtree's radiance is what the running radiance gains until the next pop at a
// position below pos (the stack is last in, first out).
struct Record {
    uint32_t pos;
    V3 mark;
    Sink sink;
};
constexpr uint32_t MaxRecords = 34;

// The adjoint workgroup's gradients of the small levels (TexLevel::lds), rgb a texel.
__device__ __forceinline__ float* adjointLds() {
    __shared__ float g[AdjLdsTexels * 3];
    return g;
}
__device__ __forceinline__ float* levelGradient(const RenderParams& P, const TexDesc& d, uint32_t level) {
    return level == 0 ? P.gradTextures + d.gradOffset : P.gradPyramid + (d.gradLevels + (long long)d.levels[level].offset);
}
__device__ void adjointLdsClear() {
    float* g = adjointLds();
    for (uint32_t i = threadIdx.x; i < AdjLdsTexels * 3; i += blockDim.x) g[i] = 0.0f;
    __syncthreads();
}
// At the workgroup's end: its small levels' gradients into the global ones, one atomic a nonzero value.
__device__ void adjointLdsFlush(const RenderParams& P) {
    __syncthreads();
    const float* g = adjointLds();
    for (uint32_t t = 0; t < P.textureCount; ++t) {
        const TexDesc& d = P.textures[t];
        for (uint32_t k = 0; k < d.count; ++k) {
            const TexLevel& l = d.levels[k];
            if (l.lds == NoLds) continue;
            float* out = levelGradient(P, d, k);
            const uint32_t n = l.width * l.height * 3;
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const float v = g[l.lds * 3 + i];
                if (v != 0.0f) atomicAdd(out + i, v);
            }
        }
    }
}
// A contribution R of a lobe whose albedo the map scales has d R / d value = R / value per channel, and the value is
// linear in the texels its lookup reads: each gets its weight's share, level 0's into the caller's gradient, the
// levels above into the pyramid gradient (pushed down to level 0 after the frame), the small levels by way of the
// workgroup's LDS.
__device__ void scatterTexture(const RenderParams& P, const Sink& s, V3 R, V3 delta) {
    const TexDesc& d = P.textures[s.tex];
    const float cs[3] = {s.color.x > 0.0f ? delta.x * R.x / s.color.x : 0.0f,
                         s.color.y > 0.0f ? delta.y * R.y / s.color.y : 0.0f,
                         s.color.z > 0.0f ? delta.z * R.z / s.color.z : 0.0f};
    if (cs[0] == 0.0f && cs[1] == 0.0f && cs[2] == 0.0f) return;
    TexPlan p;
    if (!planLookup(d, s.mode, s.f, true, p)) return;
    auto add = [&](uint32_t level, uint32_t x, uint32_t y, float w) {
        const TexLevel& l = d.levels[level];
        if (l.lds != NoLds) {
            float* q = adjointLds() + (l.lds + y * l.width + x) * 3;
            atomicAdd(q + 0, cs[0] * w);
            atomicAdd(q + 1, cs[1] * w);
            atomicAdd(q + 2, cs[2] * w);
            return;
        }
        float* q = levelGradient(P, d, level) + (size_t(y) * l.width + x) * 3;
        atomicAdd(q + 0, cs[0] * w);
        atomicAdd(q + 1, cs[1] * w);
        atomicAdd(q + 2, cs[2] * w);
    };
    for (int k = 0; k < 2; ++k) {
        if (p.lw[k] == 0) continue;
        const uint32_t level = uint32_t(p.levels[k]);
        const TexLevel& l = d.levels[level];
        if (p.n == 0) {
            uint32_t x, y;
            if (closestAt(l, p.s, p.t, x, y)) add(level, x, y, p.lw[k]);
            continue;
        }
        for (int i = 0; i < p.n; ++i) {
            float u = (p.s + p.pos[i] * p.sm) * float(l.width) - .5f, v = (p.t + p.pos[i] * p.tm) * float(l.height) - .5f;
            if (!finite1(u) || !finite1(v) || fabsf(u) > 1.e8f || fabsf(v) > 1.e8f) continue;
            const int x = int(floorf(u)), y = int(floorf(v));
            u -= float(x);
            v -= float(y);
            const uint32_t x0 = wrapTexel(x, l.width), x1 = wrapTexel(x + 1, l.width);
            const uint32_t y0 = wrapTexel(y, l.height), y1 = wrapTexel(y + 1, l.height);
            const float w = p.weight[i] * p.lw[k];
            add(level, x0, y0, (1 - u) * (1 - v) * w);
            add(level, x1, y0, u * (1 - v) * w);
            add(level, x0, y1, (1 - u) * v * w);
            add(level, x1, y1, u * v * w);
        }
    }
}
__device__ __forceinline__ void addRoughness(const RenderParams& P, uint32_t m, float v, float* gr) {
    if (m < RoughnessSlots) gr[m] += v;
    else atomicAdd(P.gradRoughness + m, v);
}
__device__ __forceinline__ void scatterSink(const RenderParams& P, const Sink& s, V3 R, V3 delta, float* gr) {
    if (s.kind == SinkTexture) scatterTexture(P, s, R, delta);
    else if (s.kind == SinkRoughness) addRoughness(P, s.material, (delta.x * R.x + delta.y * R.y + delta.z * R.z) * s.g, gr);
}

struct Pending {
    V3 origin, dir, T;
    uint64_t branch;
    uint32_t depths;  // depth | nonMirrorDepth << 8 | diffuseDepth << 16 | glossyDepth << 24
    uint32_t heldLight;
    float heldFar;
    V3 held;
};
constexpr uint32_t MaxPending = 32, PendingReserve = 16;
static_assert(sizeof(Pending) == PendingBytes && MaxPending == MaxPendingRecords, "the slot arrays' pending records");
__device__ __forceinline__ float bounceThinning(uint32_t top, uint32_t planned) {
    const uint32_t room = top < PendingReserve ? PendingReserve - top : 1u;
    return planned > room ? float(room) / float(planned) : 1.0f;
}
__device__ __forceinline__ bool roomForChild(uint32_t top, float u, float& weight) {
    if (top >= MaxPending) return false;
    const float chance = float(MaxPending - top) / float(PendingReserve + 1);
    if (u >= chance) return false;
    weight = 1.0f / chance;
    return true;
}
__device__ __forceinline__ bool roulette(float lum, float threshold, float u, float& scale) {
    scale = 1;
    if (lum >= threshold) return true;
    if (lum < u * threshold) return false;
    scale = threshold / lum;
    return true;
}
__device__ __forceinline__ V3 smartClamp(V3 a, float clamp) {
    if (clamp <= 0) return a;
    const float m = fmaxf(a.x, fmaxf(a.y, a.z));
    return m > clamp ? a * (clamp / m) : a;
}
__device__ __forceinline__ uint32_t surfaceFlags(V3 Ng, V3 wo, V3 wi) {
    return dot(Ng, wo) * dot(Ng, wi) > 0.0f ? (AllLobeTypes & ~Transmission) : (AllLobeTypes & ~Reflection);
}

// The hit record of a closest-hit answer (tri: the triangle record, ~0u: none). False: an invalid value.
__device__ __forceinline__ bool hitRecord(const RenderParams& P, const Ray& ray, uint32_t tri, Hit& h) {
    h.geom = ~0u;
    if (tri == ~0u) return true;
    const TriRecord& r = P.tris[tri];
    h.geom = r.geom;
    h.tri = tri;
    return geometryHit(h, ray, V3{r.ax, r.ay, r.az}, V3{r.bx, r.by, r.bz}, V3{r.cx, r.cy, r.cz});
}
__device__ bool traceHit(const RenderParams& P, const Ray& ray, Hit& h, Counters& k, TravStack& st) {
    ++k.traces;
    float t;
    const uint32_t tri = closestHit(P, ray.origin, ray.direction, ray.nearT, ray.farT, t, k, st);
    return hitRecord(P, ray, tri, h);
}
__device__ __forceinline__ bool visible(const RenderParams& P, V3 o, V3 d, float nearT, float farT, Counters& k,
                                        TravStack& st) {
    ++k.shadows;
    return !anyHit(P, o, d, nearT, farT, k, st);
}

// ---------------------------------------------------------------------------------------------------------------
// The participating medium (the structure of computeRadianceVolume, async_integrator.h:1152-1291, over this
// device's own medium: a box-bounded brick grid of FP16 density and temperature, trilinear, grey extinction, a
// blackbody ramp for emission; shadow transmittance by ratio tracking under per-brick majorants in place of the
// reference's transmittance march, :428-476). The first march along a ray keeps no step table (a march may take
// hundreds of steps): it leaves checkpoints (MarchRecord, :622-648, 8 slots with stride doubling) and each lookup
// the light sampling needs marches again from the nearest one (ReplaySteps, :823-888), bit for bit the same steps.
constexpr float VolEpsilon = 1.1920928955078125e-07f;
constexpr float HitEpsilonStart = 1e-5f;
constexpr float OneOverFourPi = 0.0795774715459476678844f;
constexpr uint32_t MarchSteps = 1u << 16;  // lattice steps a march takes at most
__device__ __forceinline__ float halfValue(uint16_t h) { return float(__builtin_bit_cast(_Float16, h)); }
__device__ __forceinline__ float lerpf(float a, float b, float w) { return a + (b - a) * w; }
__device__ __forceinline__ float voxelAt(const VolumeDesc& v, const uint16_t* pool, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || uint32_t(i) >= v.res[0] || uint32_t(j) >= v.res[1] || uint32_t(k) >= v.res[2]) return 0.0f;
    const uint32_t e = v.table[(uint32_t(k >> 3) * v.nb[1] + uint32_t(j >> 3)) * v.nb[0] + uint32_t(i >> 3)];
    if (e == ~0u) return 0.0f;
    return halfValue(pool[size_t(e) * 512 + (uint32_t(k & 7) * 8 + uint32_t(j & 7)) * 8 + uint32_t(i & 7)]);
}
__device__ float sampleGrid(const VolumeDesc& v, const uint16_t* pool, V3 p) {
    float q[3];
    for (int a = 0; a < 3; ++a) {
        q[a] = (comp(p, a) - v.bmin[a]) * v.scaleVox[a] - 0.5f;
        if (!(q[a] > -1.0f && q[a] < float(v.res[a]))) return 0.0f;
    }
    const float fx0 = floorf(q[0]), fy0 = floorf(q[1]), fz0 = floorf(q[2]);
    const int i = int(fx0), j = int(fy0), k = int(fz0);
    const float fx = q[0] - fx0, fy = q[1] - fy0, fz = q[2] - fz0;
    float t[8];  // the 8 voxels, x fastest
    if (i >= 0 && j >= 0 && k >= 0 && (i & 7) != 7 && (j & 7) != 7 && (k & 7) != 7) {
        // all 8 in one brick (i + 1 < res as i + 1 is inside the brick): one table lookup
        const uint32_t e = v.table[(uint32_t(k >> 3) * v.nb[1] + uint32_t(j >> 3)) * v.nb[0] + uint32_t(i >> 3)];
        if (e == ~0u) return 0.0f;
        const uint16_t* b = pool + size_t(e) * 512 + (uint32_t(k & 7) * 8 + uint32_t(j & 7)) * 8 + uint32_t(i & 7);
        t[0] = halfValue(b[0]), t[1] = halfValue(b[1]), t[2] = halfValue(b[8]), t[3] = halfValue(b[9]);
        t[4] = halfValue(b[64]), t[5] = halfValue(b[65]), t[6] = halfValue(b[72]), t[7] = halfValue(b[73]);
    } else {
        t[0] = voxelAt(v, pool, i, j, k), t[1] = voxelAt(v, pool, i + 1, j, k);
        t[2] = voxelAt(v, pool, i, j + 1, k), t[3] = voxelAt(v, pool, i + 1, j + 1, k);
        t[4] = voxelAt(v, pool, i, j, k + 1), t[5] = voxelAt(v, pool, i + 1, j, k + 1);
        t[6] = voxelAt(v, pool, i, j + 1, k + 1), t[7] = voxelAt(v, pool, i + 1, j + 1, k + 1);
    }
    const float c00 = lerpf(t[0], t[1], fx), c10 = lerpf(t[2], t[3], fx);
    const float c01 = lerpf(t[4], t[5], fx), c11 = lerpf(t[6], t[7], fx);
    return lerpf(lerpf(c00, c10, fy), lerpf(c01, c11, fy), fz);
}
__device__ __forceinline__ uint32_t brickAt(const VolumeDesc& v, V3 p) {
    int b[3];
    for (int a = 0; a < 3; ++a) {
        const float q = floorf((comp(p, a) - v.bmin[a]) * v.scaleVox[a] * 0.125f);
        b[a] = int(fminf(fmaxf(q, 0.0f), float(v.nb[a] - 1)));
    }
    return (uint32_t(b[2]) * v.nb[1] + uint32_t(b[1])) * v.nb[0] + uint32_t(b[0]);
}
// The medium's tap sets, counted per lane in LDS (the roofline's voxel bytes: 8 FP16 voxels a set).
__device__ __forceinline__ uint32_t& tapCount() {
    __shared__ uint32_t taps[256];
    return taps[threadIdx.x];
}
__device__ __forceinline__ float densityAt(const VolumeDesc& v, V3 p) {
    if (!(v.majorant[brickAt(v, p)] > 0.0f)) return 0.0f;
    ++tapCount();
    return sampleGrid(v, v.density, p);
}
__device__ V3 rampAt(const VolumeDesc& v, float kelvin) {
    float x = kelvin / v.rampMax * float(v.rampSize - 1);
    x = fminf(fmaxf(x, 0.0f), float(v.rampSize - 1));
    const float fi = floorf(x);
    const uint32_t i = uint32_t(fi);
    if (i + 1 >= v.rampSize) return ld3(v.ramp + size_t(v.rampSize - 1) * 3);
    const float f = x - fi;
    const float* a = v.ramp + size_t(i) * 3;
    return {lerpf(a[0], a[3], f), lerpf(a[1], a[4], f), lerpf(a[2], a[5], f)};
}
struct VCoeff {
    float sigmaT;
    V3 sigmaS, emission;
};
__device__ VCoeff coefficientsAt(const VolumeDesc& v, V3 p, bool withEmission) {
    VCoeff k{0.0f, {0, 0, 0}, {0, 0, 0}};
    const float d = densityAt(v, p);
    if (d == 0.0f) return k;
    k.sigmaT = d * v.sigmaTScale;
    k.sigmaS = ld3(v.albedo) * k.sigmaT;
    if (withEmission && v.temperature && v.emissionScale > 0.0f) {
        ++tapCount();
        k.emission = rampAt(v, sampleGrid(v, v.temperature, p)) * (v.emissionScale * d);
    }
    return k;
}
__device__ bool clipBox(const VolumeDesc& v, V3 o, V3 d, float tnear, float tfar, float& a, float& b) {
    a = tnear;
    b = tfar;
    for (int ax = 0; ax < 3; ++ax) {
        const float oa = comp(o, ax), da = comp(d, ax), lo = v.bmin[ax], hi = v.bmax[ax];
        if (da == 0.0f) {
            if (oa < lo || oa > hi) return false;
            continue;
        }
        float t0 = (lo - oa) / da, t1 = (hi - oa) / da;
        if (t0 > t1) {
            const float s = t0;
            t0 = t1;
            t1 = s;
        }
        a = fmaxf(a, t0);
        b = fminf(b, t1);
    }
    return a < b;
}
__device__ __forceinline__ bool insideBox(const VolumeDesc& v, V3 p) {
    return p.x >= v.bmin[0] && p.x <= v.bmax[0] && p.y >= v.bmin[1] && p.y <= v.bmax[1] && p.z >= v.bmin[2] &&
           p.z <= v.bmax[2];
}
// Ratio tracking brick by brick (a 3D DDA over the brick grid): free flights at the brick's majorant extinction,
// each collision keeping 1 - density / majorant; nothing below the transmittance threshold.
__device__ float transmittance(const VolumeDesc& v, V3 o, V3 d, float tnear, float tfar, float scale, float threshold,
                               Stream& s) {
    float a, b;
    if (!clipBox(v, o, d, tnear, tfar, a, b)) return 1.0f;
    const V3 entry = o + d * a;
    int cell[3], step[3];
    float tNext[3];
    for (int ax = 0; ax < 3; ++ax) {
        const float q = floorf((comp(entry, ax) - v.bmin[ax]) * v.scaleVox[ax] * 0.125f);
        cell[ax] = int(fminf(fmaxf(q, 0.0f), float(v.nb[ax] - 1)));
        const float da = comp(d, ax);
        step[ax] = da > 0.0f ? 1 : -1;
        tNext[ax] = da == 0.0f ? FLT_MAX
                               : (v.bmin[ax] + float(cell[ax] + (da > 0.0f ? 1 : 0)) * v.brickExtent[ax] - comp(o, ax)) / da;
    }
    float T = 1.0f, t = a;
    while (t < b) {
        const int ax = tNext[0] <= tNext[1] ? (tNext[0] <= tNext[2] ? 0 : 2) : (tNext[1] <= tNext[2] ? 1 : 2);
        const float tEnd = fminf(tNext[ax], b);
        const float m = v.majorant[(uint32_t(cell[2]) * v.nb[1] + uint32_t(cell[1])) * v.nb[0] + uint32_t(cell[0])];
        if (m > 0.0f && tEnd > t) {
            const float mu = m * v.sigmaTScale * scale;
            float tt = t;
            for (;;) {
                tt = tt - logf(1.0f - s.next()) / mu;
                if (!(tt < tEnd)) break;
                T *= 1.0f - fminf(1.0f, densityAt(v, o + d * tt) / m);
                if (T < threshold) return 0.0f;
            }
        }
        t = fmaxf(t, tEnd);
        cell[ax] += step[ax];
        if (cell[ax] < 0 || uint32_t(cell[ax]) >= v.nb[ax]) break;
        const float da = comp(d, ax);
        tNext[ax] = (v.bmin[ax] + float(cell[ax] + (da > 0.0f ? 1 : 0)) * v.brickExtent[ax] - comp(o, ax)) / da;
    }
    return T;
}
// Henyey-Greenstein (volume_core.h:550-564), distance and equi-angular sampling (:566-573).
__device__ float phaseEval(float g, V3 wo, V3 wi) {
    if (fabsf(g) <= 1e-3f) return OneOverFourPi;
    const float cosTheta = dot(wo, wi), denom = 1 + g * g + 2 * g * cosTheta;
    return fabsf(denom) < VolEpsilon ? 0.0f : OneOverFourPi * (1 - g * g) / (denom * sqrtf(denom));
}
__device__ V3 phaseSample(float g, V3 wo, float u1, float u2) {
    if (fabsf(g) <= 1e-3f) {
        const float cosTheta = 1.0f - 2.0f * u1, sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
        return localSpherical(cosTheta, sinTheta, u2 * TwoPi);
    }
    u1 = fmaxf(u1, VolEpsilon);
    const float sqrTerm = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
    const float cosTheta = (1.0f + g * g - sqrTerm * sqrTerm) / (2.0f * g);
    const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
    return localToGlobal(frameOf(-wo), localSpherical(cosTheta, sinTheta, TwoPi * u2));
}
__device__ __forceinline__ float sampleDistanceExponential(float u, float sigma, float a, float b) {
    return a - logf(1.0f - u * (1.0f - expf(sigma * (a - b)))) / sigma;
}
__device__ __forceinline__ float pdfDistanceExponential(float t, float sigma, float a, float b) {
    return sigma / (expf((t - a) * sigma) - expf((t - b) * sigma));
}
__device__ __forceinline__ float pdfEquiAngular(float t, float D, float thetaA, float thetaB) {
    return D / ((thetaB - thetaA) * (D * D + t * t));
}
__device__ __forceinline__ bool finitef(float x) { return x - x == 0.0f; }

// A march: step ends on the lattice a + stepSize * (u + k), the last step ending at b, each step's coefficients at
// its end; a step of zero density is not a step (it carries nothing).
struct VStep {
    float sigmaT;
    V3 sigmaS;
    float tr, tStart, delta;
};
struct MarchCk {  // the march as a recorded step begins: its lattice index, transmittance, weights so far, count
    uint32_t k;
    float vt, partial;
    uint32_t count;
};
struct MarchRecord {
    static constexpr uint32_t Slots = 16;  // 4, 8 and 32 measured slower
    MarchCk at[Slots];
    uint32_t used = 0, stride = 8, last = 0;
    __device__ bool wants(uint32_t count) const {
        if (count == 0 || count == last || count % stride) return false;
        return used < Slots || count % (2 * stride) == 0;
    }
    __device__ MarchCk& slot(uint32_t count) {
        if (used == Slots) {
            for (uint32_t i = 0; i < Slots / 2; ++i) at[i] = at[2 * i + 1];
            used = Slots / 2;
            stride *= 2;
        }
        last = count;
        return at[used++];
    }
};
struct March {
    V3 o, d;
    float a, b, u, stepSize;
    // the first march's totals
    uint32_t count;
    float vt, total;
    bool reached, weightless;
    V3 emitted;
    float firstStart, lastEnd;
    MarchRecord rec;
    __device__ float endOf(uint32_t k) const { return a + stepSize * (u + float(k)); }
    __device__ float startOf(uint32_t k) const { return k == 0 ? a : a + stepSize * (u + float(k - 1)); }
};
// Runs the march from step k0 with the state (vt, count) there; emit(step, index) returns false to stop. first: the
// first march (records checkpoints, adds emission, sets the totals).
template <bool First, class Emit>
__device__ void runMarch(const VolumeDesc& v, const x2_config& c, March& m, uint32_t k0, float vt, float partial,
                         uint32_t count, bool sampleEmission, Emit&& emit) {
    for (uint32_t k = k0; k < MarchSteps; ++k) {
        const float e = m.endOf(k);
        const bool last = !(e < m.b);
        if (last && !(e - m.b > VolEpsilon)) break;
        if (vt < c.volume_transmittance_threshold) {
            if constexpr (First) m.reached = true;
            break;
        }
        const float at = last ? m.b : e, start = m.startOf(k);
        const VCoeff q = coefficientsAt(v, m.o + m.d * at, First && sampleEmission);
        if (q.sigmaT != 0.0f) {
            const float delta = at - start;
            const VStep s{q.sigmaT, q.sigmaS, vt, start, delta};
            const float w = (luminance(q.sigmaS) * delta) * vt;
            if constexpr (First) {
                if (m.rec.wants(count)) m.rec.slot(count) = MarchCk{k, vt, partial, count};
                if (count == 0) m.firstStart = start;
                m.lastEnd = start + delta;
                if (sampleEmission) m.emitted = m.emitted + q.emission * (vt * delta);
            }
            if (!emit(s, count, w)) return;
            partial += w;
            ++count;
            vt *= expf(-q.sigmaT * delta);
        }
        if (last) break;
    }
    if constexpr (First) {
        m.count = count;
        m.vt = vt;
        m.total = partial;
    }
}
__device__ void firstMarch(const VolumeDesc& v, const x2_config& c, March& m, bool sampleEmission) {
    m.count = 0;
    m.reached = false;
    m.emitted = {0, 0, 0};
    m.firstStart = m.lastEnd = 0.0f;
    runMarch<true>(v, c, m, 0, 1.0f, 0.0f, 0, sampleEmission, [](const VStep&, uint32_t, float) { return true; });
    m.weightless = m.total == 0.0f;
    if (m.weightless) m.total = 1.0f;
}
// The cdf as the oracle's table gives it: weights summed in order, each partial no more than the total, the last the
// total; even where none weighs.
__device__ __forceinline__ float cdfAt(const March& m, float& partial, float w, uint32_t index) {
    if (index + 1 >= m.count) return m.total;
    if (m.weightless) return float(index + 1) / float(m.count);
    partial += w;
    if (partial > m.total) partial = m.total;
    return partial;
}
__device__ __forceinline__ float cdfBefore(const March& m, const MarchCk& ck) {
    return m.weightless ? float(ck.count) / float(m.count) : ck.partial;
}
// sampleDiscrete by marching again from the last checkpoint short of r * total.
__device__ VStep pickStep(const VolumeDesc& v, const x2_config& c, March& m, float r, float& pdf, float& remapped) {
    const float rw = r * m.total;
    int from = -1;
    for (int j = int(m.rec.used) - 1; j >= 0; --j)
        if (m.rec.at[j].count < m.count && cdfBefore(m, m.rec.at[j]) < rw) {
            from = j;
            break;
        }
    const MarchCk ck = from >= 0 ? m.rec.at[from] : MarchCk{0, 1.0f, 0.0f, 0};
    float partial = ck.partial, prev = from >= 0 ? cdfBefore(m, ck) : 0.0f, found = m.total, before = 0.0f;
    VStep chosen{0.0f, {0, 0, 0}, 0.0f, 0.0f, 0.0f};
    runMarch<false>(v, c, m, ck.k, ck.vt, ck.partial, ck.count, false, [&](const VStep& s, uint32_t i, float w) {
        const float here = cdfAt(m, partial, w, i);
        chosen = s;
        if (here >= rw || i + 1 >= m.count) {
            found = here;
            before = prev;
            return false;
        }
        prev = here;
        return true;
    });
    const float value = found - before;
    pdf = value / m.total;
    remapped = (rw - before) / value;
    return chosen;
}
// findInterval by marching again: the last step starting at or before t, its pdfDiscrete, whether it holds t.
__device__ VStep stepAt(const VolumeDesc& v, const x2_config& c, March& m, float t, float& pdf, bool& holds) {
    int from = -1;
    for (int j = int(m.rec.used) - 1; j >= 0; --j)
        if (m.rec.at[j].count < m.count && m.startOf(m.rec.at[j].k) <= t) {
            from = j;
            break;
        }
    const MarchCk ck = from >= 0 ? m.rec.at[from] : MarchCk{0, 1.0f, 0.0f, 0};
    float partial = ck.partial, here = from >= 0 ? cdfBefore(m, ck) : 0.0f, kept = 0.0f, keptPrev = 0.0f;
    VStep chosen{0.0f, {0, 0, 0}, 0.0f, 0.0f, 0.0f};
    runMarch<false>(v, c, m, ck.k, ck.vt, ck.partial, ck.count, false, [&](const VStep& s, uint32_t i, float w) {
        if (i > 0 && s.tStart > t) return false;
        const float before = here;
        here = cdfAt(m, partial, w, i);
        chosen = s;
        kept = here;
        keptPrev = i > 0 ? before : 0.0f;
        return i + 1 < m.count;
    });
    pdf = (kept - keptPrev) / m.total;
    holds = chosen.tStart <= t && t <= chosen.tStart + chosen.delta;
    return chosen;
}
// The glowing medium as a light source: its emission region (volume_emission.h:13-171, MoonRay's EmissiveRegion over
// a DenseEmissionDistribution; Simon et al., EGSR 2017, for the direction pdf as a line integral of the 3D pdf).
struct Dist1D {  // one record of the region's pool: its values, cdf, integral and size
    const float* func;
    const float* cdf;
    float integral;
    uint32_t n;
};
__device__ __forceinline__ Dist1D distAt(const float* pool, uint32_t at) {
    const uint32_t n = uint32_t(pool[at + 1]);
    return Dist1D{pool + at + 2, pool + at + 2 + n, pool[at], n};
}
__device__ __forceinline__ uint32_t distFloats(uint32_t n) { return 2u * n + 3u; }
// geom::findInterval: the last cdf entry at or below u (entries are nondecreasing), kept within [0, size - 2].
__device__ __forceinline__ int cdfInterval(const float* cdf, int size, float u) {
    int lo = 0, count = size;  // entries [lo, lo + count) not yet placed; those before lo are <= u
    while (count > 0) {
        const int half = count >> 1;
        if (cdf[lo + half] <= u) {
            lo += half + 1;
            count -= half + 1;
        } else {
            count = half;
        }
    }
    return min(max(lo - 1, 0), size - 2);
}
__device__ __forceinline__ float distPdf(const Dist1D& d, int i) {
    return d.integral > 0.0f ? d.func[i] / (d.integral * float(d.n)) : 0.0f;
}
__device__ __forceinline__ int distSample(const Dist1D& d, float u, float& pdf, float& remapped) {
    const int i = cdfInterval(d.cdf, int(d.n) + 1, u);
    pdf = distPdf(d, i);
    remapped = (u - d.cdf[i]) / (d.cdf[i + 1] - d.cdf[i]);
    return i;
}
__device__ __forceinline__ V3 affinePoint(const float* m, V3 p) {
    return {p.x * m[0] + p.y * m[3] + p.z * m[6] + m[9], p.x * m[1] + p.y * m[4] + p.z * m[7] + m[10],
            p.x * m[2] + p.y * m[5] + p.z * m[8] + m[11]};
}
__device__ __forceinline__ V3 affineVector(const float* m, V3 v) {
    return {v.x * m[0] + v.y * m[3] + v.z * m[6], v.x * m[1] + v.y * m[4] + v.z * m[7], v.x * m[2] + v.y * m[5] + v.z * m[8]};
}
// The cell's discrete probability: its z plane's, its row's within the plane, its x within the row.
__device__ float cellPdf(const EmissionRegion& r, int x, int y, int z) {
    if (x < 0 || x >= r.res[0] || y < 0 || y >= r.res[1] || z < 0 || z >= r.res[2]) return 0.0f;
    const float pz = distPdf(distAt(r.pool, r.marginalZ), z);
    const float py = distPdf(distAt(r.pool, r.marginalY + uint32_t(z) * distFloats(uint32_t(r.res[1]))), y);
    const float px =
        distPdf(distAt(r.pool, r.conditionalX + uint32_t(z * r.res[1] + y) * distFloats(uint32_t(r.res[0]))), x);
    return px * py * pz;
}
// A place in the region drawn by its distribution (z, then y given z, then x given both; uniform within the cell).
__device__ V3 regionPoint(const EmissionRegion& r, float u1, float u2, float u3) {
    float pz, py, px, rz, ry, rx;
    const int z = distSample(distAt(r.pool, r.marginalZ), u1, pz, rz);
    const int y = distSample(distAt(r.pool, r.marginalY + uint32_t(z) * distFloats(uint32_t(r.res[1]))), u2, py, ry);
    const int x = distSample(
        distAt(r.pool, r.conditionalX + uint32_t(z * r.res[1] + y) * distFloats(uint32_t(r.res[0]))), u3, px, rx);
    return affinePoint(r.distToWorld, V3{float(x) + rx, float(y) + ry, float(z) + rz});
}
// The solid-angle pdf of direction wi from p: the 3D pdf integrated along the ray with the t^2 Jacobian, cell by
// cell (a DDA through the cells, each cell contributing pdf * (t1^3 - t0^3) / 3); tEnd: where the ray leaves the
// region's box (render-space distance); glow (when given): a distance past which the ray meets only cells of zero
// value up to the box's exit (a little past the last positive cell's exit), 0 when it meets none.
__device__ float regionPdf(const EmissionRegion& r, V3 pWorld, V3 wiWorld, float& tEnd, float* glow = nullptr) {
    const V3 p = affinePoint(r.worldToDist, pWorld);
    V3 w = affineVector(r.worldToDist, wiWorld);
    const float lengthScale = 1.0f / sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
    w = w * lengthScale;
    const float pa[3] = {p.x, p.y, p.z}, wa[3] = {w.x, w.y, w.z};
    float tStart = 0.0f;
    tEnd = 3.402823466e+38f;
    for (int ax = 0; ax < 3; ++ax) {
        const float inv = 1.0f / wa[ax];
        float t0 = (0.0f - pa[ax]) * inv, t1 = (float(r.res[ax]) - pa[ax]) * inv;
        if (t0 > t1) {
            const float sw = t0;
            t0 = t1;
            t1 = sw;
        }
        tStart = t0 > tStart ? t0 : tStart;
        tEnd = t1 < tEnd ? t1 : tEnd;
        if (tStart > tEnd) return 0.0f;
    }
    tEnd *= lengthScale;
    int cell[3], stepDir[3], stop[3];
    float tNext[3], tDelta[3];
    for (int ax = 0; ax < 3; ++ax) {
        const float q = pa[ax] + tStart * wa[ax];
        const int c0 = int(q);
        cell[ax] = c0 < 0 ? 0 : c0 > r.res[ax] - 1 ? r.res[ax] - 1 : c0;
        const bool up = wa[ax] >= 0;
        tNext[ax] = tStart + ((up ? float(cell[ax]) + 1.0f : float(cell[ax])) - q) / wa[ax];
        tDelta[ax] = (up ? 1.0f : -1.0f) / wa[ax];
        stepDir[ax] = up ? 1 : -1;
        stop[ax] = up ? r.res[ax] : -1;
    }
    float t = tStart, sum = 0.0f, lastLit = 0.0f;
    const int guard = (r.res[0] + r.res[1] + r.res[2]) * 2 + 8;
    for (int i = 0; i < guard; ++i) {
        const float pd = cellPdf(r, cell[0], cell[1], cell[2]);
        const int ax = (tNext[1] < tNext[0] || tNext[2] < tNext[0]) ? (tNext[2] < tNext[1] ? 2 : 1) : 0;
        if (pd > 0.0f) {
            const float t0 = t * lengthScale, t1 = tNext[ax] * lengthScale;
            sum += pd * r.invUnitVolume * (t1 * t1 * t1 - t0 * t0 * t0) / 3.0f;
            lastLit = t1;
        }
        t = tNext[ax];
        cell[ax] += stepDir[ax];
        if (cell[ax] == stop[ax]) break;
        tNext[ax] += tDelta[ax];
    }
    // past the last positive cell by 1e-4 of the distance: well inside the next cell (a cell spans 4 voxels), whose
    // zero value means no emission there
    if (glow) *glow = lastLit * 1.0001f + 1e-6f;
    return sum;
}
__device__ V3 regionSample(const EmissionRegion& r, V3 p, float u1, float u2, float u3, float& pdf, float& tEnd,
                           float* glow = nullptr) {
    V3 wi = regionPoint(r, u1, u2, u3) - p;
    wi = wi * (1.0f / sqrtf(wi.x * wi.x + wi.y * wi.y + wi.z * wi.z));
    pdf = regionPdf(r, p, wi, tEnd, glow);
    return wi;
}
// computeEmissiveVolumeIntegral (async_integrator.h:478-566) in the medium's box: the medium's light along the ray
// from tnear to tEnd or the first surface before it, through the medium's own extinction (Beer's law a step at a
// time), stepped as a march of the given depth (step ends on the lattice a + stepSize * (u + k), emission and
// extinction at each step's end) until the transmittance falls below the threshold. One draw from s when the ray
// has a stretch of the box to march. glow: a distance past which the ray meets no emission (regionPdf's); the march
// stops there, which changes nothing it adds (every step beyond adds zero emission) and saves its steps.
__device__ V3 emissiveIntegral(const RenderParams& P, Stream& s, V3 o, V3 d, float tnear, float tEnd, float glow,
                               uint32_t depth, Counters& k, TravStack& st) {
    const VolumeDesc& v = P.vol;
    const x2_config& c = P.cfg;
    {
        ++k.traces;
        float t;
        if (closestHit(P, o, d, tnear, tEnd, t, k, st) != ~0u && t < tEnd) tEnd = t;
    }
    float a, b;
    if (!clipBox(v, o, d, tnear, tEnd, a, b) || !(b - a > VolEpsilon)) return {0, 0, 0};
    const float u = s.next();
    const float stepSize = v.minFeature * c.inv_volume_quality * fmaxf(float(int(depth) - 1), 1.0f);
    V3 L{0, 0, 0};
    float tr = 1.0f;
    b = fminf(b, glow);
    for (uint32_t i = 0; i < MarchSteps; ++i) {
        const float e = a + stepSize * (u + float(i));
        const bool last = !(e < b);
        if (last && !(e - b > VolEpsilon)) break;
        if (tr < c.volume_transmittance_threshold) break;
        const float at = last ? b : e, start = i == 0 ? a : a + stepSize * (u + float(i - 1));
        const VCoeff q = coefficientsAt(v, o + d * at, true);
        if (q.sigmaT != 0.0f) {
            const float delta = at - start;
            L = L + (q.emission * tr) * delta;
            tr *= expf(-q.sigmaT * delta);
        }
        if (last) break;
    }
    return L;
}
// estimateInScatteringSourceTermDirect (:575-591).
__device__ V3 volumeSourceTerm(const RenderParams& P, Stream& s, const x2_light& L, V3 Pt, V3 rayDir, float g,
                               const float ul[3], float scale, Counters& k, TravStack& st) {
    V3 wi;
    LightHit lh;
    if (!lightSample(L, Pt, false, V3{0, 0, 0}, ul, wi, lh)) return {0, 0, 0};
    float pdfL;
    const V3 Li = lightEval(L, wi, lh, pdfL);
    if (pdfL == 0.0f || isExactlyZero(Li) || !finitef(pdfL)) return {0, 0, 0};
    const float near = HitEpsilonStart, far = fmaxf(HitEpsilonEnd * lh.distance, near);
    if (!visible(P, Pt, wi, near, far, k, st)) return {0, 0, 0};
    const float tr = transmittance(P.vol, Pt, wi, 0.0f, far, scale, P.cfg.volume_transmittance_threshold, s);
    const float ph = phaseEval(g, -rayDir, wi);
    return Li * (tr * (ph / pdfL));
}
// computeRadianceVolume for one ray: the medium's light into radiance, the march's transmittance into T; returns
// the alpha transmittance.
__device__ float volumeRadiance(const RenderParams& P, const Identity& root, V3& T, uint32_t depth,
                                uint32_t nonMirrorDepth, uint32_t& volumeDepth, uint64_t branch, V3 o, V3 d,
                                float tnear, float tfar, V3& radiance, Counters& k, TravStack& st) {
    const VolumeDesc& v = P.vol;
    const x2_config& c = P.cfg;
    March m;
    if (!clipBox(v, o, d, tnear, tfar, m.a, m.b)) return 1.0f;
    m.o = o;
    m.d = d;
    const uint64_t volumeBranch = childBranch(branch, 0xfffc, volumeDepth);
    Identity vid = root;
    vid.branch = childBranch(volumeBranch, 0, 0);
    {
        Stream marching(vid, depth);
        m.u = marching.next();
    }
    m.stepSize = v.minFeature * c.inv_volume_quality * fmaxf(float(int(depth) - 1), 1.0f);
    const bool sampleEmission = depth == 0;
    firstMarch(v, c, m, sampleEmission);
    if (m.count == 0) return 1.0f;
    const float vt = m.reached ? 0.0f : m.vt;
    const V3 T0 = T;
    if (sampleEmission && !isExactlyZero(m.emitted)) radiance = radiance + T0 * m.emitted;
    if (volumeDepth < c.max_volume_depth) {
        const float firstStart = m.firstStart, rayEnd = m.lastEnd;
        const bool highQuality = nonMirrorDepth == 0;
        const uint32_t samplesPerLight = highQuality ? c.light_samples : 1,
                       scatterCount = highQuality ? c.volume_illumination_samples : 1;
        if (scatterCount > 0) {
            const float invN = 1.0f / float(samplesPerLight * scatterCount);
            for (uint32_t li = 0; li < c.light_count; ++li) {
                const x2_light& L = waveLight(P.lights, li);
                const bool equiAngular = L.type == X2_LIGHT_RECT;
                V3 direct{0, 0, 0};
                for (uint32_t ls = 0; ls < samplesPerLight; ++ls)
                    for (uint32_t ts = 0; ts < scatterCount; ++ts) {
                        Identity sid = root;
                        sid.branch = childBranch(volumeBranch, 1 + li, ls * scatterCount + ts);
                        Stream ss(sid, depth);
                        float D = 0.0f, thetaA = 0.0f, thetaB = 0.0f, offset = 0.0f;
                        if (equiAngular) {
                            const float ue = ss.next();
                            float uel[3];
                            uel[0] = ss.next();
                            uel[1] = ss.next();
                            uel[2] = ss.next();
                            const V3 local{(0.5f - uel[0]) * 2.0f * L.half_width, (0.5f - uel[1]) * 2.0f * L.half_height, 0.0f};
                            const V3 pivot = ld3(L.position) + rows(L.local_to_render, local);
                            offset = dot(pivot - o, d);
                            const V3 foot = o + d * offset;
                            D = length(pivot - foot);
                            thetaA = atan2f(firstStart - offset, D);
                            thetaB = atan2f(rayEnd - offset, D);
                            const float te = D * tanf((1 - ue) * thetaA + ue * thetaB);
                            const float pdfTe = pdfEquiAngular(te, D, thetaA, thetaB);
                            if (pdfTe != 0.0f && finitef(pdfTe)) {
                                const float tePrime = te + offset;
                                float pdfK;
                                bool holds;
                                const VStep vp = stepAt(v, c, m, tePrime, pdfK, holds);
                                if (holds && !isExactlyZero(vp.sigmaS)) {
                                    const float trScatter = vp.tr * expf(-vp.sigmaT * (tePrime - vp.tStart));
                                    float mis = 1.0f;
                                    const float pdfTd =
                                        pdfK * pdfDistanceExponential(tePrime, vp.sigmaT, vp.tStart, vp.tStart + vp.delta);
                                    if (finitef(pdfTd)) mis = powerHeuristic(pdfTe, pdfTd);
                                    const V3 Ls = volumeSourceTerm(P, ss, L, o + d * tePrime, d, v.g, uel, 1.0f, k, st);
                                    direct = direct + T0 * vp.sigmaS * (Ls * trScatter) * (invN * mis / pdfTe);
                                }
                            }
                        }
                        const float ud = ss.next();
                        float udl[3];
                        udl[0] = ss.next();
                        udl[1] = ss.next();
                        udl[2] = ss.next();
                        float pdfDensity, udRemapped;
                        const VStep vp = pickStep(v, c, m, ud, pdfDensity, udRemapped);
                        if (pdfDensity == 0.0f || isExactlyZero(vp.sigmaS)) continue;
                        const float t0 = vp.tStart, t1 = vp.tStart + vp.delta;
                        const float td = sampleDistanceExponential(udRemapped, vp.sigmaT, t0, t1);
                        const float pdfTd = pdfDensity * pdfDistanceExponential(td, vp.sigmaT, t0, t1);
                        if (pdfTd == 0.0f || !finitef(pdfTd)) continue;
                        const float trScatter = vp.tr * expf(-vp.sigmaT * (td - t0));
                        float mis = 1.0f;
                        if (equiAngular) {
                            const float pdfTe = pdfEquiAngular(td - offset, D, thetaA, thetaB);
                            if (finitef(pdfTe)) mis = powerHeuristic(pdfTd, pdfTe);
                        }
                        const V3 Ls = volumeSourceTerm(P, ss, L, o + d * td, d, v.g, udl, 1.0f, k, st);
                        direct = direct + T0 * vp.sigmaS * (Ls * trScatter) * (invN * mis / pdfTd);
                    }
                radiance = radiance + direct;
            }
            // The glowing medium lighting the steps (volumeScatterRegions, :999-1057): a place drawn in the region,
            // an equi-angular distance toward it, the region's direction pdf from there and the medium's light along
            // that direction.
            const EmissionRegion& R = v.region;
            if (R.pool && (R.visibility & 0x100u)) {
                const float invN = 1.0f / float(samplesPerLight * scatterCount);
                V3 regionLight{0, 0, 0};
                for (uint32_t i = 0; i < samplesPerLight * scatterCount; ++i) {
                    Identity eid = root;
                    eid.branch = childBranch(volumeBranch, 0xfe00u, i);
                    Stream es(eid, depth);
                    const float u1 = es.next(), u2 = es.next(), u3 = es.next();
                    const V3 pivot = regionPoint(R, u1, u2, u3);
                    const float ue = es.next();
                    const float offset = dot(pivot - o, d);
                    const V3 foot = o + d * offset;
                    const float D = length(pivot - foot);
                    const float thetaA = atan2f(firstStart - offset, D), thetaB = atan2f(rayEnd - offset, D);
                    const float te = D * tanf((1 - ue) * thetaA + ue * thetaB);
                    const float pdfTe = pdfEquiAngular(te, D, thetaA, thetaB);
                    if (pdfTe == 0.0f || !finitef(pdfTe)) continue;
                    const float tePrime = te + offset;
                    float pdfK;
                    bool holds;
                    const VStep vp = stepAt(v, c, m, tePrime, pdfK, holds);
                    if (!holds || isExactlyZero(vp.sigmaS)) continue;
                    const float trScatter = vp.tr * expf(-vp.sigmaT * (tePrime - vp.tStart));
                    const V3 scatterPoint = o + d * tePrime;
                    V3 wi = pivot - scatterPoint;
                    wi = wi * (1.0f / length(wi));
                    float tEnd, glow;
                    const float pdfWi = regionPdf(R, scatterPoint, wi, tEnd, &glow);
                    if (pdfWi == 0.0f || !finitef(pdfWi)) continue;
                    const V3 Le = emissiveIntegral(P, es, scatterPoint, wi, 0.0f, tEnd, glow, depth + 1, k, st);
                    const float ph = phaseEval(v.g, -d, wi);
                    regionLight = regionLight + (T0 * vp.sigmaS) * (Le * trScatter) * (invN * ph / (pdfTe * pdfWi));
                }
                radiance = radiance + regionLight;
            }
        }
        ++volumeDepth;
        if (volumeDepth < c.max_volume_depth && c.volume_illumination_samples > 0) {
            const float invN = 1.0f / float(c.volume_illumination_samples);
            for (uint32_t i = 0; i < c.volume_illumination_samples; ++i) {
                Identity mid = root;
                mid.branch = childBranch(volumeBranch, 0xffff, i);
                Stream ms(mid, depth);
                float pdfDensity, utRemapped;
                const VStep vp = pickStep(v, c, m, ms.next(), pdfDensity, utRemapped);
                float sigmaT = vp.sigmaT;
                const float t0 = vp.tStart, t1 = vp.tStart + vp.delta;
                const float td = sampleDistanceExponential(utRemapped, sigmaT, t0, t1);
                const float pdfT = pdfDensity * pdfDistanceExponential(td, sigmaT, t0, t1);
                if (pdfT == 0.0f || !finitef(pdfT)) continue;
                const float trScatter = vp.tr * expf(-vp.sigmaT * (td - t0));
                V3 scatterPoint = o + d * td;
                V3 throughput = T0 * vp.sigmaS * (trScatter / pdfT);
                const float up0 = ms.next(), up1 = ms.next();
                V3 scatterDir = phaseSample(v.g, -d, up0, up1);
                for (uint32_t vd = volumeDepth; vd < c.max_volume_depth; ++vd) {
                    if (luminance(throughput) < c.volume_transmittance_threshold) break;
                    const float ai = powf(c.volume_attenuation_factor, float(vd)),
                                bi = powf(c.volume_contribution_factor, float(vd)),
                                ci = powf(c.volume_phase_attenuation_factor, float(vd));
                    const float tFreePath = -logf(1.0f - ms.next()) / sigmaT;
                    const V3 rayDir = scatterDir;
                    scatterPoint = scatterPoint + scatterDir * tFreePath;
                    if (!insideBox(v, scatterPoint)) break;
                    const VCoeff kk = coefficientsAt(v, scatterPoint, false);
                    const float g = v.g * ci;
                    sigmaT = kk.sigmaT;
                    if (fabsf(sigmaT) < VolEpsilon) break;
                    throughput = throughput * kk.sigmaS * (1.0f / sigmaT);
                    for (uint32_t li = 0; li < c.light_count; ++li) {
                        float usl[3];
                        usl[0] = ms.next();
                        usl[1] = ms.next();
                        usl[2] = ms.next();
                        const V3 Ls = volumeSourceTerm(P, ms, waveLight(P.lights, li), scatterPoint, rayDir, g, usl, ai, k, st);
                        radiance = radiance + throughput * Ls * (invN * bi);
                    }
                    // the glowing medium lighting this scatter (:1127-1143)
                    if (v.region.pool && (v.region.visibility & 0x100u)) {
                        const float ue0 = ms.next(), ue1 = ms.next(), ue2 = ms.next();
                        float pdfWi, tEnd, glow;
                        const V3 wi = regionSample(v.region, scatterPoint, ue0, ue1, ue2, pdfWi, tEnd, &glow);
                        if (pdfWi != 0.0f && finitef(pdfWi)) {
                            const V3 Le = emissiveIntegral(P, ms, scatterPoint, wi, 0.0f, tEnd, glow, depth + 1, k, st);
                            const float ph = phaseEval(g, -rayDir, wi);
                            radiance = radiance + (throughput * Le) * (invN * bi * ph / pdfWi);
                        }
                    }
                    const float uq0 = ms.next(), uq1 = ms.next();
                    scatterDir = phaseSample(g, -rayDir, uq0, uq1);
                }
            }
        }
    }
    T = T * vt;
    return vt;
}

// A sample's state between its vertices: the pending rays (a stack: a vertex's children are pushed after it, so the
// path tree is walked depth first), the radiance and alpha so far and the camera ray's clip distances; with Adj the
// adjoint's state too: the pixel's adjoint per sample, the sinks of the pending rays by stack position, and the open
// subtrees (the ancestors of the vertex being shaded). With Pg (the paged render) it keeps the sample's paging flags.
struct SampleFlagsNone {};
struct SampleFlagsKept {
    uint32_t flags;     // SampleParked, SampleProvisional (on-demand texture tiles)
    uint32_t missTile;  // the first tile it found missing, + 1 (0: none)
};
// In scenes with textures every pending ray carries its differentials (diffs, by stack position: the popped ray's
// stays at the stack's top until the vertex pushes its first child), which the map lookups' footprints come from.
struct DiffsNone {};
struct DiffsKept {
    RayDiff diffs[MaxPending];
};
template <bool Adj, bool Pg = false, bool Tx = false>
struct PathState : std::conditional_t<Pg, SampleFlagsKept, SampleFlagsNone>, std::conditional_t<Tx, DiffsKept, DiffsNone> {
    Pending stack[MaxPending];
    uint32_t top;
    V3 radiance;
    float alpha, cameraNear, cameraFar;
    V3 delta;
    Sink sinks[Adj ? MaxPending : 1];
    Record recs[Adj ? MaxRecords : 1];
    uint32_t nrec;
};

// A sample that met a missing texture tile: parked (it ends and runs again once the tile is resident), or, with
// partial shading, provisional (it finished from the coarse level, into the provisional film, and runs again).
enum : uint32_t { SampleParked = 1u, SampleProvisional = 2u };
template <class St>
struct HasSampleFlags {
    static constexpr bool value = false;
};
template <bool A, bool T>
struct HasSampleFlags<PathState<A, true, T>> {
    static constexpr bool value = true;
};
template <class St>
struct HasDiffs {
    static constexpr bool value = false;
};
template <bool A, bool G>
struct HasDiffs<PathState<A, G, true>> {
    static constexpr bool value = true;
};
template <class St>
__device__ __forceinline__ uint32_t sampleFlags(const St& S) {
    if constexpr (HasSampleFlags<St>::value) return S.flags;
    else return 0u;
}
template <class St>
__device__ __forceinline__ uint32_t sampleMissTile(const St& S) {
    if constexpr (HasSampleFlags<St>::value) return S.missTile;
    else return 0u;
}

// Starts sample (pixel, index): its camera ray becomes the one pending ray. False: an invalid value (Error).
template <bool Adj, bool Vol, class St>
__device__ __forceinline__ bool startSample(const RenderParams& P, uint32_t pixel, uint32_t index, St& S) {
    const x2_config& c = P.cfg;
    if (P.failPixel == pixel + 1) return false;  // injected failure, from the host
    const Identity root{pixel, index, c.seed, 0};
    float x, y;
    {
        float r0, r1;
        drawPair(root, 0, c.sampling, r0, r1);
        const float w = c.filter_width;
        if (!finite1(w) || w <= 0) return false;
        x = (r0 - .5f) * w + .5f;
        y = (r1 - .5f) * w + .5f;
        if (!finite1(x) || !finite1(y)) return false;
    }
    Ray camera;
    if (!cameraRay(c, float(pixel % c.width) + x, float(pixel / c.width) + y, camera)) return false;
    if constexpr (HasDiffs<St>::value) {  // the camera's record holds its auxiliary directions
        RayDiff& d = S.diffs[0];
        d.dOx = d.dOy = V3{0, 0, 0};
        d.dOx.x = notANumber();
        if (!cameraDifferentials(c, float(pixel % c.width) + x, float(pixel / c.width) + y, d.dDx, d.dDy)) return false;
    }
    S.cameraNear = camera.nearT;
    S.cameraFar = camera.farT;
    Pending& p = S.stack[0];
    p.origin = camera.origin;
    p.dir = camera.direction;
    p.T = {1, 1, 1};
    p.branch = 0;
    p.depths = 0;
    p.heldLight = Vol ? 0u : ~0u;
    p.heldFar = 0;
    p.held = {0, 0, 0};
    S.top = 1;
    S.radiance = V3{0, 0, 0};
    S.alpha = 0.0f;
    if constexpr (HasSampleFlags<St>::value) {
        S.flags = 0u;
        S.missTile = 0u;
    }
    if constexpr (Adj) {
        const float inv = 1.0f / float(c.spp);
        S.delta = V3{P.dImage[size_t(pixel) * 3] * inv, P.dImage[size_t(pixel) * 3 + 1] * inv,
                     P.dImage[size_t(pixel) * 3 + 2] * inv};
        S.sinks[0].kind = SinkNone;
        S.nrec = 0;
    }
    return true;
}

// How a vertex's shadow rays are answered: InlineTrace traces each one where it is made; the wavefront form queues
// it with the contribution it would add (WaveQueue) and adds that once the ray is traced.
struct InlineTrace {
    __device__ __forceinline__ void vertex(V3, V3) {}
    __device__ __forceinline__ bool shadow(const RenderParams& P, V3 o, V3 d, float nearT, float farT, V3, Counters& k,
                                           TravStack& st) {
        return visible(P, o, d, nearT, farT, k, st);
    }
};

// The glowing medium lighting a surface vertex (computeRadianceEmissiveRegionsBundled, async_integrator.h:2481-2613):
// the lobes the region lights as one sampler (a lobe picked in proportion to its albedo's luminance), light samples
// (a place drawn in the region) and BSDF samples, weighed against each other by the power heuristic, each carrying
// the medium's light along its direction (emissiveIntegral from the offset origin on the direction's side).
__device__ void surfaceRegionLight(const RenderParams& P, const Identity& root, uint64_t branch, uint32_t depth,
                                   bool split, const Bsdf& bsdf, V3 wo, V3 Ng, V3 Pt, V3 above, V3 below, V3 T,
                                   V3& radiance, Counters& k, TravStack& st) {
    const EmissionRegion& R = P.vol.region;
    const x2_config& c = P.cfg;
    const uint32_t nL = split ? c.light_samples : 1u, nB = split ? c.bsdf_samples : 1u;
    uint32_t recs[2], nRec = 0;
    float cdf[2], norm = 0.0f;
    for (uint32_t q = 0; q < bsdf.count; ++q) {
        if (!(lobeRayMask(bsdf.ltype[q]) & R.visibility)) continue;
        recs[nRec] = q;
        cdf[nRec] = fmaxf(fabsf(luminance(bsdf.kind[q] == 0 ? bsdf.albedo : bsdf.scale)), VolEpsilon);
        norm += cdf[nRec++];
    }
    if (!nRec) return;
    {
        float acc = 0.0f;
        for (uint32_t j = 0; j < nRec; ++j) {
            cdf[j] = cdf[j] / norm + acc;
            acc = cdf[j];
        }
    }
    V3 fs[2];
    uint32_t matched = 0;
    auto add = [&](V3 contribution) {
        for (uint32_t j = 0; j < matched; ++j) radiance = radiance + fs[j] * contribution;
    };
    auto emitted = [&](V3 wi, float tEnd, float glow, Stream& es) {
        return emissiveIntegral(P, es, dot(wi, Ng) >= 0.0f ? above : below, wi, 0.0f, tEnd, glow, depth + 1, k, st);
    };
    for (uint32_t ls = 0; ls < nL; ++ls) {
        Identity eid = root;
        eid.branch = childBranch(branch, 0xfd00u, ls);
        Stream es(eid, depth);
        const float u1 = es.next(), u2 = es.next(), u3 = es.next();
        float pdfLight, tEnd, glow;
        const V3 wi = regionSample(R, Pt, u1, u2, u3, pdfLight, tEnd, &glow);
        if (pdfLight == 0.0f) continue;
        // BsdfOneSampler::eval: the matching lobes' sum, the pdf weighted by the lobes' picks
        const uint32_t flags = surfaceFlags(Ng, wo, wi);
        V3 f{0, 0, 0};
        float pdfBsdf = 0.0f, prev = 0.0f;
        matched = 0;
        for (uint32_t j = 0; j < nRec; ++j) {
            float p;
            const V3 col = lobeEval(bsdf, recs[j], wo, wi, p);
            if (matchesFlags(bsdf.ltype[recs[j]], flags)) {
                f = f + col;
                fs[matched++] = col;
            }
            pdfBsdf += p * (cdf[j] - prev);
            prev = cdf[j];
        }
        if (isExactlyZero(f)) continue;
        V3 contribution = emitted(wi, tEnd, glow, es);
        const float mis = nB ? powerHeuristic(float(nL) * pdfLight, float(nB) * pdfBsdf) : 1.0f;
        contribution = (contribution * T) * (mis / (pdfLight * float(nL)));
        add(contribution);
    }
    for (uint32_t bs = 0; bs < nB; ++bs) {
        Identity eid = root;
        eid.branch = childBranch(branch, 0xfc00u, bs);
        Stream es(eid, depth);
        const float uLobe = es.next(), u1 = es.next(), u2 = es.next();
        // BsdfOneSampler::sample: a lobe by its pick, sampled, the other matching lobes' values added
        uint32_t pick = 0;
        while (pick + 1 < nRec && cdf[pick] <= uLobe) ++pick;
        const float lobePdf = pick == 0 ? cdf[0] : cdf[pick] - cdf[pick - 1];
        V3 wi;
        float pdfBsdf;
        V3 f = lobeSample(bsdf, recs[pick], wo, u1, u2, wi, pdfBsdf);
        if (pdfBsdf == 0.0f) continue;
        pdfBsdf *= lobePdf;
        matched = 0;
        fs[matched++] = f;
        const uint32_t flags = surfaceFlags(Ng, wo, wi);
        if (!matchesFlags(bsdf.ltype[recs[pick]], flags)) f = V3{0, 0, 0};
        float prev = 0.0f;
        for (uint32_t j = 0; j < nRec; ++j) {
            if (j != pick) {
                float p;
                const V3 col = lobeEval(bsdf, recs[j], wo, wi, p);
                if (matchesFlags(bsdf.ltype[recs[j]], flags)) {
                    f = f + col;
                    fs[matched++] = col;
                }
                pdfBsdf += p * (cdf[j] - prev);
            }
            prev = cdf[j];
        }
        if (pdfBsdf == 0.0f || isExactlyZero(f)) continue;
        float tEnd, glow;
        const float pdfLight = regionPdf(R, Pt, wi, tEnd, &glow);
        if (pdfLight == 0.0f) continue;
        V3 contribution = emitted(wi, tEnd, glow, es);
        const float mis = nL ? powerHeuristic(float(nB) * pdfBsdf, float(nL) * pdfLight) : 1.0f;
        contribution = (contribution * T) * (mis / (pdfBsdf * float(nB)));
        add(contribution);
    }
}

// Shades a traced vertex (pv, along ray, found hit): adds the light found there and pushes the rays it spawns onto
// the sample's pending stack. False: an invalid value (Error; the sample adds nothing). With Adj every
// contribution's derivative is carried to its sinks, scaled by the pixel's adjoint (gr: the lane's roughness
// gradient accumulators). St is the sample's state (PathState, or a wavefront path slot), Tr answers shadow rays:
// a shadow ray's contribution is added here only when Tr.shadow says it is unoccluded.
template <bool Adj, bool Vol, class St, class Tr>
__device__ __forceinline__ bool shadeHit(const RenderParams& P, uint32_t pixel, uint32_t index, St& S, Tr& tracer,
                                         const Pending& pv, const Ray& ray, const Hit& hit, Counters& k, TravStack& st,
                                         float* gr) {
    const x2_config& c = P.cfg;
    const uint32_t rrMin = c.rr_min_depth > 1 ? c.rr_min_depth : 1u;
    const Identity root{pixel, index, c.seed, 0};
    auto& stack = S.stack;
    uint32_t& top = S.top;
    V3& radiance = S.radiance;
    float& alpha = S.alpha;
    const V3& delta = S.delta;
    auto& sinks = S.sinks;
    const uint32_t depth = pv.depths & 0xffu, nonMirrorDepth = (pv.depths >> 8) & 0xffu,
                   diffuseDepth = (pv.depths >> 16) & 0xffu, glossyDepth = pv.depths >> 24;
    // With a medium there is no bounce-ray reuse (:2919), and heldLight holds the path's volume depth.
    uint32_t volumeDepth = Vol ? pv.heldLight : 0u;
    if (!Vol && pv.heldLight != ~0u) {
        const float kept = hit.geom != ~0u && hit.t < pv.heldFar ? 0.0f : 1.0f;
        if (kept > 0.0f) radiance = radiance + (kept < 1.0f ? pv.held * kept : pv.held);
    }
    float volumeAlpha = 1.0f;
    V3 T = pv.T;
    if constexpr (Vol)  // the medium on the ray's way (:2106-2120): its light, and the throughput through it
        volumeAlpha = volumeRadiance(P, root, T, depth, nonMirrorDepth, volumeDepth, pv.branch, ray.origin,
                                     ray.direction, ray.nearT, hit.geom == ~0u ? ray.farT : hit.t, radiance, k, st);
    const float hitAlpha = 1.0f - (0.0f + 0.0f + 0.0f) / 3.0f;
    if (hit.geom == ~0u) {
        if (depth == 0) alpha += 1.0f * (1.0f - (volumeAlpha + volumeAlpha + volumeAlpha) / 3.0f);
        return true;
    }
    // shade::hitState for a flat triangle: Ng = (b - a) x (c - a), turned to face the ray.
    const TriRecord& tr = P.tris[hit.tri];
    const V3 a{tr.ax, tr.ay, tr.az}, b{tr.bx, tr.by, tr.bz}, cc{tr.cx, tr.cy, tr.cz};
    V3 ng = cross(b - a, cc - a);
    float l = length(ng);
    if (!(l > 0)) {
        const double ux = double(b.x) - a.x, uy = double(b.y) - a.y, uz = double(b.z) - a.z;
        const double vx = double(cc.x) - a.x, vy = double(cc.y) - a.y, vz = double(cc.z) - a.z;
        const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
        const double dl = sqrt(nx * nx + ny * ny + nz * nz);
        ng = dl > 0 ? V3{float(nx / dl), float(ny / dl), float(nz / dl)} : -ray.direction;
        l = length(ng);
        if (!(l > 0)) return false;
    }
    V3 Ng = ng / l, N = Ng;
    const V3 Pt = hit.point, wo = -ray.direction;
    if (!(dot(Ng, ray.direction) < 0.0f)) {
        N = -N;
        Ng = -Ng;
    }
    if (!finite3(N) || !finite3(Ng)) return false;
    // The hit's differentials (shade_material.h:538-558), in scenes with textures; the texture coordinates of the
    // triangle's corners.
    float uv[6] = {0, 0, 1, 0, 1, 1};
    if (P.triUV)
        for (int q = 0; q < 6; ++q) uv[q] = P.triUV[size_t(hit.tri) * 6 + q];
    HitDiff hd;
    hd.dsdx = hd.dtdx = hd.dsdy = hd.dtdy = 0.0f;
    hd.valid = false;
    if constexpr (HasDiffs<St>::value)
        hitDifferentials(a, b, cc, uv, N, Ng, Pt, ray, depth == 0, S.diffs[top], c.spp, P.materials[tr.material].albedo_map != 0,
                         hd);
    Bsdf bsdf;
    float rInput = 0.0f;  // the material's input roughness (the adjoint's GGX derivative)
    Sink texSink, roughSink;
    V3 texDirect{0, 0, 0};  // the direct light through the textured Lambert lobe, scattered once at the end
    texSink.kind = SinkNone;
    roughSink.kind = SinkNone;
    {
        const x2_material& m = P.materials[tr.material];
        bsdf.count = 0;
        bsdf.type = 0;
        bsdf.frame = frameOf(N);
        V3 albedo = ld3(m.albedo);
        if (m.albedo_map) {  // the map at the hit's (s, t) (shade_material.h:121-160, shade_mapnodes.h:420-470)
            if constexpr (!HasDiffs<St>::value) {
                return false;  // a scene with textures runs the kernels that carry differentials
            } else {
                const float w = 1.0f - hit.u - hit.v;
                const float ss = w * uv[0] + hit.u * uv[2] + hit.v * uv[4], tt = w * uv[1] + hit.u * uv[3] + hit.v * uv[5];
                // textureMode (shade_mapnodes.h:344-347): camera paths anisotropic, diffuse ones one level's closest
                // texel, glossy ones trilinear closest texels.
                const uint32_t mode = diffuseDepth ? 2u : glossyDepth ? 1u : 0u;
                const Footprint f{ss, tt, hd.dsdx, hd.dtdx, hd.dsdy, hd.dtdy};
                const TexDesc& td = P.textures[m.albedo_map - 1];
                if (td.paged && !HasSampleFlags<St>::value) return false;  // paged: the paged render reads it
                V3 color;
                uint32_t* firstMiss = nullptr;
                if constexpr (HasSampleFlags<St>::value) firstMiss = &S.missTile;
                const int r = mapLookup(P, td, mode, f, true, color, firstMiss);
                if (r == TapInvalid) return false;
                if constexpr (HasSampleFlags<St>::value && !Adj) {  // the paged render (renderPagedKernel)
                    if (r == TapMiss) {  // the sample waits for its tile: it ends here, parked, and runs again
                        S.flags |= SampleParked;
                        S.top = 0;
                        return true;
                    }
                    if (r == TapCoarse) S.flags |= SampleProvisional;
                } else if (r != TapOk) {
                    return false;
                }
                albedo = albedo * color;
                if constexpr (Adj) {
                    texSink.kind = SinkTexture;
                    texSink.tex = m.albedo_map - 1;
                    texSink.mode = mode;
                    texSink.f = f;
                    texSink.color = color;
                }
            }
        }
        if constexpr (Adj) {
            if (m.specular_weight > 0.0f && m.roughness * m.roughness > 0.001f) {
                roughSink.kind = SinkRoughness;
                roughSink.material = tr.material;
            }
        }
        bsdf.albedo = albedo * m.diffuse_weight;
        rInput = m.roughness;
        bsdf.scale = {m.specular_weight, m.specular_weight, m.specular_weight};
        bsdf.roughness = fmaxf(0.001f, m.roughness * m.roughness);
        if (m.diffuse_weight > 0.0f) {
            bsdf.kind[bsdf.count] = 0;
            bsdf.ltype[bsdf.count++] = Reflection | Diffuse;
            bsdf.type |= Reflection | Diffuse;
        }
        if (m.specular_weight > 0.0f) {
            bsdf.kind[bsdf.count] = 1;
            bsdf.ltype[bsdf.count++] = Reflection | Glossy;
            bsdf.type |= Reflection | Glossy;
        }
        bsdf.emission = ld3(m.emission);
    }
    if (depth == 0) alpha += 1.0f * hitAlpha;
    radiance = radiance + T * bsdf.emission;
    if (bsdf.count == 0) return true;
    uint32_t indirectFlags = 0;
    if (c.bsdf_samples > 0 && c.max_depth > 0 && depth < c.max_depth) {
        if (diffuseDepth < c.max_diffuse_depth) indirectFlags |= AllSurfaceSides | Diffuse;
        if (glossyDepth < c.max_glossy_depth) indirectFlags |= AllSurfaceSides | Glossy;
        if (0u < c.max_mirror_depth) indirectFlags |= AllSurfaceSides | Mirror;
    }
    const bool doIndirect = indirectFlags != 0;
    const V3 cullN = (bsdf.type & Reflection) ? Ng : -Ng;
    const bool cull = (bsdf.type & AllSurfaceSides) != AllSurfaceSides;
    const bool split = nonMirrorDepth == 0;
    const uint32_t nl = split ? c.light_samples : 1;
    const bool clampHere = nonMirrorDepth >= c.clamp_depth;
    const bool rr = nonMirrorDepth >= rrMin && c.rr_threshold > 0;
    const uint32_t nk = split ? c.bsdf_samples : 1;  // no mirror lobes
    const uint32_t planned = nk * bsdf.count;
    const float thinning = bounceThinning(top, planned);
    const V3 above = hit.origin, below = hit.below;
    tracer.vertex(above, below);
    if constexpr (Vol)
        if (P.vol.region.pool)
            surfaceRegionLight(P, root, pv.branch, depth, split, bsdf, wo, Ng, Pt, above, below, T, radiance, k, st);

    // Light samples (:2615-2785).
    const uint64_t lightBranch = childBranch(pv.branch, 0xfffe, 0);
    for (uint32_t i = 0; i < c.light_count; ++i) {
        const x2_light& L = waveLight(P.lights, i);
        if (!canIlluminate(L, P.lightCorners + 12 * i, Pt, cull, cullN)) continue;
        const uint64_t lightAddress = childBranch(lightBranch, i, 0);
        Identity lid = root;
        lid.branch = lightAddress;
        Stream ls(lid, depth);
        for (uint32_t s = 0; s < nl; ++s) {
            // the sample's stream block and, at a split vertex, its stratified pair's block, computed together
            const Identity strat{pixel, uint64_t(index) * nl + s, c.seed, lightAddress};
            Identity src = lid;
            uint32_t srcGroup = 0;
            if (split) pairSource(strat, CameraDimensions + 2, c.sampling, src, srcGroup);
            Words streamBlock, pairBlock;
            block2(lid, ls.nextGroup(), src, srcGroup, streamBlock, pairBlock);
            ls.prime(streamBlock);
            float r[3];
            r[0] = ls.next();
            r[1] = ls.next();
            r[2] = ls.next();
            const float u = ls.next();
            if (split) pairFrom(strat, CameraDimensions + 2, c.sampling, pairBlock, r[0], r[1]);
            V3 wi;
            LightHit lh;
            if (!lightSample(L, Pt, cull, cullN, r, wi, lh)) continue;
            float pdfL;
            const V3 Li = lightEval(L, wi, lh, pdfL);
            if (pdfL == 0.0f || isExactlyZero(Li) || !(fabsf(pdfL) < 3.4e38f) || !finite3(Li)) continue;
            const V3 factor = T * Li * (1.0f / (float(nl) * pdfL * 1.0f));
            const uint32_t flags = surfaceFlags(Ng, wo, wi);
            V3 t{0, 0, 0};
            V3 tkq[2] = {{0, 0, 0}, {0, 0, 0}};  // each lobe's share, for the adjoint
            bool any = false;
            for (uint32_t q = 0; q < bsdf.count; ++q) {
                if (!matchesFlags(bsdf.ltype[q], flags)) continue;
                if (!(lobeRayMask(bsdf.ltype[q]) & L.visibility_mask)) continue;
                float pdf;
                const V3 f = lobeEval(bsdf, q, wo, wi, pdf);
                if (pdf == 0.0f || isExactlyZero(f)) continue;
                V3 tk = factor * f * powerHeuristic(float(nl) * pdfL, float(nk) * pdf);
                if (clampHere) tk = smartClamp(tk, c.clamp_value);
                t = t + tk;
                if constexpr (Adj) tkq[q] = tk;
                any = true;
            }
            if (!any) continue;
            float rrScale = 1.0f;
            if (rr) {
                float scale;
                if (!roulette(luminance(t), c.rr_threshold, u, scale)) continue;
                t = t * scale;
                rrScale = scale;
            }
            if (!finite3(t) || isExactlyZero(t)) continue;
            const V3 shadowOrigin = dot(wi, Ng) >= 0.0f ? above : below;
            const float shadowNear = 0.0f, shadowFar = fmaxf((lh.distance - 0.0f) * HitEpsilonEnd, 2.0f * shadowNear);
            if (tracer.shadow(P, shadowOrigin, wi, shadowNear, shadowFar, t, k, st)) {
                if constexpr (Vol) {  // the medium on the shadow ray's way (:2725-2732)
                    Identity tid = lid;
                    tid.branch = childBranch(lightAddress, s, 0xfffb);
                    Stream vs(tid, depth);
                    t = t * transmittance(P.vol, shadowOrigin, wi, shadowNear, shadowFar, 1.0f,
                                          c.volume_transmittance_threshold, vs);
                }
                radiance = radiance + t;
                if constexpr (Adj) {
                    for (uint32_t q = 0; q < bsdf.count; ++q) {
                        const V3 share = tkq[q] * rrScale;
                        if (bsdf.kind[q] == 0 && texSink.kind == SinkTexture) texDirect = texDirect + share;
                        if (bsdf.kind[q] == 1 && roughSink.kind == SinkRoughness)
                            addRoughness(P, roughSink.material, (delta.x * share.x + delta.y * share.y + delta.z * share.z) *
                                             ggxRoughnessLogDerivative(rInput, bsdf.frame.z, wo, wi), gr);
                    }
                }
            }
        }
    }

    // BSDF samples (:2787-3031).
    for (uint32_t q = 0; q < bsdf.count; ++q) {
        const uint32_t ltype = bsdf.ltype[q];
        const uint32_t lobeMask = lobeRayMask(ltype);
        const bool lobeIndirect = matchesFlags(ltype, indirectFlags);
        const uint64_t lobeBranch = childBranch(pv.branch, q, 0xffff);
        for (uint32_t i = 0; i < nk; ++i) {
            const uint64_t sampleBranch = childBranch(lobeBranch, i, 0);
            Identity bid = root;
            bid.branch = sampleBranch;
            Stream bs(bid, depth);
            // the sample's stream block and, at a split vertex, its stratified pair's block, computed together
            const Identity strat{pixel, uint64_t(index) * nk + i, c.seed, lobeBranch};
            Identity src = bid;
            uint32_t srcGroup = 0;
            if (split) pairSource(strat, CameraDimensions, c.sampling, src, srcGroup);
            Words streamBlock, pairBlock;
            block2(bid, bs.nextGroup(), src, srcGroup, streamBlock, pairBlock);
            bs.prime(streamBlock);
            float r1 = bs.next(), r2 = bs.next();
            const float choice = bs.next(), u = bs.next();
            if (split) pairFrom(strat, CameraDimensions, c.sampling, pairBlock, r1, r2);
            V3 wi;
            float pdf;
            const V3 f = lobeSample(bsdf, q, wo, r1, r2, wi, pdf);
            if (!(pdf != 0.0f && !isExactlyZero(f)) || !finite3(f) || !finite3(wi)) continue;
            if (!matchesFlags(ltype, surfaceFlags(Ng, wo, wi))) continue;
            const V3 pt = T * f * (1.0f / (float(nk) * pdf));
            uint32_t hits = 0;
            for (uint32_t j = 0; j < c.light_count; ++j) {
                LightHit lh;
                const x2_light& Lj = waveLight(P.lights, j);
                if ((lobeMask & Lj.visibility_mask) && lightIntersect(Lj, Pt, wi, MaxFloat, lh)) ++hits;
            }
            V3 direct{0, 0, 0};
            float lightDistance = 0;
            bool lightHit = false;
            uint32_t lightIndex = 0;
            if (hits) {
                const uint32_t chosen = uint32_t(choice * float(hits)) < hits ? uint32_t(choice * float(hits)) : hits - 1;
                for (uint32_t j = 0, seen = 0; j < c.light_count; ++j) {
                    LightHit lh;
                    const x2_light& Lj = waveLight(P.lights, j);
                    if (!(lobeMask & Lj.visibility_mask) || !lightIntersect(Lj, Pt, wi, MaxFloat, lh))
                        continue;
                    if (seen++ != chosen) continue;
                    float pdfL;
                    const V3 Li = lightEval(Lj, wi, lh, pdfL) * float(hits);
                    if (!(pdfL == 0.0f || isExactlyZero(Li)) && fabsf(pdfL) < 3.4e38f && finite3(Li)) {
                        direct = Li * pt * powerHeuristic(float(nk) * pdf, float(nl) * pdfL);
                        lightDistance = lh.distance;
                        lightHit = true;
                        lightIndex = j;
                    }
                    break;
                }
            }
            if (!lightHit && !doIndirect) continue;
            if (clampHere) direct = smartClamp(direct, c.clamp_value);
            V3 indirect = doIndirect ? pt : V3{0, 0, 0};
            if (rr) {
                float scale;
                if (!roulette(fmaxf(luminance(direct), luminance(indirect)), c.rr_threshold, u, scale)) continue;
                direct = direct * scale;
                indirect = indirect * scale;
            }
            const V3 origin = dot(wi, Ng) >= 0.0f ? above : below;
            float shadowFar = 0.0f, kept = 0.0f;
            const bool bounces = lobeIndirect && !isExactlyZero(indirect) && finite3(indirect);
            bool held = false;
            if (lightHit && !isBlack(direct) && finite3(direct)) {
                shadowFar = fmaxf((lightDistance - 0.0f) * HitEpsilonEnd, 2.0f * 0.0f);
                held = !Vol && bounces && thinning >= 1.0f && top + PendingReserve < MaxPending && lobeMask &&
                       (c.shadow_reuse & lobeMask) == lobeMask;
                if (!held) kept = tracer.shadow(P, origin, wi, 0.0f, shadowFar, direct, k, st) ? 1.0f : 0.0f;
            }
            const V3 directBefore = direct;
            Sink lobeSink;
            lobeSink.kind = SinkNone;
            if constexpr (Adj) {
                if (bsdf.kind[q] == 0) lobeSink = texSink;
                else if (roughSink.kind == SinkRoughness) {
                    lobeSink = roughSink;
                    lobeSink.g = ggxRoughnessLogDerivative(rInput, bsdf.frame.z, wo, wi);
                }
            }
            if (kept > 0.0f && !held) {
                if (kept < 1.0f) direct = direct * kept;
                if constexpr (Vol) {  // (:2939-2945)
                    Identity tid = bid;
                    tid.branch = childBranch(sampleBranch, 0xfffb, 0);
                    Stream vs(tid, depth);
                    direct = direct * transmittance(P.vol, origin, wi, 0.0f, shadowFar, 1.0f,
                                                    c.volume_transmittance_threshold, vs);
                }
                radiance = radiance + direct;
                if constexpr (Adj) {
                    if (lobeSink.kind == SinkTexture) texDirect = texDirect + direct;
                    else scatterSink(P, lobeSink, direct, delta, gr);
                }
            }
            if (!bounces) continue;
            float keptShare = 1.0f;
            if (thinning < 1.0f || top + PendingReserve >= MaxPending) {
                Identity kid = bid;
                kid.branch = childBranch(sampleBranch, 0xfff7, 0);
                Stream chance(kid, depth);
                if (thinning < 1.0f) {
                    if (chance.next() >= thinning) continue;
                    keptShare = 1.0f / thinning;
                }
                if (top + PendingReserve >= MaxPending) {
                    float tight = 1.0f;
                    if (!roomForChild(top, chance.next(), tight)) continue;
                    keptShare *= tight;
                }
            }
            if constexpr (Adj) sinks[top] = lobeSink;
            if constexpr (HasDiffs<St>::value)  // async_integrator.h:2985
                S.diffs[top] = scatterDifferentials(bsdf.kind[q] == 0, bsdf.frame, bsdf.roughness, hd, Pt, wo, wi, r1, r2,
                                                    origin, 1.0f / fminf(sqrtf(float(c.spp ? c.spp : 1)), 2.0f));
            Pending& child = stack[top++];
            child.origin = origin;
            child.dir = wi;
            child.T = indirect * keptShare;
            child.branch = sampleBranch;
            child.depths = (depth + 1) | ((nonMirrorDepth + 1) << 8) |
                           ((diffuseDepth + ((ltype & Diffuse) ? 1u : 0u)) << 16) |
                           ((glossyDepth + ((ltype & Glossy) ? 1u : 0u)) << 24);
            child.heldLight = Vol ? volumeDepth + 1 : held ? lightIndex : ~0u;  // (:3000)
            child.heldFar = shadowFar;
            child.held = directBefore;
        }
    }
    // The vertex's direct light through its textured lobe, scattered to the texels once (the scatter is linear in R).
    if constexpr (Adj)
        if (texSink.kind == SinkTexture && !isExactlyZero(texDirect)) scatterTexture(P, texSink, texDirect, delta);
    return true;
}

// Shades the sample's last pending ray: traces it, then shadeHit with Tr answering its shadow rays (InlineTrace:
// traced where they are made).
template <bool Adj, bool Vol, class St, class Tr>
__device__ __forceinline__ bool shadeVertexWith(const RenderParams& P, uint32_t pixel, uint32_t index, St& S, Tr& tracer,
                                                Counters& k, TravStack& st, float* gr) {
    const Pending pv = S.stack[--S.top];
    if constexpr (Adj) {
        while (S.nrec && S.recs[S.nrec - 1].pos > S.top) {
            --S.nrec;
            scatterSink(P, S.recs[S.nrec].sink, S.radiance - S.recs[S.nrec].mark, S.delta, gr);
        }
        if (S.nrec < MaxRecords) S.recs[S.nrec++] = Record{S.top, S.radiance, S.sinks[S.top]};
    }
    const bool fromCamera = (pv.depths & 0xffu) == 0;
    const Ray ray{pv.origin, pv.dir, fromCamera ? S.cameraNear : 0.0f, fromCamera ? S.cameraFar : MaxFloat};
    Hit hit;
    if (!traceHit(P, ray, hit, k, st)) return false;
    return shadeHit<Adj, Vol>(P, pixel, index, S, tracer, pv, ray, hit, k, st, gr);
}
template <bool Adj, bool Vol, class St>
__device__ __forceinline__ bool shadeVertex(const RenderParams& P, uint32_t pixel, uint32_t index, St& S, Counters& k,
                                            TravStack& st, float* gr) {
    InlineTrace tracer;
    return shadeVertexWith<Adj, Vol>(P, pixel, index, S, tracer, k, st, gr);
}

// Closes a finished sample's open subtrees; with Adj each carries its share of the radiance to its sink.
template <bool Adj, class St>
__device__ __forceinline__ void finishSample(const RenderParams& P, St& S, float* gr) {
    if constexpr (Adj) {
        while (S.nrec) {
            --S.nrec;
            scatterSink(P, S.recs[S.nrec].sink, S.radiance - S.recs[S.nrec].mark, S.delta, gr);
        }
    }
}

// One sample; 0 Done (rgba set), 2 Error.
template <bool Adj, bool Vol>
__device__ int integrate(const RenderParams& P, uint32_t pixel, uint32_t index, double rgba[4], Counters& k,
                         TravStack& st, float* gr) {
    PathState<Adj> S;
    if (!startSample<Adj, Vol>(P, pixel, index, S)) return 2;
    while (S.top)
        if (!shadeVertex<Adj, Vol>(P, pixel, index, S, k, st, gr)) return 2;
    finishSample<Adj>(P, S, gr);
    rgba[0] = S.radiance.x;
    rgba[1] = S.radiance.y;
    rgba[2] = S.radiance.z;
    rgba[3] = S.alpha;
    return 0;
}

// A kernel's RenderParams (its one argument, at the start of its argument segment) read through a pointer made
// opaque at each call. A loop that takes them so each iteration makes the compiler load each field where it is used
// (scalar loads from the argument segment) instead of holding every field it can hoist in scalar registers across
// the loop, which spills most of them to vector lanes (every reload a VALU read-lane).
__device__ __forceinline__ const RenderParams& argParams() {
    auto kp = (const __attribute__((address_space(4))) RenderParams*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(kp));
    return *(const RenderParams*)kp;
}

// The estimator over a lane's samples, one after another (path regeneration): a sample's vertices are shaded one
// per iteration, and when its last pending ray is done the lane takes its next sample, so the lanes of a wave keep
// shading vertices together instead of waiting for the wave's longest path. feed.next gives the next sample (false:
// none left), feed.done takes a finished sample's radiance and alpha, feed.failed a sample that hit an invalid value
// (Error; it adds nothing). ArgP: P is the kernel's one argument, read afresh each iteration (argParams).
template <bool Adj, bool Vol, bool Tx, bool ArgP = false, class Feed>
__device__ void runSamples(const RenderParams& P0, Feed& feed, Counters& k, TravStack& st, float* gr) {
    PathState<Adj, Feed::Paged, Tx> S;
    S.top = 0;
    uint32_t pixel = 0, index = 0;
    bool live = false;
    for (;;) {
        const RenderParams& P = ArgP ? argParams() : P0;
        const bool idle = S.top == 0;
        if (idle && live) {
            finishSample<Adj>(P, S, gr);
            feed.done(pixel, index, S.radiance, S.alpha, sampleFlags(S), sampleMissTile(S));
            live = false;
        }
        // A finished lane takes its next sample once Feed::Regen lanes of the wave are finished (or all are), so that
        // the new samples start together (consecutive items: one tile's neighbouring pixels) rather than one by one.
        bool claim = idle;
        if constexpr (Feed::Regen > 1) {
            const unsigned long long idleLanes = __ballot(idle), activeLanes = __ballot(1);
            claim = idle && (__popcll(idleLanes) >= Feed::Regen || idleLanes == activeLanes);
        }
        if (idle && !claim) continue;
        if (claim) {
            if (!feed.next(pixel, index)) break;
            if (!startSample<Adj, Vol>(P, pixel, index, S)) {
                feed.failed();
                S.top = 0;
                continue;
            }
            live = true;
        }
        if (!shadeVertex<Adj, Vol>(P, pixel, index, S, k, st, gr)) {
            feed.failed();
            S.top = 0;
            live = false;
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Shadow rays pooled across the wave. A vertex's shadow rays (at most one per light sample and one per lobe sample)
// wait in the wave's LDS pool with the radiance each adds when unoccluded; the wave's next trace phase (traceWave)
// traces them together with the lanes' next rays, each lane taking the pool's next ray whoever made it once its own
// is done, and each lane then adds its own rays' radiance in the order it made them. The traces are the same rays as
// traced where they are made, and the additions the same sums in the same order, so the sample's radiance is bit for
// bit the same; what changes is that a lane with no ray traces another lane's instead of idling.
constexpr uint32_t PoolRays = 6;  // shadow rays a lane's vertex may queue (the host checks the frame's bound)
struct ShadowPool {  // one wave's, 13,440 B: ray n of lane l at n * 64 + l (bank-conflict free)
    float dx[64 * PoolRays], dy[64 * PoolRays], dz[64 * PoolRays], far[64 * PoolRays];
    float vr[64 * PoolRays], vg[64 * PoolRays], vb[64 * PoolRays];  // the radiance each adds when unoccluded
    float org[6 * 64];                                              // each lane's origins above (0..2), below (3..5)
    uint16_t list[64 * PoolRays];                                   // the wave's rays, compacted
    uint8_t flag[64 * PoolRays];                                    // in: origin below; out: 2 unoccluded
};
static_assert(sizeof(ShadowPool) == 13440, "a wave's shadow pool");
__device__ __forceinline__ void waveSync() {  // the wave's LDS writes seen by its own later reads, lane to lane
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
struct PoolQueue {
    ShadowPool* pool;
    uint32_t lane, n;
    bool overflow;
    V3 above;
    __device__ __forceinline__ void vertex(V3 a, V3 b) {
        above = a;
        pool->org[0 * 64 + lane] = a.x, pool->org[1 * 64 + lane] = a.y, pool->org[2 * 64 + lane] = a.z;
        pool->org[3 * 64 + lane] = b.x, pool->org[4 * 64 + lane] = b.y, pool->org[5 * 64 + lane] = b.z;
    }
    // Every shadow ray of the surface estimator starts at 0 along the ray (nearT) from the vertex's origin above or
    // below the surface; it is queued, and answered as occluded for now.
    __device__ __forceinline__ bool shadow(const RenderParams&, V3 o, V3 d, float, float farT, V3 value, Counters&,
                                           TravStack&) {
        if (n >= PoolRays) {
            overflow = true;
            return false;
        }
        const uint32_t e = n++ * 64 + lane;
        pool->dx[e] = d.x, pool->dy[e] = d.y, pool->dz[e] = d.z, pool->far[e] = farT;
        pool->vr[e] = value.x, pool->vg[e] = value.y, pool->vb[e] = value.z;
        const bool isAbove = __float_as_uint(o.x) == __float_as_uint(above.x) &&
                             __float_as_uint(o.y) == __float_as_uint(above.y) &&
                             __float_as_uint(o.z) == __float_as_uint(above.z);
        pool->flag[e] = isAbove ? 0u : 1u;
        return false;
    }
};
// One trace phase of the wave: each lane's own next ray (closest hit) and the rays the wave's last vertices queued
// in the pool (any hit), in one loop. A lane starts on its own ray, or on the pool's next ray when it has none, and
// when its ray ends takes the pool's next one, until the pool is empty: the wave waits once for its longest ray
// instead of once for the longest closest-hit ray and again for the longest shadow ray, and a lane whose ray ended
// early traces another instead of idling. Each ray's answer is the one closestHit or anyHit gives it (the least
// key, whatever the order; occluded or not): a pool ray is walked as a closest-hit ray that stops at its first hit,
// with its far distance as the bound. The pool's next ray is counted in a register for the wave: the lanes that
// want a ray take consecutive ones by their rank among them. Returns the lane's own ray's triangle record (~0u:
// none). Every lane of the wave calls it together.
__device__ uint32_t traceWave(const RenderParams& P, PoolQueue& q, bool own, const Ray& ray, Counters& k,
                              TravStack& st) {
    ShadowPool& p = *q.pool;
    const uint32_t lane = q.lane, n = q.n;
    const unsigned long long lower = (1ull << lane) - 1ull;
    uint32_t first = 0, total = 0;  // the lane's first place in the compacted list, the wave's pool rays
#pragma unroll
    for (int b = 0; b < 3; ++b) {  // an exclusive prefix sum of n (< 8) over the lanes, a bit at a time
        const unsigned long long m = __ballot((n >> b) & 1u);
        first += uint32_t(__popcll(m & lower)) << b;
        total += uint32_t(__popcll(m)) << b;
    }
    for (uint32_t j = 0; j < n; ++j) p.list[first + j] = uint16_t(j * 64 + lane);
    waveSync();
    uint32_t ownBest = ~0u, next = 0;
    bool active = false, pooled = false, occluded = false;
    uint32_t e = 0;  // the current pool ray's slot
    WoopRay w;
    BoxRay br;
    V3 dir{0, 0, 0};
    float lo = 0.0f, farT = 0.0f, bestT = 0.0f, hi = 0.0f;  // hi: the bound on entry distances, bestT with slack
    auto slackAbove = [](float b) { return b > 0 ? b * Slack : b / Slack; };
    uint32_t best = ~0u, ref = Sentinel, leaf = 0;
    int sp = 0;
    // The stack's top entry held in registers (topRef, topDist) instead of st[sp - 1]: the pop that most often
    // follows a push (a node whose children all missed, a postponed leaf) then waits on no private-memory load
    // (753 against 703 Mpaths/s on the 1080p icospheres frame).
    bool cached = false;
    uint32_t topRef = 0, topDist = 0;
    auto start = [&](V3 o, V3 d, float nearT, float f) {
        w = woopSetup(o, d, nearT, f);
        br = boxRay(o, d);
        dir = d;
        lo = nearT > 0 ? nearT / Slack : nearT * Slack;
        farT = f;
        best = ~0u;
        bestT = f;
        hi = slackAbove(f);
        sp = 0;
        cached = false;
        ref = 0;
        leaf = 0;
        occluded = false;
        active = true;
    };
    auto take = [&]() {  // the lanes without a ray take the pool's next ones, in lane order
        const unsigned long long want = __ballot(!active);
        if (next >= total) return;
        if (!active) {
            const uint32_t j = next + uint32_t(__popcll(want & lower));
            if (j < total) {
                e = p.list[j];
                const uint32_t owner = e & 63u, side = p.flag[e] ? 3u : 0u;
                pooled = true;
                ++k.shadows;
                start(V3{p.org[(side + 0) * 64 + owner], p.org[(side + 1) * 64 + owner],
                         p.org[(side + 2) * 64 + owner]},
                      V3{p.dx[e], p.dy[e], p.dz[e]}, 0.0f, p.far[e]);
            }
        }
        next += uint32_t(__popcll(want));
    };
    auto pop = [&]() -> uint32_t {  // the nearest stacked entry still worth visiting, or Sentinel
        while (sp > 0) {
            --sp;
            uint2 s;
            if (cached) {
                s = make_uint2(topRef, topDist);
                cached = false;
            } else {
                s = st.get(sp);
            }
            if (__uint_as_float(s.y) <= hi) return s.x;
        }
        return Sentinel;
    };
    auto visitNode = [&](uint32_t node) -> uint32_t {  // stacks the far children, returns the nearest (or a pop)
        ++k.nodes;
        unsigned long long key[8];
        nodeChildren(P.nodes[node], br, lo, hi, key);
        unsigned long long m = key[0];
#pragma unroll
        for (int i = 1; i < 8; ++i) m = key[i] < m ? key[i] : m;
#pragma unroll
        for (int i = 7; i >= 0; --i)
            if (key[i] != ~0ull && key[i] != m) {
                if (cached) st.put(sp - 1, make_uint2(topRef, topDist));
                topRef = uint32_t(key[i]);
                topDist = uint32_t(key[i] >> 32);
                cached = true;
                ++sp;
            }
        return m != ~0ull ? uint32_t(m) : pop();
    };
    auto visitLeaf = [&](uint32_t leafRef) {  // a pool ray stops at its first hit
        const uint32_t firstTri = (leafRef & 0x7fffffffu) >> 3, count = (leafRef & 7u) + 1;
        // A leaf's first three triangles (the builder's leaves hold at most three unless the SAH prefers more) read
        // before any is tested, so their loads wait once, not one after another.
        float pre[3][9];
        const float* v0 = P.triVerts + size_t(firstTri) * 9;
#pragma unroll
        for (int q = 0; q < 9; ++q) pre[0][q] = v0[q];
        if (count > 1)
#pragma unroll
            for (int q = 0; q < 9; ++q) pre[1][q] = v0[9 + q];
        if (count > 2)
#pragma unroll
            for (int q = 0; q < 9; ++q) pre[2][q] = v0[18 + q];
        for (uint32_t i = 0; i < count; ++i) {
            ++k.triangles;
            float t;
            float v[9];
            if (i < 3) {
#pragma unroll
                for (int q = 0; q < 9; ++q) v[q] = i == 0 ? pre[0][q] : i == 1 ? pre[1][q] : pre[2][q];
            } else {
                const float* g = P.triVerts + size_t(firstTri + i) * 9;
#pragma unroll
                for (int q = 0; q < 9; ++q) v[q] = g[q];
            }
            const V3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
            if (!woopHit(w, a, b, c, t) || culled(P, firstTri + i, a, b, c, dir)) continue;
            if (pooled) {
                occluded = true;
                return;
            }
            if (best == ~0u || t < bestT) {
                best = firstTri + i;
                bestT = t;
                hi = slackAbove(t);
            } else if (t == bestT) {  // a tie in distance: the least (geom, prim) wins (closestHit's 64-bit select)
                const TriRecord &r = P.tris[firstTri + i], &qr = P.tris[best];
                const unsigned long long rk = (unsigned long long)r.geom << 32 | r.prim,
                                         qk = (unsigned long long)qr.geom << 32 | qr.prim;
                best = rk < qk ? firstTri + i : best;
            }
        }
    };
    if (own) {
        ++k.traces;
        start(ray.origin, ray.direction, ray.nearT, ray.farT);
    }
    take();
    for (;;) {
        if (active) {
            while (ref < Sentinel) {  // an inner node
                ref = visitNode(ref);
                if ((ref & LeafBit) && leaf == 0) {
                    leaf = ref;
                    ref = pop();
                }
                if (__popcll(__ballot(leaf == 0)) <= LeafWait) break;
            }
            while (leaf || (ref & LeafBit)) {
                if (!leaf) {
                    leaf = ref;
                    ref = pop();
                }
                visitLeaf(leaf);
                leaf = 0;
                if (occluded) ref = Sentinel;
            }
            if (ref == Sentinel) {  // the ray is done: its answer
                if (pooled) p.flag[e] = occluded ? 0u : 2u;
                else ownBest = best;
                active = false;
            }
        }
        take();
        if (!__ballot(active)) break;
    }
    waveSync();
    return ownBest;
}

// The forward surface estimator with the shadow rays pooled, software-pipelined: an iteration traces each lane's
// next ray together with the shadow rays the lanes' last vertices queued (traceWave), adds those rays' light to the
// samples that queued them, each lane's in the order its vertex made them, then shades the new hits, which queue
// the next iteration's shadow rays. A sample whose last vertex queued shadow rays is held until they are traced and
// its light added; its lane meanwhile starts its next sample. Every addition to a sample's radiance comes in the
// order the one-vertex-at-a-time loop makes it, so the film is bit for bit the same. Every lane stays in the loop
// until the wave's samples are all done. Only a kernel whose one argument is the RenderParams calls it: the loop
// reads them at the start of the kernel's argument segment.
constexpr uint32_t PoolRegen = 40;
template <class Feed>
__device__ void runSamplesPooled(const RenderParams& P, Feed& feed, Counters& k, TravStack& st, PoolQueue& q) {
    PathState<false, false> S;
    S.top = 0;
    uint32_t pixel = 0, index = 0, heldPixel = 0, heldIndex = 0;
    V3 heldRadiance{0, 0, 0};
    float heldAlpha = 0.0f;
    bool live = false, done = false, held = false;
    for (;;) {
        const RenderParams& PL = argParams();
        if (live && S.top == 0) {  // the sample's last vertex is shaded: it ends, or waits for its shadow rays
            if (q.n) {
                held = true;
                heldPixel = pixel;
                heldIndex = index;
                heldRadiance = S.radiance;
                heldAlpha = S.alpha;
            } else {
                feed.done(pixel, index, S.radiance, S.alpha, 0u);
            }
            live = false;
        }
        const unsigned long long idleLanes = __ballot(!live && !done), openLanes = __ballot(!done);
        if (!__ballot(!done || held)) break;  // every lane's feed is empty and its last sample done
        // finished lanes take new samples together once PoolRegen of them are idle (or all are): idle lanes still
        // trace the pool's rays here, so larger groups pay (the 1080p icospheres frame: 828 Mpaths/s at 40, 820 at
        // 32, 825 at 44, 817 at 48)
        const bool claim = !live && !done && (__popcll(idleLanes) >= PoolRegen || idleLanes == openLanes);
        if (claim) {
            if (!feed.next(pixel, index)) done = true;
            else if (!startSample<false, false>(PL, pixel, index, S)) {
                feed.failed();
                S.top = 0;
            } else {
                live = true;
            }
        }
        // the lane's next ray: its sample's last pending one (popped once traced, so that only the ray is held
        // through the trace)
        auto rayOf = [&](const Pending& r) {
            const bool fromCamera = (r.depths & 0xffu) == 0;
            return Ray{r.origin, r.dir, fromCamera ? S.cameraNear : 0.0f, fromCamera ? S.cameraFar : MaxFloat};
        };
        const uint32_t tri = traceWave(PL, q, live, live ? rayOf(S.stack[S.top - 1]) : Ray{}, k, st);
        {  // the pool rays' light, into the sample that queued them, in the order its vertex queued them
            V3 r = held ? heldRadiance : S.radiance;
            const ShadowPool& pool = *q.pool;
            for (uint32_t j = 0; j < q.n; ++j) {
                const uint32_t e = j * 64 + q.lane;
                if (pool.flag[e] == 2u) r = r + V3{pool.vr[e], pool.vg[e], pool.vb[e]};
            }
            q.n = 0;
            if (held) {
                feed.done(heldPixel, heldIndex, r, heldAlpha, 0u);
                held = false;
            } else {
                S.radiance = r;
            }
        }
        if (live) {
            const Pending pv = S.stack[--S.top];
            const Ray ray = rayOf(pv);
            Hit hit;
            if (!hitRecord(PL, ray, tri, hit) || !shadeHit<false, false>(PL, pixel, index, S, q, pv, ray, hit, k, st, nullptr) ||
                q.overflow) {
                feed.failed();  // an invalid value, or more shadow rays than the pool holds (the host prevents it)
                S.top = 0;
                live = false;
                q.n = 0;
                q.overflow = false;
            }
        }
    }
}

// The prototype estimator (async_integrator.h:61-126; mixed::Config::paths == 0): the camera hit, its emission, and
// cosine-sampled Lambert bounces (surface::lambert, surface_core.h:158-165) toward a constant sky, the
// surface_limit-th interaction looking along its sampled direction for the sky. Its albedo is the material's Lambert
// lobe (albedo x diffuse weight, times its map).
__device__ int integrateDirect(const RenderParams& P, uint32_t pixel, uint32_t index, double rgba[4], Counters& k,
                               TravStack& st) {
    const x2_config& c = P.cfg;
    if (c.surface_limit == 0 || c.surface_limit > 8) return 2;
    if (P.failPixel == pixel + 1) return 2;  // injected failure, from the host
    const Identity id{pixel, index, c.seed, 0};
    float x, y;
    {
        float r0, r1;
        drawPair(id, 0, c.sampling, r0, r1);
        const float w = c.filter_width;
        if (!finite1(w) || w <= 0) return 2;
        x = (r0 - .5f) * w + .5f;
        y = (r1 - .5f) * w + .5f;
        if (!finite1(x) || !finite1(y)) return 2;
    }
    Ray ray;
    if (!cameraRay(c, float(pixel % c.width) + x, float(pixel / c.width) + y, ray)) return 2;
    // GpuBackend::material (async_gpu_device.h:356-383): the map through the default binding (anisotropic, cross
    // terms in order) at the footprint of the hit ray's differentials; only the camera ray has them, so the bounces
    // read a zero footprint (level 0).
    V3 camDx{0, 0, 0}, camDy{0, 0, 0};
    if (P.textures && !cameraDifferentials(c, float(pixel % c.width) + x, float(pixel / c.width) + y, camDx, camDy))
        return 2;
    double radiance[3] = {0, 0, 0}, throughput[3] = {1, 1, 1};
    uint32_t depth = 0, primaryHit = 0;
    const V3 sky = ld3(c.sky);
    for (;;) {
        Hit hit;
        if (!traceHit(P, ray, hit, k, st)) return 2;
        if (hit.geom == ~0u) {
            if (depth) {
                radiance[0] += throughput[0] * sky.x;
                radiance[1] += throughput[1] * sky.y;
                radiance[2] += throughput[2] * sky.z;
            }
            break;
        }
        if (!depth) primaryHit = 1;
        const TriRecord& tr = P.tris[hit.tri];
        const x2_material& m = P.materials[tr.material];
        V3 albedo = ld3(m.albedo) * m.diffuse_weight;
        if (m.albedo_map) {
            float uv[6] = {0, 0, 1, 0, 1, 1};
            if (P.triUV)
                for (int q = 0; q < 6; ++q) uv[q] = P.triUV[size_t(hit.tri) * 6 + q];
            const float w = 1.0f - hit.u - hit.v;
            HitDiff hd;
            hd.dsdx = hd.dtdx = hd.dsdy = hd.dtdy = 0.0f;
            if (!depth)
                hitTexDerivatives(V3{tr.ax, tr.ay, tr.az}, V3{tr.bx, tr.by, tr.bz}, V3{tr.cx, tr.cy, tr.cz}, uv, ray, camDx,
                                  camDy, c.spp, hd);
            const Footprint f{w * uv[0] + hit.u * uv[2] + hit.v * uv[4], w * uv[1] + hit.u * uv[3] + hit.v * uv[5],
                              hd.dsdx, hd.dtdx, hd.dsdy, hd.dtdy};
            const TexDesc& td = P.textures[m.albedo_map - 1];
            if (td.paged) return 2;  // the prototype estimator reads resident textures
            V3 color;
            if (mapLookup(P, td, 0, f, false, color) != TapOk) return 2;
            albedo = albedo * color;
        }
        const V3 emission = ld3(m.emission), N = hit.normal;
        const uint32_t dimension = CameraDimensions + 2 * depth;
        const float u = draw(id, dimension, c.sampling), v = draw(id, dimension + 1, c.sampling);
        if (!finite3(albedo) || albedo.x < 0 || albedo.y < 0 || albedo.z < 0 || !finite3(N) ||
            fabsf(dot(N, N) - 1) > 2e-5f || !finite1(u) || !finite1(v) || u < 0 || u > 1 || v < 0 || v > 1)
            return 2;
        float sn, cs;
        sincosPoly(v * (2 * Pi), sn, cs);
        const float sr = sqrtf(1 - u);
        const V3 wi = localToGlobal(frameOf(N), V3{sr * cs, sr * sn, sqrtf(u)});
        const float cosine = fmaxf(dot(N, wi), 0.0f), pdf = cosine * OneOverPi;
        const V3 f = (albedo * OneOverPi) * cosine;
        radiance[0] += throughput[0] * emission.x;
        radiance[1] += throughput[1] * emission.y;
        radiance[2] += throughput[2] * emission.z;
        if (pdf <= 0) break;
        throughput[0] *= f.x / pdf;
        throughput[1] *= f.y / pdf;
        throughput[2] *= f.z / pdf;
        if (++depth == c.surface_limit) {
            if (visible(P, hit.origin, wi, 0.0f, MaxFloat, k, st)) {
                radiance[0] += throughput[0] * sky.x;
                radiance[1] += throughput[1] * sky.y;
                radiance[2] += throughput[2] * sky.z;
            }
            break;
        }
        ray = Ray{hit.origin, wi, 0.0f, MaxFloat};
    }
    rgba[0] = radiance[0];
    rgba[1] = radiance[1];
    rgba[2] = radiance[2];
    rgba[3] = primaryHit;
    return 0;
}

__device__ __forceinline__ uint32_t tilePixel(uint32_t tile, uint32_t position, const x2_config& c) {
    const uint32_t tilesX = (c.width + 7) / 8;
    const uint32_t x = (tile % tilesX) * 8 + position % 8, y = (tile / tilesX) * 8 + position / 8;
    if (!(x < c.width && y < c.height)) return ~0u;
    if (int(x) < c.viewport[0] || int(y) < c.viewport[1] || int(x) >= c.viewport[2] || int(y) >= c.viewport[3])
        return ~0u;
    return y * c.width + x;
}

__device__ __forceinline__ unsigned long long waveSum(uint32_t v) {
    unsigned long long s = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// Launch shape knobs (measured on MI355X on the 1080p 64-spp icospheres frame):
// persistent lanes claiming samples one at a time, and an occupancy target of 3 waves per SIMD (168 VGPRs) were the
// fastest of the variants tried; an LDS traversal stack was slower than private memory at that occupancy.
#define X2_OCCUPANCY __attribute__((amdgpu_waves_per_eu(3)))

// One sample per item of the launch (async_gpu.cu:387-473 Feed phase 1): item = chunk * itemsPerChunk + s * 64 +
// position; the chunk comes from the mailbox ring. The forward render adds the sample to the film; the adjoint
// replays it and scatters its gradients.
// An item's sample (false: none — outside the image or viewport, past the chunk's count, or, in an adaptive frame,
// a pixel its pass leaves out). An adaptive chunk's first is its pass and offset; the sample index starts from the
// pixel's own base, so a pixel's indices stay contiguous (async_gpu.cu:440-446).
__device__ __forceinline__ bool decodeItem(const RenderParams& P, unsigned long long item, uint32_t& pixel,
                                           uint32_t& index) {
    // a power-of-two chunk size and ring (the usual launch shape) by shift and mask, not a 64-bit division
    const unsigned long long ipc = P.itemsPerChunk, cap = P.ringCapacity;
    const unsigned long long chunkIndex = ipc && (ipc & (ipc - 1)) == 0 ? item >> __builtin_ctzll(ipc) : item / ipc;
    const uint32_t offset = uint32_t(item - chunkIndex * ipc);
    const x2_chunk chunk = P.ring[cap && (cap & (cap - 1)) == 0 ? chunkIndex & (cap - 1) : chunkIndex % cap];
    const uint32_t s = offset / TilePixels, position = offset % TilePixels;
    if (s >= chunk.count) return false;
    pixel = tilePixel(chunk.tile, position, P.cfg);
    if (pixel == ~0u) return false;
    if (P.adaptive) {
        const uint32_t slot = (chunk.first >> PassShift) % PassSlots;
        if (!(P.passMask[slot][chunk.tile] >> position & 1ull)) return false;
        index = P.passBase[slot][pixel] + (chunk.first & ((1u << PassShift) - 1u)) + s;
    } else {
        index = chunk.first + s;
    }
    return true;
}
// A finished sample into the film (async_gpu.cu:331-345): rgb + alpha, the count, and in an adaptive frame the
// odd-index sums (MoonRay's beauty odd, the error's second half-image).
__device__ __forceinline__ void filmAdd(const RenderParams& P, uint32_t pixel, uint32_t index, double r, double g,
                                        double b, double a) {
    double* f = P.sums + size_t(pixel) * 4;
    atomicAdd(f + 0, r);
    atomicAdd(f + 1, g);
    atomicAdd(f + 2, b);
    atomicAdd(f + 3, a);
    atomicAdd(P.counts + pixel, 1u);
    if (P.odd && (index & 1u)) {
        double* o = P.odd + size_t(pixel) * 4;
        atomicAdd(o + 0, r);
        atomicAdd(o + 1, g);
        atomicAdd(o + 2, b);
        atomicAdd(o + 3, a);
    }
}

template <bool Adj, bool Vol, bool Direct = false>
__device__ __forceinline__ void renderItem(const RenderParams& P, unsigned long long item, Counters& k, TravStack& st,
                                           uint32_t& errors, float* gr) {
    uint32_t pixel, index;
    if (!decodeItem(P, item, pixel, index)) return;
    double rgba[4];
    int status;
    if constexpr (Direct) status = integrateDirect(P, pixel, index, rgba, k, st);
    else status = integrate<Adj, Vol>(P, pixel, index, rgba, k, st, gr);
    if (status == 0) {
        if constexpr (!Adj) filmAdd(P, pixel, index, rgba[0], rgba[1], rgba[2], rgba[3]);
    } else {
        ++errors;
    }
}

// A persistent lane's samples: it claims the launch's items one at a time (the reference's Feed loop), skipping
// positions outside the image or the viewport, and adds each finished sample to the film (the adjoint adds none).
// The launch's next sample: claims items until one lies in the image and the viewport (false: none left).
// The device pass chain's render: its items, fixed by the compaction before the render (0: the chain has stopped).
__device__ __forceinline__ unsigned long long chainItems(const RenderParams& P) {
    if (!P.chain || P.chain->stop) return 0;
    return (unsigned long long)P.chain->activeTiles * P.chainSize * TilePixels;
}
// With Pg, *from receives the parked slot a sample came from (~0u: a chunk's item).
constexpr uint32_t FromRefine = 0x80000000u;  // a claimed item's origin: a provisional sample rendered again (k)
template <bool Pg = false>
__device__ __forceinline__ bool claimItem(const RenderParams& P, uint32_t& pixel, uint32_t& index,
                                          unsigned long long chainEnd = 0, uint32_t* from = nullptr) {
    if constexpr (Pg) *from = ~0u;
    for (;;) {
        if (P.chain) {  // the device pass chain's render: its active tiles' pixels, chainSize samples each
            const unsigned long long perTile = (unsigned long long)P.chainSize * TilePixels;
            const unsigned long long item = atomicAdd(&P.chain->claimed, 1ull);
            if (item >= chainEnd) return false;
            const uint32_t t = uint32_t(item / perTile), rest = uint32_t(item % perTile);
            const uint32_t position = rest % TilePixels;
            if (!(P.chainMasks[t] >> position & 1ull)) continue;
            pixel = tilePixel(P.chainTiles[t], position, P.cfg);
            if (pixel == ~0u) continue;
            index = P.chainBase[pixel] + rest / TilePixels;
            return true;
        }
        unsigned long long k = atomicAdd(&P.state->claimed, 1ull);
        if constexpr (Pg) {
            if (k < P.refineInCount) {  // provisional samples rendered again, first (async_gpu.cu:394-405)
                const x2_refine& r = P.refineIn[k];
                pixel = r.pixel;
                index = r.index;
                *from = FromRefine | uint32_t(k);
                return true;
            }
            k -= P.refineInCount;
            if (k < P.parkedCount) {  // the parked samples this launch runs again, ahead of its chunks
                const uint32_t at = uint32_t((P.parkedBegin + k) % P.parkCap);
                const x2_parked r = P.parked[at];
                pixel = r.pixel;
                index = r.index;
                *from = at;
                return true;
            }
            k -= P.parkedCount;
        }
        const unsigned long long item = P.itemBegin + k;
        if (item >= P.itemEnd) return false;
        if (decodeItem(P, item, pixel, index)) return true;
    }
}
// A parked sample onto the ring (async_gpu.cu:350-357): the host sizes launches so the ring never overwrites a record
// not yet run again; past that bound a sample is counted as lost (a failure of the frame), never written. A
// provisional sample's record keeps its value (v: rgb + alpha) beside it, for the provisional film to give back.
// The launch's own parked records [parkedBegin, parkedBegin + parkedCount) are below the bound, so none is
// overwritten while the launch reads it.
// A provisional sample's record for the host (host refines; async_gpu.cu:276-279): past the ring's room it is
// counted, never written (the frame then fails).
__device__ __forceinline__ void refineRecord(const RenderParams& P, uint32_t pixel, uint32_t index, uint32_t tile,
                                             const float* v) {
    const unsigned long long at = atomicAdd(&P.paging->refineTail, 1ull);
    if (at >= P.refineHead + P.refineCap) {
        atomicAdd(&P.paging->refineOverflow, 1u);
        return;
    }
    x2_refine& r = P.refineRing[at % P.refineCap];
    r.pixel = pixel;
    r.index = index;
    r.pass = 0u;
    r.tile = tile;
    *reinterpret_cast<float4*>(r.rgba) = float4{v[0], v[1], v[2], v[3]};
}
__device__ __forceinline__ void parkSample(const RenderParams& P, uint32_t pixel, uint32_t index,
                                           const float* v = nullptr) {
    const unsigned long long at = atomicAdd(&P.paging->parkTail, 1ull);
    if (at >= P.parkedBegin + P.parkCap) {
        atomicAdd(&P.paging->parkOverflow, 1u);
        return;
    }
    const uint32_t slot = uint32_t(at % P.parkCap);
    P.parked[slot] = x2_parked{pixel, index, 0u, 0u};
    if (v) *reinterpret_cast<float4*>(P.parkedRgba + size_t(slot) * 4) = float4{v[0], v[1], v[2], v[3]};
}
template <bool Adj, bool Pg = false>
struct LaneFeed {
    static constexpr uint32_t Regen = 32;  // finished lanes of a wave that take new samples together
    static constexpr bool Paged = Pg;                  // texture tiles on demand: parked and provisional samples
    const RenderParams& P;
    uint32_t& errors;
    unsigned long long chainEnd = chainItems(P);  // read once: the compaction wrote it before this render
    uint32_t from = ~0u;  // where the lane's sample came from (Pg): a parked slot, FromRefine | k, or ~0u: a chunk's item
    __device__ bool next(uint32_t& pixel, uint32_t& index) { return claimItem<Pg>(P, pixel, index, chainEnd, &from); }
    // A sample that read a coarse level (partial shading) goes into the provisional film, in place of its old value
    // when it was itself a provisional sample run again, and parks, or with host refines goes to the host as a record;
    // one run again in full takes its old value out (async_gpu.cu:282-295), so the provisional film holds exactly the
    // samples still parked or held by the host.
    __device__ void done(uint32_t pixel, uint32_t index, V3 r, float a, uint32_t flags, uint32_t missTile = 0) {
        if constexpr (!Adj) {
            if constexpr (Pg) {
                const bool refined = from != ~0u && (from & FromRefine);
                const bool old = refined || (P.partial && !P.hostRefines && from != ~0u);
                float4 was{0.0f, 0.0f, 0.0f, 0.0f};
                if (refined) was = *reinterpret_cast<const float4*>(P.refineIn[from & ~FromRefine].rgba);
                else if (old) was = *reinterpret_cast<const float4*>(P.parkedRgba + size_t(from) * 4);
                double* f = P.provSums + size_t(pixel) * 4;
                if (flags & SampleProvisional) {
                    const float v[4] = {r.x, r.y, r.z, a};
                    atomicAdd(f + 0, double(v[0]) - double(was.x));
                    atomicAdd(f + 1, double(v[1]) - double(was.y));
                    atomicAdd(f + 2, double(v[2]) - double(was.z));
                    atomicAdd(f + 3, double(v[3]) - double(was.w));
                    if (!old) atomicAdd(P.provCounts + pixel, 1u);
                    atomicAdd(&P.paging->provisional, 1u);
                    if (P.hostRefines) refineRecord(P, pixel, index, missTile, v);
                    else parkSample(P, pixel, index, v);
                    return;
                }
                if (flags) {  // parked on a missing tile: it runs again once the tile is resident
                    if (refined) {  // a provisional sample that read no coarser level either: it stays as it was
                        const float v[4] = {was.x, was.y, was.z, was.w};
                        refineRecord(P, pixel, index, missTile, v);
                        return;
                    }
                    parkSample(P, pixel, index);
                    return;
                }
                if (old) {
                    atomicAdd(f + 0, -double(was.x));
                    atomicAdd(f + 1, -double(was.y));
                    atomicAdd(f + 2, -double(was.z));
                    atomicAdd(f + 3, -double(was.w));
                    atomicSub(P.provCounts + pixel, 1u);
                }
            }
            filmAdd(P, pixel, index, r.x, r.y, r.z, a);
        }
    }
    __device__ void failed() { ++errors; }
};

// A lane per item (the 64 lanes of a wave are one 8x8 tile at one sample index, so camera rays start coherent), or
// persistent lanes claiming items one at a time (the reference's Feed loop): with
// Regen a lane claims its next item when its sample's last vertex is shaded (runSamples over LaneFeed), without it
// when the whole sample is (renderItem). Every lane reaches the end so the wave's counter sums see all 64 lanes.
// Only kernels whose one argument is the RenderParams call it (the loops read them through argParams).
template <bool Adj, bool Vol, bool Direct = false, bool Regen = false, bool Paged = false, bool Tx = false>
__device__ __forceinline__ void runLaunch(const RenderParams& P) {
    TravStack st;
    Counters k;
    uint32_t errors = 0;
    if constexpr (Vol) tapCount() = 0;
    float gr[Adj ? RoughnessSlots : 1];
    for (uint32_t i = 0; i < (Adj ? RoughnessSlots : 1); ++i) gr[i] = 0.0f;
    if constexpr (Regen) {
        LaneFeed<Adj, Paged> feed{P, errors};
        // The forward textured renders hold P across the loop: read afresh they measured 7 % slower (the adjoint,
        // the medium and the untextured renders 10-28 % faster).
        runSamples<Adj, Vol, Tx, Adj || !Tx>(P, feed, k, st, gr);
    } else {
        for (;;) {
            const unsigned long long item = P.itemBegin + atomicAdd(&P.state->claimed, 1ull);
            if (item >= P.itemEnd) break;
            renderItem<Adj, Vol, Direct>(argParams(), item, k, st, errors, gr);
        }
    }
    if (errors) atomicAdd(&P.state->errors, errors);
    if constexpr (Adj) {
        for (uint32_t m = 0; m < RoughnessSlots; ++m) {  // one atomic per wave and material
            float v = gr[m];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if ((threadIdx.x & 63u) == 0 && v != 0.0f) atomicAdd(P.gradRoughness + m, v);
        }
    }
    if (P.counters) {
        const unsigned long long t = waveSum(k.traces), sh = waveSum(k.shadows), n = waveSum(k.nodes),
                                 tr = waveSum(k.triangles);
        if ((threadIdx.x & 63u) == 0) {
            atomicAdd(&P.state->traces, t);
            atomicAdd(&P.state->shadows, sh);
            atomicAdd(&P.state->nodes, n);
            atomicAdd(&P.state->triangles, tr);
        }
        if constexpr (Vol) {
            const unsigned long long taps = waveSum(tapCount());
            if ((threadIdx.x & 63u) == 0) atomicAdd(&P.state->voxels, taps);
        }
    }
}

// The surface estimator and its adjoint, with path regeneration (a lane takes its next sample as soon as its last
// one is done). Measured against a lane per sample on the same frames: 1.01x on 32 icospheres at 1080p64, 1.55x
// forward and 1.52x adjoint on the textured Cornell box (RR off, long uneven paths).
__global__ void __launch_bounds__(256) X2_OCCUPANCY renderKernel(const RenderParams P) {
    runLaunch<false, false, false, true>(P);
}
// The same render with texture tiles on demand: samples that meet a missing tile park (or, with partial shading,
// finish provisionally) and the launch runs the parked samples of earlier launches ahead of its chunks. Its own
// instantiation, so the render above carries none of it.
__global__ void __launch_bounds__(256) X2_OCCUPANCY renderPagedKernel(const RenderParams P) {
    runLaunch<false, false, false, true, true, true>(P);
}
// The render of a scene with textures: every pending ray carries its differentials, for the maps' footprints. Its
// own instantiation, so the render of untextured scenes carries none of it.
__global__ void __launch_bounds__(256) X2_OCCUPANCY renderTexKernel(const RenderParams P) {
    runLaunch<false, false, false, true, false, true>(P);
}
// The forward surface estimator with the wave's shadow rays pooled (the frames whose vertices make at most PoolRays
// shadow rays: light_count x light_samples + 2 x bsdf_samples <= PoolRays).
__global__ void __launch_bounds__(256) X2_OCCUPANCY renderPoolKernel(const RenderParams P) {
    __shared__ ShadowPool pools[4];
    TravStack st;
    Counters k;
    uint32_t errors = 0;
    LaneFeed<false, false> feed{P, errors};
    PoolQueue q{&pools[threadIdx.x >> 6], threadIdx.x & 63u, 0u, false, V3{0, 0, 0}};
    runSamplesPooled(P, feed, k, st, q);
    if (errors) atomicAdd(&P.state->errors, errors);
    if (P.counters) {
        const unsigned long long t = waveSum(k.traces), sh = waveSum(k.shadows), n = waveSum(k.nodes),
                                 tr = waveSum(k.triangles);
        if ((threadIdx.x & 63u) == 0) {
            atomicAdd(&P.state->traces, t);
            atomicAdd(&P.state->shadows, sh);
            atomicAdd(&P.state->nodes, n);
            atomicAdd(&P.state->triangles, tr);
        }
    }
}
// The adjoint at 2 waves a SIMD: its sinks and open-subtree records need 248 registers, which 3 waves could not give
// without 141 of them spilled (the textured Cornell box's adjoint frame: 478 ms against 658 ms).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) adjointKernel(const RenderParams P) {
    adjointLdsClear();
    runLaunch<true, false, false, true, false, true>(P);
    adjointLdsFlush(P);
}
// ---------------------------------------------------------------------------------------------------------------
// The wavefront form of the surface estimator: the same samples, vertex for vertex, with the traversal taken out of
// the shading kernel so that each kind of work runs at its own occupancy. An iteration is two launches.
// waveShadeKernel, one thread a path slot: adds the unoccluded contributions of the slot's last vertex's shadow rays,
// shades the vertex whose closest hit the last trace pass found (its shadow rays become queue entries, each with the
// contribution it adds), then moves the slot's next pending ray into the closest-hit queue — finishing the sample
// into the film and starting the launch's next item when the stack is empty. waveTraceKernel traces both queues.
// Contributions are added in the order the per-lane loop adds them, so every sample's radiance is bit for bit the
// per-lane loop's. waveDrainKernel ends a launch: it finishes whatever the iterations left (slots still busy and
// unclaimed items) with the per-lane loop, so a launch is complete whatever the number of iterations.

// A slot's pending records in the slot arrays: level i at base[i * slots] (base: the slot's level 0).
struct LevelStack {
    Pending* base;
    uint32_t slots;
    __device__ __forceinline__ Pending& operator[](uint32_t i) const { return base[size_t(i) * slots]; }
};
// A path slot's sample while a pass works on it: loaded from the slot arrays, stored back at the pass's end.
struct WaveSlot {
    LevelStack stack;
    uint32_t top;
    V3 radiance;
    float alpha, cameraNear, cameraFar;
    V3 delta;  // the adjoint's (unused here)
    Sink sinks[1];
    Record recs[1];
    uint32_t nrec;
};
__device__ __forceinline__ void loadSlot(const WaveParams& W, uint32_t slot, WaveSlot& S, uint32_t& pixel,
                                         uint32_t& index) {
    const uint2 id = reinterpret_cast<const uint2*>(W.sample)[slot];
    pixel = id.x;
    index = id.y;
    S.top = W.top[slot];
    const float4* a = reinterpret_cast<const float4*>(W.accum) + size_t(slot) * 2;
    const float4 a0 = a[0], a1 = a[1];
    S.radiance = V3{a0.x, a0.y, a0.z};
    S.alpha = a0.w;
    S.cameraNear = a1.x;
    S.cameraFar = a1.y;
}
__device__ __forceinline__ void storeSlot(const WaveParams& W, uint32_t slot, const WaveSlot& S, uint32_t pixel,
                                          uint32_t index) {
    reinterpret_cast<uint2*>(W.sample)[slot] = make_uint2(pixel, index);
    W.top[slot] = S.top;
    float4* a = reinterpret_cast<float4*>(W.accum) + size_t(slot) * 2;
    a[0] = make_float4(S.radiance.x, S.radiance.y, S.radiance.z, S.alpha);
    a[1] = make_float4(S.cameraNear, S.cameraFar, 0.0f, 0.0f);
}
// Shadow rays as queue entries at the slot's positions: the ray and the contribution it adds when unoccluded. A
// vertex makes at most light_count x light_samples + 2 x bsdf_samples of them (W.perVertex).
struct WaveQueue {
    const WaveParams& W;
    uint32_t parity, slot, n;
    __device__ __forceinline__ void vertex(V3, V3) {}
    __device__ __forceinline__ bool shadow(const RenderParams&, V3 o, V3 d, float nearT, float farT, V3 value, Counters&,
                                           TravStack&) {
        if (n >= W.perVertex) return false;  // beyond the bound the queues are sized for: never
        const size_t e = size_t(slot) * W.perVertex + n++;
        float4* r = reinterpret_cast<float4*>(W.shadowRays[parity]) + e * 2;
        r[0] = make_float4(o.x, o.y, o.z, nearT);
        r[1] = make_float4(d.x, d.y, d.z, farT);
        reinterpret_cast<float4*>(W.shadowValue[parity])[e] = make_float4(value.x, value.y, value.z, 0.0f);
        return false;
    }
};
// The launch still has unclaimed items.
__device__ __forceinline__ bool feedOpen(const RenderParams& P) {
    const unsigned long long claimed = *reinterpret_cast<volatile unsigned long long*>(&P.state->claimed);
    return P.itemBegin + claimed < P.itemEnd;
}
// An exclusive prefix sum over the wave's 64 lanes (every lane active); total = the wave's sum.
__device__ __forceinline__ uint32_t waveExclusive(uint32_t v, uint32_t& total) {
    uint32_t x = v;
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= uint32_t(o)) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// The slot's traced vertex: its closest hit (found by the trace pass) reconstructed and shaded with Tr answering
// the shadow rays. False: an invalid value (Error).
template <class Tr>
__device__ __forceinline__ bool shadeTraced(const RenderParams& P, const WaveParams& W, uint32_t slot, uint32_t pixel,
                                            uint32_t index, WaveSlot& S, Tr& tracer, Counters& k, TravStack& st) {
    const Pending pv = static_cast<const Pending*>(W.current)[slot];
    const bool fromCamera = (pv.depths & 0xffu) == 0;
    const Ray ray{pv.origin, pv.dir, fromCamera ? S.cameraNear : 0.0f, fromCamera ? S.cameraFar : MaxFloat};
    Hit hit;
    hit.geom = ~0u;
    const uint32_t tri = W.hitTri[slot];
    if (tri != ~0u) {
        const TriRecord& r = P.tris[tri];
        hit.geom = r.geom;
        hit.tri = tri;
        if (!geometryHit(hit, ray, V3{r.ax, r.ay, r.az}, V3{r.bx, r.by, r.bz}, V3{r.cx, r.cy, r.cz})) return false;
    }
    float gr[1] = {0.0f};
    return shadeHit<false, false>(P, pixel, index, S, tracer, pv, ray, hit, k, st, gr);
}
// Adds the unoccluded contributions of the slot's queued shadow rays (queue parity q), in the order they were made.
__device__ __forceinline__ void resolveShadows(const WaveParams& W, uint32_t slot, uint32_t q, WaveSlot& S) {
    const uint32_t n = W.shadowCount[slot];
    for (uint32_t j = 0; j < n; ++j) {
        const size_t e = size_t(slot) * W.perVertex + j;
        if (W.shadowVis[q][e]) {
            const float4 v = reinterpret_cast<const float4*>(W.shadowValue[q])[e];
            S.radiance = S.radiance + V3{v.x, v.y, v.z};
        }
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
waveShadeKernel(const RenderParams P, const WaveParams W, uint32_t p) {
    const uint32_t slot = blockIdx.x * 256u + threadIdx.x;  // slots come in whole workgroups
    uint32_t errors = 0, queuedRay = 0, queuedShadows = 0;
    {
        uint32_t fl = W.flags[slot];
        if (fl || feedOpen(P)) {
            WaveSlot S;
            S.stack = LevelStack{static_cast<Pending*>(W.stack) + slot, W.slots};
            S.top = 0;
            uint32_t pixel = 0, index = 0;
            if (fl & WaveLive) loadSlot(W, slot, S, pixel, index);
            if (fl & WaveShadows) {
                resolveShadows(W, slot, p ^ 1u, S);
                fl &= ~WaveShadows;
            }
            if (fl & WaveTraced) {
                fl &= ~WaveTraced;
                Counters k;
                TravStack st;
                WaveQueue queue{W, p, slot, 0};
                if (!shadeTraced(P, W, slot, pixel, index, S, queue, k, st)) {
                    ++errors;
                    fl = 0;
                    S.top = 0;
                } else if (queue.n) {
                    W.shadowCount[slot] = queue.n;
                    queuedShadows = queue.n;
                    fl |= WaveShadows;
                }
            }
            for (;;) {
                if (fl & WaveLive) {
                    if (S.top) {  // the next pending ray into the closest-hit queue
                        const Pending pv = S.stack[--S.top];
                        static_cast<Pending*>(W.current)[slot] = pv;
                        const bool fromCamera = (pv.depths & 0xffu) == 0;
                        float4* ray = reinterpret_cast<float4*>(W.rays[p]) + size_t(slot) * 2;
                        ray[0] = make_float4(pv.origin.x, pv.origin.y, pv.origin.z, fromCamera ? S.cameraNear : 0.0f);
                        ray[1] = make_float4(pv.dir.x, pv.dir.y, pv.dir.z, fromCamera ? S.cameraFar : MaxFloat);
                        queuedRay = 1;
                        fl |= WaveTraced;
                        break;
                    }
                    if (fl & WaveShadows) break;  // the last vertex's shadow rays are still being traced
                    filmAdd(P, pixel, index, S.radiance.x, S.radiance.y, S.radiance.z, S.alpha);
                    fl = 0;
                }
                if (!claimItem(P, pixel, index)) break;
                if (!startSample<false, false>(P, pixel, index, S)) {
                    ++errors;
                    continue;
                }
                fl = WaveLive;
            }
            W.flags[slot] = fl;
            if (fl) storeSlot(W, slot, S, pixel, index);
        }
    }
    if (errors) atomicAdd(&P.state->errors, errors);
    // The workgroup's queued rays into the compacted lists: one atomic per list and workgroup.
    __shared__ uint32_t waveRays[4], waveShadows[4], base[2];
    uint32_t totalRays, totalShadows;
    const uint32_t atRay = waveExclusive(queuedRay, totalRays), atShadow = waveExclusive(queuedShadows, totalShadows);
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        waveRays[wave] = totalRays;
        waveShadows[wave] = totalShadows;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0, sh = 0;
        for (int w = 0; w < 4; ++w) {
            const uint32_t a = waveRays[w], b = waveShadows[w];
            waveRays[w] = r;
            waveShadows[w] = sh;
            r += a;
            sh += b;
        }
        base[0] = r ? atomicAdd(&W.counts->rays[p], r) : 0u;
        base[1] = sh ? atomicAdd(&W.counts->shadows[p], sh) : 0u;
    }
    __syncthreads();
    if (queuedRay) W.rayList[p][base[0] + waveRays[wave] + atRay] = slot;
    for (uint32_t j = 0; j < queuedShadows; ++j)
        W.shadowList[p][base[1] + waveShadows[wave] + atShadow + j] = slot * W.perVertex + j;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
waveTraceKernel(const RenderParams P, const WaveParams W, uint32_t p) {
    TravStack st;
    Counters k;
    const uint32_t nr = W.counts->rays[p], ns = W.counts->shadows[p];
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the other parity's lists are free: the next shading pass fills them
        W.counts->rays[p ^ 1u] = 0;
        W.counts->shadows[p ^ 1u] = 0;
        W.counts->traced += nr;
    }
    const uint32_t total = nr + ns, stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < total; i += stride) {
        if (i < nr) {
            const uint32_t slot = W.rayList[p][i];
            const float4* r = reinterpret_cast<const float4*>(W.rays[p]) + size_t(slot) * 2;
            const float4 a = r[0], b = r[1];
            float t;
            ++k.traces;
            W.hitTri[slot] = closestHit(P, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, a.w, b.w, t, k, st);
        } else {
            const uint32_t e = W.shadowList[p][i - nr];
            const float4* r = reinterpret_cast<const float4*>(W.shadowRays[p]) + size_t(e) * 2;
            const float4 a = r[0], b = r[1];
            ++k.shadows;
            W.shadowVis[p][e] = anyHit(P, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, a.w, b.w, k, st) ? 0u : 1u;
        }
    }
    if (P.counters && blockIdx.x < W.partialBlocks) {  // per-workgroup partial sums, added up at the launch's end
        __shared__ unsigned long long sums[4][4];
        const unsigned long long v[4] = {waveSum(k.traces), waveSum(k.shadows), waveSum(k.nodes), waveSum(k.triangles)};
        if ((threadIdx.x & 63u) == 0)
            for (int c = 0; c < 4; ++c) sums[threadIdx.x >> 6][c] = v[c];
        __syncthreads();
        if (threadIdx.x < 4) {
            unsigned long long* q = W.partials + size_t(blockIdx.x) * 4 + threadIdx.x;
            *q += sums[0][threadIdx.x] + sums[1][threadIdx.x] + sums[2][threadIdx.x] + sums[3][threadIdx.x];
        }
    }
}

// The launch's end (parity p: the last trace pass's queues): each busy slot's sample is finished with shadow rays
// traced inline, then the lane takes the launch's remaining items with the per-lane loop.
__global__ void __launch_bounds__(256) X2_OCCUPANCY waveDrainKernel(const RenderParams P, const WaveParams W,
                                                                     uint32_t p) {
    const uint32_t slot = blockIdx.x * 256u + threadIdx.x;  // every lane reaches the end (the counter sums)
    const uint32_t fl = slot < W.slots ? W.flags[slot] : 0u;
    TravStack st;
    Counters k;
    uint32_t errors = 0;
    if (fl) {
        WaveSlot S;
        S.stack = LevelStack{static_cast<Pending*>(W.stack) + slot, W.slots};
        uint32_t pixel, index;
        loadSlot(W, slot, S, pixel, index);
        if (fl & WaveShadows) resolveShadows(W, slot, p, S);
        float gr[1] = {0.0f};
        bool ok = true;
        if (fl & WaveTraced) {
            InlineTrace inl;
            ok = shadeTraced(P, W, slot, pixel, index, S, inl, k, st);
        }
        while (ok && S.top) ok = shadeVertex<false, false>(P, pixel, index, S, k, st, gr);
        if (ok) filmAdd(P, pixel, index, S.radiance.x, S.radiance.y, S.radiance.z, S.alpha);
        else ++errors;
        W.flags[slot] = 0;
    }
    if (slot < W.slots && feedOpen(P)) {
        float gr[1] = {0.0f};
        LaneFeed<false> feed{P, errors};
        runSamples<false, false, false>(P, feed, k, st, gr);
    }
    if (errors) atomicAdd(&P.state->errors, errors);
    const unsigned long long drained = waveSum(k.traces);
    if ((threadIdx.x & 63u) == 0 && drained) atomicAdd(&W.counts->traced, drained);
    if (P.counters) {
        const unsigned long long t = waveSum(k.traces), sh = waveSum(k.shadows), n = waveSum(k.nodes),
                                 tr = waveSum(k.triangles);
        if ((threadIdx.x & 63u) == 0) {
            atomicAdd(&P.state->traces, t);
            atomicAdd(&P.state->shadows, sh);
            atomicAdd(&P.state->nodes, n);
            atomicAdd(&P.state->triangles, tr);
        }
    }
}

// Every pixel's adaptive error (async_gpu.cu:364-369, pixelError async_integrator.h:3109-3125): the luma of the
// difference between the even- and odd-index half-images over the square root of the pixel's luma, plus the alpha
// difference; -1 until both halves have a sample. Each product and sum rounds on its own.
__global__ void errorsKernel(const double* sums, const double* odd, const uint32_t* counts, uint32_t pixels,
                             float* errors) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < pixels) errors[i] = adaptive::pixelErrorOf(sums + size_t(i) * 4, odd + size_t(i) * 4, counts[i]);
}

// ---------------------------------------------------------------------------------------------------------------
// The device pass chain (async_adaptive.cu:22-177): a frame's adaptive passes queued at once. Each pass: every
// pixel's error (chainErrorsKernel), each region's tree rebuilt from them (chainTreesKernel: the summed-area table
// row by row then column by column, then the nodes a level at a time, each judged with the host's arithmetic,
// x2_adaptive_core.h), each tile's active pixels and their sample bases (chainMasksKernel), the active tiles in
// the work grid's cell order (chainCompactKernel), then the render over them. Once the trees have converged or a
// pass comes out empty, every later step does nothing.
__device__ __forceinline__ bool chainHalted(const ChainParams& C) {
    return *reinterpret_cast<const volatile uint32_t*>(&C.state->stop) != 0;
}

__global__ void chainErrorsKernel(const ChainParams C) {
    if (chainHalted(C)) return;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < C.width * C.height) C.errors[i] = adaptive::pixelErrorOf(C.sums + size_t(i) * 4, C.odd + size_t(i) * 4, C.counts[i]);
}

__global__ void __launch_bounds__(512) chainTreesKernel(const ChainParams C) {
    if (chainHalted(C)) return;
    const uint32_t r = blockIdx.x;
    const ChainRegion R = C.regions[r];
    const adaptive::IBox root{R.x0, R.y0, R.x1, R.y1};
    const int w = R.x1 - R.x0, h = R.y1 - R.y0, w1 = w + 1;
    double* sat = C.sat + R.satOffset;
    adaptive::Node* nodes = static_cast<adaptive::Node*>(C.nodes) + R.nodeOffset;
    __shared__ int notReady, open, count;
    if (threadIdx.x == 0) {
        notReady = 0;
        open = 0;
        count = 1;
        nodes[0] = adaptive::Node{{float(R.x0), float(R.y0), float(R.x1), float(R.y1)}, -1, 0};
        if (r == 0) ++C.state->treeUpdates;
    }
    __syncthreads();
    for (int x = threadIdx.x; x <= w; x += blockDim.x) sat[x] = 0;
    for (int y = threadIdx.x; y < h; y += blockDim.x) {  // each row's running sums
        const float* e = C.errors + size_t(R.y0 + y) * C.width + R.x0;
        double* out = sat + size_t(y + 1) * w1;
        double run = 0;
        bool missing = false;
        out[0] = 0;
        for (int x = 0; x < w; ++x) {
            const float v = e[x];
            missing |= v < 0;
            run += double(v);
            out[x + 1] = run;
        }
        if (missing) notReady = 1;
    }
    __syncthreads();
    if (notReady) {  // a pixel without both halves yet: the root alone, open
        if (threadIdx.x == 0) C.regionDone[r] = 0;
        return;
    }
    for (int x = threadIdx.x + 1; x <= w; x += blockDim.x)  // then down each column
        for (int y = 1; y <= h; ++y) sat[size_t(y) * w1 + x] += sat[size_t(y - 1) * w1 + x];
    __syncthreads();
    const adaptive::Table t{sat, R.x0, R.y0, w1};
    const int capacity = int(R.nodeCapacity);
    int begin = 0, end = 1;
    while (begin < end) {
        for (int i = begin + int(threadIdx.x); i < end; i += blockDim.x) {
            adaptive::FBox a, b;
            const adaptive::Fate f = adaptive::judge(t, root, nodes[i].bounds, C.target, a, b);
            if (f == adaptive::Converged) {
                nodes[i].complete = 1;
            } else if (f == adaptive::Split) {
                const int c = atomicAdd(&count, 2);
                if (c + 2 <= capacity) {
                    nodes[c] = adaptive::Node{a, -1, 0};
                    nodes[c + 1] = adaptive::Node{b, -1, 0};
                    nodes[i].children = c;
                } else {
                    open = 1;  // beyond MoonRay's own bound on a region's nodes: never
                }
            } else {
                open = 1;
            }
        }
        __syncthreads();
        begin = end;
        end = count < capacity ? count : capacity;
        __syncthreads();
    }
    if (threadIdx.x == 0) C.regionDone[r] = open ? 0u : 1u;
}

// The tile's pixels under an open leaf of its region's tree, filled to their bounding rectangle (inner nodes are
// never marked complete here: a subtree is skipped when its node converged, or when it misses the tile).
__device__ uint64_t openArea(const adaptive::Node* nodes, int tx, int ty) {
    const float t0 = float(tx), t1 = float(ty), t2 = float(tx + 8), t3 = float(ty + 8);
    int stack[128], top = 0;
    stack[top++] = 0;
    uint64_t bits = 0;
    while (top) {
        const adaptive::Node n = nodes[stack[--top]];
        if (n.complete) continue;
        const adaptive::FBox c{n.bounds.x0 < t0 ? t0 : n.bounds.x0, n.bounds.y0 < t1 ? t1 : n.bounds.y0,
                               t2 < n.bounds.x1 ? t2 : n.bounds.x1, t3 < n.bounds.y1 ? t3 : n.bounds.y1};
        if (c.x1 <= c.x0 || c.y1 <= c.y0) continue;
        if (n.children >= 0) {
            if (top + 2 > 128) return ~0ull;  // deeper than the stack: the whole tile
            stack[top++] = n.children;
            stack[top++] = n.children + 1;
            continue;
        }
        const adaptive::IBox b = adaptive::outward(c);
        for (int y = adaptive::max2(b.y0, ty); y < adaptive::min2(b.y1, ty + 8); ++y)
            for (int x = adaptive::max2(b.x0, tx); x < adaptive::min2(b.x1, tx + 8); ++x)
                bits |= 1ull << ((y - ty) * 8 + (x - tx));
    }
    return adaptive::fillRectangle(bits);
}

// One thread a tile: its active pixels for the pass (all of them in the uniform stage), each pixel's first sample
// index and its planned count. Once every region's tree has converged, the chain stops instead.
__global__ void chainMasksKernel(const ChainParams C, uint32_t uniform, uint32_t size) {
    if (chainHalted(C)) return;
    if (!uniform) {
        bool converged = true;
        for (uint32_t r = 0; r < C.regionCount; ++r) converged &= C.regionDone[r] != 0;
        if (converged) {
            if (blockIdx.x == 0 && threadIdx.x == 0) C.state->stop = 1;
            return;
        }
    }
    const uint32_t tile = blockIdx.x * 256u + threadIdx.x;
    if (tile >= C.tiles) return;
    unsigned long long mask = C.inImage[tile];
    const int tx = int(tile % C.tilesX) * 8, ty = int(tile / C.tilesX) * 8;
    if (!uniform && mask) {
        const uint32_t rx = uint32_t(tx) / C.regionW < C.perDim - 1 ? uint32_t(tx) / C.regionW : C.perDim - 1;
        const uint32_t ry = uint32_t(ty) / C.regionH < C.perDim - 1 ? uint32_t(ty) / C.regionH : C.perDim - 1;
        mask &= openArea(static_cast<const adaptive::Node*>(C.nodes) + C.regions[ry * C.perDim + rx].nodeOffset, tx, ty);
    }
    C.tileMask[tile] = mask;
    for (uint32_t q = 0; q < TilePixels; ++q)
        if (mask >> q & 1ull) {
            const uint32_t pixel = (uint32_t(ty) + q / 8) * C.width + uint32_t(tx) + q % 8;
            C.base[pixel] = C.planned[pixel];
            C.planned[pixel] += size;
        }
}

// One workgroup: the pass's active tiles in cell order, for its render. A pass with no active pixel ends the chain.
__global__ void __launch_bounds__(1024) chainCompactKernel(const ChainParams C) {
    if (chainHalted(C)) return;
    __shared__ uint32_t scan[1024];
    const uint32_t tid = threadIdx.x, per = (C.tiles + 1023u) / 1024u, first = tid * per;
    const uint32_t last = first + per < C.tiles ? first + per : C.tiles;
    uint32_t mine = 0;
    for (uint32_t k = first; k < last; ++k) mine += C.tileMask[C.tileOrder[k]] ? 1u : 0u;
    scan[tid] = mine;
    __syncthreads();
    for (uint32_t offset = 1; offset < 1024u; offset <<= 1) {
        const uint32_t v = tid >= offset ? scan[tid - offset] : 0u;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    uint32_t at = scan[tid] - mine;
    for (uint32_t k = first; k < last; ++k) {
        const uint32_t tile = C.tileOrder[k];
        const unsigned long long m = C.tileMask[tile];
        if (m) {
            C.activeTiles[at] = tile;
            C.activeMasks[at] = m;
            ++at;
        }
    }
    if (tid == 1023u) {
        const uint32_t total = scan[1023];
        if (total) {
            C.state->activeTiles = total;
            C.state->claimed = 0;
            ++C.state->passes;
        } else {
            C.state->stop = 1;
        }
    }
}

__global__ void waveBeginKernel(const WaveParams W) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        W.counts->rays[0] = W.counts->rays[1] = 0;
        W.counts->shadows[0] = W.counts->shadows[1] = 0;
        W.counts->traced = 0;
    }
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < W.partialBlocks * 4u; i += gridDim.x * 256u) W.partials[i] = 0;
}
// The trace passes' per-workgroup counters into the launch's totals (one workgroup).
__global__ void waveEndKernel(const RenderParams P, const WaveParams W) {
    __shared__ unsigned long long sums[256][4];
    unsigned long long v[4] = {0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < W.partialBlocks; b += 256u)
        for (int c = 0; c < 4; ++c) v[c] += W.partials[size_t(b) * 4 + c];
    for (int c = 0; c < 4; ++c) sums[threadIdx.x][c] = v[c];
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int c = 0; c < 4; ++c) sums[threadIdx.x][c] += sums[threadIdx.x + h][c];
        __syncthreads();
    }
    if (threadIdx.x == 0 && P.counters) {
        atomicAdd(&P.state->traces, sums[0][0]);
        atomicAdd(&P.state->shadows, sums[0][1]);
        atomicAdd(&P.state->nodes, sums[0][2]);
        atomicAdd(&P.state->triangles, sums[0][3]);
    }
}

// A scene with a medium: the same estimator with the medium's light and transmittance on every ray.
// 256 VGPRs at 2 waves/SIMD; persistent lanes, one sample at a time (path regeneration measured 0.98x here).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) renderVolumeKernel(const RenderParams P) {
    runLaunch<false, true, false, false>(P);
}
// The prototype estimator's kernel (x2_config.direct_only).
__global__ void __launch_bounds__(256) directKernel(const RenderParams P) { runLaunch<false, false, true>(P); }

// ---------------------------------------------------------------------------------------------------------------
// Batched jobs: many independent frames (each job its own scene, settings and film, a RenderParams in device
// memory) in one persistent launch. The device queue is the jobs' items laid end to end in queue order (first[p]
// = the items before position p); each wave claims `claim` items (a multiple of 64) with one atomic, finds each
// 64-item group's job by binary search, and renders it as the frame kernel renders an item: a group is one 8x8 tile
// at one sample index of one job (item = (sample * tiles + tile) * 64 + position). A job's last group stamps the
// wall clock, the job's completion time.
template <int Waves>
__device__ __forceinline__ void runBatch(const BatchParams& B) {
    TravStack st;
    Counters k;
    uint32_t errors = 0;
    float gr[1] = {0.0f};
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long begin = B.first[B.posBegin], end = B.first[B.posEnd];
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(B.head, (unsigned long long)B.claim);
        base = __shfl(base, 0, 64) + begin;
        if (base >= end) break;
        for (uint32_t g = 0; g < B.claim; g += 64) {
            const unsigned long long group = base + g;
            if (group >= end) break;
            uint32_t lo = B.posBegin, hi = B.posEnd - 1;  // the last position whose first item is <= group
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) / 2;
                if (B.first[mid] <= group) lo = mid;
                else hi = mid - 1;
            }
            const uint32_t job = B.order[lo];
            const RenderParams& P = B.jobs[job];
            const unsigned long long local = group - B.first[lo];
            const uint32_t tiles = ((P.cfg.width + 7) / 8) * ((P.cfg.height + 7) / 8);
            const uint32_t sample = uint32_t(local / (64ull * tiles)), tile = uint32_t((local / 64) % tiles);
            const uint32_t pixel = tilePixel(tile, lane, P.cfg);
            if (pixel != ~0u) {
                double rgba[4];
                if (integrate<false, false>(P, pixel, sample, rgba, k, st, gr) == 0) {
                    double* f = P.sums + size_t(pixel) * 4;
                    atomicAdd(f + 0, rgba[0]);
                    atomicAdd(f + 1, rgba[1]);
                    atomicAdd(f + 2, rgba[2]);
                    atomicAdd(f + 3, rgba[3]);
                    atomicAdd(P.counts + pixel, 1u);
                } else {
                    ++errors;
                }
            }
            if (lane == 0) {
                const unsigned long long items = B.first[lo + 1] - B.first[lo];
                if (atomicAdd(B.jobDone + job, 64ull) + 64 == items) B.jobFinish[job] = wall_clock64();
            }
        }
    }
    if (errors) atomicAdd(&B.state->errors, errors);
    const unsigned long long t = waveSum(k.traces), sh = waveSum(k.shadows);
    if (lane == 0) {
        atomicAdd(&B.state->traces, t);
        atomicAdd(&B.state->shadows, sh);
    }
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) batchKernel2(const BatchParams B) {
    runBatch<2>(B);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) batchKernel3(const BatchParams B) {
    runBatch<3>(B);
}
// A texture's next mip level: each texel the mean of the 2 x 2 below it, ((a + b) + (c + d)) * 0.25 per channel with a,
// b the upper row, coordinates clamped to the level below (one thread a texel).
__global__ void pyramidKernel(const float* in, uint32_t w, uint32_t h, float* out, uint32_t ow, uint32_t oh) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ow * oh) return;
    const uint32_t x = i % ow, y = i / ow;
    const uint32_t x0 = min(2 * x, w - 1), x1 = min(2 * x + 1, w - 1), y0 = min(2 * y, h - 1), y1 = min(2 * y + 1, h - 1);
    const float* a = in + (size_t(y0) * w + x0) * 3;
    const float* b = in + (size_t(y0) * w + x1) * 3;
    const float* c = in + (size_t(y1) * w + x0) * 3;
    const float* d = in + (size_t(y1) * w + x1) * 3;
    float* o = out + size_t(i) * 3;
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = ((a[k] + b[k]) + (c[k] + d[k])) * 0.25f;
}
// The transpose of one pyramid step: a level's texel gradients go a quarter each to the texels below that made it
// (one thread a texel below, gathering from its parent; a level one texel wide or high made its parent from that
// texel twice).
__global__ void pushDownKernel(const float* g, uint32_t uw, uint32_t uh, float* below, uint32_t w, uint32_t h) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const float mx = w == 1 ? 2.0f : x < 2 * uw ? 1.0f : 0.0f, my = h == 1 ? 2.0f : y < 2 * uh ? 1.0f : 0.0f;
    if (mx == 0.0f || my == 0.0f) return;
    const float* p = g + (size_t(y >> 1) * uw + (x >> 1)) * 3;
    float* q = below + size_t(i) * 3;
    const float s = 0.25f * mx * my;
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] += s * p[k];
}

// Zeroes a list of regions (films, counters) with plain stores, so a captured graph can replay it.
__global__ void fillKernel(const FillRegion* regions, uint32_t n) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t* w = regions[r].words;
        const unsigned long long count = regions[r].count;
        for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
            w[i] = 0u;
    }
}
__global__ void stampKernel(unsigned long long* at) {
    if (threadIdx.x == 0) *at = wall_clock64();
}

// The RNG twin (diagnostics): the device's RNG functions on a list of identities.
__global__ void rngTwinKernel(const x2_rng_query* q, uint32_t n, uint32_t draws, uint32_t* words, float* out,
                              uint64_t* branches) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const x2_rng_query r = q[i];
    const Identity id{r.pixel, r.index, r.stream, r.branch};
    words[i] = samplingWord(id, r.dimension, r.policy);
    Stream s(id, r.bounce);
    for (uint32_t k = 0; k < draws; ++k) out[size_t(i) * draws + k] = s.next();
    branches[i] = childBranch(r.branch, r.a, r.b);
}

// The film's mean rgb per pixel (sums / counts).
__global__ void resolveKernel(const double* sums, const uint32_t* counts, uint32_t pixels, float* rgb) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= pixels) return;
    const double n = double(counts[i]);
    for (int c = 0; c < 3; ++c) rgb[size_t(i) * 3 + c] = n > 0 ? float(sums[size_t(i) * 4 + c] / n) : 0.0f;
}

__global__ void clearState(LaunchState* s) {
    if (threadIdx.x == 0) {
        s->traces = s->shadows = s->nodes = s->triangles = 0;
        s->claimed = 0;
        s->voxels = 0;
        s->errors = 0;
    }
}

}  // namespace x2

// Host-side launchers the lane calls (x2_lane.hip).
namespace x2 {
template <int Kind>  // 0 render, 1 adjoint, 2 render with a medium, 3 the prototype estimator, 4 paged, 5 pooled,
                    // 6 textured
hipError_t launchFrame(const RenderParams& p, hipStream_t stream) {
    const auto kernel = Kind == 0   ? renderKernel
                        : Kind == 1 ? adjointKernel
                        : Kind == 2 ? renderVolumeKernel
                        : Kind == 3 ? directKernel
                        : Kind == 5 ? renderPoolKernel
                        : Kind == 6 ? renderTexKernel
                                    : renderPagedKernel;
    static int resident = 0;  // persistent lanes: as many workgroups as the chip holds at once
    if (!resident) {
        int dev = 0, perCu = 0;
        hipDeviceProp_t prop;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, 256, 0);
        if (e != hipSuccess) return e;
        resident = std::max(1, perCu) * prop.multiProcessorCount;
    }
    // The device pass chain's render learns its items on the device: a full persistent grid.
    // A launch's items: its chunks' and the parked samples it runs again (a launch may hold parked samples alone).
    const unsigned long long items = p.chain ? ~0ull >> 1 : p.itemEnd - p.itemBegin + p.parkedCount + p.refineInCount;
    unsigned long long blocks = (items + 255) / 256;
    if (blocks > (unsigned long long)resident) blocks = resident;
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, p);
    return hipGetLastError();
}
hipError_t launchRender(const RenderParams& p, hipStream_t stream) {
    if (p.cfg.direct_only) return launchFrame<3>(p, stream);
    if (p.tileSlot) return launchFrame<4>(p, stream);  // texture tiles on demand
    if (p.vol.table) return launchFrame<2>(p, stream);
    if (p.textures) return launchFrame<6>(p, stream);
    // The pool holds a vertex's shadow rays when every light's samples and every lobe's (two at most) fit. With the
    // pool's rays traced in the same loop as the next rays it pays with one light too (the Cornell box at 1024^2 x 64:
    // 91.4 ms a frame pooled against 94.1 ms traced where they are made).
    const uint32_t nl = p.cfg.light_samples > 1 ? p.cfg.light_samples : 1u, nk = p.cfg.bsdf_samples > 1 ? p.cfg.bsdf_samples : 1u;
    if (uint64_t(p.cfg.light_count) * nl + 2ull * nk <= PoolRays) return launchFrame<5>(p, stream);
    return launchFrame<0>(p, stream);
}
hipError_t launchAdjoint(const RenderParams& p, hipStream_t stream) { return launchFrame<1>(p, stream); }
// One step of the device pass chain: errors and trees (adaptive passes), masks and compaction, the render.
hipError_t launchChainStep(const ChainParams& c, const RenderParams& p, bool uniform, hipStream_t stream) {
    if (!uniform) {
        hipLaunchKernelGGL(chainErrorsKernel, dim3((c.width * c.height + 255) / 256), dim3(256), 0, stream, c);
        hipLaunchKernelGGL(chainTreesKernel, dim3(c.regionCount), dim3(512), 0, stream, c);
    }
    hipLaunchKernelGGL(chainMasksKernel, dim3((c.tiles + 255) / 256), dim3(256), 0, stream, c, uniform ? 1u : 0u,
                       p.chainSize);
    hipLaunchKernelGGL(chainCompactKernel, dim3(1), dim3(1024), 0, stream, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return p.tileSlot ? launchFrame<4>(p, stream) : p.textures ? launchFrame<6>(p, stream) : launchFrame<0>(p, stream);
}
hipError_t launchErrors(const double* sums, const double* odd, const uint32_t* counts, uint32_t pixels, float* errors,
                        hipStream_t stream) {
    hipLaunchKernelGGL(errorsKernel, dim3((pixels + 255) / 256), dim3(256), 0, stream, sums, odd, counts, pixels, errors);
    return hipGetLastError();
}
// The wavefront form over one launch's items: its counters cleared, `iterations` shading and trace passes, the drain.
hipError_t launchWave(const RenderParams& p, const WaveParams& w, uint32_t iterations, hipStream_t stream) {
    static int traceBlocks = 0;  // the trace pass is persistent: as many workgroups as the chip holds at once
    if (!traceBlocks) {
        int dev = 0, perCu = 0;
        hipDeviceProp_t prop;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, waveTraceKernel, 256, 0);
        if (e != hipSuccess) return e;
        traceBlocks = std::max(1, perCu) * prop.multiProcessorCount;
    }
    const unsigned shadeBlocks = w.slots / 256u, blocks = unsigned(std::min<int>(traceBlocks, int(w.partialBlocks)));
    hipLaunchKernelGGL(waveBeginKernel, dim3(16), dim3(256), 0, stream, w);
    for (uint32_t i = 0; i < iterations; ++i) {
        hipLaunchKernelGGL(waveShadeKernel, dim3(shadeBlocks), dim3(256), 0, stream, p, w, i & 1u);
        hipLaunchKernelGGL(waveTraceKernel, dim3(blocks), dim3(256), 0, stream, p, w, i & 1u);
    }
    hipLaunchKernelGGL(waveDrainKernel, dim3(shadeBlocks), dim3(256), 0, stream, p, w, (iterations + 1u) & 1u);
    hipLaunchKernelGGL(waveEndKernel, dim3(1), dim3(256), 0, stream, p, w);
    return hipGetLastError();
}
hipError_t launchResolve(const double* sums, const uint32_t* counts, uint32_t pixels, float* rgb, hipStream_t stream) {
    hipLaunchKernelGGL(resolveKernel, dim3((pixels + 255) / 256), dim3(256), 0, stream, sums, counts, pixels, rgb);
    return hipGetLastError();
}
hipError_t launchClearState(LaunchState* s, hipStream_t stream) {
    hipLaunchKernelGGL(clearState, dim3(1), dim3(64), 0, stream, s);
    return hipGetLastError();
}
// The batched queue's launchers (x2_batch.hip): the persistent grid is what the chip holds of the chosen kernel at
// the chosen workgroup size.
hipError_t batchResident(int waves, uint32_t threads, int& blocks) {
    int dev = 0, perCu = 0;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, waves == 2 ? batchKernel2 : batchKernel3, int(threads), 0);
    if (e != hipSuccess) return e;
    blocks = std::max(1, perCu) * prop.multiProcessorCount;
    return hipSuccess;
}
hipError_t launchBatch(const BatchParams& b, int waves, uint32_t threads, int blocks, hipStream_t stream) {
    if (waves == 2) hipLaunchKernelGGL(batchKernel2, dim3(unsigned(blocks)), dim3(threads), 0, stream, b);
    else hipLaunchKernelGGL(batchKernel3, dim3(unsigned(blocks)), dim3(threads), 0, stream, b);
    return hipGetLastError();
}
hipError_t launchPyramid(const float* in, uint32_t w, uint32_t h, float* out, hipStream_t stream) {
    const uint32_t ow = w > 1 ? w >> 1 : 1u, oh = h > 1 ? h >> 1 : 1u;
    hipLaunchKernelGGL(pyramidKernel, dim3((ow * oh + 255) / 256), dim3(256), 0, stream, in, w, h, out, ow, oh);
    return hipGetLastError();
}
hipError_t launchPushDown(const float* g, uint32_t uw, uint32_t uh, float* below, uint32_t w, uint32_t h, hipStream_t stream) {
    hipLaunchKernelGGL(pushDownKernel, dim3((w * h + 255) / 256), dim3(256), 0, stream, g, uw, uh, below, w, h);
    return hipGetLastError();
}
hipError_t launchFill(const FillRegion* regions, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(fillKernel, dim3(1024), dim3(256), 0, stream, regions, n);
    return hipGetLastError();
}
hipError_t launchStamp(unsigned long long* at, hipStream_t stream) {
    hipLaunchKernelGGL(stampKernel, dim3(1), dim3(64), 0, stream, at);
    return hipGetLastError();
}
hipError_t launchRngTwin(const x2_rng_query* q, uint32_t n, uint32_t draws, uint32_t* words, float* out,
                         uint64_t* branches, hipStream_t stream) {
    hipLaunchKernelGGL(rngTwinKernel, dim3((n + 255) / 256), dim3(256), 0, stream, q, n, draws, words, out, branches);
    return hipGetLastError();
}
hipError_t uploadSobol(const uint32_t* table) {
    return hipMemcpyToSymbol(HIP_SYMBOL(kSobolDirections), table, sizeof(uint32_t) * 17 * 32, 0, hipMemcpyHostToDevice);
}
}  // namespace x2
