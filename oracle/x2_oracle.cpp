// x2_oracle.cpp — TEST INFRASTRUCTURE ONLY (the oracle). A CPU restatement of the reference's shared path-tracing
// integrator subset that the HIP device implements: the sample-identity RNG, the pinhole camera with a box filter,
// the watertight triangle test with FP64 hit reconstruction, rectangle and constant-environment lights, Lambert and
// GGX lobes, and MoonRay's multi-sampler estimator with its pending-ray stack, Russian roulette, MIS and bounce-ray
// reuse. Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline load it; the product never does.
//
// Built with -O3 -ffp-contract=off -fno-fast-math (the reference's CPU parity flags, native_build.py:40), so every
// float expression rounds as written. Each function cites the reference file:line it restates; citations are into
// the reference tree (xpu2/interactive/ unless a path says otherwise).
#include "x2_oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {
using U32 = uint32_t;
using U64 = uint64_t;

// ---------------------------------------------------------------------------------------------------------------
// RNG: sample_stream.h
struct Words { U32 v[4]; };
// Philox4x32-10 (sample_stream.h:31-38, Random123's constants).
Words philox(Words c, U32 ka, U32 kb) {
    for (int round = 0; round < 10; ++round) {
        const U64 a = U64(c.v[0]) * 0xd2511f53u, b = U64(c.v[2]) * 0xcd9e8d57u;
        Words n;
        n.v[0] = U32(b >> 32) ^ c.v[1] ^ ka;
        n.v[1] = U32(b);
        n.v[2] = U32(a >> 32) ^ c.v[3] ^ kb;
        n.v[3] = U32(a);
        c = n;
        ka += 0x9e3779b9u;
        kb += 0xbb67ae85u;
    }
    return c;
}
struct Identity { U32 pixel; U64 index, stream, branch; };
// key (sample_stream.h:43-48) and block (:49-51).
Words block(const Identity& id, U32 group) {
    const U64 k = id.stream ^ (id.branch * 0x9e3779b97f4a7c15ull);
    Words c{{id.pixel, U32(id.index), U32(id.index >> 32), group}};
    return philox(c, U32(k), U32(k >> 32));
}
U32 word(const Identity& id, U32 dimension) { return block(id, dimension >> 2).v[dimension & 3]; }  // :52-54
float unit(U32 bits) { return float(bits >> 8) * 0x1p-24f; }                                      // :56

// Sobol direction numbers (sample_stream.h:82-100), regenerated here from the Joe-Kuo primitive polynomials
// (S. Joe and F. Y. Kuo, new-joe-kuo-6.21201: degree s, coefficient bits a, initial m_1..m_s) by the standard
// recurrence; row 0 is the van der Corput sequence, rows 1..16 are Joe-Kuo dimensions 2..17.
struct JoeKuo { U32 s, a, m[6]; };
const JoeKuo kJoeKuo[16] = {
    {1, 0, {1}},           {2, 1, {1, 3}},          {3, 1, {1, 3, 1}},        {3, 2, {1, 1, 1}},
    {4, 1, {1, 1, 3, 3}},  {4, 4, {1, 3, 5, 13}},   {5, 2, {1, 1, 5, 5, 17}}, {5, 4, {1, 1, 5, 5, 5}},
    {5, 7, {1, 1, 7, 11, 19}}, {5, 11, {1, 1, 5, 1, 1}}, {5, 13, {1, 1, 1, 3, 11}}, {5, 14, {1, 3, 5, 5, 31}},
    {6, 1, {1, 3, 3, 9, 7, 49}}, {6, 13, {1, 1, 1, 15, 21, 21}}, {6, 16, {1, 3, 1, 13, 27, 49}},
    {6, 19, {1, 1, 1, 15, 7, 5}},
};
struct Directions {
    U32 v[17][32];
    Directions() {
        for (U32 k = 0; k < 32; ++k) v[0][k] = 1u << (31 - k);
        for (U32 d = 1; d < 17; ++d) {
            const JoeKuo& j = kJoeKuo[d - 1];
            U32 m[33] = {};
            for (U32 i = 1; i <= j.s; ++i) m[i] = j.m[i - 1];
            for (U32 i = j.s + 1; i <= 32; ++i) {
                U32 x = m[i - j.s] ^ (m[i - j.s] << j.s);
                for (U32 k = 1; k < j.s; ++k)
                    if ((j.a >> (j.s - 1 - k)) & 1u) x ^= m[i - k] << k;
                m[i] = x;
            }
            for (U32 k = 1; k <= 32; ++k) v[d][k - 1] = m[k] << (32 - k);
        }
    }
};
const Directions kDirections;
U32 sobol(U32 index, U32 dimension) {  // sample_stream.h:102-107
    U32 gray = index ^ (index >> 1), result = 0, bit = 0;
    while (gray) {
        if (gray & 1) result ^= kDirections.v[dimension][bit];
        gray >>= 1;
        ++bit;
    }
    return result;
}
U32 reverseBits(U32 x) {  // sample_stream.h:108-116
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
// Burley 2020 hash-based Owen scrambling (sample_stream.h:122-126).
U32 owenScramble(U32 x, U32 seed) {
    x = reverseBits(x);
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return reverseBits(x);
}
// Policy dimensions (sample_stream.h:127-144).
U32 samplingWord(const Identity& id, U32 dimension, U32 policy) {
    const bool camera = dimension < 5, diffuse = dimension == 16 || dimension == 17,
               light = dimension == 18 || dimension == 19, subsurface = dimension >= 20 && dimension <= 27;
    if ((!camera || !(policy & 1)) && (!diffuse || !(policy & 2)) && (!light || !(policy & 4)) &&
        (!subsurface || !(policy & 8)))
        return word(id, dimension);
    const U32 d = camera ? (dimension < 2 ? dimension : dimension < 4 ? dimension + 12 : 16) : dimension - 14;
    Identity shift = id;
    shift.index = id.index >> 32;
    if (camera && dimension >= 2) {
        const Words seeds = block(shift, dimension < 4 ? 0x50414444u : 0x50414445u);
        return owenScramble(sobol(owenScramble(U32(id.index), seeds.v[0]), dimension < 4 ? dimension - 2 : 0u),
                            seeds.v[dimension < 4 ? dimension - 1 : 1]);
    }
    return owenScramble(sobol(U32(id.index), d), block(shift, 0x534f424fu + (d >> 2)).v[d & 3]);
}
float draw(const Identity& id, U32 dimension, U32 policy) { return unit(samplingWord(id, dimension, policy)); }

constexpr U32 CameraDimensions = 16, DimensionsPerBounce = 256;  // sample_stream.h:149
// paths::Stream (async_integrator.h:151-176): one Philox block per four draws along a vertex's dimensions.
struct Stream {
    Identity id;
    U32 dimension, group = ~0u;
    U32 w[4] = {0, 0, 0, 0};
    Stream(const Identity& i, U32 bounce, U32 offset)
        : id(i), dimension(CameraDimensions + bounce * DimensionsPerBounce + offset) {}
    float next() {
        if ((dimension >> 2) != group) {
            group = dimension >> 2;
            const Words b = block(id, group);
            for (int k = 0; k < 4; ++k) w[k] = b.v[k];
        }
        return unit(w[dimension++ & 3]);
    }
};
// paths::childBranch (async_integrator.h:281-286).
U64 childBranch(U64 parent, U32 a, U32 b) {
    U64 z = parent + 0x9e3779b97f4a7c15ull * (U64(a) * 0x10001u + U64(b) + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------------------------------------------
// Vectors (surface_core.h:10-18, shade_math.h:15-26,65-69,113-120): every operation in the reference's order.
struct V3 { float x, y, z; };
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
V3 operator*(V3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
V3 operator*(float b, V3 a) { return {a.x * b, a.y * b, a.z * b}; }
V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
V3 operator/(V3 a, float b) { return {a.x / b, a.y / b, a.z / b}; }
float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
bool finite(float a) { return a == a && a <= 3.402823466e38f && a >= -3.402823466e38f; }
bool finite3(V3 a) { return finite(a.x) && finite(a.y) && finite(a.z); }
bool normalizeTo(V3 a, V3& b) {  // surface::normalize
    const float l = dot(a, a);
    if (!finite(l) || l <= 0) return false;
    b = a * (1.0f / sqrtf(l));
    return true;
}
float length(V3 a) { return sqrtf(dot(a, a)); }
float luminance(V3 c) { return 0.299f * c.x + 0.587f * c.y + 0.114f * c.z; }
bool isExactlyZero(V3 c) { return c.x == 0 && c.y == 0 && c.z == 0; }
bool isEqual0(float x) { return fabsf(x - 0.0f) <= fmaxf(fabsf(x), 1.0f) * 1e-6f; }
bool isBlack(V3 c) { return isEqual0(c.x) && isEqual0(c.y) && isEqual0(c.z); }
V3 load3(const float* p) { return {p[0], p[1], p[2]}; }
void store3(float* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

constexpr float Pi = 3.14159265358979323846f, TwoPi = 6.28318530717958647692f;
constexpr float OneOverPi = 0.318309886183790671538f, Epsilon = 1e-6f;

// The shared sine/cosine polynomial (surface_core.h:116-126; ISPC 1.31's coefficients), used for every sin/cos on
// the path so that both devices round alike. The reference's lobes call libm (CPU) or sincosf (CUDA) instead
// (shade_math.h:123-130); this build uses the polynomial on both.
void sincosPoly(float angle, float& sn, float& cs) {
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
// Camera (camera_core.h:53-67 transforms, :217-250 perspectiveRay without depth of field; surface_core.h:51-52 box
// filter). The raster-to-camera and camera-to-render matrices are row-vector (scene_rdl2) matrices.
struct Ray { V3 origin, direction; float nearT, farT; };
V3 vectorM(const float* m, V3 a) {  // surface::vector
    return {(a.x * m[0] + a.y * m[4]) + a.z * m[8], (a.x * m[1] + a.y * m[5]) + a.z * m[9],
            (a.x * m[2] + a.y * m[6]) + a.z * m[10]};
}
V3 transformH(const float* m, V3 p) {
    const V3 v = vectorM(m, p) + V3{m[12], m[13], m[14]};
    const float w = ((p.x * m[3] + p.y * m[7]) + p.z * m[11]) + m[15];
    if (w == 0.0f) return {0, 0, 0};
    return {v.x / w, v.y / w, v.z / w};
}
V3 transformPoint(const float* m, V3 p) { return vectorM(m, p) + V3{m[12], m[13], m[14]}; }
bool boxFilter(float width, float r, float& out) {
    if (!finite(width) || width <= 0 || !finite(r) || r < 0 || r > 1) return false;
    out = (r - .5f) * width + .5f;
    return finite(out);
}
bool cameraRay(const x2_config& c, float x, float y, Ray& out) {
    const V3 dir = transformH(c.raster_to_camera, {x, y, -1.0f});  // minus the focal point, 0
    V3 n;
    if (!normalizeTo(dir, n) || n.z >= 0) return false;
    Ray r;
    r.origin = transformPoint(c.camera_to_render, {0, 0, 0});
    r.nearT = (-c.near_clip - 0.0f) / n.z;
    r.farT = (-c.far_clip - 0.0f) / n.z;
    if (!normalizeTo(vectorM(c.camera_to_render, dir), r.direction) || !finite3(r.origin) || !finite(r.nearT) ||
        !finite(r.farT))
        return false;
    out = r;
    return true;
}
// The camera ray's auxiliary directions through the next pixel in x and in y (camera_core.h:222-224, :245-246).
bool cameraDifferentials(const x2_config& c, float x, float y, V3& dx, V3& dy) {
    const V3 dirX = transformH(c.raster_to_camera, {x + 1, y, -1.0f}), dirY = transformH(c.raster_to_camera, {x, y + 1, -1.0f});
    return normalizeTo(vectorM(c.camera_to_render, dirX), dx) && normalizeTo(vectorM(c.camera_to_render, dirY), dy);
}

// ---------------------------------------------------------------------------------------------------------------
// The watertight ray/triangle test (S. Woop, C. Benthin, I. Wald, "Watertight Ray/Triangle Intersection", JCGT 2013):
// the membership decision the reference leaves to OptiX/Embree. Exact distance ties go to the lower (geom, prim),
// as triangle_hit.h:26 orders them, so the closest hit is a function of the scene alone, not of the BVH.
struct WoopRay {
    V3 org;
    int kx, ky, kz;
    float sx, sy, sz;
    float nearT, farT;
};
float comp(V3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; }
WoopRay woopSetup(V3 org, V3 dir, float nearT, float farT) {
    WoopRay w;
    w.org = org;
    const float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
    w.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    w.kx = w.kz == 2 ? 0 : w.kz + 1;
    w.ky = w.kx == 2 ? 0 : w.kx + 1;
    if (comp(dir, w.kz) < 0.0f) std::swap(w.kx, w.ky);
    const float dz = comp(dir, w.kz);
    w.sx = comp(dir, w.kx) / dz;
    w.sy = comp(dir, w.ky) / dz;
    w.sz = 1.0f / dz;
    w.nearT = nearT;
    w.farT = farT;
    return w;
}
float edgeFunction(float a, float b, float c, float d) {  // a*b - c*d, again in double when it rounds to zero
    const float e = a * b - c * d;
    if (e != 0.0f) return e;
    return float(double(a) * double(b) - double(c) * double(d));
}
bool woopHit(const WoopRay& w, V3 a, V3 b, V3 c, float& t) {
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

// Hit reconstruction (resume_core.h:13-47 geometryHit, triangle_hit.h:6-21 coordinates): t, u, v in FP64 from the
// triangle's float vertices and the ray, the point rounded once, the conservative offset origins.
struct Hit {
    U32 geom = ~0u, prim = 0;
    float t = 0, u = 0, v = 0;  // u, v: the FP64 barycentrics rounded once (resume_core.h:30)
    V3 normal{0, 0, 0}, origin{0, 0, 0}, below{0, 0, 0}, point{0, 0, 0};
};
bool coordinates(const Ray& ray, V3 a, V3 b, V3 c, double& t, double& u, double& v) {
    const double ax = a.x, ay = a.y, az = a.z;
    const double ex = double(b.x) - ax, ey = double(b.y) - ay, ez = double(b.z) - az;
    const double fx = double(c.x) - ax, fy = double(c.y) - ay, fz = double(c.z) - az;
    const double dx = ray.direction.x, dy = ray.direction.y, dz = ray.direction.z;
    const double px = dy * fz - dz * fy, py = dz * fx - dx * fz, pz = dx * fy - dy * fx;
    const double determinant = (ex * px + ey * py) + ez * pz;
    if (determinant == 0) return false;
    const double tx = double(ray.origin.x) - ax, ty = double(ray.origin.y) - ay, tz = double(ray.origin.z) - az;
    const double qx = ty * ez - tz * ey, qy = tz * ex - tx * ez, qz = tx * ey - ty * ex;
    u = ((tx * px + ty * py) + tz * pz) / determinant;
    v = ((dx * qx + dy * qy) + dz * qz) / determinant;
    t = ((fx * qx + fy * qy) + fz * qz) / determinant;
    return finite(float(t)) && finite(float(u)) && finite(float(v));
}
bool geometryHit(Hit& h, const Ray& ray, V3 a, V3 b, V3 c) {
    const V3 e1 = b + (a * -1.0f), e2 = c + (a * -1.0f);
    if (!normalizeTo(cross(e1, e2), h.normal)) return false;
    if (dot(h.normal, ray.direction) > 0) h.normal = h.normal * -1.0f;
    double pt, pu, pv;
    if (!coordinates(ray, a, b, c, pt, pu, pv)) return false;
    h.t = float(pt);
    h.u = float(pu);
    h.v = float(pv);
    const V3 p{float(double(a.x) + ((double(b.x) - a.x) * pu + (double(c.x) - a.x) * pv)),
               float(double(a.y) + ((double(b.y) - a.y) * pu + (double(c.y) - a.y) * pv)),
               float(double(a.z) + ((double(b.z) - a.z) * pu + (double(c.z) - a.z) * pv))};
    h.point = p;
    const V3 error{fabsf(a.x) + fabsf(e1.x) + fabsf(e2.x), fabsf(a.y) + fabsf(e1.y) + fabsf(e2.y),
                   fabsf(a.z) + fabsf(e1.z) + fabsf(e2.z)};
    const float extent = fmaxf(fmaxf(fabsf(e1.x), fabsf(e1.y)),
                               fmaxf(fabsf(e1.z), fmaxf(fabsf(e2.x), fmaxf(fabsf(e2.y), fabsf(e2.z)))));
    const float d = 16 * 0x1p-23f *
                    (fabsf(h.normal.x) * error.x + fabsf(h.normal.y) * error.y + fabsf(h.normal.z) * error.z + extent);
    h.origin = p + h.normal * d;
    if (h.normal.x != 0) h.origin.x = nextafterf(h.origin.x, h.normal.x > 0 ? FLT_MAX : -FLT_MAX);
    if (h.normal.y != 0) h.origin.y = nextafterf(h.origin.y, h.normal.y > 0 ? FLT_MAX : -FLT_MAX);
    if (h.normal.z != 0) h.origin.z = nextafterf(h.origin.z, h.normal.z > 0 ? FLT_MAX : -FLT_MAX);
    h.below = p + h.normal * -d;
    if (h.normal.x != 0) h.below.x = nextafterf(h.below.x, h.normal.x > 0 ? -FLT_MAX : FLT_MAX);
    if (h.normal.y != 0) h.below.y = nextafterf(h.below.y, h.normal.y > 0 ? -FLT_MAX : FLT_MAX);
    if (h.normal.z != 0) h.below.z = nextafterf(h.below.z, h.normal.z > 0 ? -FLT_MAX : FLT_MAX);
    return finite3(h.origin) && finite3(h.below);
}

// ---------------------------------------------------------------------------------------------------------------
// The oracle's own acceleration structure: a binary BVH, binned SAH, at most 4 triangles a leaf. It shares nothing
// with the device's BVH8; the closest hit both find is the same only because the membership test and tie-break are.
struct BNode {
    float lo[3], hi[3];
    U32 a, n;  // n > 0: leaf of triangles order[a, a+n); n == 0: children a and a + 1
};
// ---------------------------------------------------------------------------------------------------------------
// Textures: mip pyramids and the path tracer's map lookups. The filters are texture_core.h:86-206's, which adapt
// OpenImageIO 3.1.17's texturesys.cpp (Apache-2.0, copyright Contributors to the OpenImageIO project): the
// anisotropic filter for camera paths, trilinear closest texels for glossy ones, one level's closest texel for
// diffuse ones (shade_mapnodes.h:344-347), read as MoonRay's ImageMap reads a map (shade_mapnodes.h:420-470: unit
// scale, no offset or bias, periodic, the t flip, and the footprint's cross terms swapped as XPU's vector shading
// passes them, async_cpu.h:928). Two deviations, applied alike on the device: theta's sine and cosine come from the
// shared polynomial, and hypot(A - C, B) is the square root of the FP64 sum of squares.
constexpr U32 MipMax = 16;
struct MipLevel {
    U32 width = 0, height = 0;
    std::vector<float> texels;  // rgb, row-major
};
struct Texture {
    U32 width = 0, height = 0;   // level 0
    std::vector<MipLevel> levels;
    bool half = false;           // every level rounded to half floats (the paged tiles' values)
    U32 coarseFrom = MipMax;     // a fetch at a finer level reads its ancestor at this level (partial shading
                                 // with no tile resident: every parent read from the always-resident levels)
    size_t gradOffset = 0;       // level 0's texels in the adjoint's concatenated gradient
    size_t levelGrad[MipMax] = {};  // each level's texels in the adjoint's pyramid gradient (floats)
};
float roundHalf(float v) {  // to the nearest IEEE binary16 (ties to even) and back
    U32 x;
    std::memcpy(&x, &v, 4);
    const U32 sign = (x >> 16) & 0x8000u, e = (x >> 23) & 0xffu;
    U32 m = x & 0x7fffffu;
    uint16_t h;
    if (e == 0xff) h = uint16_t(sign | 0x7c00u | (m >> 13) | (m && !(m >> 13)));
    else {
        const int exp = int(e) - 127 + 15;
        if (exp >= 31) h = uint16_t(sign | 0x7c00u);
        else if (exp <= 0) {
            if (exp < -10) h = uint16_t(sign);
            else {
                m |= 0x800000u;
                const int shift = 14 - exp;
                U32 q = m >> shift;
                const U32 rest = m & ((1u << shift) - 1), halfway = 1u << (shift - 1);
                if (rest > halfway || (rest == halfway && (q & 1u))) ++q;
                h = uint16_t(sign | q);
            }
        } else {
            U32 q = (U32(exp) << 10) | (m >> 13);
            const U32 rest = m & 0x1fffu;
            if (rest > 0x1000u || (rest == 0x1000u && (q & 1u))) ++q;
            h = uint16_t(sign | q);
        }
    }
    const U32 s = U32(h & 0x8000u) << 16, he = (h >> 10) & 31u, hm = h & 1023u;
    if (he == 0) {
        const float f = float(hm) * (1.0f / 16777216.0f);
        return s ? -f : f;
    }
    const U32 u = s | (he == 31 ? 0x7F800000u | (hm << 13) : ((he + 112u) << 23) | (hm << 13));
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// The pyramid of a texture (the reference reads a file's own levels, down to 1 x 1, host_texture.h:330-380; here
// they are made): level k + 1 is max(1, w >> 1) x max(1, h >> 1), each texel the mean of the 2 x 2 block below it,
// ((a + b) + (c + d)) * 0.25 with a, b the upper row, coordinates clamped to the level below. Paged textures then
// round every level to half floats, as the reference holds a float file's levels (texture_core.h:28-31).
void buildPyramid(Texture& t) {
    t.levels.resize(1);
    while (t.levels.size() < MipMax && (t.levels.back().width > 1 || t.levels.back().height > 1)) {
        const MipLevel& in = t.levels.back();
        MipLevel out;
        out.width = std::max<U32>(1, in.width >> 1);
        out.height = std::max<U32>(1, in.height >> 1);
        out.texels.resize(size_t(out.width) * out.height * 3);
        for (U32 y = 0; y < out.height; ++y)
            for (U32 x = 0; x < out.width; ++x) {
                const U32 x0 = std::min(2 * x, in.width - 1), x1 = std::min(2 * x + 1, in.width - 1);
                const U32 y0 = std::min(2 * y, in.height - 1), y1 = std::min(2 * y + 1, in.height - 1);
                for (int c = 0; c < 3; ++c) {
                    const float a = in.texels[(size_t(y0) * in.width + x0) * 3 + c], b = in.texels[(size_t(y0) * in.width + x1) * 3 + c];
                    const float d0 = in.texels[(size_t(y1) * in.width + x0) * 3 + c], d1 = in.texels[(size_t(y1) * in.width + x1) * 3 + c];
                    out.texels[(size_t(y) * out.width + x) * 3 + c] = ((a + b) + (d0 + d1)) * 0.25f;
                }
            }
        t.levels.push_back(std::move(out));
    }
    if (t.half)
        for (MipLevel& l : t.levels)
            for (float& v : l.texels) v = roundHalf(v);
}
void setLevel0(Texture& t, const float* texels) {
    t.levels.assign(1, MipLevel{});
    t.levels[0].width = t.width;
    t.levels[0].height = t.height;
    t.levels[0].texels.assign(texels, texels + size_t(t.width) * t.height * 3);
    buildPyramid(t);
}
// A box-bounded brick grid (the medium; see the volume section below for how it is read).
struct Volume {
    bool present = false, hasTemperature = false;
    V3 bmin{0, 0, 0}, bmax{0, 0, 0}, albedo{0, 0, 0};
    U32 res[3] = {0, 0, 0}, nb[3] = {0, 0, 0};
    float scaleVox[3] = {0, 0, 0}, brickExtent[3] = {0, 0, 0}, minFeature = 0;
    float sigmaTScale = 0, g = 0, emissionScale = 0, rampMax = 1;
    U32 rampSize = 0;
    std::vector<float> ramp, majorant;
    std::vector<U32> table;  // per brick: its pool index, ~0u when empty
    std::vector<uint16_t> density, temperature;  // pools of 512 halves a brick, z, y, x within
    // The glowing medium as a light source (see emissionRegion in the volume section).
    bool hasRegion = false;
    int regionRes[3] = {0, 0, 0};
    U32 regionVisibility = 0;
    float distToWorld[12] = {}, worldToDist[12] = {};
    float invUnitVolume = 0;
    U32 conditionalX = 0, marginalY = 0, marginalZ = 0;
    std::vector<float> pool;
};
struct Scene {
    std::vector<V3> tri;  // 3 vertices per triangle, scene order
    std::vector<float> triUV;  // 6 per triangle, or empty: (0,0), (1,0), (1,1)
    std::vector<Texture> textures;
    std::vector<U32> geomOf, primOf, material;
    std::vector<uint8_t> singleSided;  // per scene triangle, empty when every mesh is two-sided
    std::vector<x2_material> materials;
    // Test hook: per-material roughness used only for the value of the GGX lobe (not for sampling, pdfs or MIS),
    // negative = the material's own. Central differences over it isolate d f / d roughness with the sampling fixed.
    std::vector<float> evalRoughness;
    std::vector<x2_light> lights;
    Volume volume;  // volume.present: the scene has a medium
    std::vector<BNode> nodes;
    std::vector<U32> order;
    U32 triangles = 0;
};
struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float* l, const float* h) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow(V3 p) {
        const float q[3] = {p.x, p.y, p.z};
        grow(q, q);
    }
    float area() const {
        if (lo[0] > hi[0]) return 0.0f;
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};
void buildBvh(Scene& s) {
    const U32 n = s.triangles;
    std::vector<Box> boxes(n);
    std::vector<float> centroid(size_t(n) * 3);
    for (U32 i = 0; i < n; ++i) {
        for (int v = 0; v < 3; ++v) boxes[i].grow(s.tri[size_t(i) * 3 + v]);
        for (int k = 0; k < 3; ++k) centroid[size_t(i) * 3 + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
    }
    s.order.resize(n);
    for (U32 i = 0; i < n; ++i) s.order[i] = i;
    s.nodes.clear();
    s.nodes.reserve(size_t(n) * 2 + 1);
    s.nodes.push_back(BNode{});
    struct Task { U32 node, first, count; };
    std::vector<Task> tasks{{0, 0, n}};
    constexpr int Bins = 16;
    while (!tasks.empty()) {
        const Task t = tasks.back();
        tasks.pop_back();
        Box bounds, cb;
        for (U32 i = t.first; i < t.first + t.count; ++i) {
            const U32 p = s.order[i];
            bounds.grow(boxes[p].lo, boxes[p].hi);
            cb.grow(&centroid[size_t(p) * 3], &centroid[size_t(p) * 3]);
        }
        BNode& node = s.nodes[t.node];
        for (int k = 0; k < 3; ++k) {
            node.lo[k] = t.count ? bounds.lo[k] : 0.0f;
            node.hi[k] = t.count ? bounds.hi[k] : 0.0f;
        }
        if (t.count <= 4) {
            s.nodes[t.node].a = t.count ? t.first : ~0u;  // only an empty scene's root holds nothing
            s.nodes[t.node].n = t.count;
            continue;
        }
        int axis = 0;
        float best = FLT_MAX;
        int bestSplit = -1;
        for (int k = 0; k < 3; ++k) {
            const float extent = cb.hi[k] - cb.lo[k];
            if (!(extent > 0.0f)) continue;
            Box bin[Bins];
            U32 cnt[Bins] = {};
            const float scale = float(Bins) / extent;
            for (U32 i = t.first; i < t.first + t.count; ++i) {
                const U32 p = s.order[i];
                int b = int((centroid[size_t(p) * 3 + k] - cb.lo[k]) * scale);
                b = std::min(std::max(b, 0), Bins - 1);
                ++cnt[b];
                bin[b].grow(boxes[p].lo, boxes[p].hi);
            }
            float leftArea[Bins];
            U32 leftCount[Bins];
            Box acc;
            U32 c = 0;
            for (int b = 0; b < Bins - 1; ++b) {
                acc.grow(bin[b].lo, bin[b].hi);
                c += cnt[b];
                leftArea[b] = acc.area();
                leftCount[b] = c;
            }
            acc = Box{};
            c = 0;
            for (int b = Bins - 1; b > 0; --b) {
                acc.grow(bin[b].lo, bin[b].hi);
                c += cnt[b];
                const U32 l = leftCount[b - 1];
                if (l == 0 || c == 0) continue;
                const float cost = leftArea[b - 1] * float(l) + acc.area() * float(c);
                if (cost < best) {
                    best = cost;
                    axis = k;
                    bestSplit = b;
                }
            }
        }
        U32 mid;
        if (bestSplit < 0) {
            mid = t.first + t.count / 2;  // all centroids alike: split the list in half
        } else {
            const float extent = cb.hi[axis] - cb.lo[axis], scale = float(Bins) / extent;
            auto it = std::partition(s.order.begin() + t.first, s.order.begin() + t.first + t.count, [&](U32 p) {
                int b = int((centroid[size_t(p) * 3 + axis] - cb.lo[axis]) * scale);
                b = std::min(std::max(b, 0), Bins - 1);
                return b < bestSplit;
            });
            mid = U32(it - s.order.begin());
            if (mid == t.first || mid == t.first + t.count) mid = t.first + t.count / 2;
        }
        const U32 left = U32(s.nodes.size());
        s.nodes.push_back(BNode{});
        s.nodes.push_back(BNode{});
        s.nodes[t.node].a = left;
        s.nodes[t.node].n = 0;
        tasks.push_back({left, t.first, mid - t.first});
        tasks.push_back({left + 1, mid, t.first + t.count - mid});
    }
}
// A slab test that never culls a box holding a triangle the exact ray meets (Ize 2013: the far distance scaled by
// 1 + 2 gamma(3)), with a further relative slack on the distances to compare against, so that triangles whose
// computed distance ties or nearly ties the best are still tested.
constexpr float FarScale = 1.0f + 2.0f * (3.0f * 0x1p-24f) / (1.0f - 3.0f * 0x1p-24f);
constexpr float Slack = 1.0f + 0x1p-12f;
bool boxVisit(const BNode& b, V3 org, V3 inv, const bool zero[3], float lo, float hi) {
    float tn = -FLT_MAX, tf = FLT_MAX;
    const float o[3] = {org.x, org.y, org.z}, iv[3] = {inv.x, inv.y, inv.z};
    for (int k = 0; k < 3; ++k) {
        if (zero[k]) {
            if (o[k] < b.lo[k] || o[k] > b.hi[k]) return false;
            continue;
        }
        float t0 = (b.lo[k] - o[k]) * iv[k], t1 = (b.hi[k] - o[k]) * iv[k];
        if (t0 > t1) std::swap(t0, t1);
        tn = std::max(tn, t0);
        tf = std::min(tf, t1);
    }
    tf = tf * (tf > 0 ? FarScale : 1.0f / FarScale);
    const float loS = lo > 0 ? lo / Slack : lo * Slack, hiS = hi > 0 ? hi * Slack : hi / Slack;
    return tn <= tf && tn <= hiS && tf >= loS;
}
bool keyLess(float t, U32 g, U32 p, float bt, U32 bg, U32 bp) {
    return t < bt || (t == bt && (g < bg || (g == bg && p < bp)));
}
// Closest hit (or any hit with anyHit) among triangles with t in [near, far]. Returns the scene triangle index.
U32 traverse(const Scene& s, V3 org, V3 dir, float nearT, float farT, bool anyHit, bool brute, float& tOut) {
    const WoopRay w = woopSetup(org, dir, nearT, farT);
    U32 bestTri = ~0u;
    float bestT = farT;
    U32 bg = ~0u, bp = ~0u;
    auto test = [&](U32 tri) {
        float t;
        const V3 a = s.tri[size_t(tri) * 3], b = s.tri[size_t(tri) * 3 + 1], c = s.tri[size_t(tri) * 3 + 2];
        if (!woopHit(w, a, b, c, t)) return false;
        // a single-sided mesh's back face is no hit (backFaceCullingFilter, async_cpu.h:122-133)
        if (!s.singleSided.empty() && s.singleSided[tri] && dot(dir, cross(b - a, c - a)) >= 0.0f) return false;
        if (bestTri == ~0u || keyLess(t, s.geomOf[tri], s.primOf[tri], bestT, bg, bp)) {
            bestTri = tri;
            bestT = t;
            bg = s.geomOf[tri];
            bp = s.primOf[tri];
        }
        return true;
    };
    if (brute || s.nodes.empty()) {
        for (U32 i = 0; i < s.triangles; ++i)
            if (test(i) && anyHit) break;
        tOut = bestT;
        return bestTri;
    }
    const bool zero[3] = {dir.x == 0.0f, dir.y == 0.0f, dir.z == 0.0f};
    const V3 inv{1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
    U32 stack[128];
    int top = 0;
    stack[top++] = 0;
    while (top) {
        const BNode& b = s.nodes[stack[--top]];
        if (b.n == 0 && b.a == ~0u) continue;
        if (!boxVisit(b, org, inv, zero, nearT, bestTri == ~0u ? farT : bestT)) continue;
        if (b.n) {
            for (U32 i = b.a; i < b.a + b.n; ++i)
                if (test(s.order[i]) && anyHit) {
                    tOut = bestT;
                    return bestTri;
                }
        } else {
            if (top + 2 > 128) return ~0u;  // cannot happen for the depths the builder makes
            stack[top++] = b.a + 1;
            stack[top++] = b.a;
        }
    }
    tOut = bestT;
    return bestTri;
}

// ---------------------------------------------------------------------------------------------------------------
// Lights (light_core.h): the rectangle and the constant environment, no texture, no spread, no clear radius.
constexpr float InfiniteDistance = 3.4028234663852886e+38f - 1e32f;
constexpr float EnvDistance = InfiniteDistance * 0.9f;
constexpr float HitEpsilonEnd = 1.0f - 1e-3f;
struct LightHit { V3 N; float distance; };
V3 rows(const float* m, V3 v) {  // light_core.h:197-199
    return ((load3(m) * v.x) + (load3(m + 3) * v.y)) + (load3(m + 6) * v.z);
}
V3 polar(float cosTheta, float sinTheta, float phi) {  // light_core.h:221-225
    float s, c;
    sincosPoly(phi, s, c);
    return {sinTheta * c, sinTheta * s, cosTheta};
}
float planeDistance(const x2_light& l, V3 p) {  // light_core.h:255-260
    float d = dot(load3(l.direction), p) + l.plane_offset;
    if (l.sidedness == X2_SIDE_REVERSE) d = -d;
    else if (l.sidedness == X2_SIDE_TWO) d = fabsf(d);
    return d;
}
float areaToSolidAngleScale(V3 wi, V3 n, float dist) {  // light_core.h:215-219
    const float denom = -dot(wi, n);
    return denom < 0.00001f ? 0.0f : (dist * dist) / denom;
}
// Light::sample (light_core.h:476-500 env, :523-574 rect).
bool lightSample(const x2_light& l, V3 p, const V3* n, const float r[3], V3& wi, LightHit& h) {
    if (l.type == X2_LIGHT_ENV) {
        if (l.hemispherical) {
            const float cosTheta = r[0] > 1.0f ? 1.0f : r[0];
            wi = rows(l.local_to_render, polar(cosTheta, sqrtf(1.0f - cosTheta * cosTheta), r[1] * TwoPi));
        } else {
            const float cosTheta = 1.0f - 2.0f * r[0];
            wi = polar(cosTheta, sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta)), r[1] * TwoPi);
        }
        if (n && dot(*n, wi) < Epsilon) return false;
        h.N = wi * -1.0f;
        h.distance = EnvDistance;
        return true;
    }
    if (l.type != X2_LIGHT_RECT) return false;
    const float distance = planeDistance(l, p);
    if (l.sidedness != X2_SIDE_TWO && distance < Epsilon) return false;
    const V3 local{(0.5f - r[0]) * (2 * l.half_width), (0.5f - r[1]) * (2 * l.half_height), 0.0f};
    wi = (rows(l.local_to_render, local) + load3(l.position)) - p;
    h.distance = sqrtf(dot(wi, wi));
    if (h.distance < Epsilon) return false;
    wi = wi * (1.0f / h.distance);
    if (n && dot(*n, wi) < Epsilon) return false;
    const V3 dir = load3(l.direction);
    h.N = dot(dir, wi) < 0.0f ? dir : dir * -1.0f;
    return true;
}
// Light::eval (light_core.h:740-752 env, :789-799 rect).
V3 lightEval(const x2_light& l, V3 wi, const LightHit& h, float& pdf) {
    const V3 radiance = load3(l.radiance) * V3{1, 1, 1};
    if (l.type == X2_LIGHT_ENV) {
        pdf = l.inv_area;
        return radiance;
    }
    pdf = l.inv_area;
    pdf *= areaToSolidAngleScale(wi, h.N, h.distance);
    return radiance;
}
// Light::intersect (light_core.h:813-847).
bool lightIntersect(const x2_light& l, V3 p, V3 wi, float maxDistance, LightHit& h) {
    if (l.type == X2_LIGHT_ENV) {
        if (l.hemispherical && dot(wi, load3(l.direction)) < 0.0f) return false;
        if (EnvDistance > maxDistance) return false;
        h.N = wi * -1.0f;
        h.distance = EnvDistance;
        return true;
    }
    if (l.type != X2_LIGHT_RECT) return false;
    const V3 localWi = rows(l.render_to_local_rot, wi), localP = rows(l.render_to_local, p - load3(l.position));
    // rejectPosDirSidedness (light_core.h:261-264)
    if ((localP.z < 0.0f) == (localWi.z < 0.0f)) return false;
    if (localP.z < 0.0f ? l.sidedness == X2_SIDE_REGULAR : l.sidedness == X2_SIDE_REVERSE) return false;
    const float localDistance = -localP.z / localWi.z, renderDistance = l.scale * localDistance;
    if (renderDistance > maxDistance) return false;
    const V3 hit = localP + localWi * localDistance;
    if (fabsf(hit.x) >= l.half_width || fabsf(hit.y) >= l.half_height) return false;
    const V3 dir = load3(l.direction);
    h.N = localWi.z < 0.0f ? dir : dir * -1.0f;
    h.distance = renderDistance;
    return true;
}
// Light::canIlluminate (light_core.h:1024-1060), no subsurface radius.
bool canIlluminate(const x2_light& l, V3 p, const V3* n) {
    const float threshold = Epsilon - 0.0f;
    if (l.type == X2_LIGHT_ENV) return true;
    if (l.sidedness != X2_SIDE_TWO && planeDistance(l, p) < threshold) return false;
    if (!n) return true;
    const float c[4][2] = {{l.half_width, l.half_height}, {l.half_width, -l.half_height},
                           {-l.half_width, l.half_height}, {-l.half_width, -l.half_height}};
    for (int i = 0; i < 4; ++i)
        if (dot(*n, (rows(l.local_to_render, V3{c[i][0], c[i][1], 0.0f}) + load3(l.position)) - p) > threshold)
            return true;
    return false;
}
// lightShadowSpan (async_integrator.h:305-311), no clear radius, min or max shadow distance.
void lightShadowSpan(float distance, float& nearT, float& farT) {
    nearT = fmaxf(nearT, 0.0f);
    farT = fmaxf((distance - 0.0f) * HitEpsilonEnd, 2.0f * nearT);
}
float powerHeuristic(float f, float g) { return (f * f) / (f * f + g * g); }  // light_core.h:211

// ---------------------------------------------------------------------------------------------------------------
// Lobes (shade_lobes.h:205-282, 481-556; shade_lobes_eval.h:11, 93-146, 609-668; shade_math.h:135-262).
enum : U32 { Reflection = 1u, Transmission = 2u, Diffuse = 4u, Glossy = 16u, Mirror = 64u };
constexpr U32 AllSurfaceSides = Reflection | Transmission, AllLobes = Diffuse | Glossy | Mirror;
constexpr U32 AllLobeTypes = AllSurfaceSides | AllLobes;
bool matchesFlags(U32 type, U32 flags) { return (type & AllSurfaceSides & flags) && (type & AllLobes & flags); }
U32 lobeRayMask(U32 t) {  // light_core.h:189-194
    U32 mask = 0;
    if (t & 1u) mask |= t & 0x54u;
    if (t & 2u) mask |= (t & 0x54u) << 1;
    return mask;
}
struct Frame { V3 x, y, z; };
Frame frameOf(V3 N) {  // shade_math.h:138-141 (Duff et al.)
    const float sign = copysignf(1.0f, N.z), a = -1.0f / (sign + N.z), b = N.x * N.y * a;
    return {{1.0f + sign * N.x * N.x * a, sign * b, -sign * N.x}, {b, sign + N.y * N.y * a, -N.y}, N};
}
V3 localToGlobal(const Frame& f, V3 v) { return (f.x * v.x + f.y * v.y) + f.z * v.z; }
V3 localSpherical(float cosTheta, float sinTheta, float phi) {
    float s, c;
    sincosPoly(phi, s, c);
    return {sinTheta * c, sinTheta * s, cosTheta};
}
enum LobeKind : U32 { KindLambert = 0, KindGGX = 1 };
struct Lobe {
    U32 kind, type;
    V3 scale, albedo;
    Frame frame;
    float roughness;  // GGX: max(0.001, r^2)
    float evalRoughness;  // the same, for the lobe's value only (the evalRoughness test hook)
};
struct Slice { V3 Ng, wo; U32 flags; };
struct Bsdf {
    Lobe lobes[2];
    U32 count = 0, type = 0;
    V3 selfEmission{0, 0, 0};
};
V3 evalLambert(const Lobe& l, V3 wi, float& pdf) {  // shade_lobes_eval.h:611-620
    const V3 N = l.frame.z;
    const float cosWi = fmaxf(dot(N, wi), 0.0f);
    pdf = cosWi * OneOverPi;
    return (((1.0f * l.albedo) * l.scale) * OneOverPi) * cosWi;
}
float widenRoughness(float roughness, float cosNO) { return (1.2f - 0.2f * sqrtf(cosNO)) * roughness; }
bool halfVector(V3 wo, V3 wi, V3& H) {  // shade_math.h:222-230
    H = wo + wi;
    const float l = length(H);
    if (l > 0.001f) {
        H = H * (1.0f / l);
        return true;
    }
    return false;
}
V3 reflect(V3 N, V3 wo) {  // shade_math.h:190-194
    const float NdotWo = dot(N, wo);
    return -1.0f * wo + 2.0f * NdotWo * N;
}
V3 evalGGX(const Lobe& l, const Slice& s, V3 wi, float& pdf) {  // shade_lobes_eval.h:93-121, 131-135
    const V3 N = l.frame.z;
    const float cosNO = dot(N, s.wo), cosNI = dot(N, wi), roughness = l.roughness;
    pdf = 0.0f;
    if (cosNO <= 0.0f || cosNI <= 0.0f) return {0, 0, 0};
    V3 m;
    if (!halfVector(s.wo, wi, m)) return {0, 0, 0};
    const float cosMI = dot(m, wi);
    if (cosMI <= 0.0f) return {0, 0, 0};
    const float cosThetaM = dot(m, l.frame.z);
    if (cosThetaM <= 0.0f) return {0, 0, 0};
    float alpha2 = l.evalRoughness * l.evalRoughness;
    const float cosThetaM2 = cosThetaM * cosThetaM, piCosThetaM4 = Pi * cosThetaM2 * cosThetaM2;
    const float tanThetaM2 = (1.0f - cosThetaM2) / cosThetaM2;
    float tmp = alpha2 + tanThetaM2, D4 = 0.25f * alpha2 / (piCosThetaM4 * tmp * tmp);
    const float invCosNO = 1.0f / cosNO, invCosNI = 1.0f / cosNI;
    const float G1o = 2.0f / (1.0f + sqrtf(1.0f + alpha2 * (1.0f - cosNO * cosNO) * (invCosNO * invCosNO)));
    const float G1i = 2.0f / (1.0f + sqrtf(1.0f + alpha2 * (1.0f - cosNI * cosNI) * (invCosNI * invCosNI)));
    (void)invCosNI;
    const V3 fr = l.scale * G1o * G1i * D4 * invCosNO * 1.0f;
    const float w2 = 0.0f;  // no energy compensation: favg black (shade_lobes_eval.h:23-26)
    const float alpha = widenRoughness(roughness, cosNO);
    alpha2 = alpha * alpha;
    tmp = alpha2 + tanThetaM2;
    D4 = 0.25f * alpha2 / (piCosThetaM4 * tmp * tmp);
    const float p2 = 0.0f, p1 = D4 * cosThetaM / cosMI;
    pdf = (1.0f - w2) * p1 + w2 * p2;
    return 1.0f * fr;  // shadowTerminatorFix off
}
V3 sampleGGXDir(const Slice& s, float r1, float r2, float roughness, float cosNO, const Frame& f) {  // :122-130
    const float alpha = widenRoughness(roughness, cosNO), alpha2 = alpha * alpha;
    const float tanThetaM2 = alpha2 * r1 / (1.0f - r1), cosThetaM = 1.0f / sqrtf(1.0f + tanThetaM2);
    const float sinThetaM = cosThetaM * sqrtf(tanThetaM2);
    const V3 m = localToGlobal(f, localSpherical(cosThetaM, sinThetaM, 2.0f * Pi * r2));
    return reflect(m, s.wo);
}
V3 lobeEval(const Lobe& l, const Slice& s, V3 wi, float& pdf) {
    return l.kind == KindLambert ? evalLambert(l, wi, pdf) : evalGGX(l, s, wi, pdf);
}
V3 lobeSample(const Lobe& l, const Slice& s, float r1, float r2, V3& wi, float& pdf) {  // :657-690
    if (l.kind == KindLambert) {
        const float cosTheta = r1 > 1.0f ? 1.0f : sqrtf(r1), sinTheta = sqrtf(1.0f - r1);
        wi = localToGlobal(l.frame, localSpherical(cosTheta, sinTheta, r2 * TwoPi));
        return evalLambert(l, wi, pdf);
    }
    const float cosNO = dot(l.frame.z, s.wo);
    if (cosNO <= 0.0f) {
        wi = {0, 0, 0};
        pdf = 0.0f;
        return {0, 0, 0};
    }
    const float w2 = 0.0f;
    wi = sampleGGXDir(s, (r1 - w2) / (1.0f - w2), r2, l.roughness, cosNO, l.frame);
    return evalGGX(l, s, wi, pdf);
}
U32 surfaceFlags(const Slice& s, V3 wi) {  // shade_lobes.h:323-326
    return dot(s.Ng, s.wo) * dot(s.Ng, wi) > 0.0f ? (s.flags & ~U32(Transmission)) : (s.flags & ~U32(Reflection));
}
// The material's Bsdf at a hit: a Lambert lobe and a GGX lobe about the shading normal (lambertLobe,
// cookTorranceLobe with KindGGX), in that order.
void buildBsdf(const x2_material& m, V3 N, Bsdf& b, V3 albedo, float evalRoughness = -1.0f) {
    b.count = 0;
    b.type = 0;
    const Frame f = frameOf(N);
    if (m.diffuse_weight > 0.0f) {
        Lobe& l = b.lobes[b.count++];
        l.kind = KindLambert;
        l.type = Reflection | Diffuse;
        l.frame = f;
        l.albedo = albedo * m.diffuse_weight;
        l.scale = {1, 1, 1};
        l.roughness = 1.0f;
        l.evalRoughness = 1.0f;
        b.type |= l.type;
    }
    if (m.specular_weight > 0.0f) {
        Lobe& l = b.lobes[b.count++];
        l.kind = KindGGX;
        l.type = Reflection | Glossy;
        l.frame = f;
        l.albedo = {0, 0, 0};
        l.scale = {m.specular_weight, m.specular_weight, m.specular_weight};
        l.roughness = fmaxf(0.001f, m.roughness * m.roughness);
        l.evalRoughness = evalRoughness < 0.0f ? l.roughness : fmaxf(0.001f, evalRoughness * evalRoughness);
        b.type |= l.type;
    }
    b.selfEmission = load3(m.emission);
}

U32 wrapTexel(int x, U32 n) {
    const int m = x % int(n);
    return U32(m < 0 ? m + int(n) : m);
}
V3 texelAt(const MipLevel& l, U32 x, U32 y) { return load3(&l.texels[(size_t(y) * l.width + x) * 3]); }
// A texel of level `level`, or with coarseFrom set its ancestor there (the parent walk of async_gpu_device.h:52-58,
// the coordinates halved a level at a time and clamped to the level read).
V3 fetchTexel(const Texture& t, U32 level, U32 x, U32 y) {
    if (level >= t.coarseFrom || t.coarseFrom >= t.levels.size()) return texelAt(t.levels[level], x, y);
    U32 k = level;
    while (k < t.coarseFrom) {
        ++k;
        x >>= 1;
        y >>= 1;
    }
    const MipLevel& u = t.levels[k];
    return texelAt(u, x < u.width ? x : u.width - 1, y < u.height ? y : u.height - 1);
}
// A map lookup's footprint (tex::Footprint: s, t and their screen derivatives) and mode (0 anisotropic, 1 trilinear
// closest texels, 2 one level's closest texel).
struct Footprint {
    float u, v, dudx, dvdx, dudy, dvdy;
};
// texture_core.h:144-153: the four texels about (u, v) of a level, weighted bilinearly; each fetch wraps.
bool bilinearAt(const Texture& tex, U32 level, float u, float v, V3& out) {
    const MipLevel& m = tex.levels[level];
    float s = u * float(m.width) - .5f, t = v * float(m.height) - .5f;
    if (!finite(s) || !finite(t) || fabsf(s) > 1.e8f || fabsf(t) > 1.e8f) return false;
    const int x = int(floorf(s)), y = int(floorf(t));
    s -= float(x);
    t -= float(y);
    const U32 x0 = wrapTexel(x, m.width), x1 = wrapTexel(x + 1, m.width);
    const U32 y0 = wrapTexel(y, m.height), y1 = wrapTexel(y + 1, m.height);
    const V3 a = fetchTexel(tex, level, x0, y0), b = fetchTexel(tex, level, x1, y0), c = fetchTexel(tex, level, x0, y1),
             d = fetchTexel(tex, level, x1, y1);
    out = ((a * (1 - s)) + (b * s)) * (1 - t) + ((c * (1 - s)) + (d * s)) * t;
    return true;
}
// texture_core.h:155-160: the texel nearest (u, v); an exact half-texel tie takes the lower one.
bool closestAt(const MipLevel& m, float u, float v, U32& x, U32& y) {
    const float s = u * float(m.width) - .5f, t = v * float(m.height) - .5f;
    if (!finite(s) || !finite(t) || fabsf(s) > 1.e8f || fabsf(t) > 1.e8f) return false;
    int ix = int(floorf(s)), iy = int(floorf(t));
    if (s - float(ix) > .5f) ++ix;
    if (t - float(iy) > .5f) ++iy;
    x = wrapTexel(ix, m.width);
    y = wrapTexel(iy, m.height);
    return true;
}
// OpenImageIO's fast_atan2 and fast_exp (fmath.h), as texture_core.h:94-97 gives them.
float oiioAtan2(float y, float x) {
    const float a = fabsf(x), b = fabsf(y);
    const float k = b == 0 ? 0.f : (a == b ? 1.f : (b > a ? a / b : b / a));
    const float s = 1.f - (1.f - k), t = s * s;
    float r = s * (0.430165678f * t + 1.f) / ((0.0579354987f * t + 0.763007998f) * t + 1.f);
    if (b > a) r = 1.5707963267948966f - r;
    if (copysignf(1.f, x) < 0) r = Pi - r;
    return copysignf(r, y);
}
float oiioExp(float a) {
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
// The filter's plan (texture_core.h:161-206 up to the reads): the levels and their weights, and for the anisotropic
// filter the taps along the major axis. Shared by the lookup and the adjoint's gradient scatter.
struct TexPlan {
    float s, t;
    int levels[2];
    float lw[2];
    int n;           // taps a level (anisotropic), or 0: one closest texel a level
    float sm, tm;    // the major axis' step
    float pos[7], weight[7];
};
// swap: the cross terms swapped (the path tracer's maps); the prototype estimator's binding passes them in order.
bool planLookup(const Texture& tex, U32 mode, const Footprint& f, TexPlan& p, bool swap = true) {
    const U32 count = U32(tex.levels.size());
    if (!count || count > MipMax) return false;
    float s = f.u, t = 1 - f.v;
    float dx = f.dudx, tx = -f.dvdx, dy = f.dudy, ty = -f.dvdy;
    if (swap) {  // swapCross (XPU's vector shading)
        const float c = tx;
        tx = dy;
        dy = c;
    }
    if (!finite(s) || !finite(t) || !finite(dx) || !finite(tx) || !finite(dy) || !finite(ty)) return false;
    if (fabsf(s) > 1024) s = s - floorf(s);
    if (fabsf(t) > 1024) t = t - floorf(t);
    p.s = s;
    p.t = t;
    p.levels[0] = p.levels[1] = -1;
    p.lw[0] = p.lw[1] = 0;
    p.n = 0;
    p.sm = p.tm = 0;
    if (mode != 0) {  // texture_core.h:162-177
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
        const float filt = fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fmaxf(fabsf(tx), fabsf(ty)));
        for (U32 m = 0; m < count; ++m) {
            const MipLevel& l = tex.levels[m];
            const float w = filt * float(l.width < l.height ? l.width : l.height);
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
        } else if (mode == 2) {
            p.levels[0] = p.levels[1];
            p.lw[0] = 1;
            p.lw[1] = 0;
        }
        return true;
    }
    // The anisotropic filter (texture_core.h:187-206).
    const double A = double(tx * tx + ty * ty), B = -2.0 * double(dx * tx + dy * ty), C = double(dx * dx + dy * dy);
    const double root = sqrt((A - C) * (A - C) + B * B);
    float major = fminf(sqrtf(fmaxf(float((A + C + root) * .5), 0.f)), 1000.f);
    float minor = fminf(sqrtf(fmaxf(float((A + C - root) * .5), 0.f)), 1000.f);
    const float theta = oiioAtan2(float(B), float(A - C)) * .5f + Pi * .5f;
    float aspect = minor > 0 ? fminf(fmaxf(major / minor, 1.f), 1.e6f) : (major > 0 ? 1.e6f : 1.f);
    if (aspect > 4) {
        aspect = 4;
        major = .5f * (major + minor * 4);
        minor = major / 4;
    }
    for (U32 m = 0; m < count; ++m) {
        const MipLevel& l = tex.levels[m];
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
        const MipLevel& l = tex.levels[0];
        const U32 r = l.width > l.height ? l.width : l.height;
        if (minor * float(r) < .5f) aspect = fminf(fmaxf(major * float(r) * 2, 1.f), 4.f);
    }
    int n = int(2 * aspect - 1);
    if (n < 1) n = 1;
    if (n > 7) return false;
    const float LL = 2 * (major - minor);
    float sn, cs;
    sincosPoly(theta, sn, cs);
    p.sm = cs * LL * .5f;
    p.tm = sn * LL * .5f;
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
// The map's value (texture_core.h:161-206 after the plan): false, an invalid value (the sample is an Error).
bool mapLookup(const Texture& tex, U32 mode, const Footprint& f, V3& out, bool swap = true) {
    TexPlan p;
    if (!planLookup(tex, mode, f, p, swap)) return false;
    V3 result{0, 0, 0};
    for (int k = 0; k < 2; ++k) {
        if (p.lw[k] == 0) continue;
        const MipLevel& l = tex.levels[size_t(p.levels[k])];
        if (p.n == 0) {
            U32 x, y;
            if (!closestAt(l, p.s, p.t, x, y)) return false;
            const V3 value = fetchTexel(tex, U32(p.levels[k]), x, y);
            if (!finite3(value)) return false;
            result = result + value * p.lw[k];
            continue;
        }
        V3 acc{0, 0, 0};
        for (int i = 0; i < p.n; ++i) {
            V3 value;
            if (!bilinearAt(tex, U32(p.levels[k]), p.s + p.pos[i] * p.sm, p.t + p.pos[i] * p.tm, value)) return false;
            acc = acc + value * p.weight[i];
        }
        result = result + acc * p.lw[k];
    }
    if (!finite3(result)) return false;
    out = result;
    return true;
}
// Every texel the lookup reads, with its weight in the value (the adjoint's d value / d texel): visit(level, x, y, w).
template <class Visit>
void mapWeights(const Texture& tex, const TexPlan& p, Visit&& visit) {
    for (int k = 0; k < 2; ++k) {
        if (p.lw[k] == 0) continue;
        const U32 level = U32(p.levels[k]);
        const MipLevel& l = tex.levels[level];
        if (p.n == 0) {
            U32 x, y;
            if (closestAt(l, p.s, p.t, x, y)) visit(level, x, y, p.lw[k]);
            continue;
        }
        for (int i = 0; i < p.n; ++i) {
            float s = (p.s + p.pos[i] * p.sm) * float(l.width) - .5f, t = (p.t + p.pos[i] * p.tm) * float(l.height) - .5f;
            if (!finite(s) || !finite(t) || fabsf(s) > 1.e8f || fabsf(t) > 1.e8f) continue;
            const int x = int(floorf(s)), y = int(floorf(t));
            s -= float(x);
            t -= float(y);
            const U32 x0 = wrapTexel(x, l.width), x1 = wrapTexel(x + 1, l.width);
            const U32 y0 = wrapTexel(y, l.height), y1 = wrapTexel(y + 1, l.height);
            const float w = p.weight[i] * p.lw[k];
            visit(level, x0, y0, (1 - s) * (1 - t) * w);
            visit(level, x1, y0, s * (1 - t) * w);
            visit(level, x0, y1, (1 - s) * t * w);
            visit(level, x1, y1, s * t * w);
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Ray differentials (shade_state.h:30-33 RayDiff; shade_material.h:388-505 the hit's texture derivatives and the
// transfer; shade_material.h:691-705 and shade_lobes_eval.h:725-741 the scattered ones), for Lambert and GGX lobes
// on flat triangles (no vertex normals: dN/dx = dN/dy = 0).
struct RayDiff {
    V3 dOx{0, 0, 0}, dOy{0, 0, 0}, dDx{0, 0, 0}, dDy{0, 0, 0};
    U32 valid = 0;
};
float differenceOfProducts(float a, float b, float c, float d) {  // Kahan's, shade_material.h:96-101
    const float cd = c * d;
    const float err = fmaf(-c, d, cd);
    const float dop = fmaf(a, b, -cd);
    return dop + err;
}
// shade_material.h:103-115: dP/ds and dP/dt of a triangle from its corners and texture coordinates.
bool trianglePartials(V3 x0, V3 x1, V3 x2, const float uv0[2], const float uv1[2], const float uv2[2], V3& dPds, V3& dPdt) {
    const float A[4] = {uv1[0] - uv0[0], uv1[1] - uv0[1], uv2[0] - uv0[0], uv2[1] - uv0[1]};
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
V3 normalized(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }  // shade::normalize
// What a hit knows of its footprint: the texture coordinates' derivatives, dP/dx and dP/dy, whether they hold,
// and the incoming differentials (the scattered ones start from them).
struct HitDiff {
    float dsdx = 0, dtdx = 0, dsdy = 0, dtdy = 0;
    V3 dPdx{0, 0, 0}, dPdy{0, 0, 0};
    bool valid = false;
    RayDiff incoming;
};
// shade_material.h:392-432 (hitTexDerivatives): a camera ray's auxiliary rays (directions dx, dy; a pinhole's
// offsets are zero) meet the triangle's plane; their barycentric offsets map through the texture coordinates.
void hitTexDerivatives(V3 a, V3 b, V3 c, const float st[6], const Ray& ray, V3 camDx, V3 camDy, U32 spp, HitDiff& h) {
    h.dsdx = h.dsdy = h.dtdx = h.dtdy = 0.0f;
    if (isExactlyZero(camDx) || isExactlyZero(camDy) || !spp) return;
    const V3 e1 = b - a, e2 = c - a, n = cross(e1, e2);
    U32 axis = fabsf(n.x) > fabsf(n.y) ? 0 : 1;
    if (fabsf(n.z) > (axis == 0 ? fabsf(n.x) : fabsf(n.y))) axis = 2;
    auto pa = [&](V3 v) { return axis == 0 ? v.y : axis == 1 ? v.z : v.x; };
    auto pb = [&](V3 v) { return axis == 0 ? v.z : axis == 1 ? v.x : v.y; };
    const float det = pa(e1) * pb(e2) - pb(e1) * pa(e2), den = dot(n, ray.direction);
    if (!(fabsf(det) > 0) || !(fabsf(den) > 0)) return;
    const float plane = dot(n, a - ray.origin);
    const V3 primary = ray.direction * (plane / den);
    const float scale = 1.f / fminf(sqrtf(float(spp)), 2.f);
    const V3 directions[2] = {camDx, camDy};
    float ds[2], dt[2];
    for (int i = 0; i < 2; ++i) {
        V3 d = directions[i];
        if (scale < 1) d = normalized(ray.direction + (d - ray.direction) * scale);
        const float dd = dot(n, d);
        if (!(fabsf(dd) > 0)) return;
        const V3 o = V3{0, 0, 0} * scale;
        const V3 delta = o + d * (dot(n, a - ray.origin - o) / dd) - primary;
        const float x = (pa(delta) * pb(e2) - pb(delta) * pa(e2)) / det, y = (pa(e1) * pb(delta) - pb(e1) * pa(delta)) / det;
        ds[i] = (st[2] - st[0]) * x + (st[4] - st[0]) * y;
        dt[i] = (st[3] - st[1]) * x + (st[5] - st[1]) * y;
    }
    if (!(finite(ds[0]) && finite(ds[1]) && finite(dt[0]) && finite(dt[1]))) return;
    h.dsdx = ds[0];
    h.dtdx = dt[0];
    h.dsdy = ds[1];
    h.dtdy = dt[1];
}
// shade_material.h:436-451 (transferDifferentials): the incoming auxiliary rays meet the hit's plane.
void transferDifferentials(const RayDiff& d, V3 origin, V3 dir, V3 Ng, V3 P, HitDiff& h) {
    h.dPdx = h.dPdy = V3{0, 0, 0};
    h.valid = false;
    if (!d.valid) return;
    const V3 oX = origin + d.dOx, oY = origin + d.dOy, dX = dir + d.dDx, dY = dir + d.dDy;
    const float denomX = dot(Ng, dX), denomY = dot(Ng, dY);
    if (denomX == 0.0f || denomY == 0.0f) return;
    const float planeD = -dot(Ng, P);
    const float distX = -(dot(oX, Ng) + planeD) / denomX, distY = -(dot(oY, Ng) + planeD) / denomY;
    h.dPdx = (oX + dX * distX) - P;
    h.dPdy = (oY + dY * distY) - P;
    h.valid = finite3(h.dPdx) && finite3(h.dPdy);
    if (!h.valid) h.dPdx = h.dPdy = V3{0, 0, 0};
}
bool isZeroE(float x, float e) { return fabsf(x) <= e; }
// shade_material.h:454-495 (differentialTexDerivatives): the texture derivatives from dP/dx and dP/dy over dP/ds and
// dP/dt; without differentials, the coarsest level (1024).
void differentialTexDerivatives(V3 dPds, V3 dPdt, V3 Ng, HitDiff& h) {
    h.dsdx = h.dtdx = h.dsdy = h.dtdy = 0.0f;
    if (!h.valid) {
        h.dsdx = 1024.0f;
        h.dtdy = 1024.0f;
        return;
    }
    const float e = 1e-5f;
    if (isZeroE(dot(dPds, dPdt), e)) {
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
    if (isZeroE(det, e)) return;
    const float detA = differenceOfProducts(A[0], A[3], A[1], A[2]);
    if (!(fabsf(detA) >= 1.17549435e-38f) || !(fabsf(detA) <= 3.402823466e38f)) return;
    const float inv = 1.0f / detA;
    h.dsdx = differenceOfProducts(bx[0], A[3], bx[1], A[1]) * inv;
    h.dtdx = differenceOfProducts(bx[1], A[0], bx[0], A[2]) * inv;
    h.dsdy = differenceOfProducts(by[0], A[3], by[1], A[1]) * inv;
    h.dtdy = differenceOfProducts(by[1], A[0], by[0], A[2]) * inv;
}
// The hit's differentials (shade_material.h:530-558 hitDifferentials for a mesh without subdivision derivatives):
// a camera ray's texture derivatives by hitTexDerivatives, its auxiliary rays scaled for the samples per pixel as
// the incoming differentials; any other ray's transferred and turned into texture derivatives.
void hitDifferentials(V3 a, V3 b, V3 c, const float st[6], V3 N, V3 Ng, V3 P, const Ray& ray, bool camera, V3 camDx,
                      V3 camDy, const RayDiff& incoming, U32 spp, HitDiff& h) {
    if (camera && !isExactlyZero(camDx) && !isExactlyZero(camDy) && spp) {
        hitTexDerivatives(a, b, c, st, ray, camDx, camDy, spp, h);
        const float scale = 1.f / fminf(sqrtf(float(spp)), 2.f);
        RayDiff d;
        d.dOx = V3{0, 0, 0} * scale;
        d.dOy = V3{0, 0, 0} * scale;
        d.dDx = normalized(ray.direction + (camDx - ray.direction) * scale) - ray.direction;
        d.dDy = normalized(ray.direction + (camDy - ray.direction) * scale) - ray.direction;
        d.valid = 1;
        h.incoming = d;
        transferDifferentials(d, ray.origin, ray.direction, Ng, P, h);
        return;
    }
    h.incoming = incoming;
    transferDifferentials(incoming, ray.origin, ray.direction, Ng, P, h);
    V3 dPds, dPdt;
    if (!trianglePartials(a, b, c, st, st + 2, st + 4, dPds, dPdt)) {
        const Frame f = frameOf(N);
        dPds = f.x;
        dPdt = f.y;
    }
    differentialTexDerivatives(dPds, dPdt, Ng, h);
}
// shade_math.h:279-288: a cosine-sampled direction's differentials in the lobe's frame (r1 clamped to 85 degrees
// from the normal; the polynomial's sine and cosine of phi).
void localHemisphereCosineDifferentials(float r1, float r2, V3& dDdr1, V3& dDdr2) {
    const float c85 = 0.08715580403804779f;  // cosf(85 * (pi / 180)) in binary32
    const float minR1 = c85 * c85, maxR1 = 1.0f - minR1;
    r1 = fminf(fmaxf(r1, minR1), maxR1);
    const float cosTheta = r1 > 1.0f ? 1.0f : sqrtf(r1), sinTheta = sqrtf(1.0f - r1);
    float sinPhi, cosPhi;
    sincosPoly(r2 * TwoPi, sinPhi, cosPhi);
    const float scale = 1.0f / 8.0f, halfScale = 0.5f * scale, twoPiScale = TwoPi * scale;
    dDdr1 = {-halfScale * cosPhi / sinTheta, -halfScale * sinPhi / sinTheta, halfScale / cosTheta};
    dDdr2 = {-twoPiScale * sinPhi * sinTheta, twoPiScale * cosPhi * sinTheta, 0.0f};
}
// shade_math.h:185-189.
void squarifyRectangle(V3& e1, V3& e2) {
    const float l1 = length(e1), l2 = length(e2), l = sqrtf(l1 * l2);
    e1 = e1 * (l / l1);
    e2 = e2 * (l / l2);
}
// A scattered ray's differentials (shade_material.h:691-705 with shade_lobes_eval.h:725-741): Lambert's from the
// cosine sample (it ignores the incoming ones, so dP/dx is scaled by the primary scale), GGX's the incoming ones
// reflected about the half vector (shade_math.h:203-207) and widened by the lobe's dDFactor, 1 + 7 alpha
// (shade_lobes.h:479, :548). Invalid where the hit's are.
RayDiff scatterDifferentials(bool lambert, const Frame& frame, float alpha, const HitDiff& h, V3 P, V3 wo, V3 wi,
                             float r1, float r2, V3 childOrigin, float primaryScale) {
    RayDiff d;
    if (!h.valid) return d;
    V3 dDdx = h.incoming.dDx, dDdy = h.incoming.dDy;
    if (lambert) {
        localHemisphereCosineDifferentials(r1, r2, dDdx, dDdy);
        squarifyRectangle(dDdx, dDdy);
        dDdx = localToGlobal(frame, dDdx);
        dDdy = localToGlobal(frame, dDdy);
    } else {
        V3 H;
        if (!halfVector(wo, wi, H)) H = frame.z;
        const V3 dNdx{0, 0, 0}, dNdy{0, 0, 0};
        const float Cx = dot(dDdx, H) - dot(wo, dNdx), Cy = dot(dDdy, H) - dot(wo, dNdy), dDotN = -dot(wo, H);
        dDdx = dDdx - (2.0f * (dDotN * dNdx + Cx * H));
        dDdy = dDdy - (2.0f * (dDotN * dNdy + Cy * H));
        const float factor = 1.0f + alpha * 7.0f;
        dDdx = dDdx * factor;
        dDdy = dDdy * factor;
    }
    const float scale = lambert ? primaryScale : 1.0f;
    const V3 shift = P - childOrigin;
    d.dOx = h.dPdx * scale + shift;
    d.dOy = h.dPdy * scale + shift;
    d.dDx = normalized(wi + dDdx * scale) - wi;
    d.dDy = normalized(wi + dDdy * scale) - wi;
    d.valid = finite3(d.dOx) && finite3(d.dOy) && finite3(d.dDx) && finite3(d.dDy) ? 1u : 0u;
    return d;
}
// d ln f / d r of the GGX lobe for MoonRay's input roughness r (alpha = r^2, A = alpha^2 = r^4; 0 where alpha is
// clamped): D4 ~ A / (A + tan^2 theta_m)^2, G1 = 2 / (1 + sqrt(1 + A tan^2 theta)) for wo and wi.
float ggxRoughnessLogDerivative(float r, V3 N, V3 wo, V3 wi) {
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
// The adjoint's gradient sinks and open subtrees (the device's bookkeeping, accumulated here in FP64).
enum : U32 { SinkNone = 0, SinkTexture = 1, SinkRoughness = 2 };
// A texture sink keeps its lookup (texture, mode, footprint) and value; the scatter walks the lookup's texels again.
struct Sink {
    U32 kind = SinkNone, material = 0;
    float g = 0;
    U32 tex = 0, mode = 0;
    Footprint f{0, 0, 0, 0, 0, 0};
    V3 color{0, 0, 0};
};
struct Record {
    U32 pos;
    V3 mark;
    Sink sink;
};
struct Adjoint {
    const float* dImage = nullptr;  // per pixel rgb
    double* gradTextures = nullptr;  // level 0 of every texture
    double* gradLevels = nullptr;    // the levels above (Texture::levelGrad), pushed down to level 0 at the end
    double* gradRoughness = nullptr;
};
// The contribution R of a lobe whose albedo the map scales has d R / d value = R / value (per channel), and the
// value is linear in the texels the lookup reads (mapWeights).
void scatterTexture(const Scene& S, const Adjoint& A, const Sink& s, V3 R, V3 delta) {
    const Texture& d = S.textures[s.tex];
    const float cs[3] = {s.color.x > 0.0f ? delta.x * R.x / s.color.x : 0.0f,
                         s.color.y > 0.0f ? delta.y * R.y / s.color.y : 0.0f,
                         s.color.z > 0.0f ? delta.z * R.z / s.color.z : 0.0f};
    if (cs[0] == 0.0f && cs[1] == 0.0f && cs[2] == 0.0f) return;
    TexPlan p;
    if (!planLookup(d, s.mode, s.f, p)) return;
    mapWeights(d, p, [&](U32 level, U32 x, U32 y, float w) {
        const MipLevel& l = d.levels[level];
        double* q = (level == 0 ? A.gradTextures + d.gradOffset : A.gradLevels + d.levelGrad[level]) +
                    (size_t(y) * l.width + x) * 3;
        for (int c = 0; c < 3; ++c) q[c] += double(cs[c] * w);
    });
}
// The pyramid's transpose: each level's texel gradients go to the 2 x 2 texels below that made it (buildPyramid),
// a quarter each, from the coarsest level down to level 0.
void pushDownPyramid(const Texture& t, double* level0, double* levels) {
    for (size_t k = t.levels.size() - 1; k >= 1; --k) {
        const MipLevel& up = t.levels[k];
        const MipLevel& in = t.levels[k - 1];
        const double* g = levels + t.levelGrad[k];
        double* below = k == 1 ? level0 : levels + t.levelGrad[k - 1];
        for (U32 y = 0; y < up.height; ++y)
            for (U32 x = 0; x < up.width; ++x) {
                const U32 xs[2] = {std::min(2 * x, in.width - 1), std::min(2 * x + 1, in.width - 1)};
                const U32 ys[2] = {std::min(2 * y, in.height - 1), std::min(2 * y + 1, in.height - 1)};
                for (int j = 0; j < 2; ++j)
                    for (int i = 0; i < 2; ++i)
                        for (int c = 0; c < 3; ++c)
                            below[(size_t(ys[j]) * in.width + xs[i]) * 3 + c] += 0.25 * g[(size_t(y) * up.width + x) * 3 + c];
            }
    }
}
void scatterSink(const Scene& S, const Adjoint& A, const Sink& s, V3 R, V3 delta) {
    if (s.kind == SinkTexture) scatterTexture(S, A, s, R, delta);
    else if (s.kind == SinkRoughness)
        A.gradRoughness[s.material] += double((delta.x * R.x + delta.y * R.y + delta.z * R.z) * s.g);
}

// ---------------------------------------------------------------------------------------------------------------
// The estimator (async_integrator.h:1986-3042): the subset without volumes, presence, subsurface, AOVs, light
// filters, light trees, mirrors or hair. rr_min_depth extends the roulette's start (1: the reference's).
struct Pending {
    V3 origin, dir, T;
    U64 branch;
    U32 depth, nonMirrorDepth, diffuseDepth, glossyDepth, mirrorDepth;
    float nearT, farT;
    U32 heldLight;
    float heldFar;
    V3 held;
    U32 volumeDepth;  // the volume scattering the path has been through (async_integrator.h:210)
    RayDiff diff;     // its differentials (async_integrator.h:2985), in scenes with textures
};
constexpr U32 MaxPending = 32, PendingReserve = 16;  // async_integrator.h:250-267
float bounceThinning(U32 top, U32 planned) {
    const U32 room = top < PendingReserve ? PendingReserve - top : 1u;
    return planned > room ? float(room) / float(planned) : 1.0f;
}
bool roomForChild(U32 top, float u, float& weight) {
    if (top >= MaxPending) return false;
    const float chance = float(MaxPending - top) / float(PendingReserve + 1);
    if (u >= chance) return false;
    weight = 1.0f / chance;
    return true;
}
bool roulette(float lum, float threshold, float u, float& scale) {  // :141-147
    scale = 1;
    if (lum >= threshold) return true;
    if (lum < u * threshold) return false;
    scale = threshold / lum;
    return true;
}
V3 smartClamp(V3 a, float clamp) {  // :134-138
    if (clamp <= 0) return a;
    const float m = fmaxf(a.x, fmaxf(a.y, a.z));
    return m > clamp ? a * (clamp / m) : a;
}

struct Counters { U64 traces = 0, shadows = 0; };

Hit trace(const Scene& s, const Ray& ray, Counters& k) {
    ++k.traces;
    Hit h;
    float t;
    const U32 tri = traverse(s, ray.origin, ray.direction, ray.nearT, ray.farT, false, false, t);
    if (tri == ~0u) return h;
    h.geom = s.geomOf[tri];
    h.prim = s.primOf[tri];
    if (!geometryHit(h, ray, s.tri[size_t(tri) * 3], s.tri[size_t(tri) * 3 + 1], s.tri[size_t(tri) * 3 + 2]))
        h.normal = {0, 0, 0};
    h.prim = tri;  // the scene triangle, for the shading lookups below
    return h;
}
bool visible(const Scene& s, V3 o, V3 d, float nearT, float farT, Counters& k) {
    ++k.shadows;
    float t;
    return traverse(s, o, d, nearT, farT, true, false, t) == ~0u;
}

// ---------------------------------------------------------------------------------------------------------------
// Volumes. The structure is the reference's computeRadianceVolume (async_integrator.h:1152-1291): one march along the
// ray (decoupled ray marching, VolumeMarch::run :650-798) records steps; single scattering per light by equi-angular
// and distance sampling over those steps with MIS (volumeScatterDirect :917-1012); the Wrenninge approximation of
// multiple scattering (volumeMultipleScatter :1060-1149); the throughput takes the march's transmittance. The medium
// is the build's own (one box-bounded brick grid, trilinear FP16, grey extinction, a blackbody ramp for emission) and
// so is the shadow transmittance: ratio tracking under per-brick majorants in place of the reference's transmittance
// ray march (:428-476). Steps whose end point has zero density are not recorded (they carry nothing), and the
// oracle keeps its steps in a table where the device marches again from checkpoints.
constexpr float VolEpsilon = 1.1920928955078125e-07f;  // volume::Epsilon
constexpr float HitEpsilonStart = 1e-5f;               // volume::HitEpsilonStart
constexpr float OneOverFourPi = 0.0795774715459476678844f;
constexpr U32 MarchSteps = 1u << 16;  // lattice steps a march takes at most (the reference caps at volume::MarchLimit)
float halfToFloat(uint16_t h) {  // IEEE binary16, exactly
    const U32 s = U32(h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
    if (e == 0) {
        const float v = float(m) * (1.0f / 16777216.0f);
        return s ? -v : v;
    }
    const U32 u = s | (e == 31 ? 0x7F800000u | (m << 13) : ((e + 112u) << 23) | (m << 13));
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
float lerp(float a, float b, float w) { return a + (b - a) * w; }
float voxel(const Volume& v, const std::vector<uint16_t>& pool, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || U32(i) >= v.res[0] || U32(j) >= v.res[1] || U32(k) >= v.res[2]) return 0.0f;
    const U32 e = v.table[(U32(k >> 3) * v.nb[1] + U32(j >> 3)) * v.nb[0] + U32(i >> 3)];
    if (e == ~0u) return 0.0f;
    return halfToFloat(pool[size_t(e) * 512 + (U32(k & 7) * 8 + U32(j & 7)) * 8 + U32(i & 7)]);
}
// Trilinear in the voxel centres (voxel i's centre at i + 0.5 in grid units), x first, then y, then z.
float sampleGrid(const Volume& v, const std::vector<uint16_t>& pool, V3 p) {
    float q[3];
    for (int a = 0; a < 3; ++a) {
        q[a] = (comp(p, a) - comp(v.bmin, a)) * v.scaleVox[a] - 0.5f;
        if (!(q[a] > -1.0f && q[a] < float(v.res[a]))) return 0.0f;
    }
    const float fx0 = floorf(q[0]), fy0 = floorf(q[1]), fz0 = floorf(q[2]);
    const int i = int(fx0), j = int(fy0), k = int(fz0);
    const float fx = q[0] - fx0, fy = q[1] - fy0, fz = q[2] - fz0;
    const float c00 = lerp(voxel(v, pool, i, j, k), voxel(v, pool, i + 1, j, k), fx);
    const float c10 = lerp(voxel(v, pool, i, j + 1, k), voxel(v, pool, i + 1, j + 1, k), fx);
    const float c01 = lerp(voxel(v, pool, i, j, k + 1), voxel(v, pool, i + 1, j, k + 1), fx);
    const float c11 = lerp(voxel(v, pool, i, j + 1, k + 1), voxel(v, pool, i + 1, j + 1, k + 1), fx);
    return lerp(lerp(c00, c10, fy), lerp(c01, c11, fy), fz);
}
U32 brickAt(const Volume& v, V3 p) {
    int b[3];
    for (int a = 0; a < 3; ++a) {
        const float q = floorf((comp(p, a) - comp(v.bmin, a)) * v.scaleVox[a] * 0.125f);
        b[a] = int(fminf(fmaxf(q, 0.0f), float(v.nb[a] - 1)));
    }
    return (U32(b[2]) * v.nb[1] + U32(b[1])) * v.nb[0] + U32(b[0]);
}
// Density at p: zero without a look at the voxels where the brick's majorant (its voxels and a one-voxel apron) is.
float densityAt(const Volume& v, V3 p) { return v.majorant[brickAt(v, p)] > 0.0f ? sampleGrid(v, v.density, p) : 0.0f; }
V3 rampAt(const Volume& v, float kelvin) {
    float x = kelvin / v.rampMax * float(v.rampSize - 1);
    x = fminf(fmaxf(x, 0.0f), float(v.rampSize - 1));
    const float fi = floorf(x);
    const U32 i = U32(fi);
    if (i + 1 >= v.rampSize) return load3(&v.ramp[size_t(v.rampSize - 1) * 3]);
    const float f = x - fi;
    const float* a = &v.ramp[size_t(i) * 3];
    return {lerp(a[0], a[3], f), lerp(a[1], a[4], f), lerp(a[2], a[5], f)};
}
struct VCoeff {
    float sigmaT;
    V3 sigmaS, emission;
};
VCoeff coefficientsAt(const Volume& v, V3 p, bool withEmission) {
    VCoeff k{0.0f, {0, 0, 0}, {0, 0, 0}};
    const float d = densityAt(v, p);
    if (d == 0.0f) return k;
    k.sigmaT = d * v.sigmaTScale;
    k.sigmaS = v.albedo * k.sigmaT;
    if (withEmission && v.hasTemperature && v.emissionScale > 0.0f)
        k.emission = rampAt(v, sampleGrid(v, v.temperature, p)) * (v.emissionScale * d);
    return k;
}
// The ray's stretch inside the box, clipped to [tnear, tfar].
bool clipBox(const Volume& v, V3 o, V3 d, float tnear, float tfar, float& a, float& b) {
    a = tnear;
    b = tfar;
    for (int ax = 0; ax < 3; ++ax) {
        const float oa = comp(o, ax), da = comp(d, ax), lo = comp(v.bmin, ax), hi = comp(v.bmax, ax);
        if (da == 0.0f) {
            if (oa < lo || oa > hi) return false;
            continue;
        }
        float t0 = (lo - oa) / da, t1 = (hi - oa) / da;
        if (t0 > t1) std::swap(t0, t1);
        a = fmaxf(a, t0);
        b = fminf(b, t1);
    }
    return a < b;
}
bool insideBox(const Volume& v, V3 p) {
    return p.x >= v.bmin.x && p.x <= v.bmax.x && p.y >= v.bmin.y && p.y <= v.bmax.y && p.z >= v.bmin.z && p.z <= v.bmax.z;
}
// Ratio tracking (Novak et al. 2014) along [tnear, tfar] inside the box, brick by brick (a 3D DDA over the brick
// grid): free flights at the brick's majorant extinction, each collision keeping 1 - density / majorant of the
// estimate; nothing once it falls below the transmittance threshold (as the reference's transmittance gives up).
float transmittance(const Volume& v, V3 o, V3 d, float tnear, float tfar, float scale, float threshold, Stream& s) {
    float a, b;
    if (!clipBox(v, o, d, tnear, tfar, a, b)) return 1.0f;
    const V3 entry = o + d * a;
    int cell[3], step[3];
    float tNext[3];
    for (int ax = 0; ax < 3; ++ax) {
        const float q = floorf((comp(entry, ax) - comp(v.bmin, ax)) * v.scaleVox[ax] * 0.125f);
        cell[ax] = int(fminf(fmaxf(q, 0.0f), float(v.nb[ax] - 1)));
        const float da = comp(d, ax);
        step[ax] = da > 0.0f ? 1 : -1;
        tNext[ax] = da == 0.0f ? FLT_MAX
                               : (comp(v.bmin, ax) + float(cell[ax] + (da > 0.0f ? 1 : 0)) * v.brickExtent[ax] - comp(o, ax)) / da;
    }
    float T = 1.0f, t = a;
    while (t < b) {
        const int ax = tNext[0] <= tNext[1] ? (tNext[0] <= tNext[2] ? 0 : 2) : (tNext[1] <= tNext[2] ? 1 : 2);
        const float tEnd = fminf(tNext[ax], b);
        const float m = v.majorant[(U32(cell[2]) * v.nb[1] + U32(cell[1])) * v.nb[0] + U32(cell[0])];
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
        if (cell[ax] < 0 || U32(cell[ax]) >= v.nb[ax]) break;
        const float da = comp(d, ax);
        tNext[ax] = (comp(v.bmin, ax) + float(cell[ax] + (da > 0.0f ? 1 : 0)) * v.brickExtent[ax] - comp(o, ax)) / da;
    }
    return T;
}
// Henyey-Greenstein (volume_core.h:550-564).
float phaseEval(float g, V3 wo, V3 wi) {
    if (fabsf(g) <= 1e-3f) return OneOverFourPi;
    const float cosTheta = dot(wo, wi), denom = 1 + g * g + 2 * g * cosTheta;
    return fabsf(denom) < VolEpsilon ? 0.0f : OneOverFourPi * (1 - g * g) / (denom * sqrtf(denom));
}
V3 phaseSample(float g, V3 wo, float u1, float u2) {
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
// Distance sampling in exp(-sigma t) over [a, b] and equi-angular sampling (volume_core.h:566-573).
float sampleDistanceExponential(float u, float sigma, float a, float b) {
    return a - logf(1.0f - u * (1.0f - expf(sigma * (a - b)))) / sigma;
}
float pdfDistanceExponential(float t, float sigma, float a, float b) {
    return sigma / (expf((t - a) * sigma) - expf((t - b) * sigma));
}
float sampleEquiAngular(float u, float D, float thetaA, float thetaB) { return D * tanf((1 - u) * thetaA + u * thetaB); }
float pdfEquiAngular(float t, float D, float thetaA, float thetaB) { return D / ((thetaB - thetaA) * (D * D + t * t)); }
bool finitef(float x) { return x - x == 0.0f; }
// A rect light's equi-angular pivot (volume_core.h:630-660, a rectangle without a texture).
V3 lightPivot(const x2_light& l, const float r[3]) {
    if (l.type != X2_LIGHT_RECT) return load3(l.position);
    const V3 local{(0.5f - r[0]) * 2.0f * l.half_width, (0.5f - r[1]) * 2.0f * l.half_height, 0.0f};
    return load3(l.position) + rows(l.local_to_render, local);
}
// One recorded step of a march: extinction, scattering, the transmittance before it, where it starts, its length.
struct VStep {
    float sigmaT;
    V3 sigmaS;
    float tr, tStart, delta;
};
struct MarchResult {
    std::vector<VStep> steps;
    std::vector<float> cdf;
    float total = 1.0f, vt = 1.0f;
    V3 emitted{0, 0, 0};
    bool reached = false;
};
// The march over [a, b): step ends at a + stepSize * (u + k) (a lattice, each from the one before; the last step
// ends at b), each step's coefficients at its end (VolumeMarch::run's heterogeneous stretch).
void marchVolume(const Volume& v, const x2_config& c, V3 o, V3 d, float a, float b, U32 depth, bool sampleEmission,
                 Stream marching, MarchResult& m) {
    const float u = marching.next();
    const float stepSize = v.minFeature * c.inv_volume_quality * fmaxf(float(int(depth) - 1), 1.0f);
    for (U32 k = 0; k < MarchSteps; ++k) {
        const float e = a + stepSize * (u + float(k));
        const bool last = !(e < b);
        if (last && !(e - b > VolEpsilon)) break;
        if (m.vt < c.volume_transmittance_threshold) {
            m.reached = true;
            break;
        }
        const float at = last ? b : e, start = k == 0 ? a : a + stepSize * (u + float(k - 1));
        const VCoeff q = coefficientsAt(v, o + d * at, sampleEmission);
        if (q.sigmaT != 0.0f) {
            const float delta = at - start;
            m.steps.push_back(VStep{q.sigmaT, q.sigmaS, m.vt, start, delta});
            if (sampleEmission) m.emitted = m.emitted + q.emission * (m.vt * delta);
            m.vt *= expf(-q.sigmaT * delta);
        }
        if (last) break;
    }
    // tabulate (volume_core.h:595-612): each step's weight its scattering through what lies before it.
    const U32 n = U32(m.steps.size());
    if (!n) return;
    float sum = 0.0f;
    std::vector<float> w(n);
    for (U32 i = 0; i < n; ++i) {
        w[i] = (luminance(m.steps[i].sigmaS) * m.steps[i].delta) * m.steps[i].tr;
        sum += w[i];
    }
    m.cdf.assign(n, 0.0f);
    if (sum != 0.0f) {
        float partial = 0.0f;
        for (U32 i = 0; i + 1 < n; ++i) {
            partial += w[i];
            if (partial > sum) partial = sum;
            m.cdf[i] = partial;
        }
    } else {
        for (U32 i = 0; i + 1 < n; ++i) m.cdf[i] = float(i + 1) / float(n);
        sum = 1.0f;
    }
    m.cdf[n - 1] = sum;
    m.total = sum;
}
U32 pickStep(const MarchResult& m, float r, float& pdf, float& remapped) {  // sampleDiscrete
    const U32 n = U32(m.steps.size());
    const float rw = r * m.total;
    U32 i = 0;
    while (i + 1 < n && m.cdf[i] < rw) ++i;
    const float prev = i > 0 ? m.cdf[i - 1] : 0.0f, value = m.cdf[i] - prev;
    pdf = value / m.total;
    remapped = (rw - prev) / value;
    return i;
}
// findInterval: the last step starting at or before t, its pdfDiscrete, and whether it holds t (a t between
// recorded steps lies in a step of zero density).
U32 stepAt(const MarchResult& m, float t, float& pdf, bool& holds) {
    const U32 n = U32(m.steps.size());
    U32 k = 0;
    while (k + 1 < n && m.steps[k + 1].tStart <= t) ++k;
    pdf = ((m.cdf[k] - (k > 0 ? m.cdf[k - 1] : 0.0f)) / m.total);
    holds = m.steps[k].tStart <= t && t <= m.steps[k].tStart + m.steps[k].delta;
    return k;
}
// estimateInScatteringSourceTermDirect (:575-591): a light sample from P, its shadow ray, the transmittance
// through the medium on its way (ratio tracking), the phase function.
V3 volumeSourceTerm(const Scene& S, const x2_config& c, Stream& s, const x2_light& L, V3 P, V3 rayDir, float g,
                    const float ul[3], float scale, Counters& k) {
    V3 wi;
    LightHit lh;
    const float r[3] = {ul[0], ul[1], ul[2]};
    if (!lightSample(L, P, nullptr, r, wi, lh)) return {0, 0, 0};
    float pdfL;
    const V3 Li = lightEval(L, wi, lh, pdfL);
    if (pdfL == 0.0f || isExactlyZero(Li) || !finitef(pdfL)) return {0, 0, 0};
    const float near = HitEpsilonStart, far = fmaxf(HitEpsilonEnd * lh.distance, near);
    if (!visible(S, P, wi, near, far, k)) return {0, 0, 0};
    const float tr = transmittance(S.volume, P, wi, 0.0f, far, scale, c.volume_transmittance_threshold, s);
    const float ph = phaseEval(g, -rayDir, wi);
    return Li * (tr * (ph / pdfL));
}
// The medium as a light source: MoonRay's emissive region over a dense emission distribution (volume_emission.h,
// cpu_scene.h:1520-1545 for how the reference lays the distribution out). The histogram is the build's own over its
// brick grid: cells of 4^3 voxels over the medium's box, each cell's value emission scale x the largest density x
// the brightest ramp entry that a temperature within one voxel of the cell can interpolate (an upper bound of the
// emission luminance anywhere in the cell). Each (z, y) row gets its x distribution, each z the y distribution of
// its rows' integrals, and the z distribution those planes' integrals; a distribution's pool record is its integral,
// n, the n values and the n + 1 cdf entries (accumulated value / n, normalized; uniform when the values sum to 0).
constexpr int RegionCell = 4;
U32 poolDistribution(std::vector<float>& pool, const float* func, U32 n) {
    const U32 at = U32(pool.size());
    std::vector<float> cdf(size_t(n) + 1, 0.0f);
    for (U32 i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / float(n);
    const float integral = cdf[n];
    for (U32 i = 1; i <= n; ++i) cdf[i] = integral == 0.0f ? float(i) / float(n) : cdf[i] / integral;
    pool.push_back(integral);
    pool.push_back(float(n));
    for (U32 i = 0; i < n; ++i) pool.push_back(func[i]);
    for (U32 i = 0; i <= n; ++i) pool.push_back(cdf[i]);
    return at;
}
void emissionRegion(const x2_volume& in, Volume& v) {
    int n[3];
    for (int a = 0; a < 3; ++a) n[a] = int(in.res[a]) / RegionCell;
    std::vector<float> value(size_t(n[0]) * n[1] * n[2], 0.0f);
    auto rampIndex = [&](float kelvin) {
        float x = kelvin / in.ramp_max_kelvin * float(in.ramp_size - 1);
        x = fminf(fmaxf(x, 0.0f), float(in.ramp_size - 1));
        return U32(floorf(x));
    };
    bool any = false;
    const int R[3] = {int(in.res[0]), int(in.res[1]), int(in.res[2])};
    for (int cz = 0; cz < n[2]; ++cz)
        for (int cy = 0; cy < n[1]; ++cy)
            for (int cx = 0; cx < n[0]; ++cx) {
                float dMax = 0.0f, tMin = FLT_MAX, tMax = -FLT_MAX;
                for (int z = std::max(cz * RegionCell - 1, 0); z <= std::min(cz * RegionCell + RegionCell, R[2] - 1); ++z)
                    for (int y = std::max(cy * RegionCell - 1, 0); y <= std::min(cy * RegionCell + RegionCell, R[1] - 1); ++y)
                        for (int x = std::max(cx * RegionCell - 1, 0); x <= std::min(cx * RegionCell + RegionCell, R[0] - 1); ++x) {
                            const size_t i = (size_t(z) * R[1] + y) * R[0] + x;
                            dMax = fmaxf(dMax, halfToFloat(in.density[i]));
                            const float t = halfToFloat(in.temperature[i]);
                            tMin = fminf(tMin, t);
                            tMax = fmaxf(tMax, t);
                        }
                if (!(dMax > 0.0f)) continue;
                const U32 i0 = rampIndex(tMin), i1 = std::min(rampIndex(tMax) + 1, in.ramp_size - 1);
                float lum = 0.0f;
                for (U32 i = i0; i <= i1; ++i)
                    lum = fmaxf(lum, 0.299f * in.ramp[3 * i] + 0.587f * in.ramp[3 * i + 1] + 0.114f * in.ramp[3 * i + 2]);
                const float e = in.emission_scale * dMax * lum;
                value[(size_t(cz) * n[1] + cy) * n[0] + cx] = e;
                if (e > 0.0f) any = true;
            }
    if (!any) return;
    v.hasRegion = true;
    v.regionVisibility = 0x3ffu;
    for (int a = 0; a < 3; ++a) {
        v.regionRes[a] = n[a];
        const double cell = (double(in.bounds_max[a]) - double(in.bounds_min[a])) / double(n[a]);
        v.distToWorld[4 * a] = float(cell);
        v.distToWorld[9 + a] = in.bounds_min[a];
        v.worldToDist[4 * a] = float(1.0 / cell);
        v.worldToDist[9 + a] = float(-double(in.bounds_min[a]) / cell);
    }
    double volume = 1.0;
    for (int a = 0; a < 3; ++a) volume *= (double(in.bounds_max[a]) - double(in.bounds_min[a])) / double(n[a]);
    v.invUnitVolume = float(1.0 / volume);
    v.pool.clear();
    std::vector<U32> rows;
    v.conditionalX = 0;
    for (int z = 0; z < n[2]; ++z)
        for (int y = 0; y < n[1]; ++y)
            rows.push_back(poolDistribution(v.pool, &value[(size_t(z) * n[1] + y) * n[0]], U32(n[0])));
    v.marginalY = U32(v.pool.size());
    std::vector<float> rowIntegral(n[1]), planeIntegral(n[2]);
    for (int z = 0; z < n[2]; ++z) {
        for (int y = 0; y < n[1]; ++y) rowIntegral[y] = v.pool[rows[size_t(z) * n[1] + y]];
        planeIntegral[z] = v.pool[poolDistribution(v.pool, rowIntegral.data(), U32(n[1]))];
    }
    v.marginalZ = poolDistribution(v.pool, planeIntegral.data(), U32(n[2]));
}
// One distribution record: Distribution1D's pdf of entry i and its sampling (findInterval over the cdf, the entry's
// pdf, where u falls within it).
float recordPdf(const std::vector<float>& pool, U32 at, int i) {
    const float integral = pool[at], n = pool[at + 1];
    return integral > 0.0f ? pool[at + 2 + U32(i)] / (integral * n) : 0.0f;
}
int recordSample(const std::vector<float>& pool, U32 at, float u, float& pdf, float& remapped) {
    const U32 n = U32(pool[at + 1]);
    const float* cdf = &pool[at + 2 + n];
    int i = 0;  // the last entry at or below u, within [0, n - 1]
    for (U32 j = 1; j <= n; ++j)
        if (cdf[j] <= u) i = int(j);
    if (i > int(n) - 1) i = int(n) - 1;
    pdf = recordPdf(pool, at, i);
    remapped = (u - cdf[i]) / (cdf[i + 1] - cdf[i]);
    return i;
}
U32 recordFloats(int n) { return 2u * U32(n) + 3u; }
V3 affinePoint(const float* m, V3 p) {
    return {p.x * m[0] + p.y * m[3] + p.z * m[6] + m[9], p.x * m[1] + p.y * m[4] + p.z * m[7] + m[10],
            p.x * m[2] + p.y * m[5] + p.z * m[8] + m[11]};
}
V3 affineVector(const float* m, V3 w) {
    return {w.x * m[0] + w.y * m[3] + w.z * m[6], w.x * m[1] + w.y * m[4] + w.z * m[7], w.x * m[2] + w.y * m[5] + w.z * m[8]};
}
float regionCellPdf(const Volume& v, int x, int y, int z) {
    if (x < 0 || x >= v.regionRes[0] || y < 0 || y >= v.regionRes[1] || z < 0 || z >= v.regionRes[2]) return 0.0f;
    const float pz = recordPdf(v.pool, v.marginalZ, z);
    const float py = recordPdf(v.pool, v.marginalY + U32(z) * recordFloats(v.regionRes[1]), y);
    const float px = recordPdf(v.pool, v.conditionalX + U32(z * v.regionRes[1] + y) * recordFloats(v.regionRes[0]), x);
    return px * py * pz;
}
V3 regionPoint(const Volume& v, float u1, float u2, float u3) {
    float pz, py, px, rz, ry, rx;
    const int z = recordSample(v.pool, v.marginalZ, u1, pz, rz);
    const int y = recordSample(v.pool, v.marginalY + U32(z) * recordFloats(v.regionRes[1]), u2, py, ry);
    const int x = recordSample(v.pool, v.conditionalX + U32(z * v.regionRes[1] + y) * recordFloats(v.regionRes[0]), u3, px, rx);
    return affinePoint(v.distToWorld, V3{float(x) + rx, float(y) + ry, float(z) + rz});
}
// The direction pdf from p (EmissionDistribution::pdf): the 3D pdf along the ray, t^2 dt, cell by cell through a DDA
// over the region's cells; tEnd where the ray leaves the box.
float regionPdf(const Volume& v, V3 pWorld, V3 wiWorld, float& tEnd) {
    const V3 p = affinePoint(v.worldToDist, pWorld);
    V3 w = affineVector(v.worldToDist, wiWorld);
    const float lengthScale = 1.0f / sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
    w = w * lengthScale;
    float tStart = 0.0f;
    tEnd = 3.402823466e+38f;
    for (int a = 0; a < 3; ++a) {
        const float inv = 1.0f / comp(w, a);
        float t0 = (0.0f - comp(p, a)) * inv, t1 = (float(v.regionRes[a]) - comp(p, a)) * inv;
        if (t0 > t1) std::swap(t0, t1);
        tStart = t0 > tStart ? t0 : tStart;
        tEnd = t1 < tEnd ? t1 : tEnd;
        if (tStart > tEnd) return 0.0f;
    }
    tEnd *= lengthScale;
    int cell[3], step[3], stop[3];
    float next[3], delta[3];
    for (int a = 0; a < 3; ++a) {
        const float q = comp(p, a) + tStart * comp(w, a);
        cell[a] = std::min(std::max(int(q), 0), v.regionRes[a] - 1);
        if (comp(w, a) >= 0) {
            next[a] = tStart + (float(cell[a]) + 1.0f - q) / comp(w, a);
            delta[a] = 1.0f / comp(w, a);
            step[a] = 1;
            stop[a] = v.regionRes[a];
        } else {
            next[a] = tStart + (float(cell[a]) - q) / comp(w, a);
            delta[a] = -1.0f / comp(w, a);
            step[a] = -1;
            stop[a] = -1;
        }
    }
    float t = tStart, sum = 0.0f;
    for (int i = 0; i < (v.regionRes[0] + v.regionRes[1] + v.regionRes[2]) * 2 + 8; ++i) {
        const float pd = regionCellPdf(v, cell[0], cell[1], cell[2]);
        int a = 0;  // the nearest cell wall, the lower axis on a tie
        if (next[1] < next[a]) a = 1;
        if (next[2] < next[a]) a = 2;
        if (pd > 0.0f) {
            const float t0 = t * lengthScale, t1 = next[a] * lengthScale;
            sum += pd * v.invUnitVolume * (t1 * t1 * t1 - t0 * t0 * t0) / 3.0f;
        }
        t = next[a];
        cell[a] += step[a];
        if (cell[a] == stop[a]) break;
        next[a] += delta[a];
    }
    return sum;
}
V3 regionSample(const Volume& v, V3 p, float u1, float u2, float u3, float& pdf, float& tEnd) {
    V3 wi = regionPoint(v, u1, u2, u3) - p;
    wi = wi * (1.0f / sqrtf(wi.x * wi.x + wi.y * wi.y + wi.z * wi.z));
    pdf = regionPdf(v, p, wi, tEnd);
    return wi;
}
// computeEmissiveVolumeIntegral (async_integrator.h:478-566) for the box: the ray ends at tEnd or its first surface;
// over its stretch of the box (one draw from s, when there is one) a march of the given depth (marchVolume's
// lattice) adds emission x transmittance x the step's length, the transmittance falling by Beer's law a step at a
// time, until it is below the threshold.
V3 emissiveIntegral(const Scene& S, const x2_config& c, Stream& s, V3 o, V3 d, float tnear, float tEnd, U32 depth,
                    Counters& k) {
    const Volume& v = S.volume;
    {
        ++k.traces;
        float t;
        if (traverse(S, o, d, tnear, tEnd, false, false, t) != ~0u && t < tEnd) tEnd = t;
    }
    float a, b;
    if (!clipBox(v, o, d, tnear, tEnd, a, b) || !(b - a > VolEpsilon)) return {0, 0, 0};
    const float u = s.next();
    const float stepSize = v.minFeature * c.inv_volume_quality * fmaxf(float(int(depth) - 1), 1.0f);
    V3 L{0, 0, 0};
    float tr = 1.0f;
    for (U32 i = 0; i < MarchSteps; ++i) {
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
// The dense grids as bricks: a brick is kept when any of its density voxels is nonzero (temperature is kept for the
// same bricks and reads zero elsewhere); a brick's majorant is the largest density among its own voxels and its 26
// neighbours' (a point in the brick interpolates voxels at most one away from it).
bool packVolume(const x2_volume& in, Volume& v) {
    if (!in.density || in.sigma_t_scale < 0.0f) return false;
    if (in.temperature && (!in.ramp || in.ramp_size < 2 || !(in.ramp_max_kelvin > 0.0f))) return false;
    v = Volume{};
    v.bmin = load3(in.bounds_min);
    v.bmax = load3(in.bounds_max);
    float feature = FLT_MAX;
    for (int a = 0; a < 3; ++a) {
        if (in.res[a] == 0 || in.res[a] % 8 || in.res[a] > 4096) return false;
        const float ext = comp(v.bmax, a) - comp(v.bmin, a);
        if (!(ext > 0.0f)) return false;
        v.res[a] = in.res[a];
        v.nb[a] = in.res[a] / 8;
        v.scaleVox[a] = float(in.res[a]) / ext;
        v.brickExtent[a] = ext / float(v.nb[a]);
        feature = fminf(feature, ext / float(in.res[a]));
    }
    v.minFeature = feature;
    v.sigmaTScale = in.sigma_t_scale;
    v.albedo = load3(in.albedo);
    v.g = in.g;
    v.emissionScale = in.emission_scale;
    v.hasTemperature = in.temperature != nullptr;
    if (v.hasTemperature) {
        v.rampMax = in.ramp_max_kelvin;
        v.rampSize = in.ramp_size;
        v.ramp.assign(in.ramp, in.ramp + size_t(in.ramp_size) * 3);
    }
    const U32 nx = v.nb[0], ny = v.nb[1], nz = v.nb[2];
    const size_t bricks = size_t(nx) * ny * nz;
    v.table.assign(bricks, ~0u);
    std::vector<float> brickMax(bricks, 0.0f);
    U32 used = 0;
    for (U32 bz = 0; bz < nz; ++bz)
        for (U32 by = 0; by < ny; ++by)
            for (U32 bx = 0; bx < nx; ++bx) {
                const size_t b = (size_t(bz) * ny + by) * nx + bx;
                bool any = false;
                float mx = 0.0f;
                for (U32 z = 0; z < 8; ++z)
                    for (U32 y = 0; y < 8; ++y)
                        for (U32 x = 0; x < 8; ++x) {
                            const uint16_t h = in.density[(size_t(bz * 8 + z) * v.res[1] + by * 8 + y) * v.res[0] + bx * 8 + x];
                            if (h & 0x7fffu) any = true;
                            mx = fmaxf(mx, halfToFloat(h));
                        }
                if (!any) continue;
                brickMax[b] = mx;
                v.table[b] = used++;
                for (U32 z = 0; z < 8; ++z)
                    for (U32 y = 0; y < 8; ++y)
                        for (U32 x = 0; x < 8; ++x) {
                            const size_t src = (size_t(bz * 8 + z) * v.res[1] + by * 8 + y) * v.res[0] + bx * 8 + x;
                            v.density.push_back(in.density[src]);
                            if (v.hasTemperature) v.temperature.push_back(in.temperature[src]);
                        }
            }
    v.majorant.assign(bricks, 0.0f);
    for (U32 bz = 0; bz < nz; ++bz)
        for (U32 by = 0; by < ny; ++by)
            for (U32 bx = 0; bx < nx; ++bx) {
                float m = 0.0f;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int x = int(bx) + dx, y = int(by) + dy, z = int(bz) + dz;
                            if (x < 0 || y < 0 || z < 0 || U32(x) >= nx || U32(y) >= ny || U32(z) >= nz) continue;
                            m = fmaxf(m, brickMax[(size_t(z) * ny + U32(y)) * nx + U32(x)]);
                        }
                v.majorant[(size_t(bz) * ny + by) * nx + bx] = m;
            }
    if (in.temperature && in.emission_scale > 0.0f) emissionRegion(in, v);
    v.present = true;
    return true;
}
struct VolumeState {
    V3 T;
    U32 depth, nonMirrorDepth, volumeDepth;
    U64 branch;
};
// computeRadianceVolume for one ray: adds the medium's light to radiance, returns the alpha transmittance and takes
// the march's transmittance into pv.T.
float volumeRadiance(const Scene& S, const x2_config& c, const Identity& root, VolumeState& pv, V3 o, V3 d,
                     float tnear, float tfar, V3& radiance, Counters& k) {
    const Volume& v = S.volume;
    float a, b;
    if (!clipBox(v, o, d, tnear, tfar, a, b)) return 1.0f;
    const U64 volumeBranch = childBranch(pv.branch, 0xfffc, pv.volumeDepth);
    Identity vid = root;
    vid.branch = childBranch(volumeBranch, 0, 0);
    const bool sampleEmission = pv.depth == 0;
    MarchResult m;
    marchVolume(v, c, o, d, a, b, pv.depth, sampleEmission, Stream(vid, pv.depth, 0), m);
    const U32 count = U32(m.steps.size());
    if (count == 0) return 1.0f;
    const float vt = m.reached ? 0.0f : m.vt;
    const V3 T = pv.T;
    if (sampleEmission && !isExactlyZero(m.emitted)) radiance = radiance + T * m.emitted;
    if (pv.volumeDepth < c.max_volume_depth) {
        const float firstStart = m.steps[0].tStart, rayEnd = m.steps[count - 1].tStart + m.steps[count - 1].delta;
        // integrateVolumeScatteringDirect (:917-1012)
        const bool highQuality = pv.nonMirrorDepth == 0;
        const U32 samplesPerLight = highQuality ? c.light_samples : 1,
                  scatterCount = highQuality ? c.volume_illumination_samples : 1;
        if (scatterCount > 0) {
            const float invN = 1.0f / float(samplesPerLight * scatterCount);
            for (U32 li = 0; li < c.light_count; ++li) {
                const x2_light& L = S.lights[li];
                const bool equiAngular = L.type == X2_LIGHT_RECT;
                V3 direct{0, 0, 0};
                for (U32 ls = 0; ls < samplesPerLight; ++ls)
                    for (U32 ts = 0; ts < scatterCount; ++ts) {
                        Identity sid = root;
                        sid.branch = childBranch(volumeBranch, 1 + li, ls * scatterCount + ts);
                        Stream ss(sid, pv.depth, 0);
                        float D = 0.0f, thetaA = 0.0f, thetaB = 0.0f, offset = 0.0f;
                        if (equiAngular) {
                            const float ue = ss.next();
                            const float uel[3] = {ss.next(), ss.next(), ss.next()};
                            const V3 pivot = lightPivot(L, uel);
                            offset = dot(pivot - o, d);
                            const V3 foot = o + d * offset;
                            D = length(pivot - foot);
                            thetaA = atan2f(firstStart - offset, D);
                            thetaB = atan2f(rayEnd - offset, D);
                            const float te = sampleEquiAngular(ue, D, thetaA, thetaB);
                            const float pdfTe = pdfEquiAngular(te, D, thetaA, thetaB);
                            if (pdfTe != 0.0f && finitef(pdfTe)) {
                                const float tePrime = te + offset;
                                float pdfK;
                                bool holds;
                                const VStep& vp = m.steps[stepAt(m, tePrime, pdfK, holds)];
                                if (holds && !isExactlyZero(vp.sigmaS)) {
                                    const float trScatter = vp.tr * expf(-vp.sigmaT * (tePrime - vp.tStart));
                                    float mis = 1.0f;
                                    const float pdfTd =
                                        pdfK * pdfDistanceExponential(tePrime, vp.sigmaT, vp.tStart, vp.tStart + vp.delta);
                                    if (finitef(pdfTd)) mis = powerHeuristic(pdfTe, pdfTd);
                                    const V3 Ls = volumeSourceTerm(S, c, ss, L, o + d * tePrime, d, v.g, uel, 1.0f, k);
                                    direct = direct + T * vp.sigmaS * (Ls * trScatter) * (invN * mis / pdfTe);
                                }
                            }
                        }
                        // distanceVolumeScatteringDirect
                        const float ud = ss.next();
                        const float udl[3] = {ss.next(), ss.next(), ss.next()};
                        float pdfDensity, udRemapped;
                        const VStep& vp = m.steps[pickStep(m, ud, pdfDensity, udRemapped)];
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
                        const V3 Ls = volumeSourceTerm(S, c, ss, L, o + d * td, d, v.g, udl, 1.0f, k);
                        direct = direct + T * vp.sigmaS * (Ls * trScatter) * (invN * mis / pdfTd);
                    }
                radiance = radiance + direct;
            }
            // volumeScatterRegions (:999-1057): the glowing medium lighting the steps, toward a place drawn in its
            // region, at an equi-angular distance.
            if (v.hasRegion && (v.regionVisibility & 0x100u)) {
                const float invN = 1.0f / float(samplesPerLight * scatterCount);
                V3 regionLight{0, 0, 0};
                for (U32 i = 0; i < samplesPerLight * scatterCount; ++i) {
                    Identity eid = root;
                    eid.branch = childBranch(volumeBranch, 0xfe00u, i);
                    Stream es(eid, pv.depth, 0);
                    const float u1 = es.next(), u2 = es.next(), u3 = es.next();
                    const V3 pivot = regionPoint(v, u1, u2, u3);
                    const float ue = es.next();
                    const float offset = dot(pivot - o, d);
                    const V3 foot = o + d * offset;
                    const float D = length(pivot - foot);
                    const float thetaA = atan2f(firstStart - offset, D), thetaB = atan2f(rayEnd - offset, D);
                    const float te = sampleEquiAngular(ue, D, thetaA, thetaB);
                    const float pdfTe = pdfEquiAngular(te, D, thetaA, thetaB);
                    if (pdfTe == 0.0f || !finitef(pdfTe)) continue;
                    const float tePrime = te + offset;
                    float pdfK;
                    bool holds;
                    const VStep& vp = m.steps[stepAt(m, tePrime, pdfK, holds)];
                    if (!holds || isExactlyZero(vp.sigmaS)) continue;
                    const float trScatter = vp.tr * expf(-vp.sigmaT * (tePrime - vp.tStart));
                    const V3 scatterPoint = o + d * tePrime;
                    V3 wi = pivot - scatterPoint;
                    wi = wi * (1.0f / length(wi));
                    float tEnd;
                    const float pdfWi = regionPdf(v, scatterPoint, wi, tEnd);
                    if (pdfWi == 0.0f || !finitef(pdfWi)) continue;
                    const V3 Le = emissiveIntegral(S, c, es, scatterPoint, wi, 0.0f, tEnd, pv.depth + 1, k);
                    const float ph = phaseEval(v.g, -d, wi);
                    regionLight = regionLight + (T * vp.sigmaS) * (Le * trScatter) * (invN * ph / (pdfTe * pdfWi));
                }
                radiance = radiance + regionLight;
            }
        }
        ++pv.volumeDepth;
        // approximateVolumeMultipleScattering (:1060-1149)
        if (pv.volumeDepth < c.max_volume_depth && c.volume_illumination_samples > 0) {
            const float invN = 1.0f / float(c.volume_illumination_samples);
            for (U32 i = 0; i < c.volume_illumination_samples; ++i) {
                Identity mid = root;
                mid.branch = childBranch(volumeBranch, 0xffff, i);
                Stream ms(mid, pv.depth, 0);
                float pdfDensity, utRemapped;
                const VStep& vp = m.steps[pickStep(m, ms.next(), pdfDensity, utRemapped)];
                float sigmaT = vp.sigmaT;
                const float t0 = vp.tStart, t1 = vp.tStart + vp.delta;
                const float td = sampleDistanceExponential(utRemapped, sigmaT, t0, t1);
                const float pdfT = pdfDensity * pdfDistanceExponential(td, sigmaT, t0, t1);
                if (pdfT == 0.0f || !finitef(pdfT)) continue;
                const float trScatter = vp.tr * expf(-vp.sigmaT * (td - t0));
                V3 scatterPoint = o + d * td;
                V3 throughput = T * vp.sigmaS * (trScatter / pdfT);
                const float up0 = ms.next(), up1 = ms.next();
                V3 scatterDir = phaseSample(v.g, -d, up0, up1);
                for (U32 depth = pv.volumeDepth; depth < c.max_volume_depth; ++depth) {
                    if (luminance(throughput) < c.volume_transmittance_threshold) break;
                    const float ai = powf(c.volume_attenuation_factor, float(depth)),
                                bi = powf(c.volume_contribution_factor, float(depth)),
                                ci = powf(c.volume_phase_attenuation_factor, float(depth));
                    const float tFreePath = -logf(1.0f - ms.next()) / sigmaT;
                    const V3 rayDir = scatterDir;
                    scatterPoint = scatterPoint + scatterDir * tFreePath;
                    if (!insideBox(v, scatterPoint)) break;
                    const VCoeff kk = coefficientsAt(v, scatterPoint, false);
                    const float g = v.g * ci;
                    sigmaT = kk.sigmaT;
                    if (fabsf(sigmaT) < VolEpsilon) break;
                    throughput = throughput * kk.sigmaS * (1.0f / sigmaT);
                    for (U32 li = 0; li < c.light_count; ++li) {
                        const float usl[3] = {ms.next(), ms.next(), ms.next()};
                        const V3 Ls = volumeSourceTerm(S, c, ms, S.lights[li], scatterPoint, rayDir, g, usl, ai, k);
                        radiance = radiance + throughput * Ls * (invN * bi);
                    }
                    // the glowing medium lighting this scatter (:1127-1143)
                    if (v.hasRegion && (v.regionVisibility & 0x100u)) {
                        const float ue0 = ms.next(), ue1 = ms.next(), ue2 = ms.next();
                        float pdfWi, tEnd;
                        const V3 wi = regionSample(v, scatterPoint, ue0, ue1, ue2, pdfWi, tEnd);
                        if (pdfWi != 0.0f && finitef(pdfWi)) {
                            const V3 Le = emissiveIntegral(S, c, ms, scatterPoint, wi, 0.0f, tEnd, pv.depth + 1, k);
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
    pv.T = pv.T * vt;
    return vt;
}

// computeRadianceEmissiveRegionsBundled (async_integrator.h:2481-2613): the glowing medium lighting a surface vertex.
// The lobes it lights make one sampler (a lobe picked in proportion to its albedo's luminance, baseAlbedo
// shade_lobes_eval.h:1022-1028); nL light samples (a place drawn in the region) and nB BSDF samples, each weighed
// against the other strategy by the power heuristic, each carrying the medium's light along its direction from the
// offset origin on that direction's side.
void surfaceRegionLight(const Scene& S, const x2_config& c, const Identity& root, U64 branch, U32 depth, bool split,
                        const Bsdf& bsdf, const Slice& slice, V3 Ng, V3 P, V3 above, V3 below, V3 T, V3& radiance,
                        Counters& k) {
    const Volume& v = S.volume;
    const U32 nL = split ? c.light_samples : 1u, nB = split ? c.bsdf_samples : 1u;
    U32 recs[2], nRec = 0;
    float cdf[2], norm = 0.0f;
    for (U32 q = 0; q < bsdf.count; ++q) {
        const Lobe& l = bsdf.lobes[q];
        if (!(lobeRayMask(l.type) & v.regionVisibility)) continue;
        recs[nRec] = q;
        cdf[nRec] = fmaxf(fabsf(luminance(l.kind == KindLambert ? l.scale * l.albedo : l.scale)), VolEpsilon);
        norm += cdf[nRec++];
    }
    if (!nRec) return;
    float acc = 0.0f;
    for (U32 j = 0; j < nRec; ++j) {
        cdf[j] = cdf[j] / norm + acc;
        acc = cdf[j];
    }
    V3 fs[2];
    U32 matched = 0;
    auto add = [&](V3 contribution) {
        for (U32 j = 0; j < matched; ++j) radiance = radiance + fs[j] * contribution;
    };
    auto emitted = [&](V3 wi, float tEnd, Stream& es) {
        return emissiveIntegral(S, c, es, dot(wi, Ng) >= 0.0f ? above : below, wi, 0.0f, tEnd, depth + 1, k);
    };
    for (U32 ls = 0; ls < nL; ++ls) {
        Identity eid = root;
        eid.branch = childBranch(branch, 0xfd00u, ls);
        Stream es(eid, depth, 0);
        const float u1 = es.next(), u2 = es.next(), u3 = es.next();
        float pdfLight, tEnd;
        const V3 wi = regionSample(v, P, u1, u2, u3, pdfLight, tEnd);
        if (pdfLight == 0.0f) continue;
        const U32 flags = surfaceFlags(slice, wi);
        V3 f{0, 0, 0};
        float pdfBsdf = 0.0f, prev = 0.0f;
        matched = 0;
        for (U32 j = 0; j < nRec; ++j) {
            float p;
            const V3 col = lobeEval(bsdf.lobes[recs[j]], slice, wi, p);
            if (matchesFlags(bsdf.lobes[recs[j]].type, flags)) {
                f = f + col;
                fs[matched++] = col;
            }
            pdfBsdf += p * (cdf[j] - prev);
            prev = cdf[j];
        }
        if (isExactlyZero(f)) continue;
        V3 contribution = emitted(wi, tEnd, es);
        const float mis = nB ? powerHeuristic(float(nL) * pdfLight, float(nB) * pdfBsdf) : 1.0f;
        contribution = (contribution * T) * (mis / (pdfLight * float(nL)));
        add(contribution);
    }
    for (U32 bs = 0; bs < nB; ++bs) {
        Identity eid = root;
        eid.branch = childBranch(branch, 0xfc00u, bs);
        Stream es(eid, depth, 0);
        const float uLobe = es.next(), u1 = es.next(), u2 = es.next();
        U32 pick = 0;
        while (pick + 1 < nRec && cdf[pick] <= uLobe) ++pick;
        const float lobePdf = pick == 0 ? cdf[0] : cdf[pick] - cdf[pick - 1];
        V3 wi;
        float pdfBsdf;
        V3 f = lobeSample(bsdf.lobes[recs[pick]], slice, u1, u2, wi, pdfBsdf);
        if (pdfBsdf == 0.0f) continue;
        pdfBsdf *= lobePdf;
        matched = 0;
        fs[matched++] = f;
        const U32 flags = surfaceFlags(slice, wi);
        if (!matchesFlags(bsdf.lobes[recs[pick]].type, flags)) f = V3{0, 0, 0};
        float prev = 0.0f;
        for (U32 j = 0; j < nRec; ++j) {
            if (j != pick) {
                float p;
                const V3 col = lobeEval(bsdf.lobes[recs[j]], slice, wi, p);
                if (matchesFlags(bsdf.lobes[recs[j]].type, flags)) {
                    f = f + col;
                    fs[matched++] = col;
                }
                pdfBsdf += p * (cdf[j] - prev);
            }
            prev = cdf[j];
        }
        if (pdfBsdf == 0.0f || isExactlyZero(f)) continue;
        float tEnd;
        const float pdfLight = regionPdf(v, P, wi, tEnd);
        if (pdfLight == 0.0f) continue;
        V3 contribution = emitted(wi, tEnd, es);
        const float mis = nL ? powerHeuristic(float(nB) * pdfBsdf, float(nL) * pdfLight) : 1.0f;
        contribution = (contribution * T) * (mis / (pdfBsdf * float(nB)));
        add(contribution);
    }
}

// The prototype estimator (async_integrator.h:61-126; Config::paths == 0): camera hit, emission, cosine-sampled
// Lambert bounces (surface::lambert, surface_core.h:158-165) to a constant sky; the surface_limit-th interaction
// looks along its sampled direction for the sky. Albedo: the material's Lambert lobe (albedo x diffuse weight, x map).
bool surfaceLambert(V3 albedo, V3 N, float u, float v, V3& wi, V3& f, float& pdf) {
    if (!finite3(albedo) || albedo.x < 0 || albedo.y < 0 || albedo.z < 0 || !finite3(N) ||
        fabsf(dot(N, N) - 1) > 2e-5f || !finite(u) || !finite(v) || u < 0 || u > 1 || v < 0 || v > 1)
        return false;
    float sn, cs;
    sincosPoly(v * (2 * Pi), sn, cs);
    const float s = sqrtf(1 - u);
    wi = localToGlobal(frameOf(N), V3{s * cs, s * sn, sqrtf(u)});
    const float cosine = fmaxf(dot(N, wi), 0.0f);
    pdf = cosine * OneOverPi;
    f = (albedo * OneOverPi) * cosine;
    return true;
}
int integrateDirect(const Scene& S, const x2_config& c, U32 pixel, U32 index, double out[4], Counters& k) {
    if (c.surface_limit == 0 || c.surface_limit > 8 || pixel >= c.width * c.height) return 2;
    const Identity id{pixel, index, c.seed, 0};
    float x, y;
    if (!boxFilter(c.filter_width, draw(id, 0, c.sampling), x) || !boxFilter(c.filter_width, draw(id, 1, c.sampling), y))
        return 2;
    Ray ray;
    if (!cameraRay(c, float(pixel % c.width) + x, float(pixel / c.width) + y, ray)) return 2;
    // GpuBackend::material (async_gpu_device.h:356-383): the map read through the default binding (anisotropic,
    // cross terms in order) at the footprint the hit ray's differentials give; only the camera ray has them, so the
    // bounces read a zero footprint (level 0).
    V3 camDx{0, 0, 0}, camDy{0, 0, 0};
    if (!S.textures.empty() &&
        !cameraDifferentials(c, float(pixel % c.width) + x, float(pixel / c.width) + y, camDx, camDy))
        return 2;
    double radiance[3] = {0, 0, 0}, throughput[3] = {1, 1, 1};
    U32 depth = 0, primaryHit = 0;
    const V3 sky = load3(c.sky);
    for (;;) {
        const Hit hit = trace(S, ray, k);
        if (hit.geom == ~0u) {
            if (depth) {
                radiance[0] += throughput[0] * sky.x;
                radiance[1] += throughput[1] * sky.y;
                radiance[2] += throughput[2] * sky.z;
            }
            break;
        }
        if (hit.normal.x == 0 && hit.normal.y == 0 && hit.normal.z == 0) return 2;
        if (!depth) primaryHit = 1;
        const U32 tri = hit.prim;
        const x2_material& m = S.materials[S.material[tri]];
        V3 albedo = load3(m.albedo) * m.diffuse_weight;
        if (m.albedo_map) {
            float uv[6] = {0, 0, 1, 0, 1, 1};
            if (!S.triUV.empty())
                for (int q = 0; q < 6; ++q) uv[q] = S.triUV[size_t(tri) * 6 + q];
            const float w = 1.0f - hit.u - hit.v;
            HitDiff hd;
            if (!depth)
                hitTexDerivatives(S.tri[size_t(tri) * 3], S.tri[size_t(tri) * 3 + 1], S.tri[size_t(tri) * 3 + 2], uv, ray,
                                  camDx, camDy, c.spp, hd);
            const Footprint f{w * uv[0] + hit.u * uv[2] + hit.v * uv[4], w * uv[1] + hit.u * uv[3] + hit.v * uv[5],
                              hd.dsdx, hd.dtdx, hd.dsdy, hd.dtdy};
            V3 color;
            if (!mapLookup(S.textures[m.albedo_map - 1], 0, f, color, false)) return 2;
            albedo = albedo * color;
        }
        const U32 dimension = CameraDimensions + 2 * depth;
        V3 wi, f;
        float pdf;
        if (!surfaceLambert(albedo, hit.normal, draw(id, dimension, c.sampling), draw(id, dimension + 1, c.sampling), wi,
                            f, pdf))
            return 2;
        const V3 e = load3(m.emission);
        radiance[0] += throughput[0] * e.x;
        radiance[1] += throughput[1] * e.y;
        radiance[2] += throughput[2] * e.z;
        if (pdf <= 0) break;
        throughput[0] *= f.x / pdf;
        throughput[1] *= f.y / pdf;
        throughput[2] *= f.z / pdf;
        if (++depth == c.surface_limit) {
            if (visible(S, hit.origin, wi, 0.0f, 3.402823466e+38f, k)) {
                radiance[0] += throughput[0] * sky.x;
                radiance[1] += throughput[1] * sky.y;
                radiance[2] += throughput[2] * sky.z;
            }
            break;
        }
        ray = Ray{hit.origin, wi, 0.0f, 3.402823466e+38f};
    }
    out[0] = radiance[0];
    out[1] = radiance[1];
    out[2] = radiance[2];
    out[3] = primaryHit;
    return 0;
}

// A test hook: while set (on the calling thread), the stratified draws of the sample's first non-mirror vertex are
// recorded as (kind 0 light / 1 lobe, light or lobe index, sample number, r0, r1), in the order they are drawn.
struct DrawRecord {
    float kind, which, number, r0, r1;
};
thread_local std::vector<DrawRecord>* g_drawTrace = nullptr;

int integrate(const Scene& S, const x2_config& c, U32 pixel, U32 index, double out[4], Counters& k,
              const Adjoint* adj = nullptr) {
    if (c.direct_only) return adj ? 2 : integrateDirect(S, c, pixel, index, out, k);
    if (pixel >= c.width * c.height || c.bsdf_samples == 0 || c.bsdf_samples > 64 || c.light_samples == 0 ||
        c.light_samples > 64)
        return 2;
    const Identity root{pixel, index, c.seed, 0};
    float x, y;
    if (!boxFilter(c.filter_width, draw(root, 0, c.sampling), x) || !boxFilter(c.filter_width, draw(root, 1, c.sampling), y))
        return 2;
    Ray camera;
    if (!cameraRay(c, float(pixel % c.width) + x, float(pixel / c.width) + y, camera)) return 2;
    // Scenes with textures carry ray differentials (the map lookups' footprints): the camera's auxiliary directions,
    // and MoonRay's primary scale for the samples per pixel (async_integrator.h:2029).
    const bool textured = !S.textures.empty();
    V3 camDx{0, 0, 0}, camDy{0, 0, 0};
    if (textured && !cameraDifferentials(c, float(pixel % c.width) + x, float(pixel / c.width) + y, camDx, camDy)) return 2;
    const float primaryScale = 1.0f / fminf(sqrtf(float(c.spp ? c.spp : 1)), 2.0f);
    Pending stack[MaxPending];
    U32 top = 0;
    {
        Pending p{};
        p.origin = camera.origin;
        p.dir = camera.direction;
        p.T = {1, 1, 1};
        p.nearT = camera.nearT;
        p.farT = camera.farT;
        p.heldLight = ~0u;
        stack[top++] = p;
    }
    V3 radiance{0, 0, 0};
    float alpha = 0.0f;
    const U32 rrMin = c.rr_min_depth > 1 ? c.rr_min_depth : 1u;
    // The adjoint (adj): the pixel's adjoint per sample, each pending ray's sink by stack position, and the open
    // subtrees; a subtree's radiance is what the running radiance gains until a pop below its position.
    V3 delta{0, 0, 0};
    Sink sinks[MaxPending];
    std::vector<Record> recs;
    if (adj) {
        const float inv = 1.0f / float(c.spp);
        delta = V3{adj->dImage[size_t(pixel) * 3] * inv, adj->dImage[size_t(pixel) * 3 + 1] * inv,
                   adj->dImage[size_t(pixel) * 3 + 2] * inv};
    }
    while (top) {
        Pending pv = stack[--top];
        if (adj) {
            while (!recs.empty() && recs.back().pos > top) {
                scatterSink(S, *adj, recs.back().sink, radiance - recs.back().mark, delta);
                recs.pop_back();
            }
            if (recs.size() < 34) recs.push_back(Record{top, radiance, sinks[top]});
        }
        const bool fromCamera = pv.depth == 0;
        const Ray ray{pv.origin, pv.dir, fromCamera ? camera.nearT : pv.nearT, fromCamera ? camera.farT : pv.farT};
        const Hit hit = trace(S, ray, k);
        if (pv.heldLight != ~0u) {  // bounce-ray reuse (:2095-2106)
            const float kept = hit.geom != ~0u && hit.t < pv.heldFar ? 0.0f : 1.0f;
            if (kept > 0.0f) radiance = radiance + (kept < 1.0f ? pv.held * kept : pv.held);
        }
        // The medium the ray passes on its way (:2106-2120): its light, and the throughput through it.
        float volumeAlpha = 1.0f;
        if (S.volume.present) {
            if (adj) return 2;  // the adjoint does not differentiate through media
            VolumeState vs{pv.T, pv.depth, pv.nonMirrorDepth, pv.volumeDepth, pv.branch};
            volumeAlpha = volumeRadiance(S, c, root, vs, ray.origin, ray.direction, ray.nearT,
                                         hit.geom == ~0u ? ray.farT : hit.t, radiance, k);
            pv.T = vs.T;
            pv.volumeDepth = vs.volumeDepth;
        }
        const float hitAlpha = 1.0f - (0.0f + 0.0f + 0.0f) / 3.0f;
        if (hit.geom == ~0u) {  // camera pixel weight 1; a medium's opacity as the alpha (:2160-2161)
            if (pv.depth == 0) alpha += 1.0f * (1.0f - (volumeAlpha + volumeAlpha + volumeAlpha) / 3.0f);
            continue;
        }
        // shade::hitState (shade_material.h:121-289) for a flat triangle without normals: Ng = (b-a) x (c-a),
        // turned to face the ray.
        const U32 tri = hit.prim;
        const V3 a = S.tri[size_t(tri) * 3], b = S.tri[size_t(tri) * 3 + 1], cc = S.tri[size_t(tri) * 3 + 2];
        if (hit.normal.x == 0 && hit.normal.y == 0 && hit.normal.z == 0) return 2;
        V3 ng = cross(b - a, cc - a);
        float l = length(ng);
        if (!(l > 0)) {
            const double ux = double(b.x) - a.x, uy = double(b.y) - a.y, uz = double(b.z) - a.z;
            const double vx = double(cc.x) - a.x, vy = double(cc.y) - a.y, vz = double(cc.z) - a.z;
            const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
            const double dl = sqrt(nx * nx + ny * ny + nz * nz);
            ng = dl > 0 ? V3{float(nx / dl), float(ny / dl), float(nz / dl)} : -ray.direction;
            l = length(ng);
            if (!(l > 0)) return 2;
        }
        V3 Ng = ng / l, N = Ng;
        const V3 P = hit.point, wo = -ray.direction;
        if (!(dot(Ng, ray.direction) < 0.0f)) {
            N = -N;
            Ng = -Ng;
        }
        if (!finite3(N) || !finite3(Ng)) return 2;
        Bsdf bsdf;
        const x2_material& mat = S.materials[S.material[tri]];
        V3 albedo = load3(mat.albedo);
        Sink texSink, roughSink;
        float uv[6] = {0, 0, 1, 0, 1, 1};
        if (!S.triUV.empty())
            for (int q = 0; q < 6; ++q) uv[q] = S.triUV[size_t(tri) * 6 + q];
        HitDiff hd;
        if (textured) hitDifferentials(a, b, cc, uv, N, Ng, P, ray, fromCamera, camDx, camDy, pv.diff, c.spp, hd);
        if (mat.albedo_map) {  // the map at the hit's (s, t) (shade_material.h:121-160, shade_mapnodes.h:420-470)
            const float w = 1.0f - hit.u - hit.v;
            const float ss = w * uv[0] + hit.u * uv[2] + hit.v * uv[4], tt = w * uv[1] + hit.u * uv[3] + hit.v * uv[5];
            // textureMode (shade_mapnodes.h:344-347): camera paths anisotropic, diffuse ones one level's closest
            // texel, glossy and mirror ones trilinear closest texels.
            const U32 mode = pv.diffuseDepth ? 2u : (pv.glossyDepth || pv.mirrorDepth) ? 1u : 0u;
            const Footprint f{ss, tt, hd.dsdx, hd.dtdx, hd.dsdy, hd.dtdy};
            V3 color;
            if (!mapLookup(S.textures[mat.albedo_map - 1], mode, f, color)) return 2;
            albedo = albedo * color;
            texSink.kind = SinkTexture;
            texSink.tex = mat.albedo_map - 1;
            texSink.mode = mode;
            texSink.f = f;
            texSink.color = color;
        }
        if (mat.specular_weight > 0.0f && mat.roughness * mat.roughness > 0.001f) {
            roughSink.kind = SinkRoughness;
            roughSink.material = S.material[tri];
        }
        buildBsdf(mat, N, bsdf, albedo, S.evalRoughness.empty() ? -1.0f : S.evalRoughness[S.material[tri]]);
        if (pv.depth == 0) alpha += 1.0f * hitAlpha;
        radiance = radiance + pv.T * bsdf.selfEmission;
        if (bsdf.count == 0) continue;
        U32 indirectFlags = 0;
        if (c.bsdf_samples > 0 && c.max_depth > 0 && pv.depth < c.max_depth) {
            if (pv.diffuseDepth < c.max_diffuse_depth) indirectFlags |= AllSurfaceSides | Diffuse;
            if (pv.glossyDepth < c.max_glossy_depth) indirectFlags |= AllSurfaceSides | Glossy;
            if (pv.mirrorDepth < c.max_mirror_depth) indirectFlags |= AllSurfaceSides | Mirror;
        }
        const bool doIndirect = indirectFlags != 0;
        const V3 cull = (bsdf.type & Reflection) ? Ng : -Ng;
        const V3* cullN = (bsdf.type & AllSurfaceSides) != AllSurfaceSides ? &cull : nullptr;
        const Slice slice{Ng, wo, AllLobeTypes};
        const bool split = pv.nonMirrorDepth == 0;
        const U32 nl = split ? c.light_samples : 1;
        const bool clampHere = pv.nonMirrorDepth >= c.clamp_depth;
        const bool rr = pv.nonMirrorDepth >= rrMin && c.rr_threshold > 0;
        U32 lobeSamples[2];
        U32 planned = 0;
        for (U32 q = 0; q < bsdf.count; ++q) {
            lobeSamples[q] = !split ? 1 : c.bsdf_samples;
            planned += lobeSamples[q];
        }
        const float thinning = bounceThinning(top, planned);
        const V3 above = hit.origin, below = hit.below;
        if (S.volume.hasRegion)
            surfaceRegionLight(S, c, root, pv.branch, pv.depth, split, bsdf, slice, Ng, P, above, below, pv.T, radiance, k);

        // Light samples (:2615-2785).
        const U64 lightBranch = childBranch(pv.branch, 0xfffe, 0);
        for (U32 i = 0; i < c.light_count; ++i) {
            const x2_light& L = S.lights[i];
            if (!canIlluminate(L, P, cullN)) continue;
            const U64 lightAddress = childBranch(lightBranch, i, 0);
            Identity lid = root;
            lid.branch = lightAddress;
            Stream ls(lid, pv.depth, 0);
            for (U32 s = 0; s < nl; ++s) {
                float r[3] = {ls.next(), ls.next(), ls.next()};
                const float u = ls.next();
                if (split) {
                    const Identity strat{pixel, U64(index) * nl + s, c.seed, lightAddress};
                    r[0] = draw(strat, CameraDimensions + 2, c.sampling);
                    r[1] = draw(strat, CameraDimensions + 3, c.sampling);
                    if (g_drawTrace) g_drawTrace->push_back({0.0f, float(i), float(s), r[0], r[1]});
                }
                V3 wi;
                LightHit lh;
                if (!lightSample(L, P, cullN, r, wi, lh)) continue;
                float pdfL;
                const V3 Li = lightEval(L, wi, lh, pdfL);
                if (pdfL == 0.0f || isExactlyZero(Li) || !(fabsf(pdfL) < 3.4e38f) || !finite3(Li)) continue;
                const V3 factor = pv.T * Li * (1.0f / (float(nl) * pdfL * 1.0f));
                const U32 flags = surfaceFlags(slice, wi);
                V3 t{0, 0, 0};
                V3 tkq[2] = {{0, 0, 0}, {0, 0, 0}};
                bool any = false;
                for (U32 q = 0; q < bsdf.count; ++q) {
                    const Lobe& lobe = bsdf.lobes[q];
                    if (!matchesFlags(lobe.type, flags)) continue;
                    if (!(lobeRayMask(lobe.type) & L.visibility_mask)) continue;
                    float pdf;
                    const V3 f = lobeEval(lobe, slice, wi, pdf);
                    if (pdf == 0.0f || isExactlyZero(f)) continue;
                    V3 tk = factor * f * powerHeuristic(float(nl) * pdfL, float(lobeSamples[q]) * pdf);
                    if (clampHere) tk = smartClamp(tk, c.clamp_value);
                    t = t + tk;
                    tkq[q] = tk;
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
                float shadowNear = 0.0f, shadowFar;
                lightShadowSpan(lh.distance, shadowNear, shadowFar);
                const float kept = visible(S, shadowOrigin, wi, shadowNear, shadowFar, k) ? 1.0f : 0.0f;
                if (kept > 0.0f) {
                    if (kept < 1.0f) t = t * kept;
                    if (S.volume.present) {  // the medium on the shadow ray's way (:2725-2732)
                        Identity tid = lid;
                        tid.branch = childBranch(lightAddress, s, 0xfffb);
                        Stream vs(tid, pv.depth, 0);
                        t = t * transmittance(S.volume, shadowOrigin, wi, shadowNear, shadowFar, 1.0f,
                                              c.volume_transmittance_threshold, vs);
                    }
                    radiance = radiance + t;
                    if (adj)
                        for (U32 q = 0; q < bsdf.count; ++q) {
                            const V3 share = tkq[q] * rrScale;
                            if (bsdf.lobes[q].kind == KindLambert && texSink.kind == SinkTexture)
                                scatterTexture(S, *adj, texSink, share, delta);
                            if (bsdf.lobes[q].kind == KindGGX && roughSink.kind == SinkRoughness)
                                adj->gradRoughness[roughSink.material] +=
                                    double((delta.x * share.x + delta.y * share.y + delta.z * share.z) *
                                           ggxRoughnessLogDerivative(mat.roughness, bsdf.lobes[q].frame.z, wo, wi));
                        }
                }
            }
        }

        // BSDF samples (:2787-3031).
        for (U32 q = 0; q < bsdf.count; ++q) {
            const Lobe& lobe = bsdf.lobes[q];
            const U32 nk = lobeSamples[q];
            const bool mirror = (lobe.type & Mirror) != 0;
            const U32 lobeMask = lobeRayMask(lobe.type);
            const bool lobeIndirect = matchesFlags(lobe.type, indirectFlags);
            const U64 lobeBranch = childBranch(pv.branch, q, 0xffff);
            for (U32 i = 0; i < nk; ++i) {
                const U64 sampleBranch = childBranch(lobeBranch, i, 0);
                Identity bid = root;
                bid.branch = sampleBranch;
                Stream bs(bid, pv.depth, 0);
                float r1 = bs.next(), r2 = bs.next();
                const float choice = bs.next(), u = bs.next();
                if (split && !mirror) {
                    const Identity strat{pixel, U64(index) * nk + i, c.seed, lobeBranch};
                    r1 = draw(strat, CameraDimensions, c.sampling);
                    r2 = draw(strat, CameraDimensions + 1, c.sampling);
                    if (g_drawTrace) g_drawTrace->push_back({1.0f, float(q), float(i), r1, r2});
                }
                V3 wi;
                float pdf;
                const V3 f = lobeSample(lobe, slice, r1, r2, wi, pdf);
                if (!(pdf != 0.0f && !isExactlyZero(f)) || !finite3(f) || !finite3(wi)) continue;
                if (!matchesFlags(lobe.type, surfaceFlags(slice, wi))) continue;
                const V3 pt = pv.T * f * (1.0f / (float(nk) * pdf));
                const bool terminators = (lobe.type & ~indirectFlags & AllLobes) != 0;
                (void)terminators;
                U32 hits = 0;
                for (U32 j = 0; j < c.light_count; ++j) {
                    LightHit lh;
                    if ((lobeMask & S.lights[j].visibility_mask) && lightIntersect(S.lights[j], P, wi, 3.4028234663852886e+38f, lh))
                        ++hits;
                }
                V3 direct{0, 0, 0};
                float lightDistance = 0;
                bool lightHit = false;
                U32 lightIndex = 0;
                if (hits) {
                    const U32 chosen = U32(choice * float(hits)) < hits ? U32(choice * float(hits)) : hits - 1;
                    for (U32 j = 0, seen = 0; j < c.light_count; ++j) {
                        LightHit lh;
                        if (!(lobeMask & S.lights[j].visibility_mask) || !lightIntersect(S.lights[j], P, wi, 3.4028234663852886e+38f, lh))
                            continue;
                        if (seen++ != chosen) continue;
                        float pdfL;
                        const V3 Li = lightEval(S.lights[j], wi, lh, pdfL) * float(hits);
                        if (!(pdfL == 0.0f || isExactlyZero(Li)) && fabsf(pdfL) < 3.4e38f && finite3(Li)) {
                            direct = mirror ? Li * pt : Li * pt * powerHeuristic(float(nk) * pdf, float(nl) * pdfL);
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
                float shadowNear = 0.0f, shadowFar = 0.0f, kept = 0.0f;
                const bool bounces = lobeIndirect && !isExactlyZero(indirect) && finite3(indirect);
                bool held = false;
                if (lightHit && !isBlack(direct) && finite3(direct)) {
                    lightShadowSpan(lightDistance, shadowNear, shadowFar);
                    held = !S.volume.present && bounces && thinning >= 1.0f && top + PendingReserve < MaxPending &&
                           lobeMask && (c.shadow_reuse & lobeMask) == lobeMask && shadowNear == 0.0f;
                    if (!held) kept = visible(S, origin, wi, shadowNear, shadowFar, k) ? 1.0f : 0.0f;
                }
                const V3 directBefore = direct;
                Sink lobeSink;
                if (lobe.kind == KindLambert) {
                    lobeSink = texSink;
                } else if (roughSink.kind == SinkRoughness) {
                    lobeSink = roughSink;
                    lobeSink.g = ggxRoughnessLogDerivative(mat.roughness, lobe.frame.z, wo, wi);
                }
                if (kept > 0.0f && !held) {
                    if (kept < 1.0f) direct = direct * kept;
                    if (S.volume.present) {  // (:2939-2945)
                        Identity tid = bid;
                        tid.branch = childBranch(sampleBranch, 0xfffb, 0);
                        Stream vs(tid, pv.depth, 0);
                        direct = direct * transmittance(S.volume, origin, wi, shadowNear, shadowFar, 1.0f,
                                                        c.volume_transmittance_threshold, vs);
                    }
                    radiance = radiance + direct;
                    if (adj) scatterSink(S, *adj, lobeSink, direct, delta);
                }
                if (!bounces) continue;
                float keptShare = 1.0f;
                if (thinning < 1.0f || top + PendingReserve >= MaxPending) {
                    Identity kid = bid;
                    kid.branch = childBranch(sampleBranch, 0xfff7, 0);
                    Stream chance(kid, pv.depth, 0);
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
                sinks[top] = lobeSink;
                Pending child{};
                child.origin = origin;
                child.dir = wi;
                child.T = indirect * keptShare;
                child.branch = sampleBranch;
                child.depth = pv.depth + 1;
                child.diffuseDepth = pv.diffuseDepth + ((lobe.type & Diffuse) ? 1 : 0);
                child.glossyDepth = pv.glossyDepth + ((lobe.type & Glossy) ? 1 : 0);
                child.mirrorDepth = pv.mirrorDepth + (mirror ? 1 : 0);
                child.nonMirrorDepth = pv.nonMirrorDepth + (mirror ? 0 : 1);
                child.farT = 3.402823466e+38f;
                child.nearT = 0.0f;
                child.heldLight = ~0u;
                child.volumeDepth = pv.volumeDepth + 1;  // (:3000)
                if (textured)
                    child.diff = scatterDifferentials(lobe.kind == KindLambert, lobe.frame, lobe.roughness, hd, P, wo, wi,
                                                      r1, r2, origin, primaryScale);
                if (held) {
                    child.heldLight = lightIndex;
                    child.heldFar = shadowFar;
                    child.held = directBefore;
                }
                stack[top++] = child;
            }
        }
    }
    if (adj)
        while (!recs.empty()) {
            scatterSink(S, *adj, recs.back().sink, radiance - recs.back().mark, delta);
            recs.pop_back();
        }
    out[0] = radiance.x;
    out[1] = radiance.y;
    out[2] = radiance.z;
    out[3] = alpha;
    return 0;
}

// The pixel of a tile's i-th position (async_integrator.h:3128-3140), ~0u outside the image or viewport.
U32 tilePixel(U32 tile, U32 position, const x2_config& c) {
    const U32 tilesX = (c.width + 7) / 8;
    const U32 x = (tile % tilesX) * 8 + position % 8, y = (tile / tilesX) * 8 + position / 8;
    if (!(x < c.width && y < c.height)) return ~0u;
    if (int(x) < c.viewport[0] || int(y) < c.viewport[1] || int(x) >= c.viewport[2] || int(y) >= c.viewport[3])
        return ~0u;
    return y * c.width + x;
}

// Chunk rendering over worker threads, each claiming chunks by an atomic counter (async_cpu.h:1584-1673
// renderChunk: sample-major over the tile's 64 pixels into tile-local FP64, then added to the film). Chunks of the
// same tile go to the same worker in list order so each pixel's samples add in index order.
void renderChunks(const Scene& S, const x2_config& c, const std::vector<x2_chunk>& chunks, U32 threads, double* sums,
                  U32* counts, U64 counters[3]) {
    std::vector<std::vector<U32>> byTile;
    std::vector<U32> tiles;
    {
        std::vector<int> slot(((c.width + 7) / 8) * ((c.height + 7) / 8), -1);
        for (U32 i = 0; i < chunks.size(); ++i) {
            const U32 t = chunks[i].tile;
            if (t >= slot.size()) continue;
            if (slot[t] < 0) {
                slot[t] = int(byTile.size());
                byTile.emplace_back();
                tiles.push_back(t);
            }
            byTile[slot[t]].push_back(i);
        }
    }
    std::atomic<U32> next{0};
    std::atomic<U64> traces{0}, shadows{0}, errors{0};
    auto worker = [&]() {
        Counters k;
        U64 errs = 0;
        for (;;) {
            const U32 w = next.fetch_add(1);
            if (w >= byTile.size()) break;
            double local[64][4] = {};
            U32 n[64] = {};
            for (const U32 ci : byTile[w]) {
                const x2_chunk& ch = chunks[ci];
                for (U32 s = 0; s < ch.count; ++s)
                    for (U32 pos = 0; pos < 64; ++pos) {
                        const U32 pixel = tilePixel(ch.tile, pos, c);
                        if (pixel == ~0u) continue;
                        double rgba[4];
                        const int st = integrate(S, c, pixel, ch.first + s, rgba, k);
                        if (st == 0) {
                            for (int q = 0; q < 4; ++q) local[pos][q] += rgba[q];
                            ++n[pos];
                        } else if (st != 3) {  // a skipped sample adds nothing, not even its count (async_gpu.cu:347)
                            ++errs;
                        }
                    }
            }
            for (U32 pos = 0; pos < 64; ++pos) {
                const U32 pixel = tilePixel(tiles[w], pos, c);
                if (pixel == ~0u) continue;
                for (int q = 0; q < 4; ++q) sums[size_t(pixel) * 4 + q] += local[pos][q];
                counts[pixel] += n[pos];
            }
        }
        traces += k.traces;
        shadows += k.shadows;
        errors += errs;
    };
    threads = std::max<U32>(1, threads);
    std::vector<std::thread> pool;
    for (U32 t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    counters[0] += traces;
    counters[1] += shadows;
    counters[2] += errors;
}
// The adjoint over every sample of a uniform frame, tiles over worker threads, each with its own FP64 gradients,
// summed at the end.
void adjointFrame(const Scene& S, const x2_config& c, U32 threads, const float* dImage, double* gradTex, double* gradR,
                  U64 counters[3]) {
    const U32 tiles = ((c.width + 7) / 8) * ((c.height + 7) / 8);
    size_t texFloats = 0, levelFloats = 0;
    for (const Texture& t : S.textures) {
        texFloats += size_t(t.width) * t.height * 3;
        for (size_t k = 1; k < t.levels.size(); ++k) levelFloats += size_t(t.levels[k].width) * t.levels[k].height * 3;
    }
    threads = std::max<U32>(1, threads);
    std::vector<std::vector<double>> gt(threads, std::vector<double>(texFloats, 0.0)), gr(threads, std::vector<double>(S.materials.size(), 0.0));
    std::vector<std::vector<double>> gl(threads, std::vector<double>(levelFloats + 1, 0.0));
    std::atomic<U32> next{0};
    std::atomic<U64> traces{0}, shadows{0}, errors{0};
    auto worker = [&](U32 w) {
        Counters k;
        U64 errs = 0;
        Adjoint A;
        A.dImage = dImage;
        A.gradTextures = gt[w].data();
        A.gradLevels = gl[w].data();
        A.gradRoughness = gr[w].data();
        for (;;) {
            const U32 tile = next.fetch_add(1);
            if (tile >= tiles) break;
            for (U32 s = 0; s < c.spp; ++s)
                for (U32 pos = 0; pos < 64; ++pos) {
                    const U32 pixel = tilePixel(tile, pos, c);
                    if (pixel == ~0u) continue;
                    double rgba[4];
                    if (integrate(S, c, pixel, s, rgba, k, &A) != 0) ++errs;
                }
        }
        traces += k.traces;
        shadows += k.shadows;
        errors += errs;
    };
    std::vector<std::thread> pool;
    for (U32 t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& t : pool) t.join();
    for (U32 w = 1; w < threads; ++w) {
        for (size_t i = 0; i < texFloats; ++i) gt[0][i] += gt[w][i];
        for (size_t i = 0; i < levelFloats; ++i) gl[0][i] += gl[w][i];
    }
    for (const Texture& t : S.textures) pushDownPyramid(t, gt[0].data() + t.gradOffset, gl[0].data());
    for (size_t i = 0; i < texFloats; ++i) gradTex[i] += gt[0][i];
    for (U32 w = 0; w < threads; ++w)
        for (size_t i = 0; i < S.materials.size(); ++i) gradR[i] += gr[w][i];
    counters[0] += traces;
    counters[1] += shadows;
    counters[2] += errors;
}
}  // namespace

struct oracle_scene {
    Scene s;
};

extern "C" {
void oracle_philox(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    const Words w = philox(Words{{counter[0], counter[1], counter[2], counter[3]}}, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = w.v[i];
}
void oracle_block(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t group, uint32_t out[4]) {
    const Words w = block(Identity{pixel, index, stream, branch}, group);
    for (int i = 0; i < 4; ++i) out[i] = w.v[i];
}
uint32_t oracle_word(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t dimension) {
    return word(Identity{pixel, index, stream, branch}, dimension);
}
float oracle_unit(uint32_t bits) { return unit(bits); }
uint32_t oracle_sobol(uint32_t index, uint32_t dimension) { return dimension < 17 ? sobol(index, dimension) : 0u; }
void oracle_directions(uint32_t out[17 * 32]) { std::memcpy(out, kDirections.v, sizeof(kDirections.v)); }
uint32_t oracle_owen_scramble(uint32_t x, uint32_t seed) { return owenScramble(x, seed); }
uint32_t oracle_sampling_word(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t dimension,
                              uint32_t policy) {
    return samplingWord(Identity{pixel, index, stream, branch}, dimension, policy);
}
uint64_t oracle_child_branch(uint64_t parent, uint32_t a, uint32_t b) { return childBranch(parent, a, b); }
void oracle_stream(uint32_t pixel, uint64_t index, uint64_t stream, uint64_t branch, uint32_t bounce, uint32_t n,
                   float* out) {
    Stream s(Identity{pixel, index, stream, branch}, bounce, 0);
    for (uint32_t i = 0; i < n; ++i) out[i] = s.next();
}
void oracle_sincos(float angle, float* s, float* c) { sincosPoly(angle, *s, *c); }

static Lobe testLobe(uint32_t kind, const float N[3], float roughness) {
    x2_material m{};
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
    m.roughness = roughness;
    m.diffuse_weight = kind == 0 ? 1.0f : 0.0f;
    m.specular_weight = kind == 1 ? 1.0f : 0.0f;
    Bsdf b;
    buildBsdf(m, load3(N), b, load3(m.albedo));
    return b.lobes[0];
}
void oracle_lobe_eval(uint32_t kind, const float N[3], const float Ng[3], const float wo[3], const float wi[3],
                      float roughness, float f[3], float* pdf) {
    const Lobe l = testLobe(kind, N, roughness);
    const Slice s{load3(Ng), load3(wo), AllLobeTypes};
    store3(f, lobeEval(l, s, load3(wi), *pdf));
}
void oracle_lobe_sample(uint32_t kind, const float N[3], const float Ng[3], const float wo[3], float roughness,
                        float r1, float r2, float wi[3], float f[3], float* pdf) {
    const Lobe l = testLobe(kind, N, roughness);
    const Slice s{load3(Ng), load3(wo), AllLobeTypes};
    V3 w;
    store3(f, lobeSample(l, s, r1, r2, w, *pdf));
    store3(wi, w);
}
int oracle_light_sample(const x2_light* l, const float p[3], const float* n, const float r[3], float wi[3],
                        float* distance, float Li[3], float* pdf) {
    V3 nn = n ? load3(n) : V3{0, 0, 0}, w;
    LightHit h;
    if (!lightSample(*l, load3(p), n ? &nn : nullptr, r, w, h)) return 0;
    store3(Li, lightEval(*l, w, h, *pdf));
    store3(wi, w);
    *distance = h.distance;
    return 1;
}
int oracle_light_intersect(const x2_light* l, const float p[3], const float wi[3], float Li[3], float* pdf,
                           float* distance) {
    LightHit h;
    if (!lightIntersect(*l, load3(p), load3(wi), 3.4028234663852886e+38f, h)) return 0;
    store3(Li, lightEval(*l, load3(wi), h, *pdf));
    *distance = h.distance;
    return 1;
}
int oracle_light_can_illuminate(const x2_light* l, const float p[3], const float* n) {
    V3 nn = n ? load3(n) : V3{0, 0, 0};
    return canIlluminate(*l, load3(p), n ? &nn : nullptr) ? 1 : 0;
}
int oracle_triangle(const float o[3], const float d[3], float near_t, float far_t, const float a[3], const float b[3],
                    const float c[3], float* t) {
    const WoopRay w = woopSetup(load3(o), load3(d), near_t, far_t);
    return woopHit(w, load3(a), load3(b), load3(c), *t) ? 1 : 0;
}

int oracle_camera_ray(const x2_config* c, float x, float y, float origin[3], float dir[3], float* nearT, float* farT) {
    Ray r;
    if (!cameraRay(*c, x, y, r)) return 0;
    store3(origin, r.origin);
    store3(dir, r.direction);
    *nearT = r.nearT;
    *farT = r.farT;
    return 1;
}
int oracle_geometry_hit(const float o[3], const float d[3], const float a[3], const float b[3], const float c[3],
                        float out[13]) {
    Hit h;
    const Ray ray{load3(o), load3(d), 0.0f, 3.4e38f};
    if (!geometryHit(h, ray, load3(a), load3(b), load3(c))) return 0;
    out[0] = h.t;
    store3(out + 1, h.point);
    store3(out + 4, h.normal);
    store3(out + 7, h.origin);
    store3(out + 10, h.below);
    return 1;
}

oracle_scene* oracle_scene_create(const x2_scene* in) {
    if (!in || !in->positions || !in->indices || !in->geom_first || !in->geom_material || !in->materials) return nullptr;
    auto* o = new oracle_scene;
    Scene& s = o->s;
    s.triangles = in->triangle_count;
    s.tri.resize(size_t(in->triangle_count) * 3);
    s.geomOf.resize(in->triangle_count);
    s.primOf.resize(in->triangle_count);
    s.material.resize(in->triangle_count);
    for (U32 g = 0; g < in->geom_count; ++g)
        for (U32 t = in->geom_first[g]; t < in->geom_first[g + 1] && t < in->triangle_count; ++t) {
            s.geomOf[t] = g;
            s.primOf[t] = t - in->geom_first[g];
            s.material[t] = in->geom_material[g];
        }
    for (U32 t = 0; t < in->triangle_count; ++t)
        for (int v = 0; v < 3; ++v) s.tri[size_t(t) * 3 + v] = load3(in->positions + size_t(in->indices[size_t(t) * 3 + v]) * 3);
    if (in->geom_flags) {
        bool any = false;
        for (U32 g = 0; g < in->geom_count; ++g) any = any || (in->geom_flags[g] & X2_GEOM_SINGLE_SIDED);
        if (any) {
            s.singleSided.assign(in->triangle_count, 0);
            for (U32 t = 0; t < in->triangle_count; ++t) s.singleSided[t] = (in->geom_flags[s.geomOf[t]] & X2_GEOM_SINGLE_SIDED) ? 1 : 0;
        }
    }
    if (in->uvs) {
        s.triUV.resize(size_t(in->triangle_count) * 6);
        for (U32 t = 0; t < in->triangle_count; ++t)
            for (int v = 0; v < 3; ++v) {
                const float* uv = in->uvs + size_t(in->indices[size_t(t) * 3 + v]) * 2;
                s.triUV[size_t(t) * 6 + v * 2] = uv[0];
                s.triUV[size_t(t) * 6 + v * 2 + 1] = uv[1];
            }
    }
    size_t gradOffset = 0;
    size_t levelOffset = 0;
    for (U32 i = 0; i < in->texture_count; ++i) {
        const x2_texture& t = in->textures[i];
        Texture tex;
        tex.width = t.width;
        tex.height = t.height;
        setLevel0(tex, t.texels);
        tex.gradOffset = gradOffset;
        gradOffset += size_t(t.width) * t.height * 3;
        for (size_t k = 1; k < tex.levels.size(); ++k) {
            tex.levelGrad[k] = levelOffset;
            levelOffset += size_t(tex.levels[k].width) * tex.levels[k].height * 3;
        }
        s.textures.push_back(std::move(tex));
    }
    s.materials.assign(in->materials, in->materials + in->material_count);
    if (in->light_count) s.lights.assign(in->lights, in->lights + in->light_count);
    if (in->volume_count && in->volumes && !packVolume(in->volumes[0], s.volume)) {
        delete o;
        return nullptr;
    }
    buildBvh(s);
    return o;
}
void oracle_scene_destroy(oracle_scene* s) { delete s; }
void oracle_closest(const oracle_scene* o, const float org[3], const float d[3], float near_t, float far_t, int brute,
                    uint32_t* geom, uint32_t* prim, float* t) {
    const U32 tri = traverse(o->s, load3(org), load3(d), near_t, far_t, false, brute != 0, *t);
    *geom = tri == ~0u ? ~0u : o->s.geomOf[tri];
    *prim = tri == ~0u ? 0u : o->s.primOf[tri];
}
int oracle_occluded(const oracle_scene* o, const float org[3], const float d[3], float near_t, float far_t, int brute) {
    float t;
    return traverse(o->s, load3(org), load3(d), near_t, far_t, true, brute != 0, t) != ~0u ? 1 : 0;
}
int oracle_sample(const oracle_scene* o, const x2_config* c, uint32_t pixel, uint32_t index, double rgba[4],
                  uint64_t counters[2]) {
    Counters k;
    const int st = integrate(o->s, *c, pixel, index, rgba, k);
    counters[0] += k.traces;
    counters[1] += k.shadows;
    return st;
}
int oracle_first_hit_draws(const oracle_scene* o, const x2_config* c, uint32_t pixel, uint32_t index, float* out,
                           uint32_t capacity, uint32_t* count) {
    std::vector<DrawRecord> trace;
    g_drawTrace = &trace;
    Counters k;
    double rgba[4];
    const int st = integrate(o->s, *c, pixel, index, rgba, k);
    g_drawTrace = nullptr;
    *count = U32(trace.size());
    for (U32 i = 0; i < std::min<U32>(capacity, U32(trace.size())); ++i) std::memcpy(out + 5 * i, &trace[i], 5 * sizeof(float));
    return st;
}
int oracle_render(const oracle_scene* o, const x2_config* c, uint32_t threads, double* sums, uint32_t* counts,
                  uint64_t counters[3]) {
    std::vector<x2_chunk> chunks;
    const U32 tiles = ((c->width + 7) / 8) * ((c->height + 7) / 8);
    for (U32 t = 0; t < tiles; ++t) chunks.push_back({t, 0, c->spp, 0});
    renderChunks(o->s, *c, chunks, threads, sums, counts, counters);
    return 0;
}
int oracle_adjoint(const oracle_scene* o, const x2_config* c, uint32_t threads, const float* d_image,
                   double* grad_textures, double* grad_roughness, uint64_t counters[3]) {
    adjointFrame(o->s, *c, threads, d_image, grad_textures, grad_roughness, counters);
    return 0;
}
void oracle_set_texture(oracle_scene* o, uint32_t texture, const float* texels) { setLevel0(o->s.textures.at(texture), texels); }
void oracle_set_texture_coarse(oracle_scene* o, uint32_t texture, uint32_t from) {
    o->s.textures.at(texture).coarseFrom = from;
}
void oracle_set_texture_half(oracle_scene* o, uint32_t texture, int half) {
    Texture& t = o->s.textures.at(texture);
    const std::vector<float> level0 = t.levels.at(0).texels;
    t.half = half != 0;
    setLevel0(t, level0.data());
}
void oracle_set_material(oracle_scene* o, uint32_t material, const x2_material* m) { o->s.materials.at(material) = *m; }
void oracle_set_eval_roughness(oracle_scene* o, uint32_t material, float r) {
    if (o->s.evalRoughness.empty()) o->s.evalRoughness.assign(o->s.materials.size(), -1.0f);
    o->s.evalRoughness.at(material) = r;
}
float oracle_ggx_roughness_log_derivative(float r, const float N[3], const float wo[3], const float wi[3]) {
    return ggxRoughnessLogDerivative(r, load3(N), load3(wo), load3(wi));
}
int oracle_surface_lambert(const float albedo[3], const float N[3], float u, float v, float wi[3], float f[3],
                           float* pdf) {
    V3 w, ff;
    if (!surfaceLambert(load3(albedo), load3(N), u, v, w, ff, *pdf)) return 0;
    store3(wi, w);
    store3(f, ff);
    return 1;
}
int oracle_texture_lookup(const x2_texture* t, float u, float v, float rgb[3]) {
    Texture tex;
    tex.width = t->width;
    tex.height = t->height;
    tex.levels.resize(1);
    tex.levels[0].width = t->width;
    tex.levels[0].height = t->height;
    tex.levels[0].texels.assign(t->texels, t->texels + size_t(t->width) * t->height * 3);
    V3 out;
    if (!bilinearAt(tex, 0, u, 1 - v, out) || !finite3(out)) return 0;
    store3(rgb, out);
    return 1;
}
int oracle_map_lookup_swap(const x2_texture* t, int half, uint32_t mode, int swap, const float f[6], float rgb[3]) {
    Texture tex;
    tex.width = t->width;
    tex.height = t->height;
    tex.half = half != 0;
    setLevel0(tex, t->texels);
    V3 out;
    if (!mapLookup(tex, mode, Footprint{f[0], f[1], f[2], f[3], f[4], f[5]}, out, swap != 0)) return 0;
    store3(rgb, out);
    return 1;
}
int oracle_map_lookup(const x2_texture* t, int half, uint32_t mode, const float f[6], float rgb[3]) {
    Texture tex;
    tex.width = t->width;
    tex.height = t->height;
    tex.half = half != 0;
    setLevel0(tex, t->texels);
    V3 out;
    if (!mapLookup(tex, mode, Footprint{f[0], f[1], f[2], f[3], f[4], f[5]}, out)) return 0;
    store3(rgb, out);
    return 1;
}
void oracle_hit_tex_derivatives(const float a[3], const float b[3], const float c[3], const float uv[6],
                                 const float org[3], const float dir[3], const float dx[3], const float dy[3],
                                 uint32_t spp, float out[4]) {
    HitDiff h;
    const Ray ray{load3(org), load3(dir), 0.0f, 3.402823466e38f};
    hitTexDerivatives(load3(a), load3(b), load3(c), uv, ray, load3(dx), load3(dy), spp, h);
    out[0] = h.dsdx; out[1] = h.dtdx; out[2] = h.dsdy; out[3] = h.dtdy;
}
int oracle_transfer(const float diff[12], int valid, const float org[3], const float dir[3], const float Ng[3],
                    const float P[3], const float dPds[3], const float dPdt[3], float out[10]) {
    RayDiff d;
    d.dOx = load3(diff); d.dOy = load3(diff + 3); d.dDx = load3(diff + 6); d.dDy = load3(diff + 9);
    d.valid = valid ? 1u : 0u;
    HitDiff h;
    transferDifferentials(d, load3(org), load3(dir), load3(Ng), load3(P), h);
    differentialTexDerivatives(load3(dPds), load3(dPdt), load3(Ng), h);
    store3(out, h.dPdx);
    store3(out + 3, h.dPdy);
    out[6] = h.dsdx; out[7] = h.dtdx; out[8] = h.dsdy; out[9] = h.dtdy;
    return h.valid ? 1 : 0;
}
int oracle_triangle_partials(const float x0[3], const float x1[3], const float x2[3], const float uv[6], float dPds[3],
                             float dPdt[3]) {
    V3 s, t;
    if (!trianglePartials(load3(x0), load3(x1), load3(x2), uv, uv + 2, uv + 4, s, t)) return 0;
    store3(dPds, s);
    store3(dPdt, t);
    return 1;
}
void oracle_hemisphere_differentials(float r1, float r2, float out[6]) {
    V3 a, b;
    localHemisphereCosineDifferentials(r1, r2, a, b);
    store3(out, a);
    store3(out + 3, b);
}
int oracle_scatter_differentials(uint32_t kind, const float N[3], float roughness, const float incoming[12],
                                 const float dP[6], int diffValid, const float P[3], const float wo[3], const float wi[3],
                                 float r1, float r2, const float childOrigin[3], float primaryScale, float out[12]) {
    HitDiff h;
    h.incoming.dOx = load3(incoming); h.incoming.dOy = load3(incoming + 3);
    h.incoming.dDx = load3(incoming + 6); h.incoming.dDy = load3(incoming + 9);
    h.incoming.valid = 1;
    h.dPdx = load3(dP);
    h.dPdy = load3(dP + 3);
    h.valid = diffValid != 0;
    const float alpha = fmaxf(0.001f, roughness * roughness);
    const RayDiff d = scatterDifferentials(kind == 0, frameOf(load3(N)), alpha, h, load3(P), load3(wo), load3(wi), r1,
                                           r2, load3(childOrigin), primaryScale);
    store3(out, d.dOx); store3(out + 3, d.dOy); store3(out + 6, d.dDx); store3(out + 9, d.dDy);
    return int(d.valid);
}
uint32_t oracle_pyramid(const x2_texture* t, int half, float* out, uint64_t capacity, uint32_t* sizes) {
    Texture tex;
    tex.width = t->width;
    tex.height = t->height;
    tex.half = half != 0;
    setLevel0(tex, t->texels);
    uint64_t at = 0;
    for (size_t k = 0; k < tex.levels.size(); ++k) {
        sizes[2 * k] = tex.levels[k].width;
        sizes[2 * k + 1] = tex.levels[k].height;
        for (float v : tex.levels[k].texels)
            if (at < capacity) out[at++] = v;
    }
    return uint32_t(tex.levels.size());
}
float oracle_volume_density(const oracle_scene* o, const float p[3]) {
    return o->s.volume.present ? densityAt(o->s.volume, load3(p)) : 0.0f;
}
float oracle_volume_transmittance(const oracle_scene* o, const x2_config* c, const float org[3], const float dir[3],
                                  float tnear, float tfar, uint32_t pixel, uint64_t index, uint64_t branch) {
    if (!o->s.volume.present) return 1.0f;
    Stream s(Identity{pixel, index, c->seed, branch}, 0, 0);
    return transmittance(o->s.volume, load3(org), load3(dir), tnear, tfar, 1.0f, c->volume_transmittance_threshold, s);
}
float oracle_phase_eval(float g, const float wo[3], const float wi[3]) { return phaseEval(g, load3(wo), load3(wi)); }
void oracle_phase_sample(float g, const float wo[3], float u1, float u2, float wi[3]) {
    store3(wi, phaseSample(g, load3(wo), u1, u2));
}
int oracle_volume_march(const oracle_scene* o, const x2_config* c, const float org[3], const float dir[3], float tnear,
                        float tfar, uint32_t depth, uint64_t branch, float out[8]) {
    const Volume& v = o->s.volume;
    float a, b;
    if (!v.present || !clipBox(v, load3(org), load3(dir), tnear, tfar, a, b)) return 0;
    MarchResult m;
    marchVolume(v, *c, load3(org), load3(dir), a, b, depth, depth == 0, Stream(Identity{0, 0, c->seed, branch}, depth, 0), m);
    const U32 n = U32(m.steps.size());
    out[0] = m.reached ? 0.0f : m.vt;
    store3(out + 1, m.emitted);
    out[4] = n ? m.steps[0].tStart : 0.0f;
    out[5] = n ? m.steps[n - 1].tStart + m.steps[n - 1].delta : 0.0f;
    out[6] = m.total;
    out[7] = m.reached ? 1.0f : 0.0f;
    return int(n);
}
// The emission region (test hooks): its cells and cell volume (0: the medium has no region), a cell's probability,
// a drawn place, the direction pdf from a point, and the emissive integral along a ray.
int oracle_region_info(const oracle_scene* o, int res[3], float* invUnitVolume, float distToWorld[12]) {
    const Volume& v = o->s.volume;
    if (!v.present || !v.hasRegion) return 0;
    for (int a = 0; a < 3; ++a) res[a] = v.regionRes[a];
    *invUnitVolume = v.invUnitVolume;
    for (int i = 0; i < 12; ++i) distToWorld[i] = v.distToWorld[i];
    return 1;
}
float oracle_region_cell_pdf(const oracle_scene* o, int x, int y, int z) { return regionCellPdf(o->s.volume, x, y, z); }
void oracle_region_point(const oracle_scene* o, const float u[3], float out[3]) {
    store3(out, regionPoint(o->s.volume, u[0], u[1], u[2]));
}
float oracle_region_pdf(const oracle_scene* o, const float p[3], const float wi[3], float* tEnd) {
    return regionPdf(o->s.volume, load3(p), load3(wi), *tEnd);
}
void oracle_emissive_integral(const oracle_scene* o, const x2_config* c, const float org[3], const float dir[3],
                              float tnear, float tEnd, uint32_t depth, uint64_t branch, float out[3]) {
    Counters k;
    Stream s(Identity{0, 0, c->seed, branch}, depth, 0);
    store3(out, emissiveIntegral(o->s, *c, s, load3(org), load3(dir), tnear, tEnd, depth, k));
}
int oracle_render_chunks(const oracle_scene* o, const x2_config* c, const x2_chunk* chunks, uint32_t n,
                         uint32_t threads, double* sums, uint32_t* counts, uint64_t counters[3]) {
    renderChunks(o->s, *c, std::vector<x2_chunk>(chunks, chunks + n), threads, sums, counts, counters);
    return 0;
}
// An adaptive pass's samples (async_gpu.cu:331-345, :440-446): pixel p renders indices first[p] .. first[p] +
// count[p] - 1, added into the running film; the odd-index samples also into odd.
int oracle_render_pixels(const oracle_scene* o, const x2_config* c, const uint32_t* first, const uint32_t* count,
                         uint32_t threads, double* sums, double* odd, uint32_t* counts, uint64_t counters[3]) {
    const Scene& S = o->s;
    std::atomic<U32> nextRow{0};
    std::atomic<U64> traces{0}, shadows{0}, errors{0};
    auto worker = [&]() {
        Counters k;
        U64 errs = 0;
        for (;;) {
            const U32 y = nextRow.fetch_add(1);
            if (y >= c->height) break;
            for (U32 x = 0; x < c->width; ++x) {
                const U32 pixel = y * c->width + x;
                for (U32 s = 0; s < count[pixel]; ++s) {
                    const U32 index = first[pixel] + s;
                    double rgba[4];
                    const int st = integrate(S, *c, pixel, index, rgba, k);
                    if (st == 0) {
                        for (int q = 0; q < 4; ++q) sums[size_t(pixel) * 4 + q] += rgba[q];
                        if (index & 1u)
                            for (int q = 0; q < 4; ++q) odd[size_t(pixel) * 4 + q] += rgba[q];
                        ++counts[pixel];
                    } else if (st != 3) {
                        ++errs;
                    }
                }
            }
        }
        traces += k.traces;
        shadows += k.shadows;
        errors += errs;
    };
    std::vector<std::thread> pool;
    for (U32 t = 0; t < std::max<U32>(1, threads); ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    counters[0] = traces;
    counters[1] = shadows;
    counters[2] = errors;
    return 0;
}
}
