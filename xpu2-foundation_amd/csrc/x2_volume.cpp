// x2_volume.cpp — packs a medium's dense grids into 8^3 bricks for the device. A brick is kept when any of its
// density voxels is nonzero; temperature is kept for the same bricks (it reads zero elsewhere). A brick's majorant
// is the largest density among its own voxels and its 26 neighbours' (a point in the brick interpolates voxels at
// most one voxel away), the bound the device's ratio tracking samples free flights against. The derived sizes are
// computed here once, in float, so that every reader of them rounds alike. A medium that glows (a temperature grid
// and a positive emission scale) is also made a light source: its emission region (emissionRegion below).
#include "x2_volume.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace x2 {
namespace {
float halfToFloat(uint16_t h) {  // IEEE binary16, exactly
    const uint32_t s = uint32_t(h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
    if (e == 0) {
        const float v = float(m) * (1.0f / 16777216.0f);
        return s ? -v : v;
    }
    const uint32_t u = s | (e == 31 ? 0x7F800000u | (m << 13) : ((e + 112u) << 23) | (m << 13));
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
float rampLuminance(const float* ramp, uint32_t i) {
    return 0.299f * ramp[3 * i] + 0.587f * ramp[3 * i + 1] + 0.114f * ramp[3 * i + 2];
}
// The ramp's entry index below kelvin, as the device's lookup places it (rampAt's x, floored and clamped).
uint32_t rampIndex(const x2_volume& in, float kelvin) {
    float x = kelvin / in.ramp_max_kelvin * float(in.ramp_size - 1);
    x = fminf(fmaxf(x, 0.0f), float(in.ramp_size - 1));
    return uint32_t(floorf(x));
}
// One 1D distribution of the pool (Distribution1D as cpu_scene.h:1520-1525 lays it out): its integral, n, the n
// values, then the cdf's n + 1 entries, accumulated value by value over n and normalized (uniform when it sums to 0).
uint32_t appendDistribution(std::vector<float>& pool, const float* func, uint32_t n) {
    const uint32_t at = uint32_t(pool.size());
    std::vector<float> cdf(size_t(n) + 1, 0.0f);
    for (uint32_t i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / float(n);
    const float integral = cdf[n];
    for (uint32_t i = 1; i <= n; ++i) cdf[i] = integral == 0.0f ? float(i) / float(n) : cdf[i] / integral;
    pool.push_back(integral);
    pool.push_back(float(n));
    pool.insert(pool.end(), func, func + n);
    pool.insert(pool.end(), cdf.begin(), cdf.end());
    return at;
}
// The medium's emission region (the reference builds one per emissive volume from a histogram of its emission
// luminance, cpu_scene.h:1526-1545): cells of EmissionCell^3 voxels over the medium's box, each cell's value an upper
// bound of the emission luminance anywhere in it: emission scale x the largest density x the brightest ramp entry
// a temperature within one voxel of the cell can interpolate (a point interpolates voxels at most one voxel away, and
// the ramp is linear between its entries). The distribution is then the reference's: each (z, y) row's x
// distribution, each z's y marginal over the rows' integrals, the z marginal over those.
void emissionRegion(const x2_volume& in, VolumeHost& out) {
    VolumeDesc& v = out.desc;
    int n[3];
    for (int a = 0; a < 3; ++a) n[a] = int(in.res[a] / EmissionCell);
    const size_t cells = size_t(n[0]) * size_t(n[1]) * size_t(n[2]);
    std::vector<float> value(cells, 0.0f);
    bool any = false;
    const int rx = int(in.res[0]), ry = int(in.res[1]), rz = int(in.res[2]);
    const int C = int(EmissionCell);
    for (int cz = 0; cz < n[2]; ++cz)
        for (int cy = 0; cy < n[1]; ++cy)
            for (int cx = 0; cx < n[0]; ++cx) {
                float dMax = 0.0f, tMin = FLT_MAX, tMax = -FLT_MAX;
                for (int z = std::max(cz * C - 1, 0); z <= std::min(cz * C + C, rz - 1); ++z)
                    for (int y = std::max(cy * C - 1, 0); y <= std::min(cy * C + C, ry - 1); ++y)
                        for (int x = std::max(cx * C - 1, 0); x <= std::min(cx * C + C, rx - 1); ++x) {
                            const size_t i = (size_t(z) * size_t(ry) + size_t(y)) * size_t(rx) + size_t(x);
                            dMax = fmaxf(dMax, halfToFloat(in.density[i]));
                            const float t = halfToFloat(in.temperature[i]);
                            tMin = fminf(tMin, t);
                            tMax = fmaxf(tMax, t);
                        }
                if (!(dMax > 0.0f)) continue;
                const uint32_t i0 = rampIndex(in, tMin), i1 = std::min(rampIndex(in, tMax) + 1, in.ramp_size - 1);
                float lum = 0.0f;
                for (uint32_t i = i0; i <= i1; ++i) lum = fmaxf(lum, rampLuminance(in.ramp, i));
                const float e = in.emission_scale * dMax * lum;
                value[(size_t(cz) * size_t(n[1]) + size_t(cy)) * size_t(n[0]) + size_t(cx)] = e;
                any = any || e > 0.0f;
            }
    if (!any) return;
    EmissionRegion& r = v.region;
    r = EmissionRegion{};
    for (int a = 0; a < 3; ++a) r.res[a] = n[a];
    r.visibility = 0x3ffu;
    double cell[3];
    for (int a = 0; a < 3; ++a) cell[a] = (double(in.bounds_max[a]) - double(in.bounds_min[a])) / double(n[a]);
    for (int a = 0; a < 3; ++a) {
        r.distToWorld[4 * a] = float(cell[a]);  // the diagonal of the 3x3 part, rows x, y, z
        r.distToWorld[9 + a] = in.bounds_min[a];
        r.worldToDist[4 * a] = float(1.0 / cell[a]);
        r.worldToDist[9 + a] = float(-double(in.bounds_min[a]) / cell[a]);
    }
    r.invUnitVolume = float(1.0 / (cell[0] * cell[1] * cell[2]));
    std::vector<float>& pool = out.emissionPool;
    pool.clear();
    std::vector<uint32_t> rows(size_t(n[2]) * size_t(n[1]));
    r.conditionalX = uint32_t(pool.size());
    for (int z = 0; z < n[2]; ++z)
        for (int y = 0; y < n[1]; ++y)
            rows[size_t(z) * size_t(n[1]) + size_t(y)] = appendDistribution(
                pool, &value[(size_t(z) * size_t(n[1]) + size_t(y)) * size_t(n[0])], uint32_t(n[0]));
    r.marginalY = uint32_t(pool.size());
    std::vector<float> rowIntegrals(static_cast<size_t>(n[1])), planeIntegrals(static_cast<size_t>(n[2]));
    for (int z = 0; z < n[2]; ++z) {
        for (int y = 0; y < n[1]; ++y) rowIntegrals[size_t(y)] = pool[rows[size_t(z) * size_t(n[1]) + size_t(y)]];
        planeIntegrals[size_t(z)] = pool[appendDistribution(pool, rowIntegrals.data(), uint32_t(n[1]))];
    }
    r.marginalZ = appendDistribution(pool, planeIntegrals.data(), uint32_t(n[2]));
}
}  // namespace

bool packVolume(const x2_volume& in, VolumeHost& out, std::string& err) {
    if (!in.density) return err = "volume without a density grid", false;
    if (!(in.sigma_t_scale >= 0.0f)) return err = "negative extinction scale", false;
    if (in.temperature && (!in.ramp || in.ramp_size < 2 || !(in.ramp_max_kelvin > 0.0f)))
        return err = "a temperature grid needs a ramp of at least two entries over a positive range", false;
    out = VolumeHost{};
    VolumeDesc& v = out.desc;
    float feature = FLT_MAX;
    for (int a = 0; a < 3; ++a) {
        if (in.res[a] == 0 || in.res[a] % 8 || in.res[a] > 4096) return err = "volume res must be multiples of 8", false;
        const float ext = in.bounds_max[a] - in.bounds_min[a];
        if (!(ext > 0.0f)) return err = "empty volume bounds", false;
        v.bmin[a] = in.bounds_min[a];
        v.bmax[a] = in.bounds_max[a];
        v.res[a] = in.res[a];
        v.nb[a] = in.res[a] / 8;
        v.scaleVox[a] = float(in.res[a]) / ext;
        v.brickExtent[a] = ext / float(v.nb[a]);
        feature = fminf(feature, ext / float(in.res[a]));
    }
    v.minFeature = feature;
    v.sigmaTScale = in.sigma_t_scale;
    v.g = in.g;
    v.emissionScale = in.emission_scale;
    for (int a = 0; a < 3; ++a) v.albedo[a] = in.albedo[a];
    v.rampMax = 1.0f;
    v.rampSize = 0;
    if (in.temperature) {
        v.rampMax = in.ramp_max_kelvin;
        v.rampSize = in.ramp_size;
        out.ramp.assign(in.ramp, in.ramp + size_t(in.ramp_size) * 3);
    }
    const uint32_t nx = v.nb[0], ny = v.nb[1], nz = v.nb[2];
    const size_t bricks = size_t(nx) * ny * nz;
    out.table.assign(bricks, ~0u);
    std::vector<float> brickMax(bricks, 0.0f);
    uint32_t used = 0;
    for (uint32_t bz = 0; bz < nz; ++bz)
        for (uint32_t by = 0; by < ny; ++by)
            for (uint32_t bx = 0; bx < nx; ++bx) {
                const size_t b = (size_t(bz) * ny + by) * nx + bx;
                bool any = false;
                float mx = 0.0f;
                for (uint32_t z = 0; z < 8; ++z)
                    for (uint32_t y = 0; y < 8; ++y)
                        for (uint32_t x = 0; x < 8; ++x) {
                            const uint16_t h =
                                in.density[(size_t(bz * 8 + z) * v.res[1] + by * 8 + y) * v.res[0] + bx * 8 + x];
                            if (h & 0x7fffu) any = true;
                            mx = fmaxf(mx, halfToFloat(h));
                        }
                if (!any) continue;
                brickMax[b] = mx;
                out.table[b] = used++;
                for (uint32_t z = 0; z < 8; ++z)
                    for (uint32_t y = 0; y < 8; ++y) {
                        const size_t src = (size_t(bz * 8 + z) * v.res[1] + by * 8 + y) * v.res[0] + bx * 8;
                        out.density.insert(out.density.end(), in.density + src, in.density + src + 8);
                        if (in.temperature)
                            out.temperature.insert(out.temperature.end(), in.temperature + src, in.temperature + src + 8);
                    }
            }
    out.majorant.assign(bricks, 0.0f);
    for (uint32_t bz = 0; bz < nz; ++bz)
        for (uint32_t by = 0; by < ny; ++by)
            for (uint32_t bx = 0; bx < nx; ++bx) {
                float m = 0.0f;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int x = int(bx) + dx, y = int(by) + dy, z = int(bz) + dz;
                            if (x < 0 || y < 0 || z < 0 || uint32_t(x) >= nx || uint32_t(y) >= ny || uint32_t(z) >= nz)
                                continue;
                            m = fmaxf(m, brickMax[(size_t(z) * ny + uint32_t(y)) * nx + uint32_t(x)]);
                        }
                out.majorant[(size_t(bz) * ny + by) * nx + bx] = m;
            }
    if (in.temperature && in.emission_scale > 0.0f) emissionRegion(in, out);
    return true;
}
}  // namespace x2
