// x2_volume.cpp — packs a medium's dense grids into 8^3 bricks for the device. A brick is kept when any of its
// density voxels is nonzero; temperature is kept for the same bricks (it reads zero elsewhere). A brick's majorant
// is the largest density among its own voxels and its 26 neighbours' (a point in the brick interpolates voxels at
// most one voxel away), the bound the device's ratio tracking samples free flights against. The derived sizes are
// computed here once, in float, so that every reader of them rounds alike.
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
    return true;
}
}  // namespace x2
