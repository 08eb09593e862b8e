// x2_volume.h — the host side of the medium: x2_volume's dense FP16 grids packed into the device's bricks.
#pragma once
#include <string>
#include <vector>

#include "x2_layout.h"

namespace x2 {
struct VolumeHost {
    VolumeDesc desc{};  // sizes and coefficients; the pointers are set when uploaded
    std::vector<uint32_t> table;
    std::vector<uint16_t> density, temperature;
    std::vector<float> majorant, ramp;
    std::vector<float> emissionPool;  // the emission region's distributions (desc.region), empty: none
};
bool packVolume(const x2_volume& in, VolumeHost& out, std::string& err);
}  // namespace x2
