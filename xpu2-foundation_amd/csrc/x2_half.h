// x2_half.h — IEEE binary16 from binary32 on the host, rounding to nearest even (as Imath's half and OIIO convert):
// the paged texture tiles and the EXR writer's half channels.
#pragma once
#include <stdint.h>
#include <string.h>

namespace x2 {
inline uint16_t toHalf(float value) {
    uint32_t x;
    memcpy(&x, &value, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t e = (x >> 23) & 0xffu;
    uint32_t m = x & 0x7fffffu;
    if (e == 0xff) return uint16_t(sign | 0x7c00u | (m >> 13) | (m && !(m >> 13)));  // inf, or a NaN kept a NaN
    const int exp = int(e) - 127 + 15;
    if (exp >= 31) return uint16_t(sign | 0x7c00u);  // overflow
    if (exp <= 0) {                                  // a subnormal half, or zero
        if (exp < -10) return uint16_t(sign);
        m |= 0x800000u;
        const int shift = 14 - exp;  // 14..24
        uint32_t h = m >> shift;
        const uint32_t rest = m & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (rest > half || (rest == half && (h & 1u))) ++h;
        return uint16_t(sign | h);
    }
    uint32_t h = (uint32_t(exp) << 10) | (m >> 13);
    const uint32_t rest = m & 0x1fffu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ++h;  // may carry into the exponent, to infinity at most
    return uint16_t(sign | h);
}
// binary16 to binary32, exactly.
inline float fromHalf(uint16_t h) {
    const uint32_t s = uint32_t(h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
    float f;
    if (e == 0) {
        f = float(m) * (1.0f / 16777216.0f);
        return s ? -f : f;
    }
    const uint32_t u = s | (e == 31 ? 0x7f800000u | (m << 13) : ((e + 112u) << 23) | (m << 13));
    memcpy(&f, &u, 4);
    return f;
}
}  // namespace x2
