// The engine's packed shading formats, unpacked on the host exactly as the
// shaders unpack them (float operations in the shaders' order, no FMA):
//   unpackUnitVector  (common/packing.hsh:2-12, the inverse of
//                      Packing::PackSignedVector3x10_1x2, Packing.cpp:24-35):
//                      each 10-bit field k -> k / 1023 * 2 - 1
//   unpackHalf2x16    (GLSL): IEEE binary16 halves, the first in the low bits
//   unpackUnorm4x8    (GLSL): each byte k -> k / 255, the first in the low bits
#pragma once

#include <cstdint>
#include <cstring>

namespace atlas_rt {

inline uint32_t float_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

inline void unpack_unit_vector(uint32_t c, float out[3]) {
    out[0] = float((c >> 0) & 1023u) / 1023.0f * 2.0f - 1.0f;
    out[1] = float((c >> 10) & 1023u) / 1023.0f * 2.0f - 1.0f;
    out[2] = float((c >> 20) & 1023u) / 1023.0f * 2.0f - 1.0f;
}

// binary16 -> binary32, exact (subnormal halves are normal floats).
inline float half_to_float(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1Fu;
    uint32_t man = h & 0x3FFu;
    uint32_t bits;
    if (exp == 0x1Fu) {
        bits = sign | 0x7F800000u | (man << 13);
    } else if (exp != 0) {
        bits = sign | ((exp + 112u) << 23) | (man << 13);
    } else if (man == 0) {
        bits = sign;
    } else {
        int e = -1;
        do {
            man <<= 1;
            e++;
        } while (!(man & 0x400u));
        bits = sign | (uint32_t(112 - e) << 23) | ((man & 0x3FFu) << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

inline void unpack_half2(uint32_t c, float out[2]) {
    out[0] = half_to_float(uint16_t(c & 0xFFFFu));
    out[1] = half_to_float(uint16_t(c >> 16));
}

inline void unpack_unorm4x8(uint32_t c, float out[4]) {
    for (int k = 0; k < 4; k++) out[k] = float((c >> (8 * k)) & 0xFFu) / 255.0f;
}

}  // namespace atlas_rt
