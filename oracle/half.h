// ORACLE — test infrastructure only (see lbm_oracle.cpp): half-precision
// conversions and the reference's stochastic rounding, shared by the D3Q19
// and D3Q27 restatements.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle_half {

inline uint32_t fbits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
inline float bitsf(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }

// IEEE half -> float (exact).
inline float half_to_float(uint16_t h) {
    uint32_t s = (uint32_t)(h & 0x8000) << 16, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    if (e == 0) {
        float v = std::ldexp((float)m, -24);
        return s ? -v : v;
    }
    if (e == 31) return bitsf(s | 0x7f800000u | (m << 13));
    return bitsf(s | ((e + 112) << 23) | (m << 13));
}

// float -> IEEE half, round to nearest even, subnormals kept (the WGSL f16()
// conversion; the WGSL spec leaves the direction open, RNE assumed — parity
// unpinned at the bit level against the Metal compiler).
inline uint16_t float_to_half_rne(float x) {
    uint32_t b = fbits(x);
    uint32_t s = (b >> 16) & 0x8000;
    uint32_t a = b & 0x7fffffff;
    if (a >= 0x7f800000u) return (uint16_t)(s | 0x7c00 | (a > 0x7f800000u ? 0x200 : 0));
    if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00);  // rounds to inf
    if (a < 0x38800000u) {                                   // below 2^-14: subnormal half
        float v = bitsf(a) * 16777216.0f;                     // exact scale by 2^24
        uint32_t q = (uint32_t)std::nearbyint(v);             // RNE in default mode
        return (uint16_t)(s | q);
    }
    uint32_t e = (a >> 23) - 112, m = a & 0x7fffff;
    uint32_t h = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fff;
    if (rem > 0x1000 || (rem == 0x1000 && (h & 1))) h += 1;
    return (uint16_t)(s | h);
}

// float -> IEEE half, round toward zero (v_cvt_pkrtz_f16_f32).
inline uint16_t float_to_half_rtz(float x) {
    uint32_t b = fbits(x);
    uint32_t s = (b >> 16) & 0x8000;
    uint32_t a = b & 0x7fffffff;
    if (a > 0x7f800000u) return (uint16_t)(s | 0x7e00);
    if (a >= 0x477ff000u) return (uint16_t)(s | (a == 0x7f800000u ? 0x7c00 : 0x7bff));
    if (a < 0x38800000u) return (uint16_t)(s | (uint32_t)(bitsf(a) * 16777216.0f));
    uint32_t e = (a >> 23) - 112, m = a & 0x7fffff;
    return (uint16_t)(s | (e << 10) | (m >> 13));
}

// lattice.rs:172-189 f16_bits: host rounding of q (add bit 12, flush subnormals).
inline uint16_t f16_bits_host(float x) {
    uint32_t b = fbits(x);
    uint16_t sign = (uint16_t)((b >> 16) & 0x8000);
    int exp = (int)((b >> 23) & 0xff) - 127 + 15;
    uint32_t man = b & 0x7fffff;
    if (exp <= 0) return sign;
    if (exp >= 31) return sign | 0x7c00;
    uint32_t h = sign | ((uint32_t)exp << 10) | (man >> 13);
    if (man & 0x1000) h += 1;
    return (uint16_t)h;
}

// shader.rs:510-517 sround (the reference's per-store hash and dither).
inline float sround_ref(float x, uint32_t key, uint32_t seed) {
    uint32_t h = (key * 0x9E3779B1u) ^ (seed * 0x85EBCA77u + 0x632BE5ABu);
    h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
    float r = (float)(h >> 8) * (1.0f / 16777216.0f) - 0.5f;
    int e = (int)((fbits(x) >> 23) & 0xff) - 127;
    float ulp = std::ldexp(1.0f, std::max(e, -14) - 10);
    float v = x + r * ulp;
    return std::min(std::max(v, -65000.0f), 65000.0f);
}

}  // namespace oracle_half
