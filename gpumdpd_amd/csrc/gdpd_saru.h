// gdpd_saru.h -- Saru/TEA8 pair hash, host and device.
//
// Steps follow force/pair_dpd.cu:212-236 (identical copy force/pair_mdpd.cu:318-342):
// seed mixing with arithmetic right shifts on signed casts, 8 TEA rounds, the
// final avalanche, then u = (float)s * 2^-32 in FP32 in both builds (the
// reference's TWO_N32 is a float literal) and r = sqrt(3) (2u - 1).
// The call sites key it by (min(id_i,id_j), max(id_i,id_j), seed, step)
// (pair_dpd.cu:342-344).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GDPD_HD __host__ __device__ __forceinline__
#else
#define GDPD_HD inline
#endif

namespace gdpd {

GDPD_HD uint32_t sar32(uint32_t v, int n) { return (uint32_t)((int32_t)v >> n); }

// seed_plus_step = seed + step (uint32 wrap), hoisted out of the pair loop.
GDPD_HD uint32_t saru_bits(uint32_t a, uint32_t b, uint32_t seed_plus_step)
{
    uint32_t c = seed_plus_step;
    c ^= (a << 7) ^ (b >> 6);
    b += (a >> 4) ^ (c >> 15);
    a ^= (b << 9) + (c << 8);
    c ^= 0xA5366B4Du * ((b >> 11) ^ (a << 1));
    b += 0x72BE1579u * ((a << 4) ^ (c >> 16));
    a ^= 0x3F38A6EDu * ((c >> 5) ^ sar32(b, 22));
    b += a * c;
    a += c ^ (b >> 2);
    b ^= sar32(b, 17);
    uint32_t st = 0x79dedea3u * (a ^ sar32(a, 14));
    uint32_t ws = (st + b) ^ sar32(st, 8);
    st = st + ws * (ws ^ 0xdddf97f5u);
    ws = 0xABCB96F7u + (ws >> 1);
    // 8 TEA rounds; the round key sum = r * 0x9e3779b9 folds to constants.
#define GDPD_TEA(r)                                                                 \
    {                                                                               \
        const uint32_t sum = 0x9e3779b9u * (uint32_t)(r);                           \
        st += ((ws << 4) + 0xA341316Cu) ^ (ws + sum) ^ ((ws >> 5) + 0xC8013EA4u);   \
        ws += ((st << 4) + 0xAD90777Du) ^ (st + sum) ^ ((st >> 5) + 0x7E95761Eu);   \
    }
    GDPD_TEA(1) GDPD_TEA(2) GDPD_TEA(3) GDPD_TEA(4)
    GDPD_TEA(5) GDPD_TEA(6) GDPD_TEA(7) GDPD_TEA(8)
#undef GDPD_TEA
    uint32_t v = (st ^ (st >> 26)) + ws;
    return (v ^ (v >> 20)) * 0x6957f5a7u;
}

// FP32 build: r = SQRT3f * ((float)s * 2^-32 * 2 - 1)
GDPD_HD float saru_uniform_f32(uint32_t s)
{
    float u = (float)s * 2.3283064365386962890625e-10f;
    return (float)1.7320508075688772935274463 * (u * 2.0f - 1.0f);
}

// FP64 build: u stays FP32, the rest is FP64
GDPD_HD double saru_uniform_f64(uint32_t s)
{
    float u = (float)s * 2.3283064365386962890625e-10f;
    return 1.7320508075688772935274463 * ((double)u * 2.0 - 1.0);
}

template <typename real> GDPD_HD real saru_uniform(uint32_t s);
template <> GDPD_HD float saru_uniform<float>(uint32_t s) { return saru_uniform_f32(s); }
template <> GDPD_HD double saru_uniform<double>(uint32_t s) { return saru_uniform_f64(s); }

} // namespace gdpd
