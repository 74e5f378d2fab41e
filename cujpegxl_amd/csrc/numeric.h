// Float primitives with one definition on host and device.
//
// Everything here is built from IEEE-754 basic operations (add, mul,
// correctly rounded div/sqrt, fma, round-to-nearest-even) so the host and
// gfx950 evaluate them to identical bits when compiled with contraction off.
// The encoder's integer outputs (quant field, transform map, quantized
// coefficients) depend on these values, which is why no libm or ocml
// transcendental is used on the hot path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define CJX_HD __host__ __device__ __forceinline__

namespace cjx {

CJX_HD std::uint32_t f2u(float f) {
    return __builtin_bit_cast(std::uint32_t, f);
}
CJX_HD float u2f(std::uint32_t u) {
    return __builtin_bit_cast(float, u);
}

// log2 for positive normal x: x = m 2^e with m in [sqrt(1/2), sqrt(2)];
// ln m = 2 atanh(t), t = (m - 1) / (m + 1), odd series to t^9.
CJX_HD float log2_det(float x) {
    const std::uint32_t u{f2u(x)};
    int e{static_cast<int>((u >> 23) & 0xffu) - 127};
    float m{u2f((u & 0x007fffffu) | 0x3f800000u)};
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float t{(m - 1.0f) / (m + 1.0f)};
    const float t2{t * t};
    float p{0.111111111f};
    p = __builtin_fmaf(p, t2, 0.142857143f);
    p = __builtin_fmaf(p, t2, 0.2f);
    p = __builtin_fmaf(p, t2, 0.333333333f);
    p = __builtin_fmaf(p, t2, 1.0f);
    const float ln_m{(2.0f * t) * p};
    return __builtin_fmaf(ln_m, 1.44269504f, static_cast<float>(e));
}

// 2^x on [-126, 127]: x = n + f, |f| <= 1/2, degree-7 Taylor of e^(f ln 2),
// exponent inserted by construction.
CJX_HD float exp2_det(float x) {
    x = x < -126.0f ? -126.0f : x;
    x = x > 127.0f ? 127.0f : x;
    const float n{__builtin_rintf(x)};
    const float g{(x - n) * 0.693147181f};
    float p{1.98412698e-4f};
    p = __builtin_fmaf(p, g, 1.38888889e-3f);
    p = __builtin_fmaf(p, g, 8.33333333e-3f);
    p = __builtin_fmaf(p, g, 4.16666667e-2f);
    p = __builtin_fmaf(p, g, 0.166666667f);
    p = __builtin_fmaf(p, g, 0.5f);
    p = __builtin_fmaf(p, g, 1.0f);
    p = __builtin_fmaf(p, g, 1.0f);
    return p * u2f(static_cast<std::uint32_t>(static_cast<int>(n) + 127) << 23);
}

// Cube root of x >= 0: exponent-thirding seed and three Newton steps.
CJX_HD float cbrt_det(float x) {
    if (!(x > 0.0f)) {
        return 0.0f;
    }
    float y{u2f(f2u(x) / 3u + 0x2a5137a0u)};
#pragma unroll
    for (int i{0}; i < 3; ++i) {
        y = (2.0f * y + x / (y * y)) * 0.333333333f;
    }
    return y;
}

CJX_HD float pow24_det(float x) {
    return exp2_det(2.4f * log2_det(x));
}

CJX_HD float exp_det(float x) {
    return exp2_det(x * 1.44269504f);
}

}  // namespace cjx
