// Float primitives with one definition on host and device.
//
// Everything here is built from IEEE-754 basic operations (add, mul, fma,
// round-to-nearest-even, integer bit manipulation) so the host and
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

// log2 for positive normal x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// f = m - 1 (exact), log2(1 + f) = f P(f), P a degree-9 fit of
// log2(1 + f) / f on that interval (no division; <= 1.7 ulp).
CJX_HD float log2_det(float x) {
    const std::uint32_t u{f2u(x)};
    int e{static_cast<int>((u >> 23) & 0xffu) - 127};
    float m{u2f((u & 0x007fffffu) | 0x3f800000u)};
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float f{m - 1.0f};
    float p{-0.11296797543764114f};
    p = __builtin_fmaf(p, f, 0.18781770765781403f);
    p = __builtin_fmaf(p, f, -0.19061124324798584f);
    p = __builtin_fmaf(p, f, 0.20429740846157074f);
    p = __builtin_fmaf(p, f, -0.2396368831396103f);
    p = __builtin_fmaf(p, f, 0.28858521580696106f);
    p = __builtin_fmaf(p, f, -0.360696405172348f);
    p = __builtin_fmaf(p, f, 0.4808979034423828f);
    p = __builtin_fmaf(p, f, -0.7213473320007324f);
    p = __builtin_fmaf(p, f, 1.4426950216293335f);
    return __builtin_fmaf(f, p, static_cast<float>(e));
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

// Cube root of x >= 0 without division: y ~ x^(-1/3) from an exponent-
// thirding seed and three Newton steps y <- y (4 - x y^3) / 3, then c = x y^2
// and one Newton correction c <- c - (c^3 - x) y^2 / 3 (<= 0.75 ulp on
// [1e-6, 64]).
CJX_HD float cbrt_det(float x) {
    if (!(x > 0.0f)) {
        return 0.0f;
    }
    float y{u2f(0x54a23300u - f2u(x) / 3u)};
    const float x3{x * (1.0f / 3.0f)};
#pragma unroll
    for (int i{0}; i < 3; ++i) {
        const float y3{(y * y) * y};
        y = y * __builtin_fmaf(-x3, y3, 4.0f / 3.0f);
    }
    float c{(x * y) * y};
    const float r{__builtin_fmaf(c * c, c, -x)};
    c = __builtin_fmaf(-r, (y * y) * (1.0f / 3.0f), c);
    return c;
}

// a / b correctly rounded, given r = RN(1 / b): two Markstein correction
// steps (the first makes the quotient faithful, the second rounds it
// correctly).  Valid when a / b and the remainders stay normal; the sign of
// a zero quotient is not kept.
CJX_HD float div_by_recip(float a, float b, float r) {
    float q{a * r};
    float e{__builtin_fmaf(-b, q, a)};
    q = __builtin_fmaf(e, r, q);
    e = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(e, r, q);
}

CJX_HD float pow24_det(float x) {
    return exp2_det(2.4f * log2_det(x));
}

CJX_HD float exp_det(float x) {
    return exp2_det(x * 1.44269504f);
}

}  // namespace cjx
