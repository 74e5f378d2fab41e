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

// Table-driven powers of the XYB conversion.  The mantissa m in [1, 2) of
// x = 2^k m falls in segment j = its top 4 bits, with centre
// c_j = 1 + (j + 1/2) / 16 and r_j = RN(1 / c_j); t = (m - c_j) r_j,
// |t| <= 1/32, and x^a = T[k][j] (1 + t)^a with T[k][j] = RN((2^k c_j)^a)
// and the binomial series of (1 + t)^a to t^3 (truncation below 4e-8: the
// t^4 terms are 0.0336 t^4 and 0.0412 t^4).  The table holds one pair
// {T[k][j], r_j 2^-23} per (k, j), so each power reads one 8-byte entry.
// m - c_j is exactly D 2^-23 with the integer D = (low 19 mantissa bits of
// x) - 2^18, so t = D (r_j 2^-23) is the same real product as (m - c_j) r_j
// and rounds to the same float.
constexpr int kXybPow = 0;    // [5][16]: x^2.4 for k = -4 .. 0
constexpr int kXybCbrt = 80;  // [10][16]: x^(1/3) for k = -9 .. 0
constexpr int kXybTabPairs = 240;

CJX_HD float xyb_series_arg(std::uint32_t u, float r_scaled) {
    return (static_cast<float>(u & 0x0007ffffu) - 262144.0f) * r_scaled;
}

// x^2.4 for x in [2^-4, 2) (the sRGB curve's upper branch: x in [0.0905, 1]).
CJX_HD float pow24_tab(float x, const float2* tab) {
    const std::uint32_t u{f2u(x)};
    const float2 e{tab[kXybPow + ((u >> 19) - (123u << 4))]};
    const float t{xyb_series_arg(u, e.y)};
    float p{0.224f};
    p = __builtin_fmaf(p, t, 1.68f);
    p = __builtin_fmaf(p, t, 2.4f);
    p = __builtin_fmaf(p, t, 1.0f);
    return e.x * p;
}

// x^(1/3) for x in [2^-9, 2) (the opsin values: x in [0.0038, 1.004]).
CJX_HD float cbrt_tab(float x, const float2* tab) {
    const std::uint32_t u{f2u(x)};
    const float2 e{tab[kXybCbrt + ((u >> 19) - (118u << 4))]};
    const float t{xyb_series_arg(u, e.y)};
    float p{0.0617283951f};
    p = __builtin_fmaf(p, t, -0.111111111f);
    p = __builtin_fmaf(p, t, 0.333333333f);
    p = __builtin_fmaf(p, t, 1.0f);
    return e.x * p;
}

CJX_HD float exp_det(float x) {
    return exp2_det(x * 1.44269504f);
}

}  // namespace cjx
