// Host side of the constant tables and the distance calibration.
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.h"

namespace cjx {
namespace {

#include "dequant_tables.inc"

// dct.h:11-16 / transform_select.cu:39-47: A[k][t] = g_k cos(pi (t + 1/2) k / n)
// with g_0 = 1/n, g_k = sqrt(2)/n, derived in double and narrowed.
void dct_basis(int n, float* out) {
    for (int k{0}; k < n; ++k) {
        const double g{k == 0 ? 1.0 / n : 1.4142135623730950488 / n};
        for (int t{0}; t < n; ++t) {
            out[k * n + t] = static_cast<float>(g * std::cos(3.14159265358979323846 * (t + 0.5) * k / n));
        }
    }
}

// IEEE binary16 bit pattern of a float, round to nearest even (normal and
// subnormal results; no infinities arise from the basis entries).
std::uint16_t f32_to_f16(float f) {
    const std::uint32_t u{f2u(f)};
    const std::uint32_t sign{(u >> 16) & 0x8000u};
    const std::uint32_t a{u & 0x7fffffffu};
    if (a < 0x33000001u) {  // below half the smallest subnormal: +-0
        return static_cast<std::uint16_t>(sign);
    }
    const int e{static_cast<int>(a >> 23) - 127};
    std::uint32_t m{(a & 0x007fffffu) | 0x00800000u};
    // Shift so that the kept bits are the half's significand.
    const int shift{e < -14 ? 13 + (-14 - e) : 13};
    const std::uint32_t kept{m >> shift};
    const std::uint32_t rest{m & ((1u << shift) - 1u)};
    const std::uint32_t half{1u << (shift - 1)};
    std::uint32_t r{kept + ((rest > half || (rest == half && (kept & 1u))) ? 1u : 0u)};
    if (e < -14) {
        return static_cast<std::uint16_t>(sign | r);  // subnormal (a carry makes it the smallest normal)
    }
    r += static_cast<std::uint32_t>(e + 15 - 1) << 10;  // the implicit bit adds the last exponent step
    return static_cast<std::uint16_t>(sign | r);
}

float f16_to_f32(std::uint16_t h) {
    const std::uint32_t sign{static_cast<std::uint32_t>(h & 0x8000u) << 16};
    const std::uint32_t e{(h >> 10) & 31u}, m{h & 1023u};
    if (e == 0) {
        const float v{static_cast<float>(m) * (1.0f / 16777216.0f)};
        return sign ? -v : v;
    }
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

// Hi and lo FP16 parts of a float matrix (n x n, row-major).
void split_f16(const float* a, int count, std::uint16_t* hi, std::uint16_t* lo) {
    for (int i{0}; i < count; ++i) {
        hi[i] = f32_to_f16(a[i]);
        lo[i] = f32_to_f16(a[i] - f16_to_f32(hi[i]));
    }
}

// The table of pow24_tab / cbrt_tab (numeric.h): per exponent k and segment j,
// {RN((2^k c_j)^a), RN(1 / c_j)} with c_j = 1 + (j + 1/2) / 16, in double.
void xyb_power_table(float2* t) {
    for (int j{0}; j < 16; ++j) {
        const double c{1.0 + (j + 0.5) / 16.0};
        // RN(1 / c) scaled by 2^-23 (exact): the series argument takes the
        // mantissa difference as an integer count of 2^-23 steps.
        const float r{std::ldexp(static_cast<float>(1.0 / c), -23)};
        for (int k{-4}; k <= 0; ++k) {
            t[kXybPow + (k + 4) * 16 + j] = float2{static_cast<float>(std::pow(std::ldexp(c, k), 2.4)), r};
        }
        for (int k{-9}; k <= 0; ++k) {
            t[kXybCbrt + (k + 9) * 16 + j] = float2{static_cast<float>(std::cbrt(std::ldexp(c, k))), r};
        }
    }
}

// transform_select.cu:112-117: the DCT8 Y weight at the normalized frequency.
void selection_weights(int n, const float* dct8_y, float* out) {
    for (int e{0}; e < n * n; ++e) {
        long sx{std::lround(e / n * 8.0 / n)};
        long sy{std::lround(e % n * 8.0 / n)};
        sx = sx > 7 ? 7 : sx;
        sy = sy > 7 ? 7 : sy;
        out[e] = dct8_y[sx * 8 + sy];
    }
}

// coeff_order.cc:10-43: diagonal zig-zag with the LLF corner first.
void natural_order(int n, std::uint16_t* out) {
    const int cx{n / 8};
    int cur{cx * cx};
    for (int i{0}; i < n; ++i) {
        for (int j{0}; j <= i; ++j) {
            int x{j};
            int y{i - j};
            if (i & 1) {
                std::swap(x, y);
            }
            const int val{(x < cx && y < cx) ? y * cx + x : cur++};
            out[val] = static_cast<std::uint16_t>(y * n + x);
        }
    }
    for (int ip{n - 1}; ip > 0; --ip) {
        const int i{ip - 1};
        for (int j{0}; j <= i; ++j) {
            int x{n - 1 - (i - j)};
            int y{n - 1 - j};
            if (i & 1) {
                std::swap(x, y);
            }
            out[cur++] = static_cast<std::uint16_t>(y * n + x);
        }
    }
}

// vardct_layout.h:88-111 with the DCTResampleScales folded into the basis.
void llf_table(int m, double* t) {
    static const double s16[2]{1.0, 0.901764195028874394};
    static const double s32[4]{1.0, 0.974886821136879522, 0.901764195028874394,
                               0.787054918159101335};
    const double* sc{m == 2 ? s16 : s32};
    for (int f{0}; f < m; ++f) {
        const double norm{f == 0 ? std::sqrt(1.0 / m) : std::sqrt(2.0 / m)};
        for (int o{0}; o < m; ++o) {
            t[f * m + o] = norm * std::cos(3.14159265358979323846 * (o + 0.5) * f / m) * sc[f];
        }
    }
}

DeviceTables build_tables() {
    DeviceTables t{};
    dct_basis(8, t.basis8);
    dct_basis(16, t.basis16);
    dct_basis(32, t.basis32);
    const float* w{reinterpret_cast<const float*>(kDequantBits)};
    std::memcpy(t.dequant8, w, sizeof(t.dequant8));
    std::memcpy(t.dequant16, w + 3 * 64, sizeof(t.dequant16));
    std::memcpy(t.dequant32, w + 3 * 64 + 3 * 256, sizeof(t.dequant32));
    selection_weights(8, t.dequant8[1], t.wsel8);
    selection_weights(16, t.dequant8[1], t.wsel16);
    selection_weights(32, t.dequant8[1], t.wsel32);
    natural_order(8, t.order8);
    natural_order(16, t.order16);
    natural_order(32, t.order32);
    llf_table(2, t.llf2);
    llf_table(4, t.llf4);
    auto recip = [](const float* w, float* r, int n) {
        for (int i{0}; i < n; ++i) {
            r[i] = 1.0f / w[i];
        }
    };
    recip(t.wsel8, t.wsel_inv8, 64);
    recip(t.wsel16, t.wsel_inv16, 256);
    recip(t.wsel32, t.wsel_inv32, 1024);
    recip(&t.dequant8[0][0], &t.dequant_inv8[0][0], 3 * 64);
    recip(&t.dequant16[0][0], &t.dequant_inv16[0][0], 3 * 256);
    recip(&t.dequant32[0][0], &t.dequant_inv32[0][0], 3 * 1024);
    {
        float bd[256]{};
        for (int r{0}; r < 16; ++r) {
            for (int c{0}; c < 16; ++c) {
                bd[r * 16 + c] = (r >> 3) == (c >> 3) ? t.basis8[(r & 7) * 8 + (c & 7)] : 0.0f;
            }
        }
        split_f16(bd, 256, t.mf8[0], t.mf8[1]);
        split_f16(t.basis16, 256, t.mf16[0], t.mf16[1]);
        split_f16(t.basis32, 1024, t.mf32[0], t.mf32[1]);
    }
    xyb_power_table(t.xyb_tab);
    t.rate[0] = 0.0f;
    for (int k{1}; k < 256; ++k) {
        t.rate[k] = 1.0f + log2_det(1.0f + static_cast<float>(k));
    }
    return t;
}

}  // namespace

const DeviceTables& host_tables() {
    static const DeviceTables tables{build_tables()};
    return tables;
}

// quant_calibration.h:49-88 and frontend_fused.cu:214-239.
FrameParams make_frame_params(float distance) {
    FrameParams p{};
    const float quant_ac{0.725f / distance};
    const float k_dc_mul{0.3f};
    float inner{k_dc_mul * std::pow((1.0f / k_dc_mul) * distance, 0.83f)};
    inner = distance < inner ? distance : inner;
    const float bt_dc{0.5f * distance > inner ? 0.5f * distance : inner};
    float quant_dc_float{1.095924047623553f / bt_dc};
    quant_dc_float = quant_dc_float < 50.0f ? quant_dc_float : 50.0f;
    float scale{static_cast<float>(1 << 16) * quant_ac / 5.0f};
    scale = scale < 1.0f ? 1.0f : (scale > static_cast<float>(1 << 15) ? static_cast<float>(1 << 15) : scale);
    int global_scale{static_cast<int>(scale)};
    const int scaled_quant_dc{static_cast<int>(quant_dc_float * 4096 * 1.6f)};
    if (global_scale > scaled_quant_dc) {
        global_scale = scaled_quant_dc <= 0 ? 1 : scaled_quant_dc;
    }
    const float inv_global_scale{1.0f * static_cast<float>(1 << 16) / static_cast<float>(global_scale)};
    float fval{quant_dc_float * inv_global_scale + 0.5f};
    fval = fval < static_cast<float>(1 << 16) ? fval : static_cast<float>(1 << 16);
    int q{static_cast<int>(std::lround(quant_ac * inv_global_scale))};
    q = q < 1 ? 1 : (q > 256 ? 256 : q);
    p.global_scale = static_cast<std::uint32_t>(global_scale);
    p.quant_dc = static_cast<std::uint32_t>(static_cast<int>(fval));
    p.raw_quant_field = static_cast<std::uint32_t>(q);
    p.global_scale_float = static_cast<float>(global_scale) * (1.0f / (1 << 16));
    p.inv_global_scale = inv_global_scale;
    p.qgsf_select = static_cast<float>(p.raw_quant_field) * p.global_scale_float;
    p.inv_qgsf_select = 1.0f / p.qgsf_select;
    p.dc_scale = p.global_scale_float * static_cast<float>(p.quant_dc);

    const float base_level{0.48f * quant_ac};
    float dampen{1.0f};
    if (distance >= 2.0f) {
        dampen = 1.0f - (distance - 2.0f) / (14.0f - 2.0f);
        dampen = dampen < 0.0f ? 0.0f : dampen;
    }
    p.mul_pb = quant_ac * dampen;
    p.add_pb = (1.0f - dampen) * base_level;
    const float efm{distance < 2.0f ? (2.0f - distance) * 0.5f : 0.0f};
    p.km[0] = 0.125f;
    p.km[1] = 0.10f - efm * 0.10f;
    p.km[2] = 0.09f - efm * 0.09f;
    p.km[3] = 0.06f - efm * 0.06f;
    const float norm{0.29959705784054957f / (p.km[0] + p.km[1] + p.km[2] + p.km[3])};
    for (float& k : p.km) {
        k *= norm;
    }
    return p;
}

}  // namespace cjx
