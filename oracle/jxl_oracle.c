/* Test infrastructure only: CPU oracle for the VarDCT encode path.
 *
 * Restates the reference encoder's algorithm in plain C (see jxl_oracle.h for
 * the numeric contract).  Citations are file:line in m-pilia/cujpegxl.
 *
 * Pinning (what checks each part; every row is exercised under tests/):
 *   numerics      jo_cbrt/jo_pow24/jo_log2/jo_exp2 against libm double over
 *                 their domains (dense samples + every segment edge)
 *   calibration   oracle/_ref (quant_calibration.h) on a distance sweep
 *   XYB           FP64 closed form jo_nv12_to_xyb_f64, flat-tile exactness
 *   gaborish      brute-force 5x5 convolution with clamp in double
 *   AQ            closed-form invariants (flat tile) + FP64 recompute per block
 *   selection     oracle/_ref decide_region (flat plane, edges, textures)
 *   DCT           FP64 DCT-II matrix product; orthogonality round trip
 *   CfL maps      oracle/_ref cfl_maps_from_sums, slope recovery
 *   dc_from_llf   oracle/_ref dc_from_llf_strided, constant-block identity
 *   quantize      rint(coeff/step) in float64 (quant_selfconsistency_test.py)
 *   entropy/bytes oracle/_ref reference_ac_encode, reference_dc_encode,
 *                 build_ac_global, build_dc_global, build_codestream_head,
 *                 write_vardct_codestream, container_framing; golden fixtures
 */
#include "jxl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Threads for the per-row, per-tile, per-region and per-block loops (their
 * iterations write disjoint outputs, so any count gives the same bytes).
 * Default 1; the CPU baseline's all-cores leg raises it. */
static int g_threads = 1;

void jo_set_threads(int n) {
    g_threads = n < 1 ? 1 : n;
}

/* ======================================================================== */
/* Numeric primitives (shared definition with the HIP kernels).              */
/* ======================================================================== */

static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* log2 for positive normal x: split x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
 * f = m - 1 (exact), log2(1 + f) = f * P(f) with P a degree-9 polynomial fitted
 * to log2(1 + f) / f on that interval (no division; <= 1.7 ulp). */
float jo_log2(float x) {
    uint32_t u = f2u(x);
    int e = (int)((u >> 23) & 0xffu) - 127;
    float m = u2f((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float f = m - 1.0f;
    float p = -0.11296797543764114f;
    p = fmaf(p, f, 0.18781770765781403f);
    p = fmaf(p, f, -0.19061124324798584f);
    p = fmaf(p, f, 0.20429740846157074f);
    p = fmaf(p, f, -0.2396368831396103f);
    p = fmaf(p, f, 0.28858521580696106f);
    p = fmaf(p, f, -0.360696405172348f);
    p = fmaf(p, f, 0.4808979034423828f);
    p = fmaf(p, f, -0.7213473320007324f);
    p = fmaf(p, f, 1.4426950216293335f);
    return fmaf(f, p, (float)e);
}

/* 2^x for x in [-126, 127]: x = n + f, |f| <= 1/2, e^(f ln2) by a degree-7
 * Taylor polynomial, exponent inserted by bit construction. */
float jo_exp2(float x) {
    if (x < -126.0f) {
        x = -126.0f;
    }
    if (x > 127.0f) {
        x = 127.0f;
    }
    const float n = rintf(x);
    const float g = (x - n) * 0.693147181f;
    float p = 1.98412698e-4f;
    p = fmaf(p, g, 1.38888889e-3f);
    p = fmaf(p, g, 8.33333333e-3f);
    p = fmaf(p, g, 4.16666667e-2f);
    p = fmaf(p, g, 0.166666667f);
    p = fmaf(p, g, 0.5f);
    p = fmaf(p, g, 1.0f);
    p = fmaf(p, g, 1.0f);
    return p * u2f((uint32_t)((int)n + 127) << 23);
}

/* The powers of the XYB conversion from one table of 272 floats: x = 2^k m,
 * m in [1, 2), segment j = the top 4 mantissa bits with centre
 * c_j = 1 + (j + 1/2)/16 and r_j = RN(1/c_j); t = (m - c_j) r_j (|t| <= 1/32,
 * the difference exact), x^a = RN((2^k c_j)^a) * (1 + t)^a with the binomial
 * series of (1 + t)^a to t^3 (truncation below 4e-8: the t^4 terms are
 * 0.0336 t^4 and 0.0412 t^4).  A project restatement of the reference's powf
 * and cbrtf (xyb_impl.cuh:37-40, 62-65) with a fixed evaluation order, so
 * every platform computes identical bits; relative error <= 1e-6. */
static float g_xyb_tab[272];
static int g_xyb_tab_ready = 0;

static void xyb_tab_init(void) {
    if (g_xyb_tab_ready) {
        return;
    }
    for (int j = 0; j < 16; ++j) {
        const double c = 1.0 + (j + 0.5) / 16.0;
        g_xyb_tab[2 * j] = (float)c;
        g_xyb_tab[2 * j + 1] = (float)(1.0 / c);
        for (int k = -4; k <= 0; ++k) {
            g_xyb_tab[32 + (k + 4) * 16 + j] = (float)pow(ldexp(c, k), 2.4);
        }
        for (int k = -9; k <= 0; ++k) {
            g_xyb_tab[112 + (k + 9) * 16 + j] = (float)cbrt(ldexp(c, k));
        }
    }
    g_xyb_tab_ready = 1;
}

void jo_xyb_table(float* out) {
    xyb_tab_init();
    memcpy(out, g_xyb_tab, sizeof(g_xyb_tab));
}

/* x^2.4 for x in [2^-4, 2). */
float jo_pow24(float x) {
    xyb_tab_init();
    const uint32_t u = f2u(x);
    const uint32_t j = (u >> 19) & 15u;
    const float m = u2f((u & 0x007fffffu) | 0x3f800000u);
    const float t = (m - g_xyb_tab[2 * j]) * g_xyb_tab[2 * j + 1];
    float p = 0.224f;
    p = fmaf(p, t, 1.68f);
    p = fmaf(p, t, 2.4f);
    p = fmaf(p, t, 1.0f);
    return g_xyb_tab[32 + ((u >> 19) - (123u << 4))] * p;
}

/* x^(1/3) for x in [2^-9, 2). */
float jo_cbrt(float x) {
    xyb_tab_init();
    const uint32_t u = f2u(x);
    const uint32_t j = (u >> 19) & 15u;
    const float m = u2f((u & 0x007fffffu) | 0x3f800000u);
    const float t = (m - g_xyb_tab[2 * j]) * g_xyb_tab[2 * j + 1];
    float p = 0.0617283951f;
    p = fmaf(p, t, -0.111111111f);
    p = fmaf(p, t, 0.333333333f);
    p = fmaf(p, t, 1.0f);
    return g_xyb_tab[112 + ((u >> 19) - (118u << 4))] * p;
}

float jo_exp(float x) {
    return jo_exp2(x * 1.44269504f);
}

/* IEEE binary16 conversion, round to nearest even, subnormals kept. */
uint16_t jo_f32_to_f16(float f) {
    const uint32_t u = f2u(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) {
        return (uint16_t)(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
    }
    if (a >= 0x477ff000u) { /* rounds to >= 65520 -> inf */
        return (uint16_t)(sign | 0x7c00u);
    }
    if (a >= 0x38800000u) { /* normal half */
        const uint32_t mant = a & 0x7fffffu;
        const uint32_t exp = (a >> 23) - 112u;
        uint32_t h = (exp << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) {
            ++h;
        }
        return (uint16_t)(sign | h);
    }
    /* subnormal or zero half: value = a_float / 2^-24 ulps */
    const uint32_t e = a >> 23;
    if (e < 102u) { /* < 2^-25: rounds to zero (ties at exactly 2^-25 -> 0) */
        return (uint16_t)sign;
    }
    const uint32_t mant = (a & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e; /* 14..24 */
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) {
        ++h;
    }
    return (uint16_t)(sign | h);
}

float jo_f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1fu;
    const uint32_t mant = h & 0x3ffu;
    if (exp == 0) {
        const float v = (float)mant * 5.96046448e-8f; /* 2^-24, exact */
        return sign ? -v : v;
    }
    if (exp == 31) {
        return u2f(sign | 0x7f800000u | (mant << 13));
    }
    return u2f(sign | ((exp + 112u) << 23) | (mant << 13));
}

static inline float clamp01(float v) {
    return fminf(fmaxf(v, 0.0f), 1.0f);
}

/* Fixed wave64 shuffle-down tree: lane l += lane l+off, off = 32..1. */
static float tree64(float* v) {
    for (int off = 32; off > 0; off >>= 1) {
        for (int l = 0; l < off; ++l) {
            v[l] = v[l] + v[l + off];
        }
    }
    return v[0];
}

/* ======================================================================== */
/* Tables                                                                    */
/* ======================================================================== */

static float g_w8[3][64];
static float g_w16[3][256];
static float g_w32[3][1024];
static float g_wsel8[64], g_wsel16[256], g_wsel32[1024];
static void build_wsel(int n, float* t);

void jo_set_dequant(const float* w8, const float* w16, const float* w32) {
    memcpy(g_w8, w8, sizeof(g_w8));
    memcpy(g_w16, w16, sizeof(g_w16));
    memcpy(g_w32, w32, sizeof(g_w32));
    build_wsel(8, g_wsel8);
    build_wsel(16, g_wsel16);
    build_wsel(32, g_wsel32);
}

static inline float dequant_weight(int n, int c, int e) {
    return n == 32 ? g_w32[c][e] : (n == 16 ? g_w16[c][e] : g_w8[c][e]);
}

/* quant_calibration.h:49-88 */
void jo_calibrate(float distance, jo_cal* out) {
    const float quant_ac = 0.725f / distance;
    const float k_dc_mul = 0.3f;
    float inner = k_dc_mul * powf((1.0f / k_dc_mul) * distance, 0.83f);
    inner = distance < inner ? distance : inner;
    const float bt_dc = 0.5f * distance > inner ? 0.5f * distance : inner;
    float quant_dc_float = 1.095924047623553f / bt_dc;
    quant_dc_float = quant_dc_float < 50.0f ? quant_dc_float : 50.0f;

    float scale = (float)(1 << 16) * quant_ac / 5.0f;
    if (scale < 1.0f) {
        scale = 1.0f;
    }
    if (scale > (float)(1 << 15)) {
        scale = (float)(1 << 15);
    }
    int global_scale = (int)scale;
    const int scaled_quant_dc = (int)(quant_dc_float * 4096 * 1.6f);
    if (global_scale > scaled_quant_dc) {
        global_scale = scaled_quant_dc <= 0 ? 1 : scaled_quant_dc;
    }
    const float inv_global_scale = 1.0f * (float)(1 << 16) / (float)global_scale;
    float fval = quant_dc_float * inv_global_scale + 0.5f;
    fval = fval < (float)(1 << 16) ? fval : (float)(1 << 16);
    int q = (int)lroundf(quant_ac * inv_global_scale);
    q = q < 1 ? 1 : (q > 256 ? 256 : q);
    out->global_scale = (uint32_t)global_scale;
    out->quant_dc = (uint32_t)(int)fval;
    out->raw_quant_field = (uint32_t)q;
    out->global_scale_float = (float)global_scale * (1.0f / (1 << 16));
    out->inv_global_scale = inv_global_scale;
}

/* coeff_order.cc:10-43 (square DCT natural order, LLF corner first). */
void jo_natural_order(uint32_t n, uint32_t* out) {
    const uint32_t cx = n / 8;
    uint32_t cur = cx * cx;
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t j = 0; j <= i; ++j) {
            uint32_t x = j, y = i - j;
            if (i & 1u) {
                const uint32_t t = x;
                x = y;
                y = t;
            }
            const uint32_t val = (x < cx && y < cx) ? y * cx + x : cur++;
            out[val] = y * n + x;
        }
    }
    for (uint32_t ip = n - 1; ip > 0; --ip) {
        const uint32_t i = ip - 1;
        for (uint32_t j = 0; j <= i; ++j) {
            uint32_t x = n - 1 - (i - j), y = n - 1 - j;
            if (i & 1u) {
                const uint32_t t = x;
                x = y;
                y = t;
            }
            out[cur++] = y * n + x;
        }
    }
}

/* dct.h:11-16: A[k][t] = g_k cos(pi (t+1/2) k / n), g_0 = 1/n, g_k = sqrt2/n,
 * derived in double and narrowed (transform_select.cu:39-47). */
void jo_dct_basis(uint32_t n, float* out) {
    for (uint32_t k = 0; k < n; ++k) {
        const double g = k == 0 ? 1.0 / n : 1.4142135623730950488 / n;
        for (uint32_t t = 0; t < n; ++t) {
            out[k * n + t] = (float)(g * cos(3.14159265358979323846 * (t + 0.5) * k / n));
        }
    }
}

static void dct_basis_f64(uint32_t n, double* out) {
    for (uint32_t k = 0; k < n; ++k) {
        const double g = k == 0 ? 1.0 / n : 1.4142135623730950488 / n;
        for (uint32_t t = 0; t < n; ++t) {
            out[k * n + t] = g * cos(3.14159265358979323846 * (t + 0.5) * k / n);
        }
    }
}

/* ac_context.h:108-131: fixed 32-cluster map over the 7425 AC contexts. */
int jo_ac_cluster(uint32_t ctx) {
    const uint32_t count_contexts = 15u * 37u;
    if (ctx < count_contexts) {
        const uint32_t bucket = ctx / 15u;
        const int bctx = (int)(ctx % 15u);
        const int group = bctx < 7 ? 0 : 1;
        const int band = bucket < 4 ? 0 : (bucket < 12 ? 1 : (bucket < 24 ? 2 : 3));
        return group * 4 + band;
    }
    const uint32_t rel = ctx - count_contexts;
    const int bctx = (int)(rel / 458u);
    const uint32_t z = rel % 458u;
    const int group = bctx < 7 ? 0 : 1;
    static const uint32_t edges[11] = {32, 64, 96, 128, 160, 200, 240, 280, 328, 376, 424};
    int band = 11;
    for (int i = 0; i < 11; ++i) {
        if (z < edges[i]) {
            band = i;
            break;
        }
    }
    return 8 + group * 12 + band;
}

/* ======================================================================== */
/* Front end                                                                 */
/* ======================================================================== */

/* xyb_impl.cuh:17-35 opsin constants; BT.709 full-range inverse. */
static const float M00 = 0.30f, M02 = 0.078f, M01 = 1.0f - 0.078f - 0.30f;
static const float M10 = 0.23f, M12 = 0.078f, M11 = 1.0f - 0.078f - 0.23f;
static const float M20 = 0.24342268924547819f, M21 = 0.20476744424496821f;
static const float M22 = 1.0f - 0.24342268924547819f - 0.20476744424496821f;
static const float OPSIN_BIAS = 0.0037930732552754493f;
static const float CR_TO_R = 1.5748f, CB_TO_G = 0.1873242729f, CR_TO_G = 0.4681242729f;
static const float CB_TO_B = 1.8556f;

/* xyb_impl.cuh:37-40 */
static inline float srgb_to_linear(float e) {
    return e <= 0.04045f ? e * (1.0f / 12.92f) : jo_pow24((e + 0.055f) * (1.0f / 1.055f));
}

/* xyb_impl.cuh:49-70 */
static void pixel_to_xyb(float yv, float cb, float cr, float* ox, float* oy, float* ob) {
    const float r = clamp01(srgb_to_linear(clamp01(yv + CR_TO_R * cr)));
    const float g = clamp01(srgb_to_linear(clamp01(yv - CB_TO_G * cb - CR_TO_G * cr)));
    const float b = clamp01(srgb_to_linear(clamp01(yv + CB_TO_B * cb)));
    float m0 = fmaxf(M00 * r + M01 * g + M02 * b + OPSIN_BIAS, 0.0f);
    float m1 = fmaxf(M10 * r + M11 * g + M12 * b + OPSIN_BIAS, 0.0f);
    float m2 = fmaxf(M20 * r + M21 * g + M22 * b + OPSIN_BIAS, 0.0f);
    const float neg_cbrt_bias = -jo_cbrt(OPSIN_BIAS);
    m0 = jo_cbrt(m0) + neg_cbrt_bias;
    m1 = jo_cbrt(m1) + neg_cbrt_bias;
    m2 = jo_cbrt(m2) + neg_cbrt_bias;
    *ox = 0.5f * (m0 - m1);
    *oy = 0.5f * (m0 + m1);
    *ob = m2;
}

/* Centered 4:2:0 bilinear with clamp-to-edge taps (tests/reference.py:23-46):
 * luma coordinate x samples chroma at x/2 - 1/4, so the weights are 1/4, 3/4.
 * Returns the upsampled raw 8-bit-scale value (exact in float). */
static inline void chroma_taps(uint32_t x, uint32_t cw, uint32_t* i0, uint32_t* i1, float* w0,
                               float* w1) {
    const uint32_t k = x >> 1;
    if (x & 1u) {
        *i0 = k;
        *i1 = k + 1 < cw ? k + 1 : cw - 1;
        *w0 = 0.75f;
        *w1 = 0.25f;
    } else {
        *i0 = k > 0 ? k - 1 : 0;
        *i1 = k;
        *w0 = 0.25f;
        *w1 = 0.75f;
    }
}

static void upsample_at(const uint8_t* chroma, size_t cp, uint32_t w, uint32_t h, uint32_t x,
                        uint32_t y, float* cb, float* cr) {
    uint32_t x0, x1, y0, y1;
    float wx0, wx1, wy0, wy1;
    chroma_taps(x, w / 2, &x0, &x1, &wx0, &wx1);
    chroma_taps(y, h / 2, &y0, &y1, &wy0, &wy1);
    const uint8_t* r0 = chroma + (size_t)y0 * cp;
    const uint8_t* r1 = chroma + (size_t)y1 * cp;
    const float b0 = (float)r0[2 * x0] * wx0 + (float)r0[2 * x1] * wx1;
    const float b1 = (float)r1[2 * x0] * wx0 + (float)r1[2 * x1] * wx1;
    const float c0 = (float)r0[2 * x0 + 1] * wx0 + (float)r0[2 * x1 + 1] * wx1;
    const float c1 = (float)r1[2 * x0 + 1] * wx0 + (float)r1[2 * x1 + 1] * wx1;
    *cb = b0 * wy0 + b1 * wy1;
    *cr = c0 * wy0 + c1 * wy1;
}

void jo_chroma_upsample(const uint8_t* chroma, size_t cp, uint32_t w, uint32_t h, float* cb,
                        float* cr) {
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            upsample_at(chroma, cp, w, h, x, y, &cb[(size_t)y * w + x], &cr[(size_t)y * w + x]);
        }
    }
}

/* xyb.cu:13-33 with exact bilinear chroma in place of the texture unit. */
void jo_nv12_to_xyb(const uint8_t* luma, size_t lp, const uint8_t* chroma, size_t cp, uint32_t w,
                    uint32_t h, float* xyb) {
    const size_t plane = (size_t)w * h;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            float cbu, cru;
            upsample_at(chroma, cp, w, h, x, y, &cbu, &cru);
            const float yv = (float)luma[(size_t)y * lp + x] * (1.0f / 255.0f);
            const float cb = cbu * (1.0f / 255.0f) - 128.0f / 255.0f;
            const float cr = cru * (1.0f / 255.0f) - 128.0f / 255.0f;
            const size_t i = (size_t)y * w + x;
            pixel_to_xyb(yv, cb, cr, &xyb[i], &xyb[plane + i], &xyb[2 * plane + i]);
        }
    }
}

/* The same transform evaluated in double with libm pow/cbrt (pinning only). */
void jo_nv12_to_xyb_f64(const uint8_t* luma, size_t lp, const uint8_t* chroma, size_t cp,
                        uint32_t w, uint32_t h, double* xyb) {
    const size_t plane = (size_t)w * h;
    const double m[3][3] = {{0.30, 1.0 - 0.078 - 0.30, 0.078},
                            {0.23, 1.0 - 0.078 - 0.23, 0.078},
                            {0.24342268924547819, 0.20476744424496821,
                             1.0 - 0.24342268924547819 - 0.20476744424496821}};
    const double bias = 0.0037930732552754493;
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            float cbu, cru;
            upsample_at(chroma, cp, w, h, x, y, &cbu, &cru);
            const double yv = luma[(size_t)y * lp + x] / 255.0;
            const double cb = cbu / 255.0 - 128.0 / 255.0;
            const double cr = cru / 255.0 - 128.0 / 255.0;
            double rgb[3] = {yv + 1.5748 * cr, yv - 0.1873242729 * cb - 0.4681242729 * cr,
                             yv + 1.8556 * cb};
            for (int c = 0; c < 3; ++c) {
                double e = rgb[c] < 0 ? 0 : (rgb[c] > 1 ? 1 : rgb[c]);
                e = e <= 0.04045 ? e / 12.92 : pow((e + 0.055) / 1.055, 2.4);
                rgb[c] = e < 0 ? 0 : (e > 1 ? 1 : e);
            }
            double o[3];
            for (int c = 0; c < 3; ++c) {
                double v = m[c][0] * rgb[0] + m[c][1] * rgb[1] + m[c][2] * rgb[2] + bias;
                o[c] = cbrt(v < 0 ? 0 : v) - cbrt(bias);
            }
            const size_t i = (size_t)y * w + x;
            xyb[i] = 0.5 * (o[0] - o[1]);
            xyb[plane + i] = 0.5 * (o[0] + o[1]);
            xyb[2 * plane + i] = o[2];
        }
    }
}

/* gaborish.h:16-28 weights; frontend_transform.cu:19-49 tap order. */
void jo_gaborish(const float* xyb, uint32_t w, uint32_t h, float* out) {
    const float k0 = -0.09495815671340026f, k1 = -0.041031725066768575f;
    const float k2 = 0.013710004822696948f, k3 = 0.006510206083837737f;
    const float k4 = -0.0014789063378272242f;
    const float nm = 1.0f / (1.0f + 4.0f * (k0 + k1 + k2 + k4 + 2.0f * k3));
    const float wc = nm, wr = nm * k0, wr2 = nm * k2, wd = nm * k1, wl = nm * k3, wd2 = nm * k4;
    const size_t plane = (size_t)w * h;
    for (int c = 0; c < 3; ++c) {
        const float* p = xyb + c * plane;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (uint32_t y = 0; y < h; ++y) {
            for (uint32_t x = 0; x < w; ++x) {
#define AT(dx, dy)                                                                            \
    p[(size_t)((long)y + (dy) < 0 ? 0 : ((long)y + (dy) >= (long)h ? h - 1 : y + (dy))) * w + \
      (size_t)((long)x + (dx) < 0 ? 0 : ((long)x + (dx) >= (long)w ? w - 1 : x + (dx)))]
                const float e = AT(-1, 0) + AT(1, 0) + AT(0, -1) + AT(0, 1);
                const float r2 = AT(-2, 0) + AT(2, 0) + AT(0, -2) + AT(0, 2);
                const float dg = AT(-1, -1) + AT(1, -1) + AT(-1, 1) + AT(1, 1);
                const float l8 = AT(-2, -1) + AT(2, -1) + AT(-2, 1) + AT(2, 1) + AT(-1, -2) +
                                 AT(1, -2) + AT(-1, 2) + AT(1, 2);
                const float d4 = AT(-2, -2) + AT(2, -2) + AT(-2, 2) + AT(2, 2);
                out[c * plane + (size_t)y * w + x] =
                    wc * AT(0, 0) + wr * e + wr2 * r2 + wd * dg + wl * l8 + wd2 * d4;
#undef AT
            }
        }
    }
}

/* ---- adaptive quantization (adaptive_quant_impl.cuh, frontend_fused.cu) ---- */

static const float K_SG_MUL = 226.77216153508914f;
static const float K_SG_LOG2 = 0.693147181f;
static const float K_SG_V_OFFSET = 7.7825991679894591f;

/* adaptive_quant_impl.cuh:29-40 */
static inline float ratio_cbrt_gamma(float v, int invert) {
    const float k_sg_mul2 = 1.0f / 73.377132366608819f;
    const float k_sg_ret_mul = k_sg_mul2 * 18.6580932135f * K_SG_LOG2;
    const float eps = 1e-2f;
    const float num_mul = k_sg_ret_mul * 3.0f * K_SG_MUL;
    const float v_offset = K_SG_V_OFFSET * K_SG_LOG2 + eps;
    const float den_mul = K_SG_LOG2 * K_SG_MUL;
    v = v < 0.0f ? 0.0f : v;
    const float v2 = v * v;
    const float num = num_mul * v2 + eps;
    const float den = den_mul * v * v2 + v_offset;
    return invert ? num / den : den / num;
}

/* adaptive_quant_impl.cuh:42-46 */
static inline float masking_sqrt(float v) {
    const float log_offset = 27.505837037000106f;
    const float mul = 211.66567973503678f * 1e8f;
    return 0.25f * sqrtf(v * sqrtf(mul) + log_offset);
}

/* adaptive_quant_impl.cuh:48-61 */
static inline float compute_mask(float out_val) {
    const float k_base = -0.7647f, k_mul4 = 9.4708735624378946f;
    const float k_mul2 = 17.35036561631863f, k_offset2 = 302.59587815579727f;
    const float k_mul3 = 6.7943250517376494f, k_offset3 = 3.7179635626140772f;
    const float k_offset4 = 0.25f * 3.7179635626140772f, k_mul0 = 0.80061762862741759f;
    const float v1 = fmaxf(out_val * k_mul0, 1e-3f);
    const float v2 = 1.0f / (v1 + k_offset2);
    const float v3 = 1.0f / (v1 * v1 + k_offset3);
    const float v4 = 1.0f / (v1 * v1 + k_offset4);
    return k_base + k_mul4 * v4 + k_mul2 * v2 + k_mul3 * v3;
}

static inline void cswap(float* a, float* b) {
    if (*a > *b) {
        const float t = *a;
        *a = *b;
        *b = t;
    }
}

/* adaptive_quant_impl.cuh:63-121 sort4 / store_min4; frontend_fused.cu:35-52 */
static float erosion_cell(const float* pre, int cx, int cy, const float* km) {
    const int xm1 = cx > 0 ? cx - 1 : cx, xp1 = cx + 1 < 8 ? cx + 1 : cx;
    const int ym1 = cy > 0 ? cy - 1 : cy, yp1 = cy + 1 < 8 ? cy + 1 : cy;
    float m0 = pre[cy * 8 + cx], m1 = pre[cy * 8 + xm1], m2 = pre[cy * 8 + xp1];
    float m3 = pre[ym1 * 8 + xm1];
    cswap(&m0, &m1);
    cswap(&m0, &m2);
    cswap(&m0, &m3);
    cswap(&m1, &m2);
    cswap(&m1, &m3);
    cswap(&m2, &m3);
    const float extra[5] = {pre[ym1 * 8 + cx], pre[ym1 * 8 + xp1], pre[yp1 * 8 + xm1],
                            pre[yp1 * 8 + cx], pre[yp1 * 8 + xp1]};
    for (int i = 0; i < 5; ++i) {
        const float v = extra[i];
        if (v < m3) {
            if (v < m0) {
                m3 = m2;
                m2 = m1;
                m1 = m0;
                m0 = v;
            } else if (v < m1) {
                m3 = m2;
                m2 = m1;
                m1 = v;
            } else if (v < m2) {
                m3 = m2;
                m2 = v;
            } else {
                m3 = v;
            }
        }
    }
    return km[0] * m0 + km[1] * m1 + km[2] * m2 + km[3] * m3;
}

/* frontend_fused.cu:214-239 host-side constants for the AQ field. */
static void aq_constants(float distance, float* km, float* mul_pb, float* add_pb) {
    const float scale = 0.725f / distance;
    const float base_level = 0.48f * scale;
    float dampen = 1.0f;
    if (distance >= 2.0f) {
        dampen = 1.0f - (distance - 2.0f) / (14.0f - 2.0f);
        if (dampen < 0.0f) {
            dampen = 0.0f;
        }
    }
    *mul_pb = scale * dampen;
    *add_pb = (1.0f - dampen) * base_level;
    const float efm = distance < 2.0f ? (2.0f - distance) * 0.5f : 0.0f;
    km[0] = 0.125f;
    km[1] = 0.10f - efm * 0.10f;
    km[2] = 0.09f - efm * 0.09f;
    km[3] = 0.06f - efm * 0.06f;
    const float norm = 0.29959705784054957f / (km[0] + km[1] + km[2] + km[3]);
    for (int i = 0; i < 4; ++i) {
        km[i] *= norm;
    }
}

/* frontend_fused.cu:54-204: per 32x32 tile (one-pixel clamped halo), 8x8
 * pre-erosion cells, tile-local erosion, per-block modulation.  Per-block sums
 * use the tree64 order over lanes l = fy*8 + fx. */
void jo_adaptive_quant(const float* xyb, uint32_t w, uint32_t h, float distance, int32_t* qf) {
    const uint32_t bw = w / 8, bh = h / 8;
    const size_t plane = (size_t)w * h;
    jo_cal cal;
    jo_calibrate(distance, &cal);
    float km[4], mul_pb, add_pb;
    aq_constants(distance, km, &mul_pb, &add_pb);
    enum { PAD = 34 };
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (uint32_t tby = 0; tby < (bh + 3) / 4; ++tby) {
        float ty[PAD * PAD], tx_[PAD * PAD], tb[PAD * PAD], pre[64];
        for (uint32_t tbx = 0; tbx < (bw + 3) / 4; ++tbx) {
            const long px0 = (long)tbx * 32, py0 = (long)tby * 32;
            for (int ly = 0; ly < PAD; ++ly) {
                for (int lx = 0; lx < PAD; ++lx) {
                    long gx = px0 + lx - 1, gy = py0 + ly - 1;
                    gx = gx < 0 ? 0 : (gx >= (long)w ? (long)w - 1 : gx);
                    gy = gy < 0 ? 0 : (gy >= (long)h ? (long)h - 1 : gy);
                    const size_t i = (size_t)gy * w + (size_t)gx;
                    tx_[ly * PAD + lx] = xyb[i];
                    ty[ly * PAD + lx] = xyb[plane + i];
                    tb[ly * PAD + lx] = xyb[2 * plane + i];
                }
            }
            for (int ci = 0; ci < 64; ++ci) {
                const int ccx = ci % 8, ccy = ci / 8;
                float sum = 0.0f;
                for (int dy = 0; dy < 4; ++dy) {
                    for (int dx = 0; dx < 4; ++dx) {
                        const int lx = ccx * 4 + dx + 1, ly = ccy * 4 + dy + 1;
                        const float yv = ty[ly * PAD + lx];
                        const float base =
                            0.25f * (ty[(ly + 1) * PAD + lx] + ty[(ly - 1) * PAD + lx] +
                                     ty[ly * PAD + lx - 1] + ty[ly * PAD + lx + 1]);
                        float diff = ratio_cbrt_gamma(yv + 0.019f, 0) * (yv - base);
                        diff *= diff;
                        diff = fminf(diff, 0.2f);
                        sum += masking_sqrt(diff);
                    }
                }
                pre[ci] = 0.25f * sum;
            }
            for (int lb = 0; lb < 16; ++lb) {
                const int lbx = lb % 4, lby = lb / 4;
                const uint32_t gbx = tbx * 4 + lbx, gby = tby * 4 + lby;
                if (gbx >= bw || gby >= bh) {
                    continue;
                }
                const int x0 = lbx * 8 + 1, y0 = lby * 8 + 1;
                float hf[64], gm[64], bl[64], er[64];
                for (int l = 0; l < 64; ++l) {
                    const int fx = l % 8, fy = l / 8;
                    const int idx = (y0 + fy) * PAD + x0 + fx;
                    const float yv = ty[idx], xv = tx_[idx], bv = tb[idx];
                    float hfp = 0.0f;
                    if (fx < 7) {
                        hfp += fminf(0.0206f, fabsf(yv - ty[idx + 1]));
                    }
                    if (fy < 7) {
                        hfp += fminf(0.0206f, fabsf(yv - ty[idx + PAD]));
                    }
                    hf[l] = hfp;
                    const float iny = yv + 0.16f;
                    gm[l] = ratio_cbrt_gamma(iny - xv, 1) + ratio_cbrt_gamma(iny + xv, 1);
                    const float pye = yv + 0.084381641171960495f + fabsf(xv);
                    bl[l] = bv > pye ? fminf(bv - pye, 0.027121074570634722f) : 0.0f;
                    er[l] = l < 4 ? erosion_cell(pre, 2 * lbx + (l & 1), 2 * lby + (l >> 1), km)
                                  : 0.0f;
                }
                const float hf_sum = tree64(hf);
                const float gamma_overall = tree64(gm) * (0.5f / 64.0f);
                float blue_sum = tree64(bl);
                const float aq = tree64(er);
                const float k_limit = 0.027121074570634722f;
                if (blue_sum >= 32.0f * k_limit) {
                    blue_sum = 64.0f * k_limit - blue_sum;
                }
                const float k_max_limit = 15.398788439047934f;
                if (blue_sum >= k_max_limit * k_limit) {
                    blue_sum = k_max_limit * k_limit;
                }
                float ov = compute_mask(aq);
                ov += hf_sum * -0.38f + 0.42f;
                ov += 0.1005613337192697f * jo_log2(gamma_overall);
                ov += blue_sum * 0.14207000358439159f;
                const float qval = jo_exp(ov) * mul_pb + add_pb;
                int qi = (int)(qval * cal.inv_global_scale + 0.5f);
                qi = qi < 1 ? 1 : (qi > 256 ? 256 : qi);
                qf[(size_t)gby * bw + gbx] = qi;
            }
        }
    }
}

/* ---- transform selection (transform_select_impl.cuh:49-201) ---------------- */

static float g_basis8[64], g_basis16[256], g_basis32[1024];
static int g_have_basis = 0;

static void ensure_basis(void) {
    if (g_have_basis) {
        return;
    }
    jo_dct_basis(8, g_basis8);
    jo_dct_basis(16, g_basis16);
    jo_dct_basis(32, g_basis32);
    g_have_basis = 1;
}

static const float* basis_for(int n) {
    ensure_basis();
    return n == 32 ? g_basis32 : (n == 16 ? g_basis16 : g_basis8);
}

/* DCT8 Y weight sampled at the normalized frequency (transform_select.cu:112-117). */
static void build_wsel(int n, float* t) {
    for (int e = 0; e < n * n; ++e) {
        const int fx = e / n, fy = e % n;
        long sx = lround(fx * 8.0 / n), sy = lround(fy * 8.0 / n);
        sx = sx > 7 ? 7 : sx;
        sy = sy > 7 ? 7 : sy;
        t[e] = g_w8[1][sx * 8 + sy];
    }
}

static const float* wsel_for(int n) {
    return n == 32 ? g_wsel32 : (n == 16 ? g_wsel16 : g_wsel8);
}

/* Separable forward DCT of one n x n block with origin `src` (row pitch
 * `pitch`): rows(fx, r) = sum_c A[fx][c] p[r][c]; coef(fx, fy) = sum_r
 * A[fy][r] rows(fx, r); both sums are ascending fmaf chains from 0. */
static void dct_block(const float* src, size_t pitch, int n, float* coef) {
    const float* a = basis_for(n);
    float rows[1024];
    for (int fx = 0; fx < n; ++fx) {
        for (int r = 0; r < n; ++r) {
            float s = 0.0f;
            for (int c = 0; c < n; ++c) {
                s = fmaf(a[fx * n + c], src[(size_t)r * pitch + c], s);
            }
            rows[fx * n + r] = s;
        }
    }
    for (int fx = 0; fx < n; ++fx) {
        for (int fy = 0; fy < n; ++fy) {
            float s = 0.0f;
            for (int r = 0; r < n; ++r) {
                s = fmaf(a[fy * n + r], rows[fx * n + r], s);
            }
            coef[fx * n + fy] = s;
        }
    }
}

static void element_cost(int n, int e, float coef, float qgsf, const float* wsel, float* d,
                         float* r) {
    const int m = n / 8, fx = e / n, fy = e % n;
    if (fx < m && fy < m) {
        *d = 0.0f;
        *r = 0.0f;
        return;
    }
    const float w = wsel[e];
    const float q = rintf(coef * qgsf / w);
    const float err = q * w / qgsf - coef;
    *d = err * err;
    *r = q != 0.0f ? 1.0f + jo_log2(1.0f + fabsf(q)) : 0.0f;
}

/* Region cost of an n x n block, summed in the device's register layout.
 * DCT8: lane fx holds column fx (elements fx*8 + f, f = 0..7), summed in f
 * order; the block's 8 lanes combine by a shuffle-down tree over offsets 4, 2,
 * 1.  DCT16: lane L (t = L & 15, g = L >> 4) holds elements t*16 + 4g + u,
 * u = 0..3, summed in u order, then tree64.  DCT32: lane L holds, for
 * b = 0..3 and u = 0..3 in that order, element (16(b&1) + t)*32 + 16(b>>1) +
 * 4g + u, then tree64. */
static float tree8(const float* s) {
    return ((s[0] + s[4]) + (s[2] + s[6])) + ((s[1] + s[5]) + (s[3] + s[7]));
}

static float block_cost_parts(const float* y, size_t pitch, int n, float qgsf, float* dist_out,
                              float* rate_out);

static float block_cost(const float* y, size_t pitch, int n, float qgsf) {
    return block_cost_parts(y, pitch, n, qgsf, NULL, NULL);
}

static float block_cost_parts(const float* y, size_t pitch, int n, float qgsf, float* dist_out,
                              float* rate_out) {
    const float* wsel = wsel_for(n);
    float coef[1024], d[1024], r[1024];
    dct_block(y, pitch, n, coef);
    for (int e = 0; e < n * n; ++e) {
        element_cost(n, e, coef[e], qgsf, wsel, &d[e], &r[e]);
    }
    float dist, rate;
    if (n == 8) {
        float sd[8], sr[8];
        for (int fx = 0; fx < 8; ++fx) {
            sd[fx] = d[fx * 8];
            sr[fx] = r[fx * 8];
            for (int f = 1; f < 8; ++f) {
                sd[fx] += d[fx * 8 + f];
                sr[fx] += r[fx * 8 + f];
            }
        }
        dist = tree8(sd);
        rate = tree8(sr);
    } else {
        float ld[64], lr[64];
        for (int l = 0; l < 64; ++l) {
            const int t = l & 15, g = l >> 4;
            int first = 1;
            for (int b = 0; b < (n == 32 ? 4 : 1); ++b) {
                for (int u = 0; u < 4; ++u) {
                    const int e = n == 32 ? (16 * (b & 1) + t) * 32 + 16 * (b >> 1) + 4 * g + u
                                          : t * 16 + 4 * g + u;
                    if (first) {
                        ld[l] = d[e];
                        lr[l] = r[e];
                        first = 0;
                    } else {
                        ld[l] += d[e];
                        lr[l] += r[e];
                    }
                }
            }
        }
        dist = tree64(ld);
        rate = tree64(lr);
    }
    if (dist_out) {
        *dist_out = dist;
    }
    if (rate_out) {
        *rate_out = rate;
    }
    return (float)(n * n) * dist + 6.0e-4f * rate;
}

static void write_first_block(int8_t* acs, uint32_t bw, uint32_t bx, uint32_t by, int n) {
    const int s = n / 8;
    for (int dy = 0; dy < s; ++dy) {
        for (int dx = 0; dx < s; ++dx) {
            acs[(size_t)(by + dy) * bw + bx + dx] = (dy == 0 && dx == 0) ? (int8_t)n : 0;
        }
    }
}

void jo_block_costs8(const float* y, uint32_t w, uint32_t h, float distance, float* cost8, float* dist8,
                     float* rate8) {
    const uint32_t bw = w / 8, bh = h / 8;
    jo_cal cal;
    jo_calibrate(distance, &cal);
    const float qgsf = (float)cal.raw_quant_field * cal.global_scale_float;
    ensure_basis();
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (uint32_t by = 0; by < bh; ++by) {
        for (uint32_t bx = 0; bx < bw; ++bx) {
            const size_t i = (size_t)by * bw + bx;
            float d, r;
            cost8[i] = block_cost_parts(y + (size_t)by * 8 * w + bx * 8, w, 8, qgsf, &d, &r);
            if (dist8) {
                dist8[i] = d;
            }
            if (rate8) {
                rate8[i] = r;
            }
        }
    }
}

void jo_select_transforms(const float* y, uint32_t w, uint32_t h, float distance, int8_t* acs) {
    const uint32_t bw = w / 8, bh = h / 8;
    jo_cal cal;
    jo_calibrate(distance, &cal);
    const float qgsf = (float)cal.raw_quant_field * cal.global_scale_float;
    const float inf = 1.0e30f, eps = 1.0e-6f;
    ensure_basis();
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (uint32_t rby = 0; rby < bh; rby += 4) {
        for (uint32_t rbx = 0; rbx < bw; rbx += 4) {
            float c8[16], c16[4], c32;
            for (int i = 0; i < 16; ++i) {
                const uint32_t bx = rbx + i % 4, by = rby + i / 4;
                c8[i] = (bx < bw && by < bh) ? block_cost(y + (size_t)by * 8 * w + bx * 8, w, 8,
                                                          qgsf)
                                             : inf;
            }
            for (int q = 0; q < 4; ++q) {
                const uint32_t bx = rbx + 2 * (q % 2), by = rby + 2 * (q / 2);
                c16[q] = (bx + 2 <= bw && by + 2 <= bh)
                             ? block_cost(y + (size_t)by * 8 * w + bx * 8, w, 16, qgsf)
                             : inf;
            }
            const int fit32 = rbx + 4 <= bw && rby + 4 <= bh;
            c32 = fit32 ? block_cost(y + (size_t)rby * 8 * w + rbx * 8, w, 32, qgsf) : inf;

            float split = 0.0f;
            int choice[4];
            for (int q = 0; q < 4; ++q) {
                const int qx = q % 2, qy = q / 2;
                const uint32_t bx = rbx + 2 * qx, by = rby + 2 * qy;
                float sum8 = 0.0f;
                int any8 = 0;
                for (int dy = 0; dy < 2; ++dy) {
                    for (int dx = 0; dx < 2; ++dx) {
                        if (bx + dx < bw && by + dy < bh) {
                            sum8 += c8[(2 * qy + dy) * 4 + 2 * qx + dx];
                            any8 = 1;
                        }
                    }
                }
                if (!any8) {
                    choice[q] = -1;
                    continue;
                }
                const int fit16 = bx + 2 <= bw && by + 2 <= bh;
                if (fit16 && c16[q] <= sum8 + eps) {
                    choice[q] = 16;
                    split += c16[q];
                } else {
                    choice[q] = 8;
                    split += sum8;
                }
            }
            if (fit32 && c32 <= split + eps) {
                write_first_block(acs, bw, rbx, rby, 32);
                continue;
            }
            for (int q = 0; q < 4; ++q) {
                const uint32_t bx = rbx + 2 * (q % 2), by = rby + 2 * (q / 2);
                if (choice[q] == 16) {
                    write_first_block(acs, bw, bx, by, 16);
                } else if (choice[q] == 8) {
                    for (int dy = 0; dy < 2; ++dy) {
                        for (int dx = 0; dx < 2; ++dx) {
                            if (bx + dx < bw && by + dy < bh) {
                                acs[(size_t)(by + dy) * bw + bx + dx] = 8;
                            }
                        }
                    }
                }
            }
        }
    }
}

/* ---- forward DCT into the covered-block layout (vardct_layout.h:54-62) ---- */

static inline size_t covered_slot(int n, uint32_t bx, uint32_t by, uint32_t bw, uint32_t e) {
    const uint32_t side = (uint32_t)n / 8, q = e / 64;
    return ((size_t)(by + q / side) * bw + bx + q % side) * 64 + e % 64;
}

void jo_forward_dct(const float* planes, uint32_t w, uint32_t h, const int8_t* acs, float* coef) {
    const uint32_t bw = w / 8, bh = h / 8;
    const size_t plane = (size_t)w * h, cplane = (size_t)bw * bh * 64;
    ensure_basis();
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 4)
    for (uint32_t by = 0; by < bh; ++by) {
        float blk[1024];
        for (uint32_t bx = 0; bx < bw; ++bx) {
            const int n = acs ? acs[(size_t)by * bw + bx] : 8;
            if (n == 0) {
                continue;
            }
            for (int c = 0; c < 3; ++c) {
                dct_block(planes + c * plane + (size_t)by * 8 * w + bx * 8, w, n, blk);
                for (int e = 0; e < n * n; ++e) {
                    coef[c * cplane + covered_slot(n, bx, by, bw, (uint32_t)e)] = blk[e];
                }
            }
        }
    }
}

void jo_forward_dct_f64(const double* planes, uint32_t w, uint32_t h, const int8_t* acs,
                        double* coef) {
    const uint32_t bw = w / 8, bh = h / 8;
    const size_t plane = (size_t)w * h, cplane = (size_t)bw * bh * 64;
    static double a[1024];
    for (uint32_t by = 0; by < bh; ++by) {
        for (uint32_t bx = 0; bx < bw; ++bx) {
            const int n = acs ? acs[(size_t)by * bw + bx] : 8;
            if (n == 0) {
                continue;
            }
            dct_basis_f64((uint32_t)n, a);
            for (int c = 0; c < 3; ++c) {
                const double* src = planes + c * plane + (size_t)by * 8 * w + bx * 8;
                for (int fx = 0; fx < n; ++fx) {
                    for (int fy = 0; fy < n; ++fy) {
                        double s = 0.0;
                        for (int r = 0; r < n; ++r) {
                            for (int cc = 0; cc < n; ++cc) {
                                s += a[fy * n + r] * a[fx * n + cc] * src[(size_t)r * w + cc];
                            }
                        }
                        coef[c * cplane + covered_slot(n, bx, by, bw, (uint32_t)(fx * n + fy))] =
                            s;
                    }
                }
            }
        }
    }
}

/* ---- chroma from luma (frontend_quantize.cu:49-117, cfl.h:58-64) ---------- */

/* First-block owning 8x8 position (bx, by): region origin if it holds a DCT32,
 * quad origin if it holds a DCT16, else the position itself. */
static void owner_of(const int8_t* acs, uint32_t bw, uint32_t bx, uint32_t by, uint32_t* ox,
                     uint32_t* oy, int* n) {
    if (!acs) {
        *ox = bx;
        *oy = by;
        *n = 8;
        return;
    }
    const uint32_t rx = bx & ~3u, ry = by & ~3u;
    if (acs[(size_t)ry * bw + rx] == 32) {
        *ox = rx;
        *oy = ry;
        *n = 32;
        return;
    }
    const uint32_t qx = bx & ~1u, qy = by & ~1u;
    if (acs[(size_t)qy * bw + qx] == 16) {
        *ox = qx;
        *oy = qy;
        *n = 16;
        return;
    }
    *ox = bx;
    *oy = by;
    *n = 8;
}

/* Raw coefficient index e held by slot s of position (bx, by). */
static inline uint32_t raw_of(uint32_t bx, uint32_t by, uint32_t ox, uint32_t oy, int n,
                              uint32_t s) {
    const uint32_t q = (by - oy) * ((uint32_t)n / 8) + (bx - ox);
    return q * 64 + s;
}

static inline int raw_is_llf(uint32_t e, int n) {
    const uint32_t m = (uint32_t)n / 8;
    return e / (uint32_t)n < m && e % (uint32_t)n < m;
}

/* Per 32x32 region FP32 regression partials (Sxy, Syy, S(b-y)y) over non-LLF
 * FP16 coefficients, in the device layout: iteration k = 16x16 quad (raster),
 * wave w = 8x8 block within the quad (raster), lane = slot; per-thread sum
 * over k, tree64, then (P0 + P1) + (P2 + P3). */
void jo_cfl_region_partials(const uint16_t* c16, const int8_t* acs, uint32_t w, uint32_t h,
                            float* partials) {
    const uint32_t bw = w / 8, bh = h / 8, rw = (bw + 3) / 4, rh = (bh + 3) / 4;
    const size_t cplane = (size_t)bw * bh * 64;
    for (uint32_t ry = 0; ry < rh; ++ry) {
        for (uint32_t rx = 0; rx < rw; ++rx) {
            float p[3][4];
            for (uint32_t wv = 0; wv < 4; ++wv) {
                float lanes[3][64];
                for (uint32_t l = 0; l < 64; ++l) {
                    float acc[3] = {0.0f, 0.0f, 0.0f};
                    for (uint32_t k = 0; k < 4; ++k) {
                        const uint32_t bx = rx * 4 + 2 * (k % 2) + wv % 2;
                        const uint32_t by = ry * 4 + 2 * (k / 2) + wv / 2;
                        float v[3] = {0.0f, 0.0f, 0.0f};
                        if (bx < bw && by < bh) {
                            uint32_t ox, oy;
                            int n;
                            owner_of(acs, bw, bx, by, &ox, &oy, &n);
                            if (!raw_is_llf(raw_of(bx, by, ox, oy, n, l), n)) {
                                const size_t i = ((size_t)by * bw + bx) * 64 + l;
                                const float xr = jo_f16_to_f32(c16[i]);
                                const float yr = jo_f16_to_f32(c16[cplane + i]);
                                const float br = jo_f16_to_f32(c16[2 * cplane + i]);
                                v[0] = xr * yr;
                                v[1] = yr * yr;
                                v[2] = (br - yr) * yr;
                            }
                        }
                        for (int c = 0; c < 3; ++c) {
                            acc[c] = k == 0 ? v[c] : acc[c] + v[c];
                        }
                    }
                    for (int c = 0; c < 3; ++c) {
                        lanes[c][l] = acc[c];
                    }
                }
                for (int c = 0; c < 3; ++c) {
                    p[c][wv] = tree64(lanes[c]);
                }
            }
            for (int c = 0; c < 3; ++c) {
                partials[((size_t)ry * rw + rx) * 3 + c] = (p[c][0] + p[c][1]) + (p[c][2] + p[c][3]);
            }
        }
    }
}

static inline int cfl_quantize_factor(float v) {
    int m = (int)lrintf(v * 84);
    return m < -128 ? -128 : (m > 127 ? 127 : m);
}

/* Per 64x64 tile: (R00 + R10) + (R01 + R11) of its regions (absent = 0), then
 * cfl_maps_from_sums (cfl.h:58-64). */
void jo_cfl_maps(const float* partials, uint32_t w, uint32_t h, int8_t* ytox, int8_t* ytob) {
    const uint32_t bw = w / 8, bh = h / 8, rw = (bw + 3) / 4, rh = (bh + 3) / 4;
    const uint32_t cmw = (bw + 7) / 8, cmh = (bh + 7) / 8;
    for (uint32_t ty = 0; ty < cmh; ++ty) {
        for (uint32_t tx = 0; tx < cmw; ++tx) {
            float r[4][3];
            for (int i = 0; i < 4; ++i) {
                const uint32_t rx = tx * 2 + (uint32_t)(i % 2), ry = ty * 2 + (uint32_t)(i / 2);
                for (int c = 0; c < 3; ++c) {
                    r[i][c] = (rx < rw && ry < rh) ? partials[((size_t)ry * rw + rx) * 3 + c] : 0.0f;
                }
            }
            float s[3];
            for (int c = 0; c < 3; ++c) {
                s[c] = (r[0][c] + r[1][c]) + (r[2][c] + r[3][c]);
            }
            const double sxy = s[0], syy = s[1], sby = s[2];
            const float fx = syy > 0.0 ? (float)(sxy / syy) : 0.0f;
            const float db = syy > 0.0 ? (float)(sby / syy) : 0.0f;
            ytox[(size_t)ty * cmw + tx] = (int8_t)cfl_quantize_factor(fx);
            ytob[(size_t)ty * cmw + tx] = (int8_t)cfl_quantize_factor(db);
        }
    }
}

/* ---- LLF -> DC (vardct_layout.h:88-111) ------------------------------------ */

/* T[f][o] = ortho_m(f, o) * DCTResampleScales[f]; DC(ox, oy) = m * sum_fx
 * sum_fy T[fx][ox] T[fy][oy] llf[fx*m + fy] in double, fx outer. */
static void llf_table(uint32_t m, double* t) {
    static const double s16[2] = {1.0, 0.901764195028874394};
    static const double s32[4] = {1.0, 0.974886821136879522, 0.901764195028874394,
                                  0.787054918159101335};
    const double* sc = m == 2 ? s16 : s32;
    for (uint32_t f = 0; f < m; ++f) {
        const double norm = f == 0 ? sqrt(1.0 / m) : sqrt(2.0 / m);
        for (uint32_t o = 0; o < m; ++o) {
            t[f * m + o] = norm * cos(3.14159265358979323846 * (o + 0.5) * f / m) * sc[f];
        }
    }
}

void jo_dc_from_llf(uint32_t m, const float* llf, float* dc) {
    if (m == 1) {
        dc[0] = llf[0];
        return;
    }
    double t[16];
    llf_table(m, t);
    for (uint32_t oy = 0; oy < m; ++oy) {
        for (uint32_t ox = 0; ox < m; ++ox) {
            double sum = 0.0;
            for (uint32_t fx = 0; fx < m; ++fx) {
                for (uint32_t fy = 0; fy < m; ++fy) {
                    sum += t[fx * m + ox] * t[fy * m + oy] * (double)llf[fx * m + fy];
                }
            }
            dc[oy * m + ox] = (float)((double)m * sum);
        }
    }
}

/* ---- quantization (frontend_quantize.cu:125-201) -------------------------- */

static const float DC_INV_QUANT[3] = {4096.0f, 512.0f, 256.0f};

void jo_quantize(const uint16_t* c16, const int8_t* acs, const int8_t* ytox, const int8_t* ytob,
                 const int32_t* qf, uint32_t w, uint32_t h, float distance, int16_t* ac,
                 int32_t* dc, uint8_t* nz_grid) {
    const uint32_t bw = w / 8, bh = h / 8, cmw = (bw + 7) / 8;
    const size_t nblk = (size_t)bw * bh, cplane = nblk * 64;
    jo_cal cal;
    jo_calibrate(distance, &cal);
    const float gsf = cal.global_scale_float;
    const float dc_scale = cal.global_scale_float * (float)cal.quant_dc;
    for (uint32_t by = 0; by < bh; ++by) {
        for (uint32_t bx = 0; bx < bw; ++bx) {
            const size_t blk = (size_t)by * bw + bx;
            const int n = acs ? acs[blk] : 8;
            if (n == 0) {
                continue;
            }
            const uint32_t m = (uint32_t)n / 8;
            const float qgsf = (float)qf[blk] * gsf;
            const size_t tile = (size_t)(by / 8) * cmw + bx / 8;
            const float yx = 0.0f + (float)(ytox ? ytox[tile] : 0) * (1.0f / 84);
            const float yb = 1.0f + (float)(ytob ? ytob[tile] : 0) * (1.0f / 84);
            float llf[3][16];
            uint32_t nz[3] = {0, 0, 0};
            for (int e = 0; e < n * n; ++e) {
                const size_t slot = covered_slot(n, bx, by, bw, (uint32_t)e);
                const float xr = jo_f16_to_f32(c16[slot]);
                const float yr = jo_f16_to_f32(c16[cplane + slot]);
                const float br = jo_f16_to_f32(c16[2 * cplane + slot]);
                if (raw_is_llf((uint32_t)e, n)) {
                    const uint32_t lp = ((uint32_t)e / (uint32_t)n) * m + (uint32_t)e % (uint32_t)n;
                    llf[0][lp] = xr;
                    llf[1][lp] = yr;
                    llf[2][lp] = br;
                    ac[slot] = 0;
                    ac[cplane + slot] = 0;
                    ac[2 * cplane + slot] = 0;
                    continue;
                }
                const float wy = dequant_weight(n, 1, e);
                const int qy = (int)lrintf(yr * qgsf / wy);
                const float y_round = (float)qy * wy / qgsf;
                const float wx = dequant_weight(n, 0, e);
                const float wb = dequant_weight(n, 2, e);
                const int qx = (int)lrintf((xr - yx * y_round) * qgsf / wx);
                const int qb = (int)lrintf((br - yb * y_round) * qgsf / wb);
                ac[slot] = (int16_t)qx;
                ac[cplane + slot] = (int16_t)qy;
                ac[2 * cplane + slot] = (int16_t)qb;
                nz[0] += (int16_t)qx != 0;
                nz[1] += (int16_t)qy != 0;
                nz[2] += (int16_t)qb != 0;
            }
            float dcv[3][16];
            for (int c = 0; c < 3; ++c) {
                jo_dc_from_llf(m, llf[c], dcv[c]);
            }
            const float dcq_x = DC_INV_QUANT[0] * dc_scale;
            const float dcq_y = DC_INV_QUANT[1] * dc_scale;
            const float dcq_b = DC_INV_QUANT[2] * dc_scale;
            const uint32_t l2 = n == 16 ? 2 : (n == 32 ? 4 : 0);
            for (uint32_t oy = 0; oy < m; ++oy) {
                for (uint32_t ox = 0; ox < m; ++ox) {
                    const uint32_t p = oy * m + ox;
                    const size_t d = (size_t)(by + oy) * bw + bx + ox;
                    const int qdy = (int)lrintf(dcv[1][p] * dcq_y);
                    const float y_dc_round = (float)qdy / dcq_y;
                    dc[nblk + d] = qdy;
                    dc[d] = (int)lrintf(dcv[0][p] * dcq_x);
                    dc[2 * nblk + d] = (int)lrintf((dcv[2][p] - 1.0f * y_dc_round) * dcq_b);
                    if (nz_grid) {
                        for (int c = 0; c < 3; ++c) {
                            nz_grid[c * nblk + d] = (uint8_t)((nz[c] + m * m - 1) >> l2);
                        }
                    }
                }
            }
        }
    }
}

/* quant.cu:30-97: DCT8 everywhere, uniform quant field, base correlations. */
void jo_quantize_dct8_uniform(const float* coef, uint32_t w, uint32_t h, float distance,
                              int32_t* ac, int32_t* dc) {
    jo_cal cal;
    jo_calibrate(distance, &cal);
    const size_t nblk = (size_t)(w / 8) * (h / 8), plane = nblk * 64;
    const float gsf = cal.global_scale_float;
    const float dc_scale = cal.global_scale_float * (float)cal.quant_dc;
    const float ac_scale = (float)cal.raw_quant_field * gsf;
    for (size_t g = 0; g < plane; ++g) {
        const int k = (int)(g % 64);
        const size_t b = g / 64;
        float f[3];
        for (int c = 0; c < 3; ++c) {
            f[c] = k == 0 ? DC_INV_QUANT[c] * dc_scale : ac_scale / g_w8[c][k];
        }
        const int32_t qy = (int32_t)rintf(coef[plane + g] * f[1]);
        const int32_t qx = (int32_t)rintf(coef[g] * f[0]);
        const float roundtrip_y = (float)qy / f[1];
        const int32_t qb = (int32_t)rintf((coef[2 * plane + g] - 1.0f * roundtrip_y) * f[2]);
        const int32_t q[3] = {qx, qy, qb};
        for (int c = 0; c < 3; ++c) {
            if (k == 0) {
                dc[c * nblk + b] = q[c];
                ac[c * plane + g] = 0;
            } else {
                ac[c * plane + g] = q[c];
            }
        }
    }
}

/* ======================================================================== */
/* Bit writer and field coders (bit_writer.h, field_coder.h)                 */
/* ======================================================================== */

typedef struct bw_t {
    uint8_t* buf;
    size_t cap;
    size_t nbits;
} bw_t;

static void bw_init(bw_t* w) {
    w->cap = 256;
    w->buf = (uint8_t*)calloc(w->cap, 1);
    w->nbits = 0;
}

static void bw_free(bw_t* w) {
    free(w->buf);
    w->buf = NULL;
}

static void bw_reserve(bw_t* w, size_t bits) {
    const size_t need = (w->nbits + bits + 7) / 8 + 8;
    if (need > w->cap) {
        size_t nc = w->cap * 2;
        while (nc < need) {
            nc *= 2;
        }
        w->buf = (uint8_t*)realloc(w->buf, nc);
        memset(w->buf + w->cap, 0, nc - w->cap);
        w->cap = nc;
    }
}

/* LSB-first: bit i of the stream is bit (i & 7) of byte i >> 3. */
static void bw_write(bw_t* w, unsigned n, uint64_t v) {
    if (n == 0) {
        return;
    }
    bw_reserve(w, n);
    for (unsigned i = 0; i < n; ++i) {
        if ((v >> i) & 1u) {
            w->buf[(w->nbits + i) >> 3] |= (uint8_t)(1u << ((w->nbits + i) & 7));
        }
    }
    w->nbits += n;
}

static void bw_pad(bw_t* w) {
    w->nbits = (w->nbits + 7) & ~(size_t)7;
}

static size_t bw_bytes(const bw_t* w) {
    return (w->nbits + 7) / 8;
}

static void bw_append(bw_t* w, const bw_t* src) {
    for (size_t i = 0; i < src->nbits; ++i) {
        bw_write(w, 1, (src->buf[i >> 3] >> (i & 7)) & 1u);
    }
}

static size_t bw_emit(const bw_t* w, uint8_t* out, size_t cap) {
    const size_t n = bw_bytes(w);
    if (out) {
        memcpy(out, w->buf, n < cap ? n : cap);
    }
    return n;
}

typedef struct distr {
    int direct;
    uint32_t v;
    uint32_t bits;
} distr;

#define VAL(x) {1, (x), 0}
#define BITSOFF(b, o) {0, (o), (b)}

/* field_coder.h:55-83 */
static void write_u32(bw_t* w, const distr d[4], uint32_t value) {
    uint32_t best = 0;
    size_t best_bits = 64;
    for (uint32_t s = 0; s < 4; ++s) {
        if (d[s].direct) {
            if (d[s].v == value) {
                best = s;
                best_bits = 2;
                break;
            }
            continue;
        }
        if (value < d[s].v || (uint64_t)value >= (uint64_t)d[s].v + (1ull << d[s].bits)) {
            continue;
        }
        if (2 + d[s].bits < best_bits) {
            best = s;
            best_bits = 2 + d[s].bits;
        }
    }
    bw_write(w, 2, best);
    if (!d[best].direct) {
        bw_write(w, d[best].bits, value - d[best].v);
    }
}

/* field_coder.h:87-107 */
static void write_u64(bw_t* w, uint64_t value) {
    if (value == 0) {
        bw_write(w, 2, 0);
    } else if (value <= 16) {
        bw_write(w, 2, 1);
        bw_write(w, 4, value - 1);
    } else if (value <= 272) {
        bw_write(w, 2, 2);
        bw_write(w, 8, value - 17);
    } else {
        bw_write(w, 2, 3);
        value -= 273;
        bw_write(w, 12, value & 0xfff);
        value >>= 12;
        for (int i = 0; i < 8; ++i) {
            if (value == 0) {
                bw_write(w, 1, 0);
                return;
            }
            bw_write(w, 1, 1);
            const unsigned chunk = i < 7 ? 8 : 4;
            bw_write(w, chunk, value & ((1u << chunk) - 1));
            value >>= chunk;
        }
    }
}

/* hybrid_uint.h:23-40 */
static void hybrid(uint32_t se, uint32_t msb, uint32_t lsb, uint32_t value, uint32_t* tok,
                   uint32_t* nb, uint32_t* bits) {
    const uint32_t split = 1u << se;
    if (value < split) {
        *tok = value;
        *nb = 0;
        *bits = 0;
        return;
    }
    uint32_t n = 31;
    while (!((value >> n) & 1u)) {
        --n;
    }
    const uint32_t m = value - (1u << n);
    *tok = split + ((n - se) << (msb + lsb)) + ((m >> (n - msb)) << lsb) + (m & ((1u << lsb) - 1));
    *nb = n - msb - lsb;
    *bits = (value >> lsb) & ((1u << *nb) - 1);
}

static inline void hybrid420(uint32_t v, uint32_t* tok, uint32_t* nb, uint32_t* bits) {
    hybrid(4, 2, 0, v, tok, nb, bits);
}

static inline uint32_t pack_signed(int32_t v) {
    return ((uint32_t)v << 1) ^ (uint32_t)(v >> 31);
}

static uint32_t ceil_log2(size_t x) {
    if (x <= 1) {
        return 0;
    }
    uint32_t b = 0;
    size_t v = x - 1;
    while (v) {
        v >>= 1;
        ++b;
    }
    return b;
}

/* ======================================================================== */
/* Prefix codes (prefix_code.cc:36-397)                                      */
/* ======================================================================== */

typedef struct hnode {
    uint32_t total;
    int16_t left;
    int16_t right_or_value;
} hnode;

static void set_depth(int p, const hnode* pool, uint8_t* depth, uint8_t level) {
    if (pool[p].left >= 0) {
        set_depth(pool[p].left, pool, depth, (uint8_t)(level + 1));
        set_depth(pool[p].right_or_value, pool, depth, (uint8_t)(level + 1));
    } else {
        depth[pool[p].right_or_value] = level;
    }
}

/* prefix_code.cc:36-96: leaves inserted in descending symbol order, stable sort
 * by count, two-queue merge, count floor doubled until depth <= limit. */
static void create_huffman_tree(const uint32_t* data, size_t length, int limit, uint8_t* depth) {
    hnode* tree = (hnode*)malloc(sizeof(hnode) * (2 * length + 4));
    for (uint32_t count_limit = 1;; count_limit *= 2) {
        size_t n = 0;
        for (size_t i = length; i != 0;) {
            --i;
            if (data[i]) {
                const uint32_t c = data[i] > count_limit - 1 ? data[i] : count_limit - 1;
                tree[n].total = c;
                tree[n].left = -1;
                tree[n].right_or_value = (int16_t)i;
                ++n;
            }
        }
        if (n == 1) {
            depth[tree[0].right_or_value] = 1;
            break;
        }
        for (size_t i = 1; i < n; ++i) { /* stable insertion sort */
            const hnode key = tree[i];
            size_t j = i;
            while (j > 0 && tree[j - 1].total > key.total) {
                tree[j] = tree[j - 1];
                --j;
            }
            tree[j] = key;
        }
        const hnode sentinel = {0xffffffffu, -1, -1};
        size_t size = n;
        tree[size++] = sentinel;
        tree[size++] = sentinel;
        size_t i = 0, j = n + 1;
        for (size_t k = n - 1; k != 0; --k) {
            size_t left, right;
            if (tree[i].total <= tree[j].total) {
                left = i++;
            } else {
                left = j++;
            }
            if (tree[i].total <= tree[j].total) {
                right = i++;
            } else {
                right = j++;
            }
            const size_t j_end = size - 1;
            tree[j_end].total = tree[left].total + tree[right].total;
            tree[j_end].left = (int16_t)left;
            tree[j_end].right_or_value = (int16_t)right;
            tree[size++] = sentinel;
        }
        set_depth((int)(2 * n - 1), tree, depth, 0);
        int maxd = 0;
        for (size_t s = 0; s < length; ++s) {
            maxd = depth[s] > maxd ? depth[s] : maxd;
        }
        if (maxd <= limit) {
            break;
        }
    }
    free(tree);
}

static uint16_t reverse_bits(int num_bits, uint16_t bits) {
    uint16_t r = 0;
    for (int i = 0; i < num_bits; ++i) {
        r = (uint16_t)((r << 1) | ((bits >> i) & 1u));
    }
    return r;
}

/* prefix_code.cc:111-130: canonical codes, LSB-first code words. */
static void depths_to_codes(const uint8_t* depth, size_t len, uint16_t* bits) {
    uint16_t bl_count[16] = {0}, next_code[16] = {0};
    for (size_t i = 0; i < len; ++i) {
        ++bl_count[depth[i]];
    }
    bl_count[0] = 0;
    int code = 0;
    for (int i = 1; i < 16; ++i) {
        code = (code + bl_count[i - 1]) << 1;
        next_code[i] = (uint16_t)code;
    }
    for (size_t i = 0; i < len; ++i) {
        if (depth[i]) {
            bits[i] = reverse_bits(depth[i], next_code[depth[i]]++);
        }
    }
}

/* prefix_code.cc:211-354: code-length sequence (one symbol per depth, no RLE
 * for these alphabets), its 18-symbol code limited to 5 bits, stored in the
 * fixed permuted order with the static length code. */
static void store_huffman_tree(const uint8_t* depths, size_t num, bw_t* w) {
    size_t new_len = num;
    while (new_len > 0 && depths[new_len - 1] == 0) {
        --new_len;
    }
    uint32_t hist[18] = {0};
    for (size_t i = 0; i < new_len; ++i) {
        ++hist[depths[i]];
    }
    int num_codes = 0, code = 0;
    for (int i = 0; i < 18; ++i) {
        if (hist[i]) {
            if (num_codes == 0) {
                code = i;
                num_codes = 1;
            } else if (num_codes == 1) {
                num_codes = 2;
                break;
            }
        }
    }
    uint8_t cl_depth[18] = {0};
    uint16_t cl_bits[18] = {0};
    create_huffman_tree(hist, 18, 5, cl_depth);
    depths_to_codes(cl_depth, 18, cl_bits);

    static const uint8_t order[18] = {1, 2, 3, 4, 0, 5, 17, 6, 16, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    static const uint8_t sym[6] = {0, 7, 3, 2, 1, 15};
    static const uint8_t len[6] = {2, 4, 3, 2, 2, 4};
    size_t to_store = 18;
    if (num_codes > 1) {
        for (; to_store > 0; --to_store) {
            if (cl_depth[order[to_store - 1]] != 0) {
                break;
            }
        }
    }
    size_t skip = 0;
    if (cl_depth[order[0]] == 0 && cl_depth[order[1]] == 0) {
        skip = 2;
        if (cl_depth[order[2]] == 0) {
            skip = 3;
        }
    }
    bw_write(w, 2, skip);
    for (size_t i = skip; i < to_store; ++i) {
        const uint8_t l = cl_depth[order[i]];
        bw_write(w, len[l], sym[l]);
    }
    if (num_codes == 1) {
        cl_depth[code] = 0;
    }
    for (size_t i = 0; i < new_len; ++i) {
        bw_write(w, cl_depth[depths[i]], cl_bits[depths[i]]);
    }
}

/* prefix_code.cc:358-397 */
static void build_and_store_huffman(const uint32_t* hist, size_t length, uint8_t* depth,
                                    uint16_t* bits, bw_t* w) {
    memset(depth, 0, length);
    memset(bits, 0, length * sizeof(uint16_t));
    size_t count = 0, s4[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < length; ++i) {
        if (hist[i]) {
            if (count < 4) {
                s4[count] = i;
            } else if (count > 4) {
                break;
            }
            ++count;
        }
    }
    size_t max_bits = 0;
    for (size_t c = length - 1; c; c >>= 1) {
        ++max_bits;
    }
    if (count <= 1) {
        bw_write(w, 4, 1);
        bw_write(w, (unsigned)max_bits, s4[0]);
        return;
    }
    create_huffman_tree(hist, length, 15, depth);
    depths_to_codes(depth, length, bits);
    if (count <= 4) {
        bw_write(w, 2, 1);
        bw_write(w, 2, count - 1);
        for (size_t i = 0; i < count; ++i) {
            for (size_t j = i + 1; j < count; ++j) {
                if (depth[s4[j]] < depth[s4[i]]) {
                    const size_t t = s4[j];
                    s4[j] = s4[i];
                    s4[i] = t;
                }
            }
        }
        for (size_t i = 0; i < count; ++i) {
            bw_write(w, (unsigned)max_bits, s4[i]);
        }
        if (count == 4) {
            bw_write(w, 1, depth[s4[0]] == 1 ? 1 : 0);
        }
    } else {
        store_huffman_tree(depth, length, w);
    }
}

size_t jo_huffman(const uint32_t* hist, uint32_t len, uint8_t* depth, uint16_t* bits,
                  uint8_t* header, size_t cap, size_t* header_bits) {
    bw_t w;
    bw_init(&w);
    build_and_store_huffman(hist, len, depth, bits, &w);
    *header_bits = w.nbits;
    const size_t n = bw_emit(&w, header, cap);
    bw_free(&w);
    return n;
}

/* ======================================================================== */
/* Histogram headers and context maps (histogram_writer.cc, context_map.cc)  */
/* ======================================================================== */

static void encode_uint_config(bw_t* w, uint32_t se, uint32_t msb, uint32_t lsb) {
    bw_write(w, ceil_log2(15 + 1), se);
    if (se == 15) {
        return;
    }
    bw_write(w, ceil_log2(se + 1), msb);
    bw_write(w, ceil_log2(se - msb + 1), lsb);
}

static void store_var_len_u16(bw_t* w, size_t n) {
    if (n == 0) {
        bw_write(w, 1, 0);
        return;
    }
    bw_write(w, 1, 1);
    unsigned nb = 0;
    for (size_t v = n; v > 1; v >>= 1) {
        ++nb;
    }
    bw_write(w, 4, nb);
    bw_write(w, nb, n - ((size_t)1 << nb));
}

/* histogram_writer.cc:59-83 (single histogram, simple zero context map) */
static void write_prefix_histograms(bw_t* w, const uint32_t* hist, size_t length,
                                    size_t num_contexts, uint32_t se, uint32_t msb, uint32_t lsb,
                                    uint8_t* depth, uint16_t* bits) {
    bw_write(w, 1, 0);
    if (num_contexts > 1) {
        bw_write(w, 1, 1);
        bw_write(w, 2, 0);
    }
    bw_write(w, 1, 1);
    encode_uint_config(w, se, msb, lsb);
    store_var_len_u16(w, length - 1);
    memset(depth, 0, length);
    memset(bits, 0, length * sizeof(uint16_t));
    if (length > 1) {
        build_and_store_huffman(hist, length, depth, bits, w);
    }
}

static void cmap_symbols(const uint8_t* map, size_t n, int mtf, uint8_t* out) {
    if (!mtf) {
        memcpy(out, map, n);
        return;
    }
    uint8_t s[256];
    for (int i = 0; i < 256; ++i) {
        s[i] = (uint8_t)i;
    }
    for (size_t i = 0; i < n; ++i) {
        size_t r = 0;
        while (s[r] != map[i]) {
            ++r;
        }
        out[i] = (uint8_t)r;
        const uint8_t v = s[r];
        while (r > 0) {
            s[r] = s[r - 1];
            --r;
        }
        s[0] = v;
    }
}

void jo_move_to_front(const uint8_t* in, uint32_t n, uint8_t* out) {
    if (n) {
        cmap_symbols(in, n, 1, out);
    }
}

/* context_map.cc:101-130 */
static void write_complex_cmap(bw_t* w, const uint8_t* map, size_t n, int mtf) {
    uint8_t* sym = (uint8_t*)malloc(n);
    cmap_symbols(map, n, mtf, sym);
    uint32_t hist[16] = {0};
    for (size_t i = 0; i < n; ++i) {
        uint32_t t, nb, b;
        hybrid(2, 0, 1, sym[i], &t, &nb, &b);
        ++hist[t];
    }
    size_t alpha = 1;
    for (size_t i = 0; i < 16; ++i) {
        if (hist[i]) {
            alpha = i + 1;
        }
    }
    bw_write(w, 1, 0);
    bw_write(w, 1, mtf ? 1 : 0);
    uint8_t depth[16];
    uint16_t bits[16];
    write_prefix_histograms(w, hist, alpha, 1, 2, 0, 1, depth, bits);
    for (size_t i = 0; i < n; ++i) {
        uint32_t t, nb, b;
        hybrid(2, 0, 1, sym[i], &t, &nb, &b);
        bw_write(w, depth[t], bits[t]);
        bw_write(w, nb, b);
    }
    free(sym);
}

/* context_map.cc:135-168: smallest of complex, complex+MTF, simple. */
static int write_best_cmap(bw_t* w, const uint8_t* map, size_t n, size_t nclusters) {
    bw_t cx, cm, sp;
    bw_init(&cx);
    bw_init(&cm);
    write_complex_cmap(&cx, map, n, 0);
    write_complex_cmap(&cm, map, n, 1);
    const bw_t* sel = &cx;
    int enc = 1;
    if (cm.nbits < sel->nbits) {
        sel = &cm;
        enc = 2;
    }
    bw_init(&sp);
    if (nclusters <= 8) {
        const uint32_t bpe = ceil_log2(nclusters);
        bw_write(&sp, 1, 1);
        bw_write(&sp, 2, bpe);
        for (size_t i = 0; i < n; ++i) {
            bw_write(&sp, bpe, map[i]);
        }
        if (sp.nbits <= sel->nbits) {
            sel = &sp;
            enc = 0;
        }
    }
    bw_append(w, sel);
    bw_free(&cx);
    bw_free(&cm);
    bw_free(&sp);
    return enc;
}

size_t jo_best_context_map(const uint8_t* map, uint32_t n, uint32_t nclusters, uint8_t* out,
                           size_t cap, size_t* nbits, int* encoding) {
    bw_t w;
    bw_init(&w);
    *encoding = write_best_cmap(&w, map, n, nclusters);
    *nbits = w.nbits;
    const size_t r = bw_emit(&w, out, cap);
    bw_free(&w);
    return r;
}

/* histogram_writer.cc:85-126 */
static void write_clustered_histograms(bw_t* w, const uint8_t* cmap, size_t num_contexts,
                                       size_t nclusters, const uint32_t* hists, size_t stride,
                                       uint8_t* depth, uint16_t* bits) {
    bw_write(w, 1, 0);
    if (num_contexts > 1) {
        write_best_cmap(w, cmap, num_contexts, nclusters);
    }
    bw_write(w, 1, 1);
    for (size_t c = 0; c < nclusters; ++c) {
        encode_uint_config(w, 4, 2, 0);
    }
    size_t alpha[256];
    for (size_t c = 0; c < nclusters; ++c) {
        size_t a = 1;
        for (size_t s = 0; s < stride; ++s) {
            if (hists[c * stride + s]) {
                a = s + 1;
            }
        }
        alpha[c] = a;
        store_var_len_u16(w, a - 1);
    }
    memset(depth, 0, nclusters * stride);
    memset(bits, 0, nclusters * stride * sizeof(uint16_t));
    for (size_t c = 0; c < nclusters; ++c) {
        if (alpha[c] > 1) {
            build_and_store_huffman(hists + c * stride, alpha[c], depth + c * stride,
                                    bits + c * stride, w);
        }
    }
}

/* ======================================================================== */
/* VarDCT frame sections (vardct_frame.cc, frame_assembly.cc, entropy.cu)    */
/* ======================================================================== */

enum { NUM_AC_CONTEXTS = 15 * (37 + 458), AC_STRIDE = 256, NUM_CLUSTERS = 32 };

static const uint8_t CTX_MAP[39] = {0, 1, 2, 2, 3,  3,  4,  5,  6,  6,  6,  6,  6,
                                    7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14,
                                    7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14};
static const uint16_t FREQ_CTX[64] = {
    0,  0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 15, 16, 16, 17, 17,
    18, 18, 19, 19, 20, 20, 21, 21, 22, 22, 23, 23, 23, 23, 24, 24, 24, 24, 25, 25, 25, 25,
    26, 26, 26, 26, 27, 27, 27, 27, 28, 28, 28, 28, 29, 29, 29, 29, 30, 30, 30, 30};
static const uint16_t NZ_CTX[64] = {
    0,   0,   31,  62,  62,  93,  93,  93,  93,  123, 123, 123, 123, 152, 152, 152,
    152, 152, 152, 152, 152, 180, 180, 180, 180, 180, 180, 180, 180, 180, 180, 180,
    180, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
    206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206};

static inline int strategy_order(int n) {
    return n == 16 ? 2 : (n == 32 ? 3 : 0);
}
/* ac_context.h:54-57 */
static inline int block_context(int c, int ord) {
    return CTX_MAP[(c < 2 ? (c ^ 1) : 2) * 13 + ord];
}
/* ac_context.h:61-69 */
static inline uint32_t nonzero_context(uint32_t pred, int bctx) {
    uint32_t nz = pred >= 64 ? 64 : pred;
    const uint32_t ctx = nz < 8 ? nz : 4 + nz / 2;
    return ctx * 15 + (uint32_t)bctx;
}
/* ac_context.h:78-84 */
static inline uint32_t zero_density_context(uint32_t left, uint32_t k, uint32_t cb, uint32_t l2,
                                            uint32_t prev) {
    left = (left + cb - 1) >> l2;
    k >>= l2;
    return (NZ_CTX[left] + FREQ_CTX[k]) * 2 + prev;
}
/* ac_context.h:136-145 */
static inline uint32_t predict_nonzeros(int has_left, int has_top, int32_t top, int32_t left) {
    if (!has_left) {
        return has_top ? (uint32_t)top : 32u;
    }
    if (!has_top) {
        return (uint32_t)left;
    }
    return (uint32_t)((top + left + 1) / 2);
}

static const uint32_t* order_for(int n) {
    static uint32_t o8[64], o16[256], o32[1024];
    static int init = 0;
    if (!init) {
        jo_natural_order(8, o8);
        jo_natural_order(16, o16);
        jo_natural_order(32, o32);
        init = 1;
    }
    return n == 32 ? o32 : (n == 16 ? o16 : o8);
}

static inline int side_of(const jo_frame* f, size_t blk) {
    return f->acs ? f->acs[blk] : 8;
}

typedef struct actok {
    uint8_t cluster;
    uint8_t nbits;
    uint16_t symbol;
    uint32_t bits;
} actok;

typedef struct toklist {
    actok* t;
    size_t n, cap;
} toklist;

static void tl_push(toklist* l, actok t) {
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 4096;
        l->t = (actok*)realloc(l->t, l->cap * sizeof(actok));
    }
    l->t[l->n++] = t;
}

/* vardct_frame.cc:268-366 tokenize_ac_group_clustered (group-local grid). */
static void tokenize_ac_group(const jo_frame* f, uint32_t bx0, uint32_t by0, uint32_t gbw,
                              uint32_t gbh, uint32_t* hist, toklist* out) {
    const uint32_t bw = f->width / 8, bh = f->height / 8;
    const size_t plane = (size_t)bw * bh * 64;
    static const int corder[3] = {1, 0, 2};
    int32_t* nzn = (int32_t*)calloc((size_t)3 * gbw * gbh, sizeof(int32_t));
    for (uint32_t sy = 0; sy < gbh; ++sy) {
        for (uint32_t sx = 0; sx < gbw; ++sx) {
            const int n = side_of(f, (size_t)(by0 + sy) * bw + bx0 + sx);
            if (n == 0) {
                continue;
            }
            const uint32_t cx = (uint32_t)n / 8, cb = cx * cx, sz = (uint32_t)(n * n);
            const uint32_t l2 = n == 16 ? 2 : (n == 32 ? 4 : 0);
            const uint32_t* ord = order_for(n);
            for (int c = 0; c < 3; ++c) {
                int32_t nz = 0;
                for (uint32_t k = cb; k < sz; ++k) {
                    nz += f->ac[c * plane + covered_slot(n, bx0 + sx, by0 + sy, bw, ord[k])] != 0;
                }
                const int32_t norm = (int32_t)(((uint32_t)nz + cb - 1) >> l2);
                for (uint32_t dy = 0; dy < cx; ++dy) {
                    for (uint32_t dx = 0; dx < cx; ++dx) {
                        nzn[(size_t)c * gbw * gbh + (sy + dy) * gbw + sx + dx] = norm;
                    }
                }
            }
        }
    }
    (void)bh;
    for (uint32_t sy = 0; sy < gbh; ++sy) {
        for (uint32_t sx = 0; sx < gbw; ++sx) {
            const uint32_t gx = bx0 + sx, gy = by0 + sy;
            const int n = side_of(f, (size_t)gy * bw + gx);
            if (n == 0) {
                continue;
            }
            const uint32_t cb = (uint32_t)(n / 8) * (uint32_t)(n / 8), sz = (uint32_t)(n * n);
            const uint32_t l2 = n == 16 ? 2 : (n == 32 ? 4 : 0);
            const uint32_t* ord = order_for(n);
            for (int ci = 0; ci < 3; ++ci) {
                const int c = corder[ci];
                const int32_t* pl = f->ac + c * plane;
                uint32_t nz = 0;
                for (uint32_t k = cb; k < sz; ++k) {
                    nz += pl[covered_slot(n, gx, gy, bw, ord[k])] != 0;
                }
                const int bctx = block_context(c, strategy_order(n));
                const int32_t* g = nzn + (size_t)c * gbw * gbh;
                const int hl = sx > 0, ht = sy > 0;
                const uint32_t pred = predict_nonzeros(hl, ht, ht ? g[(sy - 1) * gbw + sx] : 0,
                                                       hl ? g[sy * gbw + sx - 1] : 0);
                actok t;
                uint32_t sym, nb, bits;
                uint32_t ctx = nonzero_context(pred, bctx);
                hybrid420(nz, &sym, &nb, &bits);
                t.cluster = (uint8_t)jo_ac_cluster(ctx);
                t.symbol = (uint16_t)sym;
                t.nbits = (uint8_t)nb;
                t.bits = bits;
                ++hist[t.cluster * AC_STRIDE + sym];
                if (out) {
                    tl_push(out, t);
                }
                const uint32_t off = 15u * 37u + 458u * (uint32_t)bctx;
                uint32_t rem = nz, prev = nz > sz / 16 ? 0u : 1u;
                for (uint32_t k = cb; k < sz && rem > 0; ++k) {
                    const int32_t v = pl[covered_slot(n, gx, gy, bw, ord[k])];
                    ctx = off + zero_density_context(rem, k, cb, l2, prev);
                    hybrid420(pack_signed(v), &sym, &nb, &bits);
                    t.cluster = (uint8_t)jo_ac_cluster(ctx);
                    t.symbol = (uint16_t)sym;
                    t.nbits = (uint8_t)nb;
                    t.bits = bits;
                    ++hist[t.cluster * AC_STRIDE + sym];
                    if (out) {
                        tl_push(out, t);
                    }
                    prev = v != 0;
                    rem -= prev;
                }
            }
        }
    }
    free(nzn);
}

static void emit_tokens(bw_t* w, const actok* t, size_t n, const uint8_t* depth,
                        const uint16_t* bits) {
    for (size_t i = 0; i < n; ++i) {
        const size_t idx = (size_t)t[i].cluster * AC_STRIDE + t[i].symbol;
        bw_write(w, depth[idx], bits[idx]);
        bw_write(w, t[i].nbits, t[i].bits);
    }
}

static void ac_context_map(uint8_t* map) {
    for (uint32_t i = 0; i < NUM_AC_CONTEXTS; ++i) {
        map[i] = (uint8_t)jo_ac_cluster(i);
    }
}

void jo_ac_histogram(const jo_frame* f, uint32_t* hist) {
    const uint32_t bw = f->width / 8, bh = f->height / 8;
    const uint32_t xg = (bw + 31) / 32, yg = (bh + 31) / 32;
    memset(hist, 0, sizeof(uint32_t) * NUM_CLUSTERS * AC_STRIDE);
    for (uint32_t g = 0; g < xg * yg; ++g) {
        const uint32_t bx0 = (g % xg) * 32, by0 = (g / xg) * 32;
        tokenize_ac_group(f, bx0, by0, bw - bx0 < 32 ? bw - bx0 : 32, bh - by0 < 32 ? bh - by0 : 32,
                          hist, NULL);
    }
}

/* Per AC group: token stream with the given codes, zero-padded to a byte. */
size_t jo_ac_groups(const jo_frame* f, const uint8_t* depth, const uint16_t* bits,
                    uint32_t* group_sizes, uint8_t* out, size_t cap) {
    const uint32_t bw = f->width / 8, bh = f->height / 8;
    const uint32_t xg = (bw + 31) / 32, yg = (bh + 31) / 32;
    uint32_t* scratch = (uint32_t*)calloc(NUM_CLUSTERS * AC_STRIDE, sizeof(uint32_t));
    size_t total = 0;
    for (uint32_t g = 0; g < xg * yg; ++g) {
        const uint32_t bx0 = (g % xg) * 32, by0 = (g / xg) * 32;
        toklist tl = {NULL, 0, 0};
        tokenize_ac_group(f, bx0, by0, bw - bx0 < 32 ? bw - bx0 : 32, bh - by0 < 32 ? bh - by0 : 32,
                          scratch, &tl);
        bw_t w;
        bw_init(&w);
        emit_tokens(&w, tl.t, tl.n, depth, bits);
        bw_pad(&w);
        const size_t nb = bw_bytes(&w);
        group_sizes[g] = (uint32_t)nb;
        if (out && total + nb <= cap) {
            memcpy(out + total, w.buf, nb);
        }
        total += nb;
        bw_free(&w);
        free(tl.t);
    }
    free(scratch);
    return total;
}

/* dc_predict.h:22-49 */
static int32_t gradient_predict(int hl, int ht, int32_t wv, int32_t nv, int32_t nwv) {
    const int32_t left = hl ? wv : (ht ? nv : 0);
    const int32_t top = ht ? nv : left;
    const int32_t tl = (hl && ht) ? nwv : left;
    const int32_t lo = left < top ? left : top, hi = left < top ? top : left;
    if (tl < lo) {
        return hi;
    }
    if (tl > hi) {
        return lo;
    }
    return (int32_t)((int64_t)left + top - tl);
}

typedef struct tok {
    uint32_t sym, nb, bits;
} tok;

typedef struct dcgroup {
    uint32_t bx0, by0, dgw, dgh;
    tok* dct;
    size_t ndc;
    tok* mt;
    size_t nmt;
    uint32_t count;
    uint32_t dc_hist[AC_STRIDE];
    uint32_t mt_hist[AC_STRIDE];
} dcgroup;

static void tok_push(tok** t, size_t* n, uint32_t v, uint32_t* hist) {
    tok x;
    hybrid420(v, &x.sym, &x.nb, &x.bits);
    (*t)[(*n)++] = x;
    ++hist[x.sym];
}

static int32_t raw_strategy(int n) {
    return n == 16 ? 4 : (n == 32 ? 5 : 0);
}

/* vardct_frame.cc:233-319 + entropy.cu:445-523: DC channels (physical Y, X,
 * B, gradient predictor) and AcMetadata channels (YtoX, YtoB, ACS+QF, EPF,
 * zero predictor) of one DC group, with their token histograms. */
static void dc_group_tokens(const jo_frame* f, uint32_t g, dcgroup* d) {
    const uint32_t bw = f->width / 8, bh = f->height / 8;
    const uint32_t xdg = (bw + 255) / 256, cmw = (bw + 7) / 8;
    const size_t nblk = (size_t)bw * bh;
    d->bx0 = (g % xdg) * 256;
    d->by0 = (g / xdg) * 256;
    d->dgw = bw - d->bx0 < 256 ? bw - d->bx0 : 256;
    d->dgh = bh - d->by0 < 256 ? bh - d->by0 : 256;
    const size_t n = (size_t)d->dgw * d->dgh;
    const uint32_t cw = (d->dgw + 7) / 8, ch = (d->dgh + 7) / 8;
    memset(d->dc_hist, 0, sizeof(d->dc_hist));
    memset(d->mt_hist, 0, sizeof(d->mt_hist));
    d->dct = (tok*)malloc(sizeof(tok) * 3 * n);
    d->ndc = 0;
    static const int corder[3] = {1, 0, 2};
    for (int ci = 0; ci < 3; ++ci) {
        const int32_t* p = f->dc + corder[ci] * nblk;
        for (uint32_t y = 0; y < d->dgh; ++y) {
            for (uint32_t x = 0; x < d->dgw; ++x) {
                const size_t b = (size_t)(d->by0 + y) * bw + d->bx0 + x;
                const int hl = x > 0, ht = y > 0;
                const int32_t pred = gradient_predict(hl, ht, hl ? p[b - 1] : 0, ht ? p[b - bw] : 0,
                                                      (hl && ht) ? p[b - bw - 1] : 0);
                tok_push(&d->dct, &d->ndc, pack_signed(p[b] - pred), d->dc_hist);
            }
        }
    }
    d->mt = (tok*)malloc(sizeof(tok) * (2 * (size_t)cw * ch + 3 * n));
    d->nmt = 0;
    for (int map = 0; map < 2; ++map) {
        const int8_t* src = map == 0 ? f->ytox : f->ytob;
        for (uint32_t y = 0; y < ch; ++y) {
            for (uint32_t x = 0; x < cw; ++x) {
                const int32_t v = src ? src[(size_t)(d->by0 / 8 + y) * cmw + d->bx0 / 8 + x] : 0;
                tok_push(&d->mt, &d->nmt, pack_signed(v), d->mt_hist);
            }
        }
    }
    size_t* fb = (size_t*)malloc(sizeof(size_t) * n);
    uint32_t count = 0;
    for (uint32_t y = 0; y < d->dgh; ++y) {
        for (uint32_t x = 0; x < d->dgw; ++x) {
            const size_t b = (size_t)(d->by0 + y) * bw + d->bx0 + x;
            if (side_of(f, b) != 0) {
                fb[count++] = b;
            }
        }
    }
    d->count = count;
    for (uint32_t i = 0; i < count; ++i) {
        tok_push(&d->mt, &d->nmt, pack_signed(raw_strategy(side_of(f, fb[i]))), d->mt_hist);
    }
    for (uint32_t i = 0; i < count; ++i) {
        tok_push(&d->mt, &d->nmt, pack_signed(f->qf[fb[i]] - 1), d->mt_hist);
    }
    for (size_t i = 0; i < n; ++i) {
        tok_push(&d->mt, &d->nmt, 0, d->mt_hist);
    }
    free(fb);
}

static size_t trimmed_alpha(const uint32_t* h) {
    size_t a = 0;
    for (size_t s = 0; s < AC_STRIDE; ++s) {
        if (h[s]) {
            a = s + 1;
        }
    }
    return a == 0 ? 1 : a;
}

/* modular.cc:79-129 / frame_assembly.cc:70-101: GroupHeader, single-leaf tree
 * with `predictor`, then the data histogram description. */
static void write_modular_header(bw_t* w, const uint32_t* data_hist, int predictor,
                                 uint8_t* depth, uint16_t* bits) {
    bw_write(w, 1, 0);
    bw_write(w, 1, 1);
    bw_write(w, 2, 0);
    const uint32_t tree_tokens[5] = {0, (uint32_t)predictor, 0, 0, 0};
    uint32_t th[16] = {0};
    size_t tl = 1;
    for (int i = 0; i < 5; ++i) {
        ++th[tree_tokens[i]];
        tl = tree_tokens[i] + 1 > tl ? tree_tokens[i] + 1 : tl;
    }
    uint8_t td[16];
    uint16_t tb[16];
    write_prefix_histograms(w, th, tl, 6, 4, 2, 0, td, tb);
    for (int i = 0; i < 5; ++i) {
        bw_write(w, td[tree_tokens[i]], tb[tree_tokens[i]]);
    }
    const size_t alpha = trimmed_alpha(data_hist);
    static const uint32_t one = 1;
    int empty = 1;
    for (size_t s = 0; s < AC_STRIDE; ++s) {
        if (data_hist[s]) {
            empty = 0;
        }
    }
    write_prefix_histograms(w, empty ? &one : data_hist, alpha, 1, 4, 2, 0, depth, bits);
}

static void write_toks(bw_t* w, const tok* t, size_t n, const uint8_t* depth,
                       const uint16_t* bits) {
    for (size_t i = 0; i < n; ++i) {
        bw_write(w, depth[t[i].sym], bits[t[i].sym]);
        bw_write(w, t[i].nb, t[i].bits);
    }
}

/* vardct_frame.cc:543-554: one DcGroup section body (not padded). */
static void write_dc_group(bw_t* w, const jo_frame* f, uint32_t g) {
    dcgroup d;
    dc_group_tokens(f, g, &d);
    uint8_t depth[AC_STRIDE];
    uint16_t bits[AC_STRIDE];
    bw_write(w, 2, 0);
    write_modular_header(w, d.dc_hist, 5, depth, bits);
    write_toks(w, d.dct, d.ndc, depth, bits);
    bw_write(w, ceil_log2((size_t)d.dgw * d.dgh), d.count - 1);
    write_modular_header(w, d.mt_hist, 0, depth, bits);
    write_toks(w, d.mt, d.nmt, depth, bits);
    free(d.dct);
    free(d.mt);
}

void jo_dc_histograms(const jo_frame* f, uint32_t* dc_hist, uint32_t* meta_hist) {
    const uint32_t bw = f->width / 8, bh = f->height / 8;
    const uint32_t ng = ((bw + 255) / 256) * ((bh + 255) / 256);
    for (uint32_t g = 0; g < ng; ++g) {
        dcgroup d;
        dc_group_tokens(f, g, &d);
        memcpy(dc_hist + (size_t)g * AC_STRIDE, d.dc_hist, sizeof(d.dc_hist));
        memcpy(meta_hist + (size_t)g * AC_STRIDE, d.mt_hist, sizeof(d.mt_hist));
        free(d.dct);
        free(d.mt);
    }
}

size_t jo_dc_groups(const jo_frame* f, uint32_t* group_sizes, uint8_t* out, size_t cap) {
    const uint32_t bw = f->width / 8, bh = f->height / 8;
    const uint32_t ng = ((bw + 255) / 256) * ((bh + 255) / 256);
    size_t total = 0;
    for (uint32_t g = 0; g < ng; ++g) {
        bw_t w;
        bw_init(&w);
        write_dc_group(&w, f, g);
        bw_pad(&w);
        const size_t nb = bw_bytes(&w);
        group_sizes[g] = (uint32_t)nb;
        if (out && total + nb <= cap) {
            memcpy(out + total, w.buf, nb);
        }
        total += nb;
        bw_free(&w);
    }
    return total;
}

static const distr GLOBAL_SCALE_ENC[4] = {BITSOFF(11, 1), BITSOFF(11, 2049), BITSOFF(12, 4097),
                                          BITSOFF(16, 8193)};
static const distr QUANT_DC_ENC[4] = {VAL(16), BITSOFF(5, 1), BITSOFF(8, 1), BITSOFF(16, 1)};
static const distr ORDER_ENC[4] = {VAL(0x5F), VAL(0x13), VAL(0), BITSOFF(13, 0)};
static const distr DIM_ENC[4] = {BITSOFF(9, 1), BITSOFF(13, 1), BITSOFF(18, 1), BITSOFF(30, 1)};
static const distr TOC_DIST[4] = {BITSOFF(10, 0), BITSOFF(14, 1024), BITSOFF(22, 17408),
                                  BITSOFF(30, 4211712)};

/* vardct_frame.cc:533-540 (not padded) */
static void write_dc_global_bits(bw_t* w, uint32_t gs, uint32_t qdc) {
    bw_write(w, 1, 1);
    write_u32(w, GLOBAL_SCALE_ENC, gs);
    write_u32(w, QUANT_DC_ENC, qdc);
    bw_write(w, 1, 1);
    bw_write(w, 1, 1);
    bw_write(w, 1, 0);
}

size_t jo_dc_global(uint32_t gs, uint32_t qdc, uint8_t* out, size_t cap) {
    bw_t w;
    bw_init(&w);
    write_dc_global_bits(&w, gs, qdc);
    bw_pad(&w);
    const size_t n = bw_emit(&w, out, cap);
    bw_free(&w);
    return n;
}

/* frame_assembly.cc:131-146 (multi-group AcGlobal) */
size_t jo_ac_global(const uint32_t* hist, uint32_t num_ac_groups, uint8_t* depth,
                    uint16_t* bits, uint8_t* out, size_t cap) {
    uint8_t* cmap = (uint8_t*)malloc(NUM_AC_CONTEXTS);
    ac_context_map(cmap);
    bw_t w;
    bw_init(&w);
    bw_write(&w, 1, 1);
    bw_write(&w, ceil_log2(num_ac_groups), 0);
    write_u32(&w, ORDER_ENC, 0);
    write_clustered_histograms(&w, cmap, NUM_AC_CONTEXTS, NUM_CLUSTERS, hist, AC_STRIDE, depth,
                               bits);
    bw_pad(&w);
    const size_t n = bw_emit(&w, out, cap);
    bw_free(&w);
    free(cmap);
    return n;
}

/* codestream.cc:39-65 */
static void write_headers(bw_t* w, uint32_t xs, uint32_t ys) {
    bw_write(w, 8, 0xff);
    bw_write(w, 8, 0x0a);
    const int small = xs % 8 == 0 && ys % 8 == 0 && xs <= 256 && ys <= 256;
    bw_write(w, 1, small);
    if (small) {
        bw_write(w, 5, ys / 8 - 1);
    } else {
        write_u32(w, DIM_ENC, ys);
    }
    bw_write(w, 3, 0);
    if (small) {
        bw_write(w, 5, xs / 8 - 1);
    } else {
        write_u32(w, DIM_ENC, xs);
    }
    bw_write(w, 1, 1);
    bw_write(w, 1, 1);
    bw_pad(w);
}

/* codestream.cc:67-85 */
static void write_frame_header(bw_t* w) {
    static const distr FT[4] = {VAL(0), VAL(1), VAL(2), VAL(3)};
    static const distr UPS[4] = {VAL(1), VAL(2), VAL(4), VAL(8)};
    static const distr PASSES[4] = {VAL(1), VAL(2), VAL(3), BITSOFF(3, 4)};
    static const distr BLEND[4] = {VAL(0), VAL(1), VAL(2), BITSOFF(2, 3)};
    static const distr NAME[4] = {VAL(0), BITSOFF(4, 0), BITSOFF(5, 16), BITSOFF(10, 48)};
    bw_write(w, 1, 0);
    write_u32(w, FT, 0);
    bw_write(w, 1, 0);
    write_u64(w, 0);
    write_u32(w, UPS, 1);
    bw_write(w, 3, 2);
    bw_write(w, 3, 2);
    write_u32(w, PASSES, 1);
    bw_write(w, 1, 0);
    write_u32(w, BLEND, 0);
    bw_write(w, 1, 1);
    write_u32(w, NAME, 0);
    bw_write(w, 1, 1);
    write_u64(w, 0);
}

/* codestream.cc:87-98 */
static void write_toc(bw_t* w, const uint32_t* sizes, uint32_t n) {
    bw_write(w, 1, 0);
    bw_pad(w);
    for (uint32_t i = 0; i < n; ++i) {
        write_u32(w, TOC_DIST, sizes[i]);
    }
    bw_pad(w);
}

size_t jo_codestream_head(uint32_t w_, uint32_t h, const uint32_t* sizes, uint32_t n,
                          uint8_t* out, size_t cap) {
    bw_t w;
    bw_init(&w);
    write_headers(&w, w_, h);
    write_frame_header(&w);
    write_toc(&w, sizes, n);
    const size_t r = bw_emit(&w, out, cap);
    bw_free(&w);
    return r;
}

/* container.cc:8-54 */
size_t jo_container_framing(uint64_t cs, uint8_t* out, size_t cap) {
    static const uint8_t hdr[32] = {0,   0,   0,   0x0c, 'J', 'X', 'L', ' ', 0x0d, 0x0a, 0x87,
                                    0x0a, 0,   0,   0,    0x14, 'f', 't', 'y', 'p',  'j',  'x',
                                    'l', ' ', 0,   0,    0,    0,   'j', 'x', 'l',  ' '};
    uint8_t buf[48];
    memcpy(buf, hdr, 32);
    const int large = cs >= 0x100000000ull - 8;
    const size_t hs = large ? 16 : 8;
    const uint64_t box = cs + hs;
    const uint32_t field = large ? 1u : (uint32_t)box;
    for (int i = 0; i < 4; ++i) {
        buf[32 + i] = (uint8_t)(field >> (8 * (3 - i)));
    }
    memcpy(buf + 36, "jxlc", 4);
    if (large) {
        for (int i = 0; i < 8; ++i) {
            buf[40 + i] = (uint8_t)(box >> (8 * (7 - i)));
        }
    }
    if (out) {
        memcpy(out, buf, 32 + hs < cap ? 32 + hs : cap);
    }
    return 32 + hs;
}

/* vardct_frame.cc:414-449 (one AC group: one combined section) and
 * 451-547 (DcGlobal, DcGroups, AcGlobal, AcGroups) with the clustered AC
 * contexts and the per-block quant field. */
size_t jo_write_codestream(const jo_frame* f, uint8_t* out, size_t cap) {
    const uint32_t bw = f->width / 8, bh = f->height / 8;
    const uint32_t xg = (bw + 31) / 32, yg = (bh + 31) / 32, nac = xg * yg;
    const uint32_t ndc = ((bw + 255) / 256) * ((bh + 255) / 256);
    uint32_t* hist = (uint32_t*)calloc(NUM_CLUSTERS * AC_STRIDE, sizeof(uint32_t));
    uint8_t* depth = (uint8_t*)calloc(NUM_CLUSTERS * AC_STRIDE, 1);
    uint16_t* bits = (uint16_t*)calloc(NUM_CLUSTERS * AC_STRIDE, sizeof(uint16_t));
    uint8_t* cmap = (uint8_t*)malloc(NUM_AC_CONTEXTS);
    ac_context_map(cmap);
    jo_ac_histogram(f, hist);
    bw_t w;
    bw_init(&w);
    size_t total;
    if (nac == 1) {
        bw_t body;
        bw_init(&body);
        write_dc_global_bits(&body, f->global_scale, f->quant_dc);
        write_dc_group(&body, f, 0);
        bw_write(&body, 1, 1);
        write_u32(&body, ORDER_ENC, 0);
        write_clustered_histograms(&body, cmap, NUM_AC_CONTEXTS, NUM_CLUSTERS, hist, AC_STRIDE,
                                   depth, bits);
        toklist tl = {NULL, 0, 0};
        uint32_t* scratch = (uint32_t*)calloc(NUM_CLUSTERS * AC_STRIDE, sizeof(uint32_t));
        tokenize_ac_group(f, 0, 0, bw, bh, scratch, &tl);
        emit_tokens(&body, tl.t, tl.n, depth, bits);
        free(scratch);
        free(tl.t);
        bw_pad(&body);
        write_headers(&w, f->width, f->height);
        write_frame_header(&w);
        const uint32_t sz = (uint32_t)bw_bytes(&body);
        write_toc(&w, &sz, 1);
        bw_append(&w, &body);
        bw_free(&body);
        total = bw_emit(&w, out, cap);
    } else {
        const uint32_t nsec = 2 + ndc + nac;
        uint32_t* sizes = (uint32_t*)malloc(sizeof(uint32_t) * nsec);
        uint8_t dcg[16];
        sizes[0] = (uint32_t)jo_dc_global(f->global_scale, f->quant_dc, dcg, sizeof(dcg));
        size_t dcb_cap = jo_max_output_size(f->width, f->height);
        uint8_t* dcb = (uint8_t*)malloc(dcb_cap);
        const size_t dct = jo_dc_groups(f, sizes + 1, dcb, dcb_cap);
        uint8_t* acg = (uint8_t*)malloc(65536);
        sizes[1 + ndc] = (uint32_t)jo_ac_global(hist, nac, depth, bits, acg, 65536);
        uint8_t* acb = (uint8_t*)malloc(dcb_cap);
        const size_t act = jo_ac_groups(f, depth, bits, sizes + 2 + ndc, acb, dcb_cap);
        /* Headers under 1 KB plus at most 4 bytes per TOC entry. */
        const size_t head_cap = 1024 + 4 * (size_t)nsec;
        uint8_t* head = (uint8_t*)malloc(head_cap);
        const size_t hn = jo_codestream_head(f->width, f->height, sizes, nsec, head, head_cap);
        total = hn + sizes[0] + dct + sizes[1 + ndc] + act;
        if (out && total <= cap) {
            size_t at = 0;
            memcpy(out + at, head, hn);
            at += hn;
            memcpy(out + at, dcg, sizes[0]);
            at += sizes[0];
            memcpy(out + at, dcb, dct);
            at += dct;
            memcpy(out + at, acg, sizes[1 + ndc]);
            at += sizes[1 + ndc];
            memcpy(out + at, acb, act);
        }
        free(head);
        free(sizes);
        free(dcb);
        free(acg);
        free(acb);
    }
    bw_free(&w);
    free(hist);
    free(depth);
    free(bits);
    free(cmap);
    return total;
}

/* cujpegxl.cc:50-59 worst case bound. */
size_t jo_max_output_size(uint32_t w, uint32_t h) {
    const size_t bw = w / 8, bh = h / 8;
    const size_t nac = ((bw + 31) / 32) * ((bh + 31) / 32);
    const size_t ndc = ((bw + 255) / 256) * ((bh + 255) / 256);
    return 3 * bw * bh * 384 + 48 * bw * bh + (nac + ndc) * 2048 + 16384;
}

static size_t wrap_container(const uint8_t* cs, size_t n, uint8_t* out, size_t cap) {
    uint8_t fr[48];
    const size_t fn = jo_container_framing(n, fr, sizeof(fr));
    if (out && fn + n <= cap) {
        memcpy(out, fr, fn);
        memcpy(out + fn, cs, n);
    }
    return fn + n;
}

/* frame_encoder.cu:299-352 pipeline: XYB -> AQ field; gaborish -> selection
 * -> DCT (FP16 storage) -> CfL -> quantize -> entropy -> codestream. */
size_t jo_encode_nv12(const uint8_t* luma, size_t lp, const uint8_t* chroma, size_t cp,
                      uint32_t w, uint32_t h, float distance, uint8_t* out, size_t cap) {
    const uint32_t bw = w / 8, bh = h / 8;
    const size_t plane = (size_t)w * h, nblk = (size_t)bw * bh, cplane = nblk * 64;
    const uint32_t rw = (bw + 3) / 4, rh = (bh + 3) / 4;
    const size_t tiles = (size_t)((bw + 7) / 8) * ((bh + 7) / 8);
    float* xyb = (float*)malloc(sizeof(float) * 3 * plane);
    float* sharp = (float*)malloc(sizeof(float) * 3 * plane);
    int32_t* qf = (int32_t*)malloc(sizeof(int32_t) * nblk);
    int8_t* acs = (int8_t*)malloc(nblk);
    float* coef = (float*)malloc(sizeof(float) * 3 * cplane);
    uint16_t* c16 = (uint16_t*)malloc(sizeof(uint16_t) * 3 * cplane);
    float* part = (float*)malloc(sizeof(float) * 3 * rw * rh);
    int8_t* ytox = (int8_t*)malloc(tiles);
    int8_t* ytob = (int8_t*)malloc(tiles);
    int16_t* ac16 = (int16_t*)malloc(sizeof(int16_t) * 3 * cplane);
    int32_t* ac = (int32_t*)malloc(sizeof(int32_t) * 3 * cplane);
    int32_t* dc = (int32_t*)malloc(sizeof(int32_t) * 3 * nblk);

    jo_nv12_to_xyb(luma, lp, chroma, cp, w, h, xyb);
    jo_adaptive_quant(xyb, w, h, distance, qf);
    jo_gaborish(xyb, w, h, sharp);
    jo_select_transforms(sharp + plane, w, h, distance, acs);
    jo_forward_dct(sharp, w, h, acs, coef);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (size_t i = 0; i < 3 * cplane; ++i) {
        c16[i] = jo_f32_to_f16(coef[i]);
    }
    jo_cfl_region_partials(c16, acs, w, h, part);
    jo_cfl_maps(part, w, h, ytox, ytob);
    jo_quantize(c16, acs, ytox, ytob, qf, w, h, distance, ac16, dc, NULL);
    for (size_t i = 0; i < 3 * cplane; ++i) {
        ac[i] = ac16[i];
    }
    jo_cal cal;
    jo_calibrate(distance, &cal);
    const jo_frame f = {w, h, cal.global_scale, cal.quant_dc, dc, ac, acs, qf, ytox, ytob};
    const size_t bound = jo_max_output_size(w, h);
    uint8_t* cs = (uint8_t*)malloc(bound);
    const size_t n = jo_write_codestream(&f, cs, bound);
    const size_t total = wrap_container(cs, n, out, cap);
    free(cs);
    free(xyb);
    free(sharp);
    free(qf);
    free(acs);
    free(coef);
    free(c16);
    free(part);
    free(ytox);
    free(ytob);
    free(ac16);
    free(ac);
    free(dc);
    return total;
}

size_t jo_encode_dct8_uniform(const uint8_t* luma, size_t lp, const uint8_t* chroma, size_t cp,
                              uint32_t w, uint32_t h, float distance, uint8_t* out, size_t cap) {
    const uint32_t bw = w / 8, bh = h / 8;
    const size_t plane = (size_t)w * h, nblk = (size_t)bw * bh, cplane = nblk * 64;
    float* xyb = (float*)malloc(sizeof(float) * 3 * plane);
    float* sharp = (float*)malloc(sizeof(float) * 3 * plane);
    float* coef = (float*)malloc(sizeof(float) * 3 * cplane);
    int32_t* ac = (int32_t*)malloc(sizeof(int32_t) * 3 * cplane);
    int32_t* dc = (int32_t*)malloc(sizeof(int32_t) * 3 * nblk);
    int32_t* qf = (int32_t*)malloc(sizeof(int32_t) * nblk);
    jo_cal cal;
    jo_calibrate(distance, &cal);
    jo_nv12_to_xyb(luma, lp, chroma, cp, w, h, xyb);
    jo_gaborish(xyb, w, h, sharp);
    jo_forward_dct(sharp, w, h, NULL, coef);
    jo_quantize_dct8_uniform(coef, w, h, distance, ac, dc);
    for (size_t b = 0; b < nblk; ++b) {
        qf[b] = (int32_t)cal.raw_quant_field;
    }
    const jo_frame f = {w, h, cal.global_scale, cal.quant_dc, dc, ac, NULL, qf, NULL, NULL};
    const size_t bound = jo_max_output_size(w, h);
    uint8_t* cs = (uint8_t*)malloc(bound);
    const size_t n = jo_write_codestream(&f, cs, bound);
    const size_t total = wrap_container(cs, n, out, cap);
    free(cs);
    free(xyb);
    free(sharp);
    free(coef);
    free(ac);
    free(dc);
    free(qf);
    return total;
}
