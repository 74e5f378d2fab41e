// Front-end kernels (included into device.hip).
//
// region_kernel: one 256-thread workgroup per 32x32-pixel selection region.
// The region's NV12 pixels plus a 2-pixel clamped halo are converted to XYB
// once into LDS; everything downstream reads LDS: inverse gaborish, the
// adaptive-quantization field, the RD transform selection over the 21 DCT8/16/32
// candidates of the sharpened Y, the chosen DCTs of all three channels and the
// CfL regression partials.  Only the FP16 coefficients, the transform map,
// the quant field and three floats per region reach HBM.  The DC chain of the
// entropy stage starts after this kernel (with dc_maps_kernel: LLF -> DC of
// every block, and the 64x64 tiles' colour maps from the partials).
//
// tile_quantize_kernel: one workgroup per 64x64 CfL tile: the tile's colour
// maps (the same function), quantization of every coefficient with the CfL
// residual and the per-block non-zero counts the entropy coder predicts from.
//
// Element layout used throughout a region: thread (wave w, lane L) at
// iteration k owns element i = 256k + 64w + L of [0, 1024).  Element ids nest:
// a DCT32 owns [0,1024), the DCT16 on quad q owns [256q, 256q+256), the DCT8 on
// block j of quad q owns [256q + 64j, +64).  Position (quad k, block w) slot L
// holds raw coefficient e = 64 q + L of its first-block (q = covered ordinal).

namespace cjx {

constexpr int kHalo = 2;
constexpr int kTileW = 32 + 2 * kHalo;  // 36
// Chroma samples a tile's bilinear taps can reach: (32 + 2 halo) / 2 + 1 on
// each side.
constexpr int kChromaW = kTileW / 2 + 2;  // 20

// Grid-strided zeroing of the Clears buffers by every thread of a launch
// (of the frame's workgroups, in a batched launch).
__device__ __forceinline__ void clear_words(const Clears& c, const Wg& wg) {
    const std::uint32_t stride{wg.nx * wg.ny * blockDim.x};
    const std::uint32_t t{(wg.y * wg.nx + wg.x) * blockDim.x + threadIdx.x};
#pragma unroll
    for (int k{0}; k < 4; ++k) {
        for (std::uint32_t i{t}; i < c.words[k]; i += stride) {
            c.ptr[k][i] = 0;
        }
    }
}

// Cross-lane sums without the LDS crossbar.  add_down_row<n>: lane i adds
// lane i + n of its row of 16 lanes (a DPP row shift; n = 1, 2, 4, 8);
// add_down_16: lanes of rows 0 and 2 add rows 1 and 3 (v_permlane16_swap);
// add_down_32: lanes 0-31 add lanes 32-63 (v_permlane32_swap).  Callers read
// only lanes whose partners lie in the same row (or the upper row / half),
// where the result equals a shfl_down (or shfl_xor) tree's term for term.
template <int kN>
__device__ __forceinline__ float add_down_row(float v) {
    static_assert(kN == 1 || kN == 2 || kN == 4 || kN == 8, "row shift");
    return v + u2f(static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(f2u(v)), 0x100 + kN,
                                                                          0xf, 0xf, true)));
}

__device__ __forceinline__ float add_down_16(float v) {
    const auto r{__builtin_amdgcn_permlane16_swap(f2u(v), f2u(v), false, false)};
    return v + u2f(r[1]);
}

__device__ __forceinline__ float add_down_32(float v) {
    const auto r{__builtin_amdgcn_permlane32_swap(f2u(v), f2u(v), false, false)};
    return v + u2f(r[1]);
}

// Lane i - 1 of the same row of 16 lanes (DPP row shift right).
__device__ __forceinline__ float from_lane_below(float v) {
    return u2f(static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(f2u(v)), 0x111, 0xf, 0xf,
                                                                      true)));
}

// Four wave sums at once, each the shfl_down tree of its 64 values (offsets
// 32, 16, ..., 1; lane i adds lane i + offset) with the same operands in the
// same order: one permlane32 swap and add carry out offset 32 for two
// values (a in lanes 0-31, b in lanes 32-63; c and d likewise), one
// permlane16 swap and add offset 16 for all four (rows: a, c, b, d), then
// the row shifts 8..1 for all four.  The result holds sum(a) in lane 0,
// sum(c) in lane 16, sum(b) in lane 32 and sum(d) in lane 48.
__device__ __forceinline__ float wave_sum4(float a, float b, float c, float d) {
    const auto ab{__builtin_amdgcn_permlane32_swap(f2u(a), f2u(b), false, false)};
    const float x{u2f(ab[0]) + u2f(ab[1])};
    const auto cd{__builtin_amdgcn_permlane32_swap(f2u(c), f2u(d), false, false)};
    const float y{u2f(cd[0]) + u2f(cd[1])};
    const auto xy{__builtin_amdgcn_permlane16_swap(f2u(x), f2u(y), false, false)};
    float v{u2f(xy[0]) + u2f(xy[1])};
    v = add_down_row<8>(v);
    v = add_down_row<4>(v);
    v = add_down_row<2>(v);
    return add_down_row<1>(v);
}

// Lane l's value of v (a wave-uniform read).
__device__ __forceinline__ float lane_value(float v, int l) {
    return u2f(static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(f2u(v)), l)));
}

__device__ __forceinline__ float clamp01(float v) {
    return fminf(fmaxf(v, 0.0f), 1.0f);
}

// xyb_impl.cuh:37-70 on the shared numeric primitives (the powers from the
// table at `tab`, staged in LDS).
__device__ __forceinline__ float srgb_to_linear(float e, const float2* tab) {
    return e <= 0.04045f ? e * (1.0f / 12.92f) : pow24_tab((e + 0.055f) * (1.0f / 1.055f), tab);
}

// Opsin absorbance bias of the XYB transform.
constexpr float kOpsinBias{0.0037930732552754493f};

// `neg_cbrt_bias` is -cbrt_tab(kOpsinBias, tab), evaluated once by the caller.
__device__ __forceinline__ void pixel_to_xyb(float yv, float cb, float cr, float& ox, float& oy, float& ob,
                                             const float2* tab, float neg_cbrt_bias) {
    constexpr float m00{0.30f}, m02{0.078f}, m01{1.0f - 0.078f - 0.30f};
    constexpr float m10{0.23f}, m12{0.078f}, m11{1.0f - 0.078f - 0.23f};
    constexpr float m20{0.24342268924547819f}, m21{0.20476744424496821f};
    constexpr float m22{1.0f - 0.24342268924547819f - 0.20476744424496821f};
    constexpr float bias{kOpsinBias};
    const float r{clamp01(srgb_to_linear(clamp01(yv + 1.5748f * cr), tab))};
    const float g{clamp01(srgb_to_linear(clamp01(yv - 0.1873242729f * cb - 0.4681242729f * cr), tab))};
    const float b{clamp01(srgb_to_linear(clamp01(yv + 1.8556f * cb), tab))};
    float q0{fmaxf(m00 * r + m01 * g + m02 * b + bias, 0.0f)};
    float q1{fmaxf(m10 * r + m11 * g + m12 * b + bias, 0.0f)};
    float q2{fmaxf(m20 * r + m21 * g + m22 * b + bias, 0.0f)};
    q0 = cbrt_tab(q0, tab) + neg_cbrt_bias;
    q1 = cbrt_tab(q1, tab) + neg_cbrt_bias;
    q2 = cbrt_tab(q2, tab) + neg_cbrt_bias;
    ox = 0.5f * (q0 - q1);
    oy = 0.5f * (q0 + q1);
    ob = q2;
}

// a / b and sqrt(x), correctly rounded, for normal operands of moderate
// size with normal results (the AQ terms: 0.01 <= a, b <= 2e4, 27 <= x <=
// 5e5): the compiler's correctly rounded sequences (a reciprocal, one
// Newton step and two Markstein corrections; a square root and its
// one-ulp neighbour tests) without their operand scaling and special-case
// steps, which are identities for such operands.  check_aq_math_kernel
// compares both against the compiler's forms over every reachable input.
__device__ __forceinline__ float div_moderate(float a, float b) {
    float r{__builtin_amdgcn_rcpf(b)};
    r = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
    float q{a * r};
    q = __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
    return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
}

__device__ __forceinline__ float sqrt_moderate(float x) {
    const float s{__builtin_amdgcn_sqrtf(x)};
    const float below{__builtin_bit_cast(float, __builtin_bit_cast(std::int32_t, s) - 1)};
    const float above{__builtin_bit_cast(float, __builtin_bit_cast(std::int32_t, s) + 1)};
    const float r{__builtin_fmaf(-below, s, x) <= 0.0f ? below : s};
    return __builtin_fmaf(-above, s, x) > 0.0f ? above : r;
}

// adaptive_quant_impl.cuh:29-61
template <bool kModerate = true>
__device__ __forceinline__ float ratio_cbrt_gamma(float v, bool invert) {
    constexpr float k_sg_mul{226.77216153508914f};
    constexpr float k_sg_log2{0.693147181f};
    constexpr float k_sg_mul2{1.0f / 73.377132366608819f};
    constexpr float k_sg_ret_mul{k_sg_mul2 * 18.6580932135f * k_sg_log2};
    constexpr float eps{1e-2f};
    constexpr float num_mul{k_sg_ret_mul * 3.0f * k_sg_mul};
    constexpr float v_offset{7.7825991679894591f * k_sg_log2 + eps};
    constexpr float den_mul{k_sg_log2 * k_sg_mul};
    v = v < 0.0f ? 0.0f : v;
    const float v2{v * v};
    const float num{num_mul * v2 + eps};
    const float den{den_mul * v * v2 + v_offset};
    if constexpr (kModerate) {
        return invert ? div_moderate(num, den) : div_moderate(den, num);
    } else {
        return invert ? num / den : den / num;
    }
}

template <bool kModerate = true>
__device__ __forceinline__ float masking_sqrt(float v) {
    constexpr float mul{211.66567973503678f * 1e8f};
    const float x{v * __builtin_sqrtf(mul) + 27.505837037000106f};
    if constexpr (kModerate) {
        return 0.25f * sqrt_moderate(x);
    } else {
        return 0.25f * __builtin_sqrtf(x);
    }
}

// The AQ terms' division and square root in both forms over every float
// input of their domains (ratio_cbrt_gamma: v in [0, 4), both directions;
// masking_sqrt: v in [0, 0.25)): mismatching bit patterns are counted.
__global__ void check_aq_math_kernel(std::uint32_t first, std::uint32_t count, std::uint32_t sqrt_limit,
                                     unsigned long long* mismatches) {
    unsigned int bad{0};
    for (std::uint32_t i{blockIdx.x * blockDim.x + threadIdx.x}; i < count; i += gridDim.x * blockDim.x) {
        const std::uint32_t bits{first + i};
        const float v{__builtin_bit_cast(float, bits)};
        bad += __builtin_bit_cast(std::uint32_t, ratio_cbrt_gamma<true>(v, true)) !=
               __builtin_bit_cast(std::uint32_t, ratio_cbrt_gamma<false>(v, true));
        bad += __builtin_bit_cast(std::uint32_t, ratio_cbrt_gamma<true>(v, false)) !=
               __builtin_bit_cast(std::uint32_t, ratio_cbrt_gamma<false>(v, false));
        if (bits < sqrt_limit) {
            bad += __builtin_bit_cast(std::uint32_t, masking_sqrt<true>(v)) !=
                   __builtin_bit_cast(std::uint32_t, masking_sqrt<false>(v));
        }
    }
    if (bad != 0) {
        atomicAdd(mismatches, static_cast<unsigned long long>(bad));
    }
}

__device__ __forceinline__ float compute_mask(float out_val) {
    const float v1{fmaxf(out_val * 0.80061762862741759f, 1e-3f)};
    const float v2{1.0f / (v1 + 302.59587815579727f)};
    const float v3{1.0f / (v1 * v1 + 3.7179635626140772f)};
    const float v4{1.0f / (v1 * v1 + 0.25f * 3.7179635626140772f)};
    return -0.7647f + 9.4708735624378946f * v4 + 17.35036561631863f * v2 + 6.7943250517376494f * v3;
}

__device__ __forceinline__ void cswap(float& a, float& b) {
    if (a > b) {
        const float t{a};
        a = b;
        b = t;
    }
}

// frontend_fused.cu:35-52 with adaptive_quant_impl.cuh:63-121: weighted sum
// of the four smallest cells of the clamped 3x3 neighbourhood.
__device__ float erosion_cell(const float* pre, int cx, int cy, const FrameParams& p) {
    const int xm1{cx > 0 ? cx - 1 : cx}, xp1{cx + 1 < 8 ? cx + 1 : cx};
    const int ym1{cy > 0 ? cy - 1 : cy}, yp1{cy + 1 < 8 ? cy + 1 : cy};
    float m0{pre[cy * 8 + cx]}, m1{pre[cy * 8 + xm1]}, m2{pre[cy * 8 + xp1]}, m3{pre[ym1 * 8 + xm1]};
    cswap(m0, m1);
    cswap(m0, m2);
    cswap(m0, m3);
    cswap(m1, m2);
    cswap(m1, m3);
    cswap(m2, m3);
    const float extra[5]{pre[ym1 * 8 + cx], pre[ym1 * 8 + xp1], pre[yp1 * 8 + xm1],
                         pre[yp1 * 8 + cx], pre[yp1 * 8 + xp1]};
#pragma unroll
    for (int i{0}; i < 5; ++i) {
        const float v{extra[i]};
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
    return p.km[0] * m0 + p.km[1] * m1 + p.km[2] * m2 + p.km[3] * m3;
}

// Separable DCT building block.  One lane owns one line of N inputs (a pixel
// row for the row pass, a row-pass output line for the column pass) and
// produces F consecutive frequencies k0..k0+F-1 of it:
//   out[f] = sum_{c = 0..N-1} A[k0 + f][c] * in[c]   (fmaf chain in c order).
// k0 is wave-uniform, so every basis value is a scalar operand; the line is
// read with 16-byte LDS loads.
template <int N>
__device__ __forceinline__ const float* basis_n() {
    if constexpr (N == 8) {
        return g_tables.basis8;
    } else if constexpr (N == 16) {
        return g_tables.basis16;
    } else {
        return g_tables.basis32;
    }
}

// Columns col0..col0 + N - 1 of the row at `src`, whose 16-byte groups are
// permuted by `swz` (group g stored at g ^ swz; 0 for the unpermuted row-pass
// lines).
template <int N, int F>
__device__ __forceinline__ void dct_line(const float* __restrict__ src, int k0, float (&out)[F], int col0 = 0,
                                         int swz = 0) {
    float p[N];
#pragma unroll
    for (int c{0}; c < N; c += 4) {
        const float4 v{*reinterpret_cast<const float4*>(src + ((((col0 + c) >> 2) ^ swz) << 2))};
        p[c] = v.x;
        p[c + 1] = v.y;
        p[c + 2] = v.z;
        p[c + 3] = v.w;
    }
    const float* A{basis_n<N>() + k0 * N};
#pragma unroll
    for (int f{0}; f < F; ++f) {
        float s{0.0f};
#pragma unroll
        for (int c{0}; c < N; ++c) {
            s = __builtin_fmaf(A[f * N + c], p[c], s);
        }
        out[f] = s;
    }
}

// Separable DCT lines on the matrix cores.  One K=1 step of
// v_mfma_f32_16x16x1_4b_f32 adds A[b][i] * B[b][j] to D[b][i][j] for four
// independent 16x16 blocks b, rounding like fmaf (the hardware's result for
// every K=1 step equals fmaf(a, b, d) bit for bit), so N steps over c
// reproduce dct_line's fmaf chain exactly.  Lanes 16b..16b+15 supply block b:
// lane 16b + t gives A[b][t] = a_row[c] and B[b][t] = b_row[c].  On return
// lane l holds D[b][4 (l / 16) + u][l % 16] in element 4 b + u.  The B row
// is read like dct_line's source (columns from col0, groups permuted by swz).
typedef float v16f __attribute__((ext_vector_type(16)));

template <int N>
__device__ __forceinline__ v16f mfma_lines(const float* __restrict__ a_row, const float* __restrict__ b_row,
                                           int col0 = 0, int swz = 0) {
    // The lane's whole basis row is read before the chain starts (one
    // memory latency instead of one per four steps).
    float4 ar[N / 4];
#pragma unroll
    for (int c{0}; c < N / 4; ++c) {
        ar[c] = *reinterpret_cast<const float4*>(a_row + 4 * c);
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the chain
    v16f acc;
#pragma unroll
    for (int i{0}; i < 16; ++i) {
        acc[i] = 0.0f;
    }
#pragma unroll
    for (int c{0}; c < N; c += 4) {
        const float4 a{ar[c / 4]};
        const float4 b{*reinterpret_cast<const float4*>(b_row + ((((col0 + c) >> 2) ^ swz) << 2))};
        acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.x, b.x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.y, b.y, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.z, b.z, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x1f32(a.w, b.w, acc, 0, 0, 0);
    }
    return acc;
}

// Row-pass buffer geometry (floats): line (block b, frequency k) of size n
// starts at b * BS + k * RS.  RS is a multiple of 4 (16-byte loads) chosen so
// that eight consecutive lines start in eight different 4-bank groups.
template <int N>
struct RowBuf;
template <>
struct RowBuf<8> {
    static constexpr int RS = 12, BS = 104;
};
template <>
struct RowBuf<16> {
    static constexpr int RS = 20, BS = 320;
};
template <>
struct RowBuf<32> {
    static constexpr int RS = 36, BS = 1152;
};

// Sharpened planes: rows of 32 floats without padding, the row's eight
// 16-byte groups permuted by (row / 2) mod 8, so that the 16 consecutive rows
// a matrix-core row pass reads at one column fall into 16 different 4-bank
// groups.  Planes are stored in the order X, B, Y.
constexpr int kSP = 32;
__device__ __forceinline__ int sp_swz(int row) {
    return (row >> 1) & 7;
}
__device__ __forceinline__ int sp_index(int row, int col) {
    return row * kSP + (((col >> 2) ^ sp_swz(row)) << 2) + (col & 3);
}
// Row-pass lines of the three selection sizes side by side.
constexpr int kLines32 = 0;
constexpr int kLines16 = kLines32 + RowBuf<32>::BS;
constexpr int kLines8 = kLines16 + 4 * RowBuf<16>::BS;
constexpr int kScratch = kLines8 + 16 * RowBuf<8>::BS;

// Element arrays are written in transposed (column) order by the transforms
// and read in element order: one pad slot per 32 elements keeps the
// column-order writes of 16 lanes in distinct banks.
constexpr int kCoefPadded = 1024 + 2 * 32;  // FP16 arrays: two halfs (one bank) per 32
__device__ __forceinline__ int cpad(int e) {
    return e + 2 * (e >> 5);
}

// Arrays whose lifetimes do not overlap share storage, which keeps the
// workgroup under 32 KB of LDS (five resident workgroups per CU):
// - steps 1-3: the XYB tile, then the AQ terms (step 3) past its end;
//   steps 4-5: the row-pass lines;
// - the powers' table and the chroma window in step 1, then the
//   sharpened planes, from step 2 until their last row pass (Y in
//   step 4, X and B in step 5); each plane's space then holds coefficients:
//   X's and B's the chosen transforms (step 5 on), Y's (the last plane, and
//   256 bytes past it) the DCT16 and DCT32 candidates (step 4 on).  The DCT8
//   candidates of Y have their own array.
constexpr int kPlaneHalfs = 2 * 32 * kSP;
static_assert(kCoefPadded <= kPlaneHalfs, "a coefficient array per plane");

struct RegionShared {
    union {
        struct {
            float xyb[3][kTileW * kTileW];  // NV12 -> XYB tile with halo (steps 1-3)
            struct {
                float cells[64];
                float ero[64];     // fuzzy erosion of each cell
                float blk[16][3];  // per block: HF sum, gamma, blue sum
            } aq;                  // AQ terms (step 3)
        } f;
        float scratch[kScratch];  // row-pass lines (steps 4-5)
    } u;
    union {
        alignas(16) float sharp[3][32 * kSP];
        struct {
            _Float16 x[kPlaneHalfs];      // X coefficients of the chosen transforms (cpad)
            _Float16 b[kPlaneHalfs];      // B coefficients of the chosen transforms
            _Float16 y[2 * kCoefPadded];  // Y coefficients of the DCT16 candidates, then the DCT32 candidate
        } coef;
    } p;
    _Float16 coefy8[kCoefPadded];  // Y coefficients of the DCT8 candidates
    float cost8[16];       // quad-major
    float cost16[4];
    float cost32;
    float cost32p[4];      // matrix-core form: the DCT32 cost's part from each wave
    float cfl[4][3];
    int acs[16];           // region-local raster
    float feat[112];       // network features (block sums first; learned mode and export)
    int blkqf[16];         // quant field of the region's blocks (export)
    int choice[4];         // the heuristic's choice per quad (export)
    int learned_ok;        // the network's decision was confident
};
// LDS is allocated in 1,280-byte blocks: five workgroups per CU (160 KB)
// need at most 25 blocks each.
static_assert(sizeof(RegionShared) <= 25 * 1280, "five workgroups per CU");

// Sharpened plane of channel c (0 X, 1 Y, 2 B).
__device__ __forceinline__ float* plane_of(RegionShared& S, int c) {
    return S.p.sharp[c == 0 ? 0 : (c == 1 ? 2 : 1)];
}

// Y coefficients of candidate size 8 << sz, element order (cpad).
__device__ __forceinline__ _Float16* coefy_of(RegionShared& S, int sz) {
    return sz == 0 ? S.coefy8 : S.p.coef.y + (sz - 1) * kCoefPadded;
}

// True for the lowest (n/8)^2 frequencies (fx, fy < n/8) of element
// e = fx * n + fy of an n x n transform, n = 8, 16 or 32 (shifts, not a
// division by a runtime n).
__device__ __forceinline__ bool is_llf(int e, int n) {
    const int l2{n == 32 ? 5 : (n == 16 ? 4 : 3)};
    const int m{n >> 3};
    return (e >> l2) < m && (e & (n - 1)) < m;
}

// Selection terms of M Y coefficients of one lane (transform_select_impl.cuh:
// 49-108): squared quantization error and rate proxy; LLF coefficients
// (llf[i], from the callers' closed forms of is_llf for their lane layouts)
// contribute 0, which the terms of a zero coefficient are exactly (q = 0,
// err = 0, rate[0] = 0).  The two quotients are the correctly rounded ones,
// by reciprocals.  The weights come in as arguments (read four at a time by
// the callers, the frequencies of a lane run in fours); the rate entries of
// all M quotients are read together.
template <int M>
__device__ __forceinline__ void selection_terms(const float (&s)[M], const bool (&llf)[M], const float (&wq)[M],
                                                const float (&wi)[M], const FrameParams& prm, float (&d)[M],
                                                float (&rt)[M]) {
    float aq[M];
    int k[M];
#pragma unroll
    for (int i{0}; i < M; ++i) {
        const float sv{llf[i] ? 0.0f : s[i]};
        const float q{__builtin_rintf(div_by_recip(sv * prm.qgsf_select, wq[i], wi[i]))};
        const float err{div_by_recip(q * wq[i], prm.qgsf_select, prm.inv_qgsf_select) - sv};
        d[i] = err * err;
        aq[i] = fabsf(q);
        k[i] = aq[i] < 256.0f ? static_cast<int>(aq[i]) : 0;
    }
    float r[M];
#pragma unroll
    for (int i{0}; i < M; ++i) {
        r[i] = g_tables.rate[k[i]];
    }
#pragma unroll
    for (int i{0}; i < M; ++i) {
        if (aq[i] >= 256.0f) {
            r[i] = 1.0f + log2_det(1.0f + aq[i]);
        }
        rt[i] = r[i];
    }
}

// Weights and reciprocals of frequencies e0 .. e0 + 3 (e0 a multiple of 4)
// into wq[i0 ..], wi[i0 ..].
template <int M>
__device__ __forceinline__ void selection_weights4(const float* w, const float* w_inv, int e0, int i0,
                                                   float (&wq)[M], float (&wi)[M]) {
    const float4 a{*reinterpret_cast<const float4*>(w + e0)};
    const float4 b{*reinterpret_cast<const float4*>(w_inv + e0)};
    wq[i0] = a.x;
    wq[i0 + 1] = a.y;
    wq[i0 + 2] = a.z;
    wq[i0 + 3] = a.w;
    wi[i0] = b.x;
    wi[i0 + 1] = b.y;
    wi[i0 + 2] = b.z;
    wi[i0 + 3] = b.w;
}

// Colour maps of a 64x64 tile from its regions' partial sums (cfl.h:58-64 on
// (R00 + R10) + (R01 + R11); absent regions contribute zeros).
__device__ __forceinline__ void cfl_tile_maps(const float (&r)[4][3], int& mx, int& mb) {
    float s[3];
    for (int c{0}; c < 3; ++c) {
        s[c] = (r[0][c] + r[1][c]) + (r[2][c] + r[3][c]);
    }
    const double sxy{s[0]}, syy{s[1]}, sby{s[2]};
    const float fx{syy > 0.0 ? static_cast<float>(sxy / syy) : 0.0f};
    const float db{syy > 0.0 ? static_cast<float>(sby / syy) : 0.0f};
    mx = static_cast<int>(__builtin_rintf(fx * 84));
    mb = static_cast<int>(__builtin_rintf(db * 84));
    mx = mx < -128 ? -128 : (mx > 127 ? 127 : mx);
    mb = mb < -128 ? -128 : (mb > 127 ? 127 : mb);
}

// Partial sums of tile (tx, ty)'s regions, zeros for regions past the frame.
__device__ __forceinline__ void tile_partials(const Geometry& geo, const float* cfl_part, int tx, int ty,
                                              float (&r)[4][3]) {
    for (int i{0}; i < 4; ++i) {
        const int rx{tx * 2 + (i & 1)}, ry{ty * 2 + (i >> 1)};
        const bool ok{rx < static_cast<int>(geo.rw) && ry < static_cast<int>(geo.rh)};
        for (int c{0}; c < 3; ++c) {
            r[i][c] = ok ? cfl_part[(static_cast<std::size_t>(ry) * geo.rw + rx) * 3 + c] : 0.0f;
        }
    }
}

// Profiling builds only (CJX_REGION_CUT, never the shipped libraries): the
// region kernel returns, uniformly, at the start of step k + 1 when the
// device word says k (k = 6, 7, 8 cut the transform selection after its
// opening barrier, its row passes and its column passes), so that the time
// and instruction counts of each step are differences of two cut builds.
#ifdef CJX_REGION_CUT
__device__ int g_region_cut;
#define CJX_REGION_CUT_AT(k)                                        \
    if (__builtin_amdgcn_readfirstlane(g_region_cut) == (k)) { \
        return;                                                     \
    }
#else
#define CJX_REGION_CUT_AT(k)
#endif

// Step 1 of a region: NV12 -> XYB over the 36x36 tile (2-pixel halo,
// clamped coordinates), exact 4:2:0 bilinear chroma (centred taps clamped
// to the plane, tests/reference.py:23-46).
//
// The bilinear's weights are 3/4 on the sample of the pixel's own chroma
// cell (coordinate x >> 1) and 1/4 on its neighbour towards the pixel
// (clamped to the plane), in each direction.  Every partial sum of the
// float form (products of samples and weights, then their sums) is exact,
// so the interpolated value is exactly (9 s_kk + 3 s_kn + 3 s_nk + s_nn) / 16:
// the tile computes that integer sum, for Cb and Cr at once in the two
// halves of one word (at most 16 * 255 each), and scales it by 1/16, which
// gives the float form's bits.
__device__ __forceinline__ void xyb_tile(RegionShared& S, const std::uint8_t* luma, std::size_t luma_pitch,
                                         const std::uint8_t* chroma, std::size_t chroma_pitch, int px0, int py0,
                                         int W, int H, int tid) {
    // The powers' table, then the tile's chroma window: samples
    // (cx0 + j, cy0 + i), clamped to the plane, Cb in the low half-word and
    // Cr in the high one.  Every tap of the tile's pixels (their coordinates
    // clamped to the frame, the taps to the plane) lands in the window.  Both
    // live in the sharpened planes' space, which is free until step 2.
    float2* const tab{reinterpret_cast<float2*>(&S.p.sharp[0][0])};
    std::uint32_t* const win{reinterpret_cast<std::uint32_t*>(&S.p.sharp[0][0]) + 2 * kXybTabPairs};
    const int cw{W / 2}, ch{H / 2};
    const int cx0{(px0 >> 1) - 2}, cy0{(py0 >> 1) - 2};
    for (int i{tid}; i < kChromaW * kChromaW; i += 256) {
        int x{cx0 + i % kChromaW}, y{cy0 + i / kChromaW};
        x = x < 0 ? 0 : (x >= cw ? cw - 1 : x);
        y = y < 0 ? 0 : (y >= ch ? ch - 1 : y);
        const std::uint8_t* p{chroma + static_cast<std::size_t>(static_cast<std::uint32_t>(y)) * chroma_pitch +
                              2u * static_cast<std::uint32_t>(x)};
        win[i] = static_cast<std::uint32_t>(p[0]) | (static_cast<std::uint32_t>(p[1]) << 16);
    }
    if (tid < kXybTabPairs) {
        tab[tid] = g_tables.xyb_tab[tid];
    }
    __syncthreads();
    // Thread t converts column t % 36 of the tile, rows t / 36 + 7 k: the
    // column's luma address and horizontal chroma taps are computed once.
    if (tid < 7 * kTileW) {
        const int lx{tid % kTileW};
        int gx{px0 + lx - kHalo};
        gx = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
        const int kx{gx >> 1};
        const int nx{(gx & 1) != 0 ? (kx + 1 < cw ? kx + 1 : cw - 1) : (kx > 0 ? kx - 1 : 0)};
        const int ox{kx - cx0}, onx{nx - cx0};
        const std::uint8_t* const lcol{luma + static_cast<std::uint32_t>(gx)};
        const float neg_cbrt_bias{-cbrt_tab(kOpsinBias, tab)};
        // (1/255) / 16, exact: the 1/16 of the sum folded into the scale.
        constexpr float kScale{(1.0f / 255.0f) * 0.0625f};
        const auto row_of = [&](int ly) {
            const int gy{py0 + ly - kHalo};
            return gy < 0 ? 0 : (gy >= H ? H - 1 : gy);
        };
        const auto luma_at = [&](int gy) {
            return lcol[static_cast<std::size_t>(static_cast<std::uint32_t>(gy)) * luma_pitch];
        };
        // The next row's luma is loaded one iteration ahead (the row past
        // the tile's last is clamped to the frame, so the load is in bounds).
        int gy{row_of(tid / kTileW)};
        std::uint8_t lnext{luma_at(gy)};
        for (int ly{tid / kTileW}; ly < kTileW; ly += 7) {
            const int i{ly * kTileW + lx};
            const std::uint8_t lcur{lnext};
            const int gy_next{row_of(ly + 7)};
            lnext = luma_at(gy_next);
            const int ky{gy >> 1};
            const int ny{(gy & 1) != 0 ? (ky + 1 < ch ? ky + 1 : ch - 1) : (ky > 0 ? ky - 1 : 0)};
            const int rk{(ky - cy0) * kChromaW}, rn{(ny - cy0) * kChromaW};
            const std::uint32_t hk{__umul24(win[rk + ox], 3u) + win[rk + onx]};
            const std::uint32_t hn{__umul24(win[rn + ox], 3u) + win[rn + onx]};
            const std::uint32_t v{(hk << 1) + hk + hn};
            const float yv{static_cast<float>(lcur) * (1.0f / 255.0f)};
            const float cb{static_cast<float>(v & 0xffffu) * kScale - 128.0f / 255.0f};
            const float cr{static_cast<float>(v >> 16) * kScale - 128.0f / 255.0f};
            pixel_to_xyb(yv, cb, cr, S.u.f.xyb[0][i], S.u.f.xyb[1][i], S.u.f.xyb[2][i], tab, neg_cbrt_bias);
            gy = gy_next;
        }
    }
}

// Step 2: inverse gaborish of the three channels into the sharpened planes
// (gaborish.h:16-28; tap order frontend_transform.cu:30-48).
__device__ __forceinline__ void gaborish_tile(RegionShared& S, int tid) {
    constexpr float k0{-0.09495815671340026f}, k1{-0.041031725066768575f};
    constexpr float k2{0.013710004822696948f}, k3{0.006510206083837737f};
    constexpr float k4{-0.0014789063378272242f};
    constexpr float nm{1.0f / (1.0f + 4.0f * (k0 + k1 + k2 + k4 + 2.0f * k3))};
    constexpr float wc{nm}, wr{nm * k0}, wr2{nm * k2}, wd{nm * k1}, wl{nm * k3}, wd2{nm * k4};
    // Thread (gy, gx) of each channel: the four outputs of row gy, columns
    // 4 gx .. 4 gx + 3, from the five input rows gy - 2 .. gy + 2, columns
    // 4 gx - 2 .. 4 gx + 5 (two aligned 16-byte reads per row; the tile's
    // row pitch is 36 floats).
    const int gx{tid & 7}, gy{tid >> 3};
#pragma unroll 1
    for (int c{0}; c < 3; ++c) {
        const float* const base{&S.u.f.xyb[c][gy * kTileW + 4 * gx]};
        float r[5][8];
#pragma unroll
        for (int dy{0}; dy < 5; ++dy) {
            const float4 a{*reinterpret_cast<const float4*>(base + dy * kTileW)};
            const float4 b{*reinterpret_cast<const float4*>(base + dy * kTileW + 4)};
            r[dy][0] = a.x;
            r[dy][1] = a.y;
            r[dy][2] = a.z;
            r[dy][3] = a.w;
            r[dy][4] = b.x;
            r[dy][5] = b.y;
            r[dy][6] = b.z;
            r[dy][7] = b.w;
        }
        float o[4];
#pragma unroll
        for (int k{0}; k < 4; ++k) {
#define AT(dx, dy) r[(dy) + 2][k + 2 + (dx)]
            const float e{AT(-1, 0) + AT(1, 0) + AT(0, -1) + AT(0, 1)};
            const float r2{AT(-2, 0) + AT(2, 0) + AT(0, -2) + AT(0, 2)};
            const float dg{AT(-1, -1) + AT(1, -1) + AT(-1, 1) + AT(1, 1)};
            const float l8{AT(-2, -1) + AT(2, -1) + AT(-2, 1) + AT(2, 1) + AT(-1, -2) + AT(1, -2) +
                           AT(-1, 2) + AT(1, 2)};
            const float d4{AT(-2, -2) + AT(2, -2) + AT(-2, 2) + AT(2, 2)};
            o[k] = wc * AT(0, 0) + wr * e + wr2 * r2 + wd * dg + wl * l8 + wd2 * d4;
#undef AT
        }
        *reinterpret_cast<float4*>(&plane_of(S, c)[sp_index(gy, 4 * gx)]) = float4{o[0], o[1], o[2], o[3]};
    }
}

// ---------------------------------------------------------------------------
// Matrix-core DCT form with FP16 operands (kModeMfma16).  Every transform is
// built from v_mfma_f32_16x16x16_f16 steps: D[i][j] += sum_k A[i][k] B[k][j]
// over a 16x16x16 tile, lane l supplying A[l % 16][4 (l / 16) + 0..3] and
// B[4 (l / 16) + 0..3][l % 16] and holding D[4 (l / 16) + u][l % 16], u =
// 0..3.  A row pass T = P B^T takes the pixels as A and the basis as B; the
// column pass C = B T takes the basis as A and T, exactly as the row pass
// leaves it in the lanes, as B: no data moves between the passes.  A 16x16
// quad is one tile per pass (basis A16, or blockdiag(A8, A8) for its four
// DCT8s); a DCT32 output tile is two K = 16 steps per pass.  Every operand is
// split into FP16 hi + lo (x = hi + lo to about 2^-22 relative) and each
// product is lo*hi + hi*lo + hi*hi in FP32 (lo*lo dropped), so the DC and
// the low frequencies keep about FP32 accuracy.  Pixels are scaled by 1/4
// (exact) into |x| <= 1 (sharpened XYB stays within 2.97 * 0.9) and the
// results by 4.
typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr float kMfScale = 0.25f;

struct MfOp {
    h4v hi, lo;
};

__device__ __forceinline__ MfOp mf_split(float a, float b, float c, float d) {
    MfOp o;
    const float v[4]{a, b, c, d};
#pragma unroll
    for (int i{0}; i < 4; ++i) {
        const _Float16 h{static_cast<_Float16>(v[i])};
        o.hi[i] = h;
        o.lo[i] = static_cast<_Float16>(v[i] - static_cast<float>(h));
    }
    return o;
}

// Row r of an FP16 basis table pair (n columns), columns c0 .. c0 + 3.
__device__ __forceinline__ MfOp mf_basis(const std::uint16_t* hi, const std::uint16_t* lo, int n, int r, int c0) {
    MfOp o;
    o.hi = *reinterpret_cast<const h4v*>(hi + r * n + c0);
    o.lo = *reinterpret_cast<const h4v*>(lo + r * n + c0);
    return o;
}

// acc + A B with A = a.hi + a.lo, B = b.hi + b.lo; the small products first.
__device__ __forceinline__ f4v mf_mma(const MfOp& a, const MfOp& b, f4v acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, acc, 0, 0, 0);
}

// Pixels (row, c0 .. c0 + 3) of a sharpened plane, scaled and split.
__device__ __forceinline__ MfOp mf_pixels(const float* plane, int row, int c0) {
    const float4 p{*reinterpret_cast<const float4*>(plane + sp_index(row, c0))};
    return mf_split(p.x * kMfScale, p.y * kMfScale, p.z * kMfScale, p.w * kMfScale);
}

// The 16x16 quad at (row0, col0) of `plane` through basis tables (hi, lo):
// lane l returns C[4 (l / 16) + u][l % 16] (row = vertical frequency).
__device__ __forceinline__ f4v mf_quad(const float* plane, int row0, int col0, const std::uint16_t* hi,
                                       const std::uint16_t* lo, int lane) {
    const int m{lane & 15}, g{lane >> 4};
    const MfOp b{mf_basis(hi, lo, 16, m, 4 * g)};
    const f4v zero{0.0f, 0.0f, 0.0f, 0.0f};
    const f4v t{mf_mma(mf_pixels(plane, row0 + m, col0 + 4 * g), b, zero)};
    const f4v c{mf_mma(b, mf_split(t[0], t[1], t[2], t[3]), zero)};
    return c * (1.0f / kMfScale);
}

// DCT32 of the region's plane: output tiles (0, fx0) and (16, fx0) (vertical,
// horizontal frequency origins) into c[0], c[1]; with `both` false only
// (fy0, fx0) into c[0].
__device__ __forceinline__ void mf_dct32(const float* plane, int fx0, int fy0, bool both, int lane, f4v (&c)[2]) {
    const int m{lane & 15}, g{lane >> 4};
    const std::uint16_t* hi{g_tables.mf32[0]};
    const std::uint16_t* lo{g_tables.mf32[1]};
    const f4v zero{0.0f, 0.0f, 0.0f, 0.0f};
    MfOp t[2];
#pragma unroll
    for (int h{0}; h < 2; ++h) {  // row-pass tiles at rows 16 h, frequencies fx0..
        f4v acc{zero};
#pragma unroll
        for (int k{0}; k < 2; ++k) {
            acc = mf_mma(mf_pixels(plane, 16 * h + m, 16 * k + 4 * g), mf_basis(hi, lo, 32, fx0 + m, 16 * k + 4 * g),
                         acc);
        }
        t[h] = mf_split(acc[0], acc[1], acc[2], acc[3]);
    }
#pragma unroll
    for (int o{0}; o < 2; ++o) {
        if (o == 1 && !both) {
            break;
        }
        const int f0{both ? 16 * o : fy0};
        f4v acc{zero};
#pragma unroll
        for (int h{0}; h < 2; ++h) {
            acc = mf_mma(mf_basis(hi, lo, 32, f0 + m, 16 * h + 4 * g), t[h], acc);
        }
        c[o] = acc * (1.0f / kMfScale);
    }
}

// Lane l's four elements of a quad result, as (block j of the quad, element
// e of the block's transform): DCT16 (n = 16) e = fx * 16 + fy, j = 0; four
// DCT8s (n = 8) j = 2 (fy / 8) + fx / 8, e = (fx % 8) * 8 + fy % 8.
__device__ __forceinline__ void mf_quad_element(int lane, int u, int n, int& j, int& e) {
    const int fx{lane & 15}, fy{4 * (lane >> 4) + u};
    if (n == 16) {
        j = 0;
        e = fx * 16 + fy;
    } else {
        j = 2 * (fy >> 3) + (fx >> 3);
        e = (fx & 7) * 8 + (fy & 7);
    }
}

// ---------------------------------------------------------------------------
// The learned transform selection (kModeLearned) and the feature export of
// the plain path that trains it.  Features of a region, 6 per 8x8 block b
// (region raster, 16 blocks), all from the block's DCT8 of the sharpened Y
// (the selection's DCT8 candidate, C[e] with e = fx * 8 + fy) and its
// selection terms:
//   f[6b + 0] = log2(1 + 4096 var)     var  = sum over AC of C^2
//   f[6b + 1] = log2(1 + 1024 grad)    grad = (64 / 112) sum (l[fx] + l[fy]) C^2
//   f[6b + 2] = log2(1 + 64 cost8)     the block's DCT8 RD cost
//   f[6b + 3] = log2(1 + 262144 dist)  dist = sum of squared quantization errors
//   f[6b + 4] = log2(1 + rate)         rate = sum of the rate terms
//   f[6b + 5] = 4 C[0]                 the block's mean
// With the basis' normalization (C = O P O^T / 8, O orthonormal) var is the
// pixels' variance and grad the mean squared difference of horizontally and
// vertically adjacent pixels inside the block (112 pairs): the DCT-II
// diagonalizes the path's difference operator with eigenvalues
// l[k] = 4 sin^2(pi k / 16).  Blocks past the frame's edge have all six at 0.
struct MlpWeights {
    std::uint16_t w1[64 * 96];  // FP16 bit patterns, row-major [hidden][input]
    std::uint16_t w2[7 * 64];   // [output][hidden]
    float b1[64];
    float b2[8];
    float threshold;  // |logit| below which a decision is not confident
};

__device__ __align__(16) MlpWeights g_mlp;

__device__ __forceinline__ float h2f(std::uint16_t h) {
    return static_cast<float>(__builtin_bit_cast(_Float16, h));
}

// One thread's share of the network's weights, read at the kernel's start
// (the learned forms), so that no read waits inside the decision: layer 1's
// 24 weights of hidden unit tid / 4 (inputs 24 (tid % 4) ..) and its bias,
// and layer 2's 8 weights of output tid / 8 (hidden 8 (tid % 8) ..; threads
// below 56) and its bias.
struct MlpRegs {
    uint4 w1[3];
    uint4 w2;
    float b1, b2;
};

__device__ __forceinline__ MlpRegs mlp_load(int tid) {
    MlpRegs r;
    const int j{tid >> 2}, k{tid & 3};
    const std::uint16_t* const w{g_mlp.w1 + j * 96 + 24 * k};
#pragma unroll
    for (int i{0}; i < 3; ++i) {
        r.w1[i] = *reinterpret_cast<const uint4*>(w + 8 * i);
    }
    const int o{tid < 56 ? tid >> 3 : 0}, p{tid & 7};
    r.w2 = *reinterpret_cast<const uint4*>(g_mlp.w2 + o * 64 + 8 * p);
    r.b1 = g_mlp.b1[j];
    r.b2 = g_mlp.b2[o];
    return r;
}

// (64 / 112) l[k] of the gradient feature, k = 0..7.
__device__ __forceinline__ float grad_weight(int k) {
    constexpr float w[8]{0.0f, 0.08699481934309006f, 0.33473509550094604f, 0.7055046558380127f,
                         1.1428571939468384f, 1.5802096128463745f, 1.950979232788086f, 2.1987195014953613f};
    return w[k];
}

// Block b's raw sums (var, grad, dist, rate, mean) into S.feat[6 b ..]; the
// cost sits in S.cost8.
__device__ __forceinline__ void feature_store(RegionShared& S, int b, float var, float grad, float dist, float rate,
                                              float mean) {
    S.feat[6 * b + 0] = var;
    S.feat[6 * b + 1] = grad;
    S.feat[6 * b + 3] = dist;
    S.feat[6 * b + 4] = rate;
    S.feat[6 * b + 5] = mean;
}

// The raw sums -> the 96 features in S.feat (every thread calls this;
// barriers inside).
__device__ __forceinline__ void feature_finish(RegionShared& S, int tid, int bx0, int by0, int bw, int bh) {
    float f[6]{};
    if (tid < 16) {
        const int bx{tid & 3}, by{tid >> 2};
        if (bx0 + bx < bw && by0 + by < bh) {
            const float* s{&S.feat[6 * tid]};
            const float c8{S.cost8[((by >> 1) * 2 + (bx >> 1)) * 4 + (by & 1) * 2 + (bx & 1)]};
            f[0] = log2_det(1.0f + 4096.0f * fmaxf(s[0], 0.0f));
            f[1] = log2_det(1.0f + 1024.0f * fmaxf(s[1], 0.0f));
            f[2] = log2_det(1.0f + 64.0f * fmaxf(c8, 0.0f));
            f[3] = log2_det(1.0f + 262144.0f * fmaxf(s[3], 0.0f));
            f[4] = log2_det(1.0f + fmaxf(s[4], 0.0f));
            f[5] = 4.0f * s[5];
        }
    }
    __syncthreads();
    if (tid < 16) {
#pragma unroll
        for (int k{0}; k < 6; ++k) {
            S.feat[6 * tid + k] = f[k];
        }
    }
    __syncthreads();
}

// The network on the region's features and, when every decision it takes is
// confident (|logit| >= threshold for the root where a DCT32 fits and for
// each quad where a DCT16 fits), the transform map in S.acs.  Returns that
// confidence (the same value in every thread).  `hid` is LDS scratch of 72
// floats.  Every thread calls this; barriers inside.
__device__ __forceinline__ bool mlp_decide(RegionShared& S, const FrameParams& prm, const MlpRegs& mr, float* hid,
                                           int tid, int bx0, int by0, int bw, int bh) {
    {
        const int j{tid >> 2}, k{tid & 3};
        const float* const x{&S.feat[24 * k]};
        float s{0.0f};
#pragma unroll
        for (int i{0}; i < 24; i += 8) {
            const uint4 pw{mr.w1[i / 8]};
            const std::uint32_t word[4]{pw.x, pw.y, pw.z, pw.w};
#pragma unroll
            for (int h{0}; h < 4; ++h) {
                s = __builtin_fmaf(h2f(static_cast<std::uint16_t>(word[h] & 0xffffu)), x[i + 2 * h], s);
                s = __builtin_fmaf(h2f(static_cast<std::uint16_t>(word[h] >> 16)), x[i + 2 * h + 1], s);
            }
        }
        s = add_down_row<2>(s);
        s = add_down_row<1>(s);
        if (k == 0) {
            hid[j] = fmaxf(s + mr.b1, 0.0f);
        }
    }
    __syncthreads();
    if (tid < 56) {
        const int o{tid >> 3}, p{tid & 7};
        const std::uint32_t word[4]{mr.w2.x, mr.w2.y, mr.w2.z, mr.w2.w};
        float s{0.0f};
#pragma unroll
        for (int q{0}; q < 8; ++q) {
            const std::uint32_t w{word[q >> 1]};
            s = __builtin_fmaf(h2f(static_cast<std::uint16_t>((q & 1) != 0 ? w >> 16 : w & 0xffffu)), hid[8 * p + q], s);
        }
        s = add_down_row<4>(s);
        s = add_down_row<2>(s);
        s = add_down_row<1>(s);
        if (p == 0) {
            hid[64 + o] = s + mr.b2;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const float* z{hid + 64};
        const float th{g_mlp.threshold};
        const std::uint32_t force{(prm.mode & kModeForceMask) >> kModeForceShift};
        const bool fit32{bx0 + 4 <= bw && by0 + 4 <= bh};
        bool ok{force == 0};
        const float root{z[1] - z[0]};
        ok = ok && (!fit32 || fabsf(root) >= th);
        const bool use32{fit32 && root > 0.0f};
        int choice[4];
        for (int q{0}; q < 4; ++q) {
            const int qx{q & 1}, qy{q >> 1};
            const bool fit16{bx0 + 2 * qx + 2 <= bw && by0 + 2 * qy + 2 <= bh};
            choice[q] = fit16 && z[2 + q] > 0.0f ? 16 : 8;
            if (!use32) {
                ok = ok && (!fit16 || fabsf(z[2 + q]) >= th);
            }
        }
        if (ok) {
            if (use32) {
                for (int b{0}; b < 16; ++b) {
                    S.acs[b] = b == 0 ? 32 : 0;
                }
            } else {
                for (int q{0}; q < 4; ++q) {
                    const int ox{2 * (q & 1)}, oy{2 * (q >> 1)};
                    if (choice[q] == 16) {
                        S.acs[oy * 4 + ox] = 16;
                        S.acs[oy * 4 + ox + 1] = 0;
                        S.acs[(oy + 1) * 4 + ox] = 0;
                        S.acs[(oy + 1) * 4 + ox + 1] = 0;
                    }
                }
            }
        }
        S.learned_ok = ok ? 1 : 0;
    }
    __syncthreads();
    return S.learned_ok != 0;
}

// Training rows of the plain path: the region's 96 features, then its labels
// (12 floats: DCT32 at the root, DCT16 per quad as the heuristic chose it,
// the mean log2 quant field of the region's blocks, the root mask (a DCT32
// fits), the quad masks (a DCT16 fits), 0).
constexpr int kFeatRow = 96 + 12;

__device__ __forceinline__ void export_features(const RegionShared& S, float* out, int region, int tid, int bx0,
                                                int by0, int bw, int bh) {
    float* const row{out + static_cast<std::size_t>(region) * kFeatRow};
    if (tid < 96) {
        row[tid] = S.feat[tid];
    } else if (tid < 108) {
        const int k{tid - 96};
        float v{0.0f};
        if (k == 0) {
            v = S.acs[0] == 32 ? 1.0f : 0.0f;
        } else if (k <= 4) {
            v = S.choice[k - 1] == 16 ? 1.0f : 0.0f;
        } else if (k == 5) {
            float s{0.0f};
            int n{0};
            for (int b{0}; b < 16; ++b) {
                if (bx0 + (b & 3) < bw && by0 + (b >> 2) < bh) {
                    s += log2_det(static_cast<float>(S.blkqf[b]));
                    ++n;
                }
            }
            v = s / static_cast<float>(n);
        } else if (k == 6) {
            v = bx0 + 4 <= bw && by0 + 4 <= bh ? 1.0f : 0.0f;
        } else if (k <= 10) {
            const int q{k - 7};
            v = bx0 + 2 * (q & 1) + 2 <= bw && by0 + 2 * (q >> 1) + 2 <= bh ? 1.0f : 0.0f;
        }
        row[96 + k] = v;
    }
}

template <bool kMf16, bool kLearned>
__device__ __attribute__((always_inline)) inline void region_kernel_body(const Wg& wg, 
    Geometry geo, FrameParams prm, const std::uint8_t* __restrict__ luma, std::size_t luma_pitch,
    const std::uint8_t* __restrict__ chroma, std::size_t chroma_pitch,
    std::uint16_t* __restrict__ coeffs, std::int8_t* __restrict__ acs_out,
    std::int32_t* __restrict__ qf_out, float* __restrict__ cfl_part, std::uint32_t* __restrict__ list_count,
    Clears clr, float* __restrict__ feat_out) {
    __shared__ RegionShared S;
    clear_words(clr, wg);  // every workgroup of the grid takes part
    const int tid{static_cast<int>(threadIdx.x)};
    if (wg.x == 0 && tid < 2 && list_count != nullptr) {
        list_count[tid] = 0;  // the quantizer's item lists, appended after this kernel
    }
    const int wu{__builtin_amdgcn_readfirstlane(tid >> 6)};  // the wave (uniform)
    [[maybe_unused]] MlpRegs mlp_regs;
    if constexpr (kLearned) {
        mlp_regs = mlp_load(tid);
    }
    const int lane{tid & 63};
    // Workgroup i runs on XCD i mod 8: give each XCD a contiguous run of
    // regions so that rows shared by neighbouring regions stay in one L2.
    const int nreg{static_cast<int>(geo.rw * geo.rh)};
    const int per_xcd{(nreg + 7) / 8};
    const int region{static_cast<int>(wg.x & 7u) * per_xcd + static_cast<int>(wg.x >> 3)};
    if (region >= nreg) {
        return;
    }
    const int rx{region % static_cast<int>(geo.rw)};
    const int ry{region / static_cast<int>(geo.rw)};
    const int px0{rx * 32}, py0{ry * 32};
    const int W{static_cast<int>(geo.width)}, H{static_cast<int>(geo.height)};
    const int bw{static_cast<int>(geo.bw)}, bh{static_cast<int>(geo.bh)};
    const int bx0{rx * 4}, by0{ry * 4};

    // 1. NV12 -> XYB over the 36x36 tile (2-pixel halo, clamped coordinates).
    xyb_tile(S, luma, luma_pitch, chroma, chroma_pitch, px0, py0, W, H, tid);
    if (tid < 16) {
        S.acs[tid] = 8;
    }
    __syncthreads();

    CJX_REGION_CUT_AT(1)
    // 2. Inverse gaborish (gaborish.h:16-28; tap order frontend_transform.cu:30-48).
    {
        gaborish_tile(S, tid);
        // 3a. AQ pre-erosion terms per interior pixel (frontend_fused.cu:118-137)
        //     and their 4x4 cell sums in row order: four lanes per cell, one
        //     per row; row dy continues the running sum of row dy - 1.
        const int cell{tid >> 2}, dy{tid & 3};
        const int row{(cell >> 3) * 4 + dy}, col0{(cell & 7) * 4};
        float m[4];
#pragma unroll
        for (int dx{0}; dx < 4; ++dx) {
            const float* s{&S.u.f.xyb[1][(row + kHalo) * kTileW + col0 + dx + kHalo]};
            const float yv{s[0]};
            const float base{0.25f * (s[kTileW] + s[-kTileW] + s[-1] + s[1])};
            float diff{ratio_cbrt_gamma(yv + 0.019f, false) * (yv - base)};
            diff *= diff;
            diff = fminf(diff, 0.2f);
            m[dx] = masking_sqrt(diff);
        }
        float sum{0.0f};
#pragma unroll
        for (int k{0}; k < 4; ++k) {
            const float prev{from_lane_below(sum)};
            if (dy == k) {
                float acc{k == 0 ? 0.0f : prev};
                acc += m[0];
                acc += m[1];
                acc += m[2];
                acc += m[3];
                sum = acc;
            }
        }
        if (dy == 3) {
            S.u.f.aq.cells[cell] = 0.25f * sum;
        }
    }
    __syncthreads();

    CJX_REGION_CUT_AT(2)
    // 3b. Per-block modulation and the quant integer (frontend_fused.cu:117-203).
    //     The 64 cells' fuzzy erosion and the 16 blocks' final step each run
    //     once, in parallel lanes; the per-pixel sums are wave sums as before
    //     (a block's four erosion terms sum as (e0 + e2) + (e1 + e3), which is
    //     the wave tree over four non-zero lanes).
    if (tid < 64) {
        S.u.f.aq.ero[tid] = erosion_cell(S.u.f.aq.cells, tid & 7, tid >> 3, prm);
    }
#pragma unroll
    for (int j{0}; j < 4; ++j) {
        const int lb{j * 4 + wu};
        const int lbx{lb & 3}, lby{lb >> 2};
        const int fx{lane & 7}, fy{lane >> 3};
        const int idx{(lby * 8 + fy + kHalo) * kTileW + lbx * 8 + fx + kHalo};
        const float yv{S.u.f.xyb[1][idx]}, xv{S.u.f.xyb[0][idx]}, bv{S.u.f.xyb[2][idx]};
        float hf{0.0f};
        if (fx < 7) {
            hf += fminf(0.0206f, fabsf(yv - S.u.f.xyb[1][idx + 1]));
        }
        if (fy < 7) {
            hf += fminf(0.0206f, fabsf(yv - S.u.f.xyb[1][idx + kTileW]));
        }
        const float iny{yv + 0.16f};
        const float gm{ratio_cbrt_gamma(iny - xv, true) + ratio_cbrt_gamma(iny + xv, true)};
        const float pye{yv + 0.084381641171960495f + fabsf(xv)};
        const float bl{bv > pye ? fminf(bv - pye, 0.027121074570634722f) : 0.0f};
        const float sums{wave_sum4(hf, gm, bl, 0.0f)};  // lanes 0, 32, 16
        const float gamma_overall{lane_value(sums, 32) * (0.5f / 64.0f)};
        const float blue_sum{lane_value(sums, 16)};
        if (lane == 0) {
            S.u.f.aq.blk[lb][0] = sums;
            S.u.f.aq.blk[lb][1] = gamma_overall;
            S.u.f.aq.blk[lb][2] = blue_sum;
        }
    }
    __syncthreads();
    if (tid < 16) {
        const int lb{tid};
        const int lbx{lb & 3}, lby{lb >> 2};
        if (bx0 + lbx < bw && by0 + lby < bh) {
            const float e0{S.u.f.aq.ero[(2 * lby) * 8 + 2 * lbx]}, e1{S.u.f.aq.ero[(2 * lby) * 8 + 2 * lbx + 1]};
            const float e2{S.u.f.aq.ero[(2 * lby + 1) * 8 + 2 * lbx]}, e3{S.u.f.aq.ero[(2 * lby + 1) * 8 + 2 * lbx + 1]};
            const float aq{(e0 + e2) + (e1 + e3)};
            float blue_sum{S.u.f.aq.blk[lb][2]};
            constexpr float k_limit{0.027121074570634722f};
            if (blue_sum >= 32.0f * k_limit) {
                blue_sum = 64.0f * k_limit - blue_sum;
            }
            constexpr float k_max_limit{15.398788439047934f};
            if (blue_sum >= k_max_limit * k_limit) {
                blue_sum = k_max_limit * k_limit;
            }
            float ov{compute_mask(aq)};
            ov += S.u.f.aq.blk[lb][0] * -0.38f + 0.42f;
            ov += 0.1005613337192697f * log2_det(S.u.f.aq.blk[lb][1]);
            ov += blue_sum * 0.14207000358439159f;
            const float qval{exp_det(ov) * prm.mul_pb + prm.add_pb};
            int qi{static_cast<int>(qval * prm.inv_global_scale + 0.5f)};
            qi = qi < 1 ? 1 : (qi > 256 ? 256 : qi);
            qf_out[static_cast<std::size_t>(by0 + lby) * bw + bx0 + lbx] = qi;
            S.blkqf[lb] = qi;
        }
    }

    CJX_REGION_CUT_AT(3)
    // 4. Transform selection on the sharpened Y (transform_select_impl.cuh:49-201).
    //    Plain: the three sizes run side by side (wave 0 the DCT32 and wave 1
    //    the four DCT16s on the matrix cores, waves 2-3 the sixteen DCT8s on
    //    the VALU); each wave keeps its coefficients' selection terms in
    //    registers and sums them there (per lane in register order, then a
    //    lane tree).  Learned: the DCT8s and their costs, the region's
    //    features and the network first; the DCT16 and DCT32 candidates only
    //    where the network is not confident (the heuristic decides there).
    __syncthreads();  // the XYB tile is dead from here on; its space holds row-pass lines
    CJX_REGION_CUT_AT(6)
    float* const lines{S.u.scratch};
    const int grp{lane >> 4}, t16{lane & 15};
    const float* const yp{plane_of(S, 1)};
    const bool want_features{kLearned || feat_out != nullptr};
    float td[16], tr[16];
    // Plain pieces: row passes of the DCT32 / DCT16 candidates (waves 0-1)
    // and of the DCT8s (waves 2-3), then the column passes with the costs.
    const auto rows_big = [&]() {
        if (wu == 0) {
            // DCT32 as four 16x16 quadrants: block (fh, rh) = (grp >> 1, grp & 1).
            const int row{16 * (grp & 1) + t16};
            const v16f acc{mfma_lines<32>(g_tables.basis32 + (16 * (grp >> 1) + t16) * 32, yp + row * kSP, 0,
                                          sp_swz(row))};
#pragma unroll
            for (int b{0}; b < 4; ++b) {
#pragma unroll
                for (int u{0}; u < 4; ++u) {
                    const int f{16 * (b >> 1) + 4 * grp + u}, r{16 * (b & 1) + t16};
                    lines[kLines32 + f * RowBuf<32>::RS + r] = acc[4 * b + u];
                }
            }
        } else {
            // Block q = quad q.
            const int row{(grp >> 1) * 16 + t16};
            const v16f acc{mfma_lines<16>(g_tables.basis16 + t16 * 16, yp + row * kSP, (grp & 1) * 16, sp_swz(row))};
#pragma unroll
            for (int q{0}; q < 4; ++q) {
#pragma unroll
                for (int u{0}; u < 4; ++u) {
                    lines[kLines16 + q * RowBuf<16>::BS + (4 * grp + u) * RowBuf<16>::RS + t16] = acc[4 * q + u];
                }
            }
        }
    };
    const auto rows8 = [&]() {
        // 16 blocks x 8 rows over 128 lanes, all 8 frequencies per lane.
        const int job{((wu - 2) << 6) | lane};
        const int b{job >> 3}, r{job & 7};
        float o[8];
        const int row{(b >> 2) * 8 + r};
        dct_line<8, 8>(yp + row * kSP, 0, o, (b & 3) * 8, sp_swz(row));
#pragma unroll
        for (int f{0}; f < 8; ++f) {
            lines[kLines8 + b * RowBuf<8>::BS + f * RowBuf<8>::RS + r] = o[f];
        }
    };
    // Column passes: 16 terms per lane on waves 0-1, 8 on waves 2-3; costs
    // straight from the registers: per lane the sum of its terms in register
    // order, then the wave tree (DCT32; DCT16 per block) or an 8-lane tree
    // (DCT8: a block's 64 terms sit in 8 consecutive lanes, one column each).
    const auto cols_big = [&]() {
        if (wu == 0) {
            const v16f acc{mfma_lines<32>(g_tables.basis32 + (16 * (grp >> 1) + t16) * 32,
                                          &lines[kLines32 + (16 * (grp & 1) + t16) * RowBuf<32>::RS])};
            float v[16], wq[16], wi[16];
            int e[16];
            bool llf[16];
#pragma unroll
            for (int b{0}; b < 4; ++b) {
                const int e0{(16 * (b & 1) + t16) * 32 + 16 * (b >> 1) + 4 * grp};
                selection_weights4(g_tables.wsel32, g_tables.wsel_inv32, e0, 4 * b, wq, wi);
#pragma unroll
                for (int u{0}; u < 4; ++u) {
                    e[4 * b + u] = e0 + u;
                    v[4 * b + u] = acc[4 * b + u];
                    llf[4 * b + u] = b == 0 && lane < 4;  // fx, fy < 4
                }
            }
            selection_terms(v, llf, wq, wi, prm, td, tr);
#pragma unroll
            for (int i{0}; i < 16; ++i) {
                coefy_of(S, 2)[cpad(e[i])] = static_cast<_Float16>(v[i]);
            }
            float ad{td[0]}, ar{tr[0]};
#pragma unroll
            for (int i{1}; i < 16; ++i) {
                ad += td[i];
                ar += tr[i];
            }
            const float sums{wave_sum4(ad, ar, 0.0f, 0.0f)};  // lanes 0, 32
            const float rs{lane_value(sums, 32)};
            if (lane == 0) {
                S.cost32 = 1024.0f * sums + 6.0e-4f * rs;
            }
        } else {
            const v16f acc{mfma_lines<16>(g_tables.basis16 + t16 * 16,
                                          &lines[kLines16 + grp * RowBuf<16>::BS + t16 * RowBuf<16>::RS])};
            // Frequency (t16, 4 grp + u) in each of the four blocks.
            const int e0{t16 * 16 + 4 * grp};
            float w4[4], wi4[4];
            selection_weights4(g_tables.wsel16, g_tables.wsel_inv16, e0, 0, w4, wi4);
            float v[16], wq[16], wi[16];
            int e[16];
            bool llf[16];
#pragma unroll
            for (int i{0}; i < 16; ++i) {
                e[i] = e0 + (i & 3);
                v[i] = acc[i];
                wq[i] = w4[i & 3];
                wi[i] = wi4[i & 3];
                llf[i] = lane < 2 && (i & 3) < 2;  // fx, fy < 2
            }
            selection_terms(v, llf, wq, wi, prm, td, tr);
#pragma unroll
            for (int i{0}; i < 16; ++i) {
                coefy_of(S, 1)[cpad(256 * (i >> 2) + e[i])] = static_cast<_Float16>(v[i]);
            }
            float ad[4], ar[4];
#pragma unroll
            for (int q{0}; q < 4; ++q) {
                ad[q] = td[4 * q];
                ar[q] = tr[4 * q];
#pragma unroll
                for (int u{1}; u < 4; ++u) {
                    ad[q] += td[4 * q + u];
                    ar[q] += tr[4 * q + u];
                }
            }
#pragma unroll
            for (int h{0}; h < 2; ++h) {
                // Quads 2h (lanes 0, 32) and 2h + 1 (lanes 16, 48).
                const float sums{wave_sum4(ad[2 * h], ar[2 * h], ad[2 * h + 1], ar[2 * h + 1])};
                const float d1{lane_value(sums, 16)}, r0{lane_value(sums, 32)}, r1{lane_value(sums, 48)};
                if (lane == 0) {
                    S.cost16[2 * h] = 256.0f * sums + 6.0e-4f * r0;
                    S.cost16[2 * h + 1] = 256.0f * d1 + 6.0e-4f * r1;
                }
            }
        }
    };
    const auto cols8 = [&]() {
        const int job{((wu - 2) << 6) | lane};
        const int b{job >> 3}, fx{job & 7};
        const int bx{b & 3}, by{b >> 2};
        const int base{256 * ((by >> 1) * 2 + (bx >> 1)) + 64 * ((by & 1) * 2 + (bx & 1))};
        float o[8];
        dct_line<8, 8>(&lines[kLines8 + b * RowBuf<8>::BS + fx * RowBuf<8>::RS], 0, o);
        {
            float wq[8], wi[8], d8[8], r8[8];
            int e[8];
            bool llf[8];
            selection_weights4(g_tables.wsel8, g_tables.wsel_inv8, fx * 8, 0, wq, wi);
            selection_weights4(g_tables.wsel8, g_tables.wsel_inv8, fx * 8 + 4, 4, wq, wi);
#pragma unroll
            for (int f{0}; f < 8; ++f) {
                e[f] = fx * 8 + f;
                llf[f] = f == 0 && fx == 0;  // the DC
            }
            selection_terms(o, llf, wq, wi, prm, d8, r8);
#pragma unroll
            for (int f{0}; f < 8; ++f) {
                td[f] = d8[f];
                tr[f] = r8[f];
                coefy_of(S, 0)[cpad(base + e[f])] = static_cast<_Float16>(o[f]);
            }
        }
        float ad{td[0]}, ar{tr[0]};
#pragma unroll
        for (int f{1}; f < 8; ++f) {
            ad += td[f];
            ar += tr[f];
        }
        float fv{0.0f}, fg{0.0f};
        if (want_features) {
#pragma unroll
            for (int f{0}; f < 8; ++f) {
                const float c2{o[f] * o[f]};
                fv += (fx | f) != 0 ? c2 : 0.0f;
                fg = __builtin_fmaf(grad_weight(fx) + grad_weight(f), c2, fg);
            }
        }
        const auto sum8 = [](float x) { return add_down_row<1>(add_down_row<2>(add_down_row<4>(x))); };
        ad = sum8(ad);
        ar = sum8(ar);
        if (want_features) {
            fv = sum8(fv);
            fg = sum8(fg);
        }
        if ((lane & 7) == 0) {
            S.cost8[((by >> 1) * 2 + (bx >> 1)) * 4 + (by & 1) * 2 + (bx & 1)] = 64.0f * ad + 6.0e-4f * ar;
            if (want_features) {
                feature_store(S, b, fv, fg, ad, ar, o[0]);
            }
        }
    };
    // Matrix-core FP16 pieces: wave w transforms quad w as four DCT8s (with
    // their costs, from registers) and as one DCT16, and output tile w of the
    // DCT32 (vertical origin 16 (w / 2), horizontal 16 (w % 2)); the DCT16
    // and DCT32 results stay in registers until every wave has read the Y
    // plane, whose space then takes their coefficients.
    const auto mf_dct8_costs = [&]() {
        const int q{wu};
        const f4v c8{mf_quad(yp, (q >> 1) * 16, (q & 1) * 16, g_tables.mf8[0], g_tables.mf8[1], lane)};
        float d8{0.0f}, r8{0.0f}, fv{0.0f}, fg{0.0f};
        float v[4], wq[4], wi[4], d[4], rt[4];
        int e[4], j[4];
        bool llf[4];
#pragma unroll
        for (int u{0}; u < 4; ++u) {
            mf_quad_element(lane, u, 8, j[u], e[u]);
            v[u] = c8[u];
            llf[u] = u == 0 && (lane & 7) == 0 && (grp & 1) == 0;  // the DC
        }
        selection_weights4(g_tables.wsel8, g_tables.wsel_inv8, e[0], 0, wq, wi);
        selection_terms(v, llf, wq, wi, prm, d, rt);
#pragma unroll
        for (int u{0}; u < 4; ++u) {
            coefy_of(S, 0)[cpad(256 * q + 64 * j[u] + e[u])] = static_cast<_Float16>(c8[u]);
            d8 += d[u];
            r8 += rt[u];
            if (want_features) {
                const float c2{c8[u] * c8[u]};
                fv += e[u] != 0 ? c2 : 0.0f;
                fg = __builtin_fmaf(grad_weight(e[u] >> 3) + grad_weight(e[u] & 7), c2, fg);
            }
        }
        // The DC of block (grp / 2, column / 8) sits in lane 16 grp + 8 k,
        // element u = 0, for even grp.
        const float dc{c8[0]};
        // A DCT8 block's 64 terms sit in 16 lanes: bits 0-2 of the column and
        // bit 0 of the lane group vary within it.
        const auto sum16 = [](float x) { return add_down_16(add_down_row<4>(add_down_row<2>(add_down_row<1>(x)))); };
        d8 = sum16(d8);
        r8 = sum16(r8);
        if (want_features) {
            fv = sum16(fv);
            fg = sum16(fg);
        }
        if ((lane & 7) == 0 && (grp & 1) == 0) {
            const int j{2 * (grp >> 1) + ((lane & 15) >> 3)};
            S.cost8[q * 4 + j] = 64.0f * d8 + 6.0e-4f * r8;
            if (want_features) {
                const int b{(2 * (q >> 1) + (j >> 1)) * 4 + 2 * (q & 1) + (j & 1)};
                feature_store(S, b, fv, fg, d8, r8, dc);
            }
        }
    };
    const auto mf_big = [&]() {  // every wave; barriers inside
        const int q{wu};
        const f4v c16{mf_quad(yp, (q >> 1) * 16, (q & 1) * 16, g_tables.mf16[0], g_tables.mf16[1], lane)};
        f4v c32[2];
        mf_dct32(yp, 16 * (wu & 1), 16 * (wu >> 1), false, lane, c32);
        __syncthreads();
        float d16{0.0f}, r16{0.0f}, d32{0.0f}, r32{0.0f};
        float v16[4], w16[4], wi16[4], v32[4], w32[4], wi32[4], d[2][4], rt[2][4];
        int e16[4], e32[4];
        bool llf16[4], llf32[4];
        const int e32_0{(16 * (wu & 1) + t16) * 32 + 16 * (wu >> 1) + 4 * grp};
#pragma unroll
        for (int u{0}; u < 4; ++u) {
            int j;
            mf_quad_element(lane, u, 16, j, e16[u]);
            v16[u] = c16[u];
            e32[u] = e32_0 + u;
            v32[u] = c32[0][u];
            llf16[u] = u < 2 && lane < 2;   // fx, fy < 2
            llf32[u] = wu == 0 && lane < 4;  // fx, fy < 4
        }
        selection_weights4(g_tables.wsel16, g_tables.wsel_inv16, e16[0], 0, w16, wi16);
        selection_weights4(g_tables.wsel32, g_tables.wsel_inv32, e32_0, 0, w32, wi32);
        selection_terms(v16, llf16, w16, wi16, prm, d[0], rt[0]);
        selection_terms(v32, llf32, w32, wi32, prm, d[1], rt[1]);
#pragma unroll
        for (int u{0}; u < 4; ++u) {
            coefy_of(S, 1)[cpad(256 * q + e16[u])] = static_cast<_Float16>(c16[u]);
            d16 += d[0][u];
            r16 += rt[0][u];
            coefy_of(S, 2)[cpad(e32[u])] = static_cast<_Float16>(c32[0][u]);
            d32 += d[1][u];
            r32 += rt[1][u];
        }
        const float sums{wave_sum4(d16, r16, d32, r32)};  // lanes 0, 32, 16, 48
        const float ds16{sums}, rs16{lane_value(sums, 32)}, ds32{lane_value(sums, 16)}, rs32{lane_value(sums, 48)};
        if (lane == 0) {
            S.cost16[q] = 256.0f * ds16 + 6.0e-4f * rs16;
            S.cost32p[q] = 1024.0f * ds32 + 6.0e-4f * rs32;
        }
        __syncthreads();
        if (tid == 0) {
            S.cost32 = (S.cost32p[0] + S.cost32p[1]) + (S.cost32p[2] + S.cost32p[3]);
        }
    };
    // The heuristic's decision from the costs (thread 0).
    const auto decide = [&]() {
        constexpr float inf{1.0e30f}, eps{1.0e-6f};
        // A forced size (test modes): 1 DCT8, 2 DCT16, 3 DCT32 wherever it fits.
        const std::uint32_t force{(prm.mode & kModeForceMask) >> kModeForceShift};
        const bool fit32{bx0 + 4 <= bw && by0 + 4 <= bh};
        const float c32{fit32 ? S.cost32 : inf};
        float split{0.0f};
        int choice[4];
        for (int q{0}; q < 4; ++q) {
            const int qx{q & 1}, qy{q >> 1};
            float sum8{0.0f};
            bool any8{false};
            for (int dy{0}; dy < 2; ++dy) {
                for (int dx{0}; dx < 2; ++dx) {
                    if (bx0 + 2 * qx + dx < bw && by0 + 2 * qy + dy < bh) {
                        sum8 += S.cost8[q * 4 + dy * 2 + dx];
                        any8 = true;
                    }
                }
            }
            if (!any8) {
                choice[q] = -1;
                continue;
            }
            const bool fit16{bx0 + 2 * qx + 2 <= bw && by0 + 2 * qy + 2 <= bh};
            const float c16{fit16 ? S.cost16[q] : inf};
            if (force != 0 ? (force >= 2 && fit16) : (fit16 && c16 <= sum8 + eps)) {
                choice[q] = 16;
                split += c16;
            } else {
                choice[q] = 8;
                split += sum8;
            }
        }
        const bool root32{force != 0 ? (force == 3 && fit32) : (fit32 && c32 <= split + eps)};
        for (int q{0}; q < 4; ++q) {
            S.choice[q] = choice[q];
        }
        if (root32) {
            for (int b{0}; b < 16; ++b) {
                S.acs[b] = b == 0 ? 32 : 0;
            }
        } else {
            for (int q{0}; q < 4; ++q) {
                const int ox{2 * (q & 1)}, oy{2 * (q >> 1)};
                if (choice[q] == 16) {
                    S.acs[oy * 4 + ox] = 16;
                    S.acs[oy * 4 + ox + 1] = 0;
                    S.acs[(oy + 1) * 4 + ox] = 0;
                    S.acs[(oy + 1) * 4 + ox + 1] = 0;
                }
            }
        }
    };

    if constexpr (!kMf16) {
        if (wu < 2) {
            if constexpr (!kLearned) {
                rows_big();
            }
        } else {
            rows8();
        }
        __syncthreads();
        CJX_REGION_CUT_AT(7)
        if (wu >= 2) {
            cols8();
        } else if constexpr (!kLearned) {
            cols_big();
        }
    } else {
        mf_dct8_costs();
        if constexpr (!kLearned) {
            mf_big();
        }
    }
    __syncthreads();
    CJX_REGION_CUT_AT(8)
    if (want_features) {
        feature_finish(S, tid, bx0, by0, bw, bh);
    }
    bool fallback{!kLearned};
    if constexpr (kLearned) {
        fallback = !mlp_decide(S, prm, mlp_regs, lines, tid, bx0, by0, bw, bh);
        if (fallback) {
            // The heuristic for this region: the larger candidates and the
            // decision from the costs.
            if constexpr (!kMf16) {
                if (wu < 2) {
                    rows_big();
                }
                __syncthreads();
                if (wu < 2) {
                    cols_big();
                }
            } else {
                mf_big();
            }
            __syncthreads();
        }
    }
    if (fallback && tid == 0) {
        decide();
    }
    __syncthreads();
    if (feat_out != nullptr) {
        export_features(S, feat_out, region, tid, bx0, by0, bw, bh);
    }
    if (tid < 16) {
        const int bx{bx0 + (tid & 3)}, by{by0 + (tid >> 2)};
        if (bx < bw && by < bh) {
            acs_out[static_cast<std::size_t>(by) * bw + bx] = static_cast<std::int8_t>(S.acs[tid]);
        }
    }

    CJX_REGION_CUT_AT(4)
    // 5. Chosen DCTs of X and B (Y reuses the selection coefficients; a region
    //    the network decided transforms Y here too when it chose a DCT16 or
    //    DCT32) into the coefficient arrays in element order, then FP16 out
    //    and the CfL partials (frontend_quantize.cu:49-117 on this region).
    const bool need_y{kLearned && !fallback};
    const auto cxb = [&](int ch) { return ch == 0 ? S.p.coef.x : S.p.coef.b; };
    if constexpr (!kMf16) {
        const int acs0{__builtin_amdgcn_readfirstlane(S.acs[0])};
        if (acs0 == 32) {
            // Waves 0 and 1 transform X and B on the matrix cores (DCT32 as four
            // 16x16 quadrants), wave 2 Y when it has no coefficients yet.
            const int ch{wu};
            float* const cl{lines + ch * RowBuf<32>::BS};
            const bool active{wu < 2 || (need_y && wu == 2)};
            if (active) {
                const int row{16 * (grp & 1) + t16};
                const v16f acc{mfma_lines<32>(g_tables.basis32 + (16 * (grp >> 1) + t16) * 32,
                                              plane_of(S, ch == 0 ? 0 : (ch == 1 ? 2 : 1)) + row * kSP, 0,
                                              sp_swz(row))};
#pragma unroll
                for (int b{0}; b < 4; ++b) {
#pragma unroll
                    for (int u{0}; u < 4; ++u) {
                        cl[(16 * (b >> 1) + 4 * grp + u) * RowBuf<32>::RS + 16 * (b & 1) + t16] = acc[4 * b + u];
                    }
                }
            }
            __syncthreads();
            if (active) {
                const v16f acc{mfma_lines<32>(g_tables.basis32 + (16 * (grp >> 1) + t16) * 32,
                                              &cl[(16 * (grp & 1) + t16) * RowBuf<32>::RS])};
                _Float16* const dst{ch < 2 ? cxb(ch) : coefy_of(S, 2)};
#pragma unroll
                for (int b{0}; b < 4; ++b) {
#pragma unroll
                    for (int u{0}; u < 4; ++u) {
                        const int fy{16 * (b >> 1) + 4 * grp + u}, fx{16 * (b & 1) + t16};
                        dst[cpad(fx * 32 + fy)] = static_cast<_Float16>(acc[4 * b + u]);
                    }
                }
            }
        } else if (__builtin_amdgcn_readfirstlane(S.acs[2] & S.acs[8] & S.acs[10]) == 16 && acs0 == 16) {
            // Every quad a DCT16: wave c transforms channel c of the four quads
            // as one chain (quad q = block q): X, B, then Y when it has no
            // coefficients yet.
            constexpr int kChanLines{4 * RowBuf<16>::BS};
            static_assert(3 * kChanLines <= kScratch, "channel lines");
            const int nch{need_y ? 3 : 2};
            float* const cl{lines + wu * kChanLines};
            const int plane{wu == 0 ? 0 : (wu == 1 ? 2 : 1)};
            if (wu < nch) {
                const int row{(grp >> 1) * 16 + t16};
                const v16f acc{mfma_lines<16>(g_tables.basis16 + t16 * 16, plane_of(S, plane) + row * kSP,
                                              (grp & 1) * 16, sp_swz(row))};
#pragma unroll
                for (int q{0}; q < 4; ++q) {
#pragma unroll
                    for (int u{0}; u < 4; ++u) {
                        cl[q * RowBuf<16>::BS + (4 * grp + u) * RowBuf<16>::RS + t16] = acc[4 * q + u];
                    }
                }
            }
            __syncthreads();
            if (wu < nch) {
                const v16f acc{mfma_lines<16>(g_tables.basis16 + t16 * 16,
                                              &cl[grp * RowBuf<16>::BS + t16 * RowBuf<16>::RS])};
                _Float16* const dst{wu < 2 ? cxb(wu) : coefy_of(S, 1)};
#pragma unroll
                for (int q{0}; q < 4; ++q) {
#pragma unroll
                    for (int u{0}; u < 4; ++u) {
                        dst[cpad(256 * q + t16 * 16 + 4 * grp + u)] = static_cast<_Float16>(acc[4 * q + u]);
                    }
                }
            }
        } else {
            // Wave w transforms quad w: one DCT16 on the matrix cores (block 0 =
            // X, block 1 = B, block 2 = Y when it has no coefficients yet, else
            // unused; block 3 unused) or four DCT8s on the VALU (lanes:
            // channel, block, row).
            constexpr int kQuadLines{3 * RowBuf<16>::BS};
            static_assert(kQuadLines >= 2 * 4 * RowBuf<8>::BS && 4 * kQuadLines <= kScratch, "quad lines");
            float* const ql{lines + wu * kQuadLines};
            const int q{wu};
            const int n16{__builtin_amdgcn_readfirstlane(S.acs[(q >> 1) * 8 + (q & 1) * 2])};
            // Block grp's plane: X, B, then Y (or X again), B again; the
            // results of repeated blocks are dropped.
            const int ch16{grp == 2 ? (need_y ? 1 : 0) : (grp & 1) * 2};
            if (n16 == 16) {
                const int row{(q >> 1) * 16 + t16};
                const v16f acc{mfma_lines<16>(g_tables.basis16 + t16 * 16, plane_of(S, ch16) + row * kSP,
                                              (q & 1) * 16, sp_swz(row))};
#pragma unroll
                for (int b{0}; b < 3; ++b) {
#pragma unroll
                    for (int u{0}; u < 4; ++u) {
                        ql[b * RowBuf<16>::BS + (4 * grp + u) * RowBuf<16>::RS + t16] = acc[4 * b + u];
                    }
                }
            } else {
                const int ch{lane >> 5}, j{(lane >> 3) & 3}, r{lane & 7};
                const int bx{2 * (q & 1) + (j & 1)}, by{2 * (q >> 1) + (j >> 1)};
                float o[8];
                const int row{by * 8 + r};
                dct_line<8, 8>(plane_of(S, 2 * ch) + row * kSP, 0, o, bx * 8, sp_swz(row));
#pragma unroll
                for (int f{0}; f < 8; ++f) {
                    ql[(ch * 4 + j) * RowBuf<8>::BS + f * RowBuf<8>::RS + r] = o[f];
                }
            }
            __syncthreads();
            if (n16 == 16) {
                const v16f acc{mfma_lines<16>(g_tables.basis16 + t16 * 16,
                                              &ql[(grp < 3 ? grp : 1) * RowBuf<16>::BS + t16 * RowBuf<16>::RS])};
#pragma unroll
                for (int b{0}; b < 3; ++b) {
                    if (b == 2 && !need_y) {
                        break;
                    }
                    _Float16* const dst{b < 2 ? cxb(b) : coefy_of(S, 1)};
#pragma unroll
                    for (int u{0}; u < 4; ++u) {
                        dst[cpad(256 * q + t16 * 16 + 4 * grp + u)] = static_cast<_Float16>(acc[4 * b + u]);
                    }
                }
            } else {
                const int ch{lane >> 5}, j{(lane >> 3) & 3}, fx{lane & 7};
                float o[8];
                dct_line<8, 8>(&ql[(ch * 4 + j) * RowBuf<8>::BS + fx * RowBuf<8>::RS], 0, o);
#pragma unroll
                for (int f{0}; f < 8; ++f) {
                    cxb(ch)[cpad(256 * q + 64 * j + fx * 8 + f)] = static_cast<_Float16>(o[f]);
                }
            }
        }
        __syncthreads();
    } else {
        // Matrix-core FP16 form; results in registers until every wave has
        // read the planes, whose space then takes their coefficients.
        const int acs0{__builtin_amdgcn_readfirstlane(S.acs[0])};
        if (acs0 == 32) {
            // Wave w: channel w % 2 (X, B), horizontal frequencies 16 (w / 2)..;
            // waves 0-1 also Y's half w when it has no coefficients yet.
            const int ch{wu & 1}, fx0{16 * (wu >> 1)};
            f4v c[2], cy[2];
            mf_dct32(plane_of(S, 2 * ch), fx0, 0, true, lane, c);
            if (need_y && wu < 2) {
                mf_dct32(plane_of(S, 1), 16 * wu, 0, true, lane, cy);
            }
            __syncthreads();
#pragma unroll
            for (int o{0}; o < 2; ++o) {
#pragma unroll
                for (int u{0}; u < 4; ++u) {
                    const int fy{16 * o + 4 * grp + u}, fx{fx0 + t16};
                    cxb(ch)[cpad(fx * 32 + fy)] = static_cast<_Float16>(c[o][u]);
                    if (need_y && wu < 2) {
                        const int fxy{16 * wu + t16};
                        coefy_of(S, 2)[cpad(fxy * 32 + fy)] = static_cast<_Float16>(cy[o][u]);
                    }
                }
            }
        } else {
            // Wave w transforms quad w of X and B (and of Y, when it has no
            // coefficients yet and takes a DCT16): one DCT16 or four DCT8s.
            const int q{wu};
            const int n16{__builtin_amdgcn_readfirstlane(S.acs[(q >> 1) * 8 + (q & 1) * 2])};
            const int n{n16 == 16 ? 16 : 8};
            const std::uint16_t* hi{n == 16 ? g_tables.mf16[0] : g_tables.mf8[0]};
            const std::uint16_t* lo{n == 16 ? g_tables.mf16[1] : g_tables.mf8[1]};
            const int qr{(q >> 1) * 16}, qc{(q & 1) * 16};
            const f4v cx{mf_quad(plane_of(S, 0), qr, qc, hi, lo, lane)};
            const f4v cb{mf_quad(plane_of(S, 2), qr, qc, hi, lo, lane)};
            const bool y16{need_y && n == 16};
            f4v cy{};
            if (y16) {
                cy = mf_quad(plane_of(S, 1), qr, qc, hi, lo, lane);
            }
            __syncthreads();
#pragma unroll
            for (int u{0}; u < 4; ++u) {
                int j, e;
                mf_quad_element(lane, u, n, j, e);
                cxb(0)[cpad(256 * q + 64 * j + e)] = static_cast<_Float16>(cx[u]);
                cxb(1)[cpad(256 * q + 64 * j + e)] = static_cast<_Float16>(cb[u]);
                if (y16) {
                    coefy_of(S, 1)[cpad(256 * q + e)] = static_cast<_Float16>(cy[u]);
                }
            }
        }
        __syncthreads();
    }

    CJX_REGION_CUT_AT(5)
    const std::size_t cplane{static_cast<std::size_t>(geo.nblocks) * 64};
    float acc[3]{0.0f, 0.0f, 0.0f};
    // Wave w writes blocks (2 (k % 2) + w % 2, 2 (k / 2) + w / 2), k = 0..3,
    // one coefficient per lane; the block's transform and slot are
    // wave-uniform (scalar registers).
    const int acs_root{__builtin_amdgcn_readfirstlane(S.acs[0])};
#pragma unroll
    for (int k{0}; k < 4; ++k) {
        const int bxl{2 * (k & 1) + (wu & 1)}, byl{2 * (k >> 1) + (wu >> 1)};
        int n{8}, ox{bxl}, oy{byl};
        if (acs_root == 32) {
            n = 32;
            ox = 0;
            oy = 0;
        } else if (__builtin_amdgcn_readfirstlane(S.acs[(byl & ~1) * 4 + (bxl & ~1)]) == 16) {
            n = 16;
            ox = bxl & ~1;
            oy = byl & ~1;
        }
        const int qc{(byl - oy) * (n / 8) + (bxl - ox)};
        const int e{qc * 64 + lane};
        const int idx{n == 32 ? e : 256 * k + 64 * wu + lane};
        const int sz{n == 32 ? 2 : (n == 16 ? 1 : 0)};
        const _Float16 hx{cxb(0)[cpad(idx)]};
        const _Float16 hy{coefy_of(S, sz)[cpad(idx)]};
        const _Float16 hb{cxb(1)[cpad(idx)]};
        float v[3]{0.0f, 0.0f, 0.0f};
        if (bx0 + bxl < bw && by0 + byl < bh) {
            std::uint16_t* const out{coeffs + (static_cast<std::size_t>(by0 + byl) * bw + bx0 + bxl) * 64};
            out[lane] = __builtin_bit_cast(std::uint16_t, hx);
            out[cplane + lane] = __builtin_bit_cast(std::uint16_t, hy);
            out[2 * cplane + lane] = __builtin_bit_cast(std::uint16_t, hb);
            if (!is_llf(e, n)) {
                const float xr{static_cast<float>(hx)}, yr{static_cast<float>(hy)};
                const float br{static_cast<float>(hb)};
                v[0] = xr * yr;
                v[1] = yr * yr;
                v[2] = (br - yr) * yr;
            }
        }
#pragma unroll
        for (int c{0}; c < 3; ++c) {
            acc[c] = k == 0 ? v[c] : acc[c] + v[c];
        }
    }
    {
        const float sums{wave_sum4(acc[0], acc[1], acc[2], 0.0f)};  // lanes 0, 32, 16
        const float s1{lane_value(sums, 32)}, s2{lane_value(sums, 16)};
        if (lane == 0) {
            S.cfl[wu][0] = sums;
            S.cfl[wu][1] = s1;
            S.cfl[wu][2] = s2;
        }
    }
    __syncthreads();
    if (tid < 3) {
        cfl_part[static_cast<std::size_t>(region) * 3 + tid] =
            (S.cfl[0][tid] + S.cfl[1][tid]) + (S.cfl[2][tid] + S.cfl[3][tid]);
    }
}

template <bool kMf16, bool kLearned>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void region_kernel(
    Geometry geo, FrameParams prm, const std::uint8_t* __restrict__ luma, std::size_t luma_pitch,
    const std::uint8_t* __restrict__ chroma, std::size_t chroma_pitch,
    std::uint16_t* __restrict__ coeffs, std::int8_t* __restrict__ acs_out,
    std::int32_t* __restrict__ qf_out, float* __restrict__ cfl_part, std::uint32_t* __restrict__ list_count,
    Clears clr, float* __restrict__ feat_out) {
    region_kernel_body<kMf16, kLearned>(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, prm, luma, luma_pitch, chroma, chroma_pitch, coeffs, acs_out, qf_out, cfl_part, list_count, clr, feat_out);
}

// ---------------------------------------------------------------------------

struct TileShared {
    unsigned int nz[64][3];
    int own[64];              // packed owner: local position | (n << 8)
    float maps[2];
};

__device__ __forceinline__ int owner_in_tile(const std::int8_t* acs, int bw, int bx, int by, int& ox,
                                             int& oy) {
    const int rxb{bx & ~3}, ryb{by & ~3};
    if (acs[static_cast<std::size_t>(ryb) * bw + rxb] == 32) {
        ox = rxb;
        oy = ryb;
        return 32;
    }
    const int qx{bx & ~1}, qy{by & ~1};
    if (acs[static_cast<std::size_t>(qy) * bw + qx] == 16) {
        ox = qx;
        oy = qy;
        return 16;
    }
    ox = bx;
    oy = by;
    return 8;
}

// The first kernel of the DC chain: thread i derives the DC of block i
// (frontend_quantize.cu:181-199, dc_from_llf_strided: the owner transform's
// (n/8)^2 lowest frequencies, resampled to the block, from the FP16
// coefficients; then the DC quantization, B against Y's round trip) and,
// for i below the tile count, the colour maps of tile i for the AcMetadata
// (the quantizer derives its tile's maps with the same function).
__device__ __attribute__((always_inline)) inline void dc_maps_kernel_body(const Wg& wg, Geometry geo, FrameParams prm,
                                                      const std::uint16_t* __restrict__ coeffs,
                                                      const std::int8_t* __restrict__ acs,
                                                      const float* __restrict__ cfl_part, std::int32_t* __restrict__ dc,
                                                      std::int8_t* __restrict__ ytox, std::int8_t* __restrict__ ytob) {
    const int i{static_cast<int>(wg.x * blockDim.x + threadIdx.x)};
    if (i < static_cast<int>(geo.cmw * geo.cmh)) {
        float r[4][3];
        tile_partials(geo, cfl_part, i % static_cast<int>(geo.cmw), i / static_cast<int>(geo.cmw), r);
        int mx, mb;
        cfl_tile_maps(r, mx, mb);
        ytox[i] = static_cast<std::int8_t>(mx);
        ytob[i] = static_cast<std::int8_t>(mb);
    }
    if (i >= static_cast<int>(geo.nblocks)) {
        return;
    }
    const int bw{static_cast<int>(geo.bw)};
    const int bx{i % bw}, by{i / bw};
    int ox, oy;
    const int n{owner_in_tile(acs, bw, bx, by, ox, oy)};
    const int m{n / 8}, dx{bx - ox}, dy{by - oy};
    const std::size_t nblk{geo.nblocks};
    const std::size_t cplane{nblk * 64};
    float dcv[3];
    for (int c{0}; c < 3; ++c) {
        // Element e of the owner's transform sits in slot e % 64 of the
        // covered block e / 64 (row-major over the m x m blocks).
        const auto llf = [&](int e) {
            const int qc{e >> 6};
            const std::size_t blk{static_cast<std::size_t>(oy + qc / m) * bw + ox + qc % m};
            return static_cast<float>(__builtin_bit_cast(_Float16, coeffs[c * cplane + blk * 64 + (e & 63)]));
        };
        if (m == 1) {
            dcv[c] = llf(0);
        } else {
            const double* t{m == 2 ? g_tables.llf2 : g_tables.llf4};
            double sum{0.0};
            for (int fx{0}; fx < m; ++fx) {
                for (int fy{0}; fy < m; ++fy) {
                    sum += t[fx * m + dx] * t[fy * m + dy] * static_cast<double>(llf(fx * n + fy));
                }
            }
            dcv[c] = static_cast<float>(static_cast<double>(m) * sum);
        }
    }
    const float dcq_x{4096.0f * prm.dc_scale}, dcq_y{512.0f * prm.dc_scale};
    const float dcq_b{256.0f * prm.dc_scale};
    const int qdy{static_cast<int>(__builtin_rintf(dcv[1] * dcq_y))};
    const float y_dc_round{static_cast<float>(qdy) / dcq_y};
    dc[nblk + i] = qdy;
    dc[i] = static_cast<int>(__builtin_rintf(dcv[0] * dcq_x));
    dc[2 * nblk + i] = static_cast<int>(__builtin_rintf((dcv[2] - 1.0f * y_dc_round) * dcq_b));
}

__global__ __launch_bounds__(256) void dc_maps_kernel(Geometry geo, FrameParams prm,
                                                      const std::uint16_t* __restrict__ coeffs,
                                                      const std::int8_t* __restrict__ acs,
                                                      const float* __restrict__ cfl_part, std::int32_t* __restrict__ dc,
                                                      std::int8_t* __restrict__ ytox, std::int8_t* __restrict__ ytob) {
    dc_maps_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, prm, coeffs, acs, cfl_part, dc, ytox, ytob);
}

__device__ __attribute__((always_inline)) inline void tile_quantize_kernel_body(const Wg& wg, 
    Geometry geo, FrameParams prm, const std::uint16_t* __restrict__ coeffs,
    const std::int8_t* __restrict__ acs, const std::int32_t* __restrict__ qf,
    const float* __restrict__ cfl_part, std::int16_t* __restrict__ ac,
    std::uint8_t* __restrict__ nz_grid, std::uint16_t* __restrict__ nz_count, ItemList list) {
    __shared__ TileShared S;
    const int tid{static_cast<int>(threadIdx.x)};
    const int w{tid >> 6};
    const int lane{tid & 63};
    const int tile{static_cast<int>(wg.x)};
    const int tx{tile % static_cast<int>(geo.cmw)}, ty{tile / static_cast<int>(geo.cmw)};
    const int bw{static_cast<int>(geo.bw)}, bh{static_cast<int>(geo.bh)};
    const std::size_t nblk{geo.nblocks};
    const std::size_t cplane{nblk * 64};

    if (tid < 64) {
        S.nz[tid][0] = S.nz[tid][1] = S.nz[tid][2] = 0u;
        const int bx{tx * 8 + (tid & 7)}, by{ty * 8 + (tid >> 3)};
        int ox{bx}, oy{by}, n{8};
        if (bx < bw && by < bh) {
            n = owner_in_tile(acs, bw, bx, by, ox, oy);
        }
        S.own[tid] = ((oy - ty * 8) * 8 + (ox - tx * 8)) | (n << 8);
    }
    if (tid == 64) {  // wave 1, beside wave 0's owners
        float r[4][3];
        tile_partials(geo, cfl_part, tx, ty, r);
        int mx, mb;
        cfl_tile_maps(r, mx, mb);
        S.maps[0] = 0.0f + static_cast<float>(mx) * (1.0f / 84);
        S.maps[1] = 1.0f + static_cast<float>(mb) * (1.0f / 84);
    }
    __syncthreads();

    const float yx{S.maps[0]}, yb{S.maps[1]};
    // Lane l quantizes coefficients 4 (l % 16) .. 4 (l % 16) + 3 of block
    // p = 16 j + 4 (l / 16) + w: a wave takes four blocks per step, one
    // 8-byte read per plane and lane, the weights and their reciprocals as
    // 16-byte reads.  y_round = qy wy / qgsf is the correctly rounded
    // quotient, by the block's RN(1 / qgsf).
    const int grp{lane >> 4}, e4{4 * (lane & 15)};
    static_assert(offsetof(DeviceTables, dequant8) % 16 == 0 && offsetof(DeviceTables, dequant16) % 16 == 0 &&
                      offsetof(DeviceTables, dequant32) % 16 == 0 && offsetof(DeviceTables, dequant_inv8) % 16 == 0 &&
                      offsetof(DeviceTables, dequant_inv16) % 16 == 0 &&
                      offsetof(DeviceTables, dequant_inv32) % 16 == 0,
                  "16-byte weight reads");
    const auto f16_at = [](uint2 h, int i) {
        const std::uint32_t wd{i < 2 ? h.x : h.y};
        return static_cast<float>(
            __builtin_bit_cast(_Float16, static_cast<std::uint16_t>((i & 1) != 0 ? wd >> 16 : wd & 0xffffu)));
    };
    const auto f4_at = [](const float4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); };
#pragma unroll 1
    for (int j{0}; j < 4; ++j) {
        const int p{16 * j + 4 * grp + w};
        const int bx{tx * 8 + (p & 7)}, by{ty * 8 + (p >> 3)};
        const bool in{bx < bw && by < bh};
        const int own{S.own[p]};
        const int n{own >> 8}, op{own & 255};
        // Non-zeros among the lane's four: X | Y << 8 | B << 16.
        std::uint32_t cnt{0};
        if (in) {
            const int ox{tx * 8 + (op & 7)}, oy{ty * 8 + (op >> 3)};
            const int m{n >> 3};
            const int e0{((by - oy) * m + (bx - ox)) * 64 + e4};
            const std::size_t slot{(static_cast<std::size_t>(by) * bw + bx) * 64 + e4};
            const uint2 hx{*reinterpret_cast<const uint2*>(coeffs + slot)};
            const uint2 hy{*reinterpret_cast<const uint2*>(coeffs + cplane + slot)};
            const uint2 hb{*reinterpret_cast<const uint2*>(coeffs + 2 * cplane + slot)};
            const float qgsf{static_cast<float>(qf[static_cast<std::size_t>(oy) * bw + ox]) *
                             prm.global_scale_float};
            const float rq{1.0f / qgsf};
            const int cs{n * n};
            const float* const wt{n == 32 ? &g_tables.dequant32[0][0]
                                          : (n == 16 ? &g_tables.dequant16[0][0] : &g_tables.dequant8[0][0])};
            const float* const wit{n == 32 ? &g_tables.dequant_inv32[0][0]
                                           : (n == 16 ? &g_tables.dequant_inv16[0][0] : &g_tables.dequant_inv8[0][0])};
            const float4 wx4{*reinterpret_cast<const float4*>(wt + e0)};
            const float4 wy4{*reinterpret_cast<const float4*>(wt + cs + e0)};
            const float4 wb4{*reinterpret_cast<const float4*>(wt + 2 * cs + e0)};
            const float4 ix4{*reinterpret_cast<const float4*>(wit + e0)};
            const float4 iy4{*reinterpret_cast<const float4*>(wit + cs + e0)};
            const float4 ib4{*reinterpret_cast<const float4*>(wit + 2 * cs + e0)};
            std::uint32_t q0[2]{0u, 0u}, q1[2]{0u, 0u}, q2[2]{0u, 0u};
#pragma unroll
            for (int i{0}; i < 4; ++i) {
                std::int16_t a0{0}, a1{0}, a2{0};
                if (!is_llf(e0 + i, n)) {
                    const float wy{f4_at(wy4, i)};
                    const int qy{static_cast<int>(
                        __builtin_rintf(div_by_recip(f16_at(hy, i) * qgsf, wy, f4_at(iy4, i))))};
                    const float y_round{div_by_recip(static_cast<float>(qy) * wy, qgsf, rq)};
                    a0 = static_cast<std::int16_t>(static_cast<int>(__builtin_rintf(div_by_recip(
                        (f16_at(hx, i) - yx * y_round) * qgsf, f4_at(wx4, i), f4_at(ix4, i)))));
                    a1 = static_cast<std::int16_t>(qy);
                    a2 = static_cast<std::int16_t>(static_cast<int>(__builtin_rintf(div_by_recip(
                        (f16_at(hb, i) - yb * y_round) * qgsf, f4_at(wb4, i), f4_at(ib4, i)))));
                }
                const int sh{16 * (i & 1)};
                q0[i >> 1] |= static_cast<std::uint32_t>(static_cast<std::uint16_t>(a0)) << sh;
                q1[i >> 1] |= static_cast<std::uint32_t>(static_cast<std::uint16_t>(a1)) << sh;
                q2[i >> 1] |= static_cast<std::uint32_t>(static_cast<std::uint16_t>(a2)) << sh;
                cnt += (a0 != 0 ? 1u : 0u) | (a1 != 0 ? 0x100u : 0u) | (a2 != 0 ? 0x10000u : 0u);
            }
            *reinterpret_cast<uint2*>(ac + slot) = uint2{q0[0], q0[1]};
            *reinterpret_cast<uint2*>(ac + cplane + slot) = uint2{q1[0], q1[1]};
            *reinterpret_cast<uint2*>(ac + 2 * cplane + slot) = uint2{q2[0], q2[1]};
        }
        // The block's counts into lane 0 of its row of 16 lanes (row
        // shifts; lanes shifted in from outside the row read 0).
        cnt += static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(cnt), 0x101, 0xf, 0xf, true));
        cnt += static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(cnt), 0x102, 0xf, 0xf, true));
        cnt += static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(cnt), 0x104, 0xf, 0xf, true));
        cnt += static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(cnt), 0x108, 0xf, 0xf, true));
        if (in && (lane & 15) == 0) {
            atomicAdd(&S.nz[op][0], cnt & 0xffu);
            atomicAdd(&S.nz[op][1], (cnt >> 8) & 0xffu);
            atomicAdd(&S.nz[op][2], cnt >> 16);
        }
    }
    __syncthreads();

    if (tid < 64) {
        const int p{tid};
        const int bx{tx * 8 + (p & 7)}, by{ty * 8 + (p >> 3)};
        // This block's items for the tokenizer's lists, in item order Y, X,
        // B: with non-zeros (dense) or without (zero count).
        bool dense[3]{false, false, false}, zero[3]{false, false, false};
        if (bx < bw && by < bh) {
            const int own{S.own[p]};
            const int n{own >> 8}, op{own & 255};
            const int m{n / 8};
            const std::size_t b{static_cast<std::size_t>(by) * bw + bx};
            const unsigned int l2{n == 16 ? 2u : (n == 32 ? 4u : 0u)};
            for (int c{0}; c < 3; ++c) {
                const unsigned int cnt{S.nz[op][c]};
                nz_grid[c * nblk + b] = static_cast<std::uint8_t>((cnt + m * m - 1) >> l2);
                nz_count[c * nblk + b] = op == p ? static_cast<std::uint16_t>(cnt) : 0;
            }
            // Covered blocks' items have no tokens (emission reads no count
            // for them).
            if (op == p) {
#pragma unroll
                for (int j{0}; j < 3; ++j) {
                    const bool any{S.nz[op][j == 0 ? 1 : (j == 1 ? 0 : 2)] != 0};
                    dense[j] = any;
                    zero[j] = !any;
                }
            }
        }
        // Wave 0 appends the tile's items: prefix sums of the per-lane counts
        // (0..3, two ballots each), one reservation per list.
        if (list.items != nullptr) {
            const int nd{(dense[0] ? 1 : 0) + (dense[1] ? 1 : 0) + (dense[2] ? 1 : 0)};
            const int nzc{(zero[0] ? 1 : 0) + (zero[1] ? 1 : 0) + (zero[2] ? 1 : 0)};
            const std::uint64_t below{(1ull << p) - 1ull};
            const std::uint64_t d0{__ballot(nd & 1)}, d1{__ballot(nd & 2)};
            const std::uint64_t z0{__ballot(nzc & 1)}, z1{__ballot(nzc & 2)};
            const std::uint32_t dpre{static_cast<std::uint32_t>(__popcll(d0 & below) + 2 * __popcll(d1 & below))};
            const std::uint32_t zpre{static_cast<std::uint32_t>(__popcll(z0 & below) + 2 * __popcll(z1 & below))};
            const std::uint32_t dtot{static_cast<std::uint32_t>(__popcll(d0) + 2 * __popcll(d1))};
            const std::uint32_t ztot{static_cast<std::uint32_t>(__popcll(z0) + 2 * __popcll(z1))};
            std::uint32_t base{0};
            if (p < 2) {
                base = atomicAdd(&list.count[p], p == 0 ? dtot : ztot);
            }
            const std::uint32_t dbase{static_cast<std::uint32_t>(__shfl(static_cast<int>(base), 0, 64))};
            const std::uint32_t zbase{static_cast<std::uint32_t>(__shfl(static_cast<int>(base), 1, 64))};
            const std::uint32_t last{3u * geo.nblocks - 1u};
            const std::uint32_t b3{(static_cast<std::uint32_t>(by) * static_cast<std::uint32_t>(bw) +
                                    static_cast<std::uint32_t>(bx)) * 3u};
            std::uint32_t kd{dbase + dpre}, kz{zbase + zpre};
#pragma unroll
            for (int j{0}; j < 3; ++j) {
                if (dense[j]) {
                    list.items[kd++] = b3 + static_cast<std::uint32_t>(j);
                }
                if (zero[j]) {
                    list.items[last - kz++] = b3 + static_cast<std::uint32_t>(j);
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void tile_quantize_kernel(
    Geometry geo, FrameParams prm, const std::uint16_t* __restrict__ coeffs,
    const std::int8_t* __restrict__ acs, const std::int32_t* __restrict__ qf,
    const float* __restrict__ cfl_part, std::int16_t* __restrict__ ac,
    std::uint8_t* __restrict__ nz_grid, std::uint16_t* __restrict__ nz_count, ItemList list) {
    tile_quantize_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, prm, coeffs, acs, qf, cfl_part, ac, nz_grid, nz_count, list);
}

// ---------------------------------------------------------------------------
// Config-1 semantics (the reference's test path: dct.cu:43-101 forward DCT8
// over the whole frame, quant.cu:30-97 quantize_dct8): DCT8 everywhere, the
// uniform quant field, no colour-correlation maps; coefficient k of channel c
// is rintf(v * f) with f = ac_scale / w8[c][k] (ac_scale = raw quant field x
// global scale) or DC_INV_QUANT[c] x dc_scale at k = 0, and B is taken
// against Y's round trip qy / f_y.  Note the operation order: a product by a
// rounded quotient, unlike the VarDCT quantizer's quotient of a product.
// One workgroup per 32x32 region: XYB, gaborish, then per channel (Y first,
// whose quantized values B needs) a DCT8 row pass into LDS lines and a column
// pass that quantizes and stores.
__global__ __launch_bounds__(256) void dct8_uniform_kernel(
    Geometry geo, FrameParams prm, const std::uint8_t* __restrict__ luma, std::size_t luma_pitch,
    const std::uint8_t* __restrict__ chroma, std::size_t chroma_pitch, std::int16_t* __restrict__ ac,
    std::int32_t* __restrict__ dc, std::int8_t* __restrict__ acs_out, std::int32_t* __restrict__ qf_out,
    std::int8_t* __restrict__ ytox, std::int8_t* __restrict__ ytob) {
    __shared__ RegionShared S;
    const int tid{static_cast<int>(threadIdx.x)};
    const int nreg{static_cast<int>(geo.rw * geo.rh)};
    const int per_xcd{(nreg + 7) / 8};
    const int region{static_cast<int>(blockIdx.x & 7u) * per_xcd + static_cast<int>(blockIdx.x >> 3)};
    if (region >= nreg) {
        return;
    }
    const int rx{region % static_cast<int>(geo.rw)};
    const int ry{region / static_cast<int>(geo.rw)};
    const int W{static_cast<int>(geo.width)}, H{static_cast<int>(geo.height)};
    const int bw{static_cast<int>(geo.bw)}, bh{static_cast<int>(geo.bh)};
    const int bx0{rx * 4}, by0{ry * 4};
    const std::size_t nblk{geo.nblocks};
    const std::size_t cplane{nblk * 64};

    xyb_tile(S, luma, luma_pitch, chroma, chroma_pitch, rx * 32, ry * 32, W, H, tid);
    __syncthreads();
    gaborish_tile(S, tid);
    __syncthreads();

    float* const lines{S.u.scratch};                                     // 16 blocks x RowBuf<8>
    std::int32_t* const qy{reinterpret_cast<std::int32_t*>(S.u.scratch + 2048)};  // Y's quantized values
    const float ac_scale{prm.qgsf_select};
    const float dcq[3]{4096.0f * prm.dc_scale, 512.0f * prm.dc_scale, 256.0f * prm.dc_scale};
    const int b{tid >> 3}, sub{tid & 7};
    const int bx{bx0 + (b & 3)}, by{by0 + (b >> 2)};
    const bool inside{tid < 128 && bx < bw && by < bh};
    const std::size_t blk{inside ? static_cast<std::size_t>(by) * bw + bx : 0};
    for (int ci{0}; ci < 3; ++ci) {
        const int c{ci == 0 ? 1 : (ci == 1 ? 0 : 2)};
        if (tid < 128) {
            float o[8];
            const int row{(b >> 2) * 8 + sub};
            dct_line<8, 8>(plane_of(S, c) + row * kSP, 0, o, (b & 3) * 8, sp_swz(row));
#pragma unroll
            for (int f{0}; f < 8; ++f) {
                lines[b * RowBuf<8>::BS + f * RowBuf<8>::RS + sub] = o[f];
            }
        }
        __syncthreads();
        if (tid < 128) {
            float o[8];
            dct_line<8, 8>(&lines[b * RowBuf<8>::BS + sub * RowBuf<8>::RS], 0, o);
#pragma unroll
            for (int f{0}; f < 8; ++f) {
                const int e{sub * 8 + f};
                const float fc{e == 0 ? dcq[c] : ac_scale / g_tables.dequant8[c][e]};
                float v{o[f]};
                if (c == 2) {
                    const float fy{e == 0 ? dcq[1] : ac_scale / g_tables.dequant8[1][e]};
                    v = v - 1.0f * (static_cast<float>(qy[b * 64 + e]) / fy);
                }
                const int q{static_cast<int>(__builtin_rintf(v * fc))};
                if (c == 1) {
                    qy[b * 64 + e] = q;
                }
                if (inside) {
                    if (e == 0) {
                        dc[c * nblk + blk] = q;
                        ac[c * cplane + blk * 64] = 0;
                    } else {
                        ac[c * cplane + blk * 64 + e] = static_cast<std::int16_t>(q);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid < 16) {
        const int x{bx0 + (tid & 3)}, y{by0 + (tid >> 2)};
        if (x < bw && y < bh) {
            acs_out[static_cast<std::size_t>(y) * bw + x] = 8;
            qf_out[static_cast<std::size_t>(y) * bw + x] = static_cast<std::int32_t>(prm.raw_quant_field);
        }
    }
    if (tid == 0 && (rx & 1) == 0 && (ry & 1) == 0) {
        const std::size_t t{static_cast<std::size_t>(ry / 2) * geo.cmw + rx / 2};
        ytox[t] = 0;
        ytob[t] = 0;
    }
}

}  // namespace cjx
