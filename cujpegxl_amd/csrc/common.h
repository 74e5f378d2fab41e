// Shared definitions for the gfx950 VarDCT encoder: frame geometry, the
// per-frame parameter block, and the constant tables every kernel reads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "numeric.h"

namespace cjx {

// Geometry (vardct_layout.h, entropy.cu:21-22, frontend_quantize.cu CfL tiles):
// 8x8 blocks; 32x32 selection regions (4x4 blocks); 64x64 CfL tiles (8x8
// blocks); 256x256 AC groups (32x32 blocks); 2048x2048 DC groups (256x256).
constexpr int kRegionBlocks = 4;
constexpr int kTileBlocks = 8;
constexpr int kAcGroupBlocks = 32;
constexpr int kDcGroupBlocks = 256;
constexpr int kHistStride = 256;  // hybrid-uint token alphabet bound
constexpr int kNumClusters = 32;  // fixed AC context clustering
constexpr int kNumAcContexts = 15 * (37 + 458);

// Buffers the entropy stage accumulates into, zeroed by the kernel that
// precedes it (the quantizer, or the non-zero counter for host-supplied
// coefficients) instead of by fill launches of their own.
struct Clears {
    std::uint32_t* ptr[4];
    std::uint32_t words[4];
};

// The AC items (block, channel) of first blocks, in no particular order,
// listed by the quantizer (one reservation per workgroup and list): those
// with non-zero coefficients from the front of `items`, those without (whose
// one token is the zero count) from the back.  Covered blocks' items have no
// tokens.  The tokenizer walks the two lists only.
// `count[0]` and `count[1]` (front, back) are zeroed by the region kernel.
struct ItemList {
    std::uint32_t* items;  // item = block * 3 + (0: Y, 1: X, 2: B); 3 * blocks entries
    std::uint32_t* count;
};

// A workgroup's coordinates in its launch: blockIdx and gridDim of a frame's
// own launch, or the frame's share of a launch over a batch of frames.
struct Wg {
    std::uint32_t x, y, nx, ny;
};

struct Geometry {
    std::uint32_t width, height;
    std::uint32_t bw, bh;        // 8x8 blocks
    std::uint32_t rw, rh;        // 32x32 regions
    std::uint32_t cmw, cmh;      // 64x64 CfL tiles
    std::uint32_t xg, yg;        // AC groups
    std::uint32_t xdg, ydg;      // DC groups
    std::uint32_t nblocks;
    std::uint32_t num_ac, num_dc;

    static Geometry of(std::uint32_t w, std::uint32_t h) {
        Geometry g{};
        g.width = w;
        g.height = h;
        g.bw = w / 8;
        g.bh = h / 8;
        g.rw = (g.bw + 3) / 4;
        g.rh = (g.bh + 3) / 4;
        g.cmw = (g.bw + 7) / 8;
        g.cmh = (g.bh + 7) / 8;
        g.xg = (g.bw + 31) / 32;
        g.yg = (g.bh + 31) / 32;
        g.xdg = (g.bw + 255) / 256;
        g.ydg = (g.bh + 255) / 256;
        g.nblocks = g.bw * g.bh;
        g.num_ac = g.xg * g.yg;
        g.num_dc = g.xdg * g.ydg;
        return g;
    }
};

// Distance-derived constants (quant_calibration.h:49-88 and
// frontend_fused.cu:214-239), computed once on the host per configuration.
struct FrameParams {
    std::uint32_t global_scale;
    std::uint32_t quant_dc;
    std::uint32_t raw_quant_field;
    float global_scale_float;
    float inv_global_scale;
    float qgsf_select;  // raw_quant_field * global_scale_float
    float inv_qgsf_select;  // RN(1 / qgsf_select)
    float km[4];        // fuzzy-erosion weights
    float mul_pb, add_pb;
    float dc_scale;     // global_scale_float * quant_dc
    std::uint32_t mode;  // kMode* bits of the encoder (0: the plain path)
};

// Encoder modes (FrameParams::mode).
// kModeMfma16: the region kernel's DCTs as FP16-operand matrix products with
// FP32 accumulation and a hi + lo split of every operand.
constexpr std::uint32_t kModeMfma16 = 1u;
// kModeForce*: every region takes the largest transform of that size that
// fits instead of the RD choice (DCT error reports per transform size).
constexpr std::uint32_t kModeForceShift = 4;
constexpr std::uint32_t kModeForceMask = 3u << kModeForceShift;
constexpr std::uint32_t kModeForce8 = 1u << kModeForceShift;
constexpr std::uint32_t kModeForce16 = 2u << kModeForceShift;
constexpr std::uint32_t kModeForce32 = 3u << kModeForceShift;
// kModeLearned: the transform map from the selection network where it is
// confident, the heuristic elsewhere.
constexpr std::uint32_t kModeLearned = 2u;
constexpr std::uint32_t kModeMask = kModeMfma16 | kModeLearned | kModeForceMask;

FrameParams make_frame_params(float distance);

// Constant tables (filled once per device by init_device_tables()).
struct DeviceTables {
    float basis8[64];
    float basis16[256];
    float basis32[1024];
    float wsel8[64];       // DCT8 Y weight sampled at each DCTn frequency
    float wsel16[256];
    float wsel32[1024];
    float dequant8[3][64];
    float dequant16[3][256];
    float dequant32[3][1024];
    float wsel_inv8[64];   // RN(1 / weight) of the tables above
    float wsel_inv16[256];
    float wsel_inv32[1024];
    float dequant_inv8[3][64];
    float dequant_inv16[3][256];
    float dequant_inv32[3][1024];
    std::uint16_t order8[64];     // natural scan order, raw index per scan position
    std::uint16_t order16[256];
    std::uint16_t order32[1024];
    double llf2[4];        // T[f][o] = ortho_2(f, o) * resample_scale_2[f]
    double llf4[16];
    float rate[256];       // selection rate term 1 + log2(1 + k) for |q| = k >= 1; 0 for k = 0
    float2 xyb_tab[kXybTabPairs];  // {power, segment reciprocal} pairs of pow24_tab / cbrt_tab
    // FP16 operands of the matrix-core DCT form: [0] = RN16(a), [1] =
    // RN16(a - [0]) of every basis entry a (FP16 bit patterns, row-major):
    // blockdiag(A8, A8) (a 16x16 quad of four DCT8s), A16 and A32.
    std::uint16_t mf8[2][256];
    std::uint16_t mf16[2][256];
    std::uint16_t mf32[2][1024];
};

extern __constant__ __align__(16) DeviceTables g_tables;

// Host mirror of the tables (same values), built once.
const DeviceTables& host_tables();
bool init_device_tables();

__device__ __forceinline__ const float* basis_of(int n) {
    return n == 32 ? g_tables.basis32 : (n == 16 ? g_tables.basis16 : g_tables.basis8);
}
__device__ __forceinline__ const float* wsel_of(int n) {
    return n == 32 ? g_tables.wsel32 : (n == 16 ? g_tables.wsel16 : g_tables.wsel8);
}
__device__ __forceinline__ float dequant_of(int n, int c, int e) {
    return n == 32 ? g_tables.dequant32[c][e]
                   : (n == 16 ? g_tables.dequant16[c][e] : g_tables.dequant8[c][e]);
}
__device__ __forceinline__ const float* wsel_inv_of(int n) {
    return n == 32 ? g_tables.wsel_inv32 : (n == 16 ? g_tables.wsel_inv16 : g_tables.wsel_inv8);
}
__device__ __forceinline__ float dequant_inv_of(int n, int c, int e) {
    return n == 32 ? g_tables.dequant_inv32[c][e]
                   : (n == 16 ? g_tables.dequant_inv16[c][e] : g_tables.dequant_inv8[c][e]);
}
__device__ __forceinline__ const std::uint16_t* order_of(int n) {
    return n == 32 ? g_tables.order32 : (n == 16 ? g_tables.order16 : g_tables.order8);
}

}  // namespace cjx
