// Test infrastructure only: a C-ABI shim over the reference's own host C++
// sources (bitstream writer, host VarDCT encoder, selection/CfL/calibration
// headers), compiled from where they lie under the reference tree into
// oracle/_ref/libjxlref.so by oracle/Makefile.  It is the checker that pins
// the C restatement in oracle/jxl_oracle.c; nothing in the product links it.
//
// Every function below forwards to one reference function and flattens its
// C++ containers into caller-provided buffers.

#include <cstdint>
#include <cstring>
#include <vector>

#include "src/ac_context.h"
#include "src/bitstream/bit_writer.h"
#include "src/bitstream/container.h"
#include "src/bitstream/context_map.h"
#include "src/bitstream/frame_assembly.h"
#include "src/bitstream/histogram_writer.h"
#include "src/bitstream/prefix_code.h"
#include "src/cfl.h"
#include "src/coeff_order.h"
#include "src/quant_calibration.h"
#include "src/quant_weights_dct16.h"
#include "src/quant_weights_dct32.h"
#include "src/quant_weights_dct8.h"
#include "src/transform_select_impl.cuh"
#include "src/vardct_layout.h"
#include "tools/bitstream/vardct_frame.h"

using namespace cujpegxl;
using namespace cujpegxl::bitstream;

namespace {

// Copies `src` into `dst` (capacity `cap`) and returns the full size, so a
// caller can size a second call when `cap` was too small.
size_t put_bytes(const std::vector<std::uint8_t>& src, std::uint8_t* dst, size_t cap) {
    if (dst != nullptr) {
        std::memcpy(dst, src.data(), src.size() < cap ? src.size() : cap);
    }
    return src.size();
}

FrameCoefficients make_fc(uint32_t w, uint32_t h, uint32_t gs, uint32_t qdc, uint32_t rqf,
                          const int32_t* dc, const int32_t* ac, const int8_t* acs,
                          const int8_t* ytox, const int8_t* ytob) {
    FrameCoefficients fc{};
    fc.width = w;
    fc.height = h;
    fc.global_scale = gs;
    fc.quant_dc = qdc;
    fc.raw_quant_field = rqf;
    const size_t blocks{(w / 8u) * (h / 8u)};
    for (int c{0}; c < 3; ++c) {
        fc.dc[c].assign(dc + c * blocks, dc + (c + 1) * blocks);
        fc.ac[c].assign(ac + c * blocks * 64, ac + (c + 1) * blocks * 64);
    }
    if (acs != nullptr) {
        fc.acs.assign(acs, acs + blocks);
    }
    if (ytox != nullptr && ytob != nullptr) {
        const size_t tiles{((w / 8u + 7) / 8) * ((h / 8u + 7) / 8)};
        fc.ytox_map.assign(ytox, ytox + tiles);
        fc.ytob_map.assign(ytob, ytob + tiles);
    }
    return fc;
}

}  // namespace

extern "C" {

void ref_natural_coeff_order(uint32_t n, uint32_t* out) {
    const std::vector<std::uint32_t> o{natural_coeff_order(n)};
    std::memcpy(out, o.data(), o.size() * sizeof(uint32_t));
}

void ref_calibrate(float d, uint32_t* ints3, float* floats2) {
    const QuantCalibration c{calibrate_quant(d)};
    ints3[0] = c.global_scale;
    ints3[1] = c.quant_dc;
    ints3[2] = c.raw_quant_field;
    floats2[0] = c.global_scale_float;
    floats2[1] = c.inv_global_scale;
}

void ref_dequant_tables(float* w8, float* w16, float* w32) {
    std::memcpy(w8, DCT8_DEQUANT_WEIGHTS, sizeof(DCT8_DEQUANT_WEIGHTS));
    std::memcpy(w16, DCT16_DEQUANT_WEIGHTS, sizeof(DCT16_DEQUANT_WEIGHTS));
    std::memcpy(w32, DCT32_DEQUANT_WEIGHTS, sizeof(DCT32_DEQUANT_WEIGHTS));
}

// Whole-frame host selection: decide_region over every 32x32 region.
void ref_select_transforms(const float* y, uint32_t w, uint32_t h, float distance, int8_t* acs) {
    const size_t bw{w / 8u};
    const size_t bh{h / 8u};
    const QuantCalibration cal{calibrate_quant(distance)};
    const float qgsf{static_cast<float>(cal.raw_quant_field) * cal.global_scale_float};
    for (size_t ry{0}; ry < (bh + 3) / 4; ++ry) {
        for (size_t rx{0}; rx < (bw + 3) / 4; ++rx) {
            decide_region(y, w, bw, bh, rx * 4, ry * 4, qgsf, TRANSFORM_SELECT_LAMBDA, acs);
        }
    }
}

float ref_block_ac_cost(const float* y, uint32_t w, uint32_t px0, uint32_t py0, int n, float qgsf) {
    return block_ac_cost(y, w, px0, py0, n, qgsf, TRANSFORM_SELECT_LAMBDA);
}

void ref_cfl_maps_from_sums(double sxy, double syy, double sby, int* mx, int* mb) {
    cfl_maps_from_sums(sxy, syy, sby, mx, mb);
}

void ref_dc_from_llf(uint32_t m, const float* llf, float* dc) {
    dc_from_llf_strided(m, llf, m, dc);
}

int ref_ac_cluster(uint32_t ctx) {
    return ac_cluster(ctx);
}

uint32_t ref_zero_density_context(uint32_t left, uint32_t k, uint32_t cb, uint32_t l2,
                                  uint32_t prev) {
    return ac_zero_density_context(left, k, cb, l2, prev);
}

uint32_t ref_nonzero_context(uint32_t predicted, int block_ctx) {
    return ac_nonzero_context(predicted, block_ctx);
}

int ref_block_context(int c, int ord) {
    return ac_block_context(c, ord);
}

// build_and_store_huffman_tree: depth/bits (len entries) plus the header bits.
size_t ref_huffman(const uint32_t* hist, uint32_t len, uint8_t* depth, uint16_t* bits,
                   uint8_t* header, size_t cap, size_t* header_bits) {
    BitWriter w{};
    build_and_store_huffman_tree(hist, len, depth, bits, w);
    *header_bits = w.bits_written();
    return put_bytes(w.bytes(), header, cap);
}

size_t ref_best_context_map(const uint8_t* map, uint32_t n, uint32_t nclusters, uint8_t* out,
                            size_t cap, size_t* nbits, int* encoding) {
    BitWriter w{};
    *encoding = static_cast<int>(write_best_prefix_context_map(w, map, n, nclusters));
    *nbits = w.bits_written();
    return put_bytes(w.bytes(), out, cap);
}

void ref_move_to_front(const uint8_t* in, uint32_t n, uint8_t* out) {
    move_to_front_transform(in, n, out);
}

size_t ref_write_codestream(uint32_t w, uint32_t h, uint32_t gs, uint32_t qdc, uint32_t rqf,
                            const int32_t* dc, const int32_t* ac, const int8_t* acs,
                            const int8_t* ytox, const int8_t* ytob, int clustered, uint8_t* out,
                            size_t cap) {
    const FrameCoefficients fc{make_fc(w, h, gs, qdc, rqf, dc, ac, acs, ytox, ytob)};
    return put_bytes(write_vardct_codestream(fc, clustered != 0), out, cap);
}

// reference_ac_encode: histogram (32*256), depth/bits (32*256), group streams
// concatenated into `out` with per-group sizes.
size_t ref_ac_encode(uint32_t w, uint32_t h, const int32_t* dc, const int32_t* ac,
                     const int8_t* acs, uint32_t* hist, uint8_t* depth, uint16_t* bits,
                     uint32_t* group_sizes, uint8_t* out, size_t cap) {
    const FrameCoefficients fc{make_fc(w, h, 4096, 32, 32, dc, ac, acs, nullptr, nullptr)};
    const AcReference ref{reference_ac_encode(fc)};
    std::memcpy(hist, ref.histogram.data(), ref.histogram.size() * sizeof(uint32_t));
    std::memcpy(depth, ref.depth.data(), ref.depth.size());
    std::memcpy(bits, ref.bits.data(), ref.bits.size() * sizeof(uint16_t));
    std::vector<std::uint8_t> all{};
    for (size_t g{0}; g < ref.group_streams.size(); ++g) {
        group_sizes[g] = static_cast<uint32_t>(ref.group_streams[g].size());
        all.insert(all.end(), ref.group_streams[g].begin(), ref.group_streams[g].end());
    }
    return put_bytes(all, out, cap);
}

// reference_dc_encode: group sections concatenated with sizes.
size_t ref_dc_encode(uint32_t w, uint32_t h, uint32_t rqf, const int32_t* dc, const int8_t* acs,
                     const int8_t* ytox, const int8_t* ytob, uint32_t* group_sizes, uint8_t* out,
                     size_t cap) {
    const size_t blocks{(w / 8u) * (h / 8u)};
    const std::vector<int32_t> ac(3 * blocks * 64, 0);
    const FrameCoefficients fc{make_fc(w, h, 4096, 32, rqf, dc, ac.data(), acs, ytox, ytob)};
    const DcReference ref{reference_dc_encode(fc)};
    std::vector<std::uint8_t> all{};
    for (size_t g{0}; g < ref.groups.size(); ++g) {
        group_sizes[g] = static_cast<uint32_t>(ref.groups[g].section.size());
        all.insert(all.end(), ref.groups[g].section.begin(), ref.groups[g].section.end());
    }
    return put_bytes(all, out, cap);
}

size_t ref_dc_global(uint32_t gs, uint32_t qdc, uint8_t* out, size_t cap) {
    return put_bytes(build_dc_global(QuantParams{gs, qdc}), out, cap);
}

size_t ref_ac_global(const uint32_t* hist, uint32_t num_ac_groups, uint8_t* depth, uint16_t* bits,
                     uint8_t* out, size_t cap) {
    const AcGlobalResult r{build_ac_global(hist, num_ac_groups)};
    std::memcpy(depth, r.depth.data(), r.depth.size());
    std::memcpy(bits, r.bits.data(), r.bits.size() * sizeof(uint16_t));
    return put_bytes(r.section, out, cap);
}

size_t ref_codestream_head(uint32_t w, uint32_t h, const uint32_t* sizes, uint32_t n, uint8_t* out,
                           size_t cap) {
    const std::vector<std::uint32_t> s(sizes, sizes + n);
    return put_bytes(build_codestream_head(w, h, s), out, cap);
}

size_t ref_container_framing(uint64_t codestream_size, uint8_t* out, size_t cap) {
    return put_bytes(container_framing(codestream_size), out, cap);
}

}  // extern "C"
