// Host-callable launchers for the device kernels (device.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.h"

namespace cjx {

struct Constants;
struct CodeSet;
struct Layout;
struct DcBlobs;

std::size_t sizeof_constants();
std::size_t sizeof_codeset();
std::size_t sizeof_layout();
std::size_t sizeof_dc_blobs();
std::uint32_t dc_chunks_per_group();
std::uint32_t blob_words_per_group();
// Host view of Constants to fill before upload.
void fill_constants(void* dst, const std::uint32_t* head, std::uint32_t head_bits,
                    const std::uint32_t* dc_global, std::uint32_t dc_global_bits,
                    const std::uint32_t* ac_prefix, std::uint32_t ac_prefix_bits,
                    const std::uint32_t* dc_pre, std::uint32_t dc_pre_bits,
                    const std::uint32_t* dc_mid, std::uint32_t dc_mid_bits);
bool constants_fit(std::uint32_t head_bits, std::uint32_t ac_prefix_bits, std::uint32_t dc_pre_bits,
                   std::uint32_t dc_mid_bits);

// A launcher's `start`/`stop`/`done` events, when given, are bound to the
// kernel itself (its start or completion), with no marker packet on the
// stream between it and its neighbours.
void launch_region(hipStream_t s, const Geometry& geo, const FrameParams& prm, const std::uint8_t* luma,
                   std::size_t luma_pitch, const std::uint8_t* chroma, std::size_t chroma_pitch,
                   std::uint16_t* coeffs, std::int8_t* acs, std::int32_t* qf, float* cfl_part,
                   std::uint32_t* list_count, const Clears& clr, hipEvent_t start = nullptr,
                   hipEvent_t stop = nullptr, float* features = nullptr);
// The selection network's weights for the region kernel (FP16 W1 [64][96]
// and W2 [7][64] bit patterns, FP32 biases, the confidence threshold as a
// logit magnitude).
bool set_mlp_weights(const std::uint16_t* w1, const float* b1, const std::uint16_t* w2, const float* b2,
                     float threshold);
// The weights the library was built with (csrc/mlp_weights.inc).
bool set_default_mlp_weights();
// Training: partial-gradient floats for a batch of n regions; one step
// (gradient of the batch's mean loss into `grad`, the loss into `loss[0]`,
// and with lr > 0 an Adam step on params / m / v); forward outputs [n][7].
std::size_t mlp_partial_floats(std::uint32_t n);
void launch_mlp_step(hipStream_t s, const float* x, const float* labels, std::uint32_t n, float lambda,
                     float* partial, float* grad, float* loss, float* params, float* m, float* v, std::uint32_t step,
                     float lr);
void launch_mlp_forward(hipStream_t s, const float* params, const float* x, std::uint32_t n, float* z);
// Config-1 semantics: DCT8 everywhere, uniform quant field, no CfL maps.
void launch_dct8_uniform(hipStream_t s, const Geometry& geo, const FrameParams& prm, const std::uint8_t* luma,
                         std::size_t luma_pitch, const std::uint8_t* chroma, std::size_t chroma_pitch,
                         std::int16_t* ac, std::int32_t* dc, std::int8_t* acs, std::int32_t* qf, std::int8_t* ytox,
                         std::int8_t* ytob);
void launch_tile_quantize(hipStream_t s, const Geometry& geo, const FrameParams& prm,
                          const std::uint16_t* coeffs, const std::int8_t* acs, const std::int32_t* qf,
                          const float* cfl_part, std::int16_t* ac, std::uint8_t* nz_grid,
                          std::uint16_t* nz_count, const ItemList& list);
// LLF -> DC of every block from the FP16 coefficients, and the 64x64
// tiles' colour maps from the regions' CfL partials.
void launch_dc_maps(hipStream_t s, const Geometry& geo, const FrameParams& prm, const std::uint16_t* coeffs,
                    const std::int8_t* acs, const float* cfl_part, std::int32_t* dc, std::int8_t* ytox,
                    std::int8_t* ytob);
void launch_nz(hipStream_t s, const Geometry& geo, const std::int16_t* ac, const std::int8_t* acs,
               std::uint8_t* nz_grid, std::uint16_t* nz_count, const Clears& clr);
void launch_fill_u32(hipStream_t s, std::uint32_t* p, std::size_t n, std::uint32_t v);
// Mismatching results of the AQ terms' moderate-operand division and square
// root against the compiler's correctly rounded forms, over every float
// input of their domains (check_aq_math_kernel); false on a HIP error.
bool check_aq_math(unsigned long long* mismatches);
void launch_ac_tokens(hipStream_t s, const Geometry& geo, const std::int16_t* ac, const std::int8_t* acs,
                      const std::uint8_t* nz_grid, const std::uint16_t* nz_count, std::uint32_t* tokens,
                      std::uint16_t* item_tokens, std::uint32_t* partial, std::uint32_t* hist,
                      const ItemList& list);
std::size_t ac_partial_words();
void launch_dc_hist(hipStream_t s, const Geometry& geo, const std::int32_t* dc, const std::int8_t* acs,
                    const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                    std::uint32_t* dc_hist, std::uint32_t* meta_hist, std::uint32_t* fb_count);
// Builds the codes of `count` host histograms on the current device and copies
// the results back (stage harness; allocates).
bool build_codes_host(const std::uint32_t* hists, std::uint32_t count, std::uint32_t* alpha, std::uint8_t* depth,
                      std::uint16_t* bits, std::uint32_t* hdr_words, std::uint32_t* hdr_bits);
void launch_huffman(hipStream_t s, int first, int count, const std::uint32_t* hists, CodeSet* cs);
// Words of one AC group's staging slot (bound on its token stream).
std::size_t ac_group_slot_words();
// AC groups are emitted in this many parts (row bands).
std::size_t ac_group_parts();
void launch_ac_group_emit(hipStream_t s, const Geometry& geo, const std::int8_t* acs,
                          const std::uint16_t* item_tokens, const std::uint32_t* tokens, const CodeSet* cs,
                          std::uint32_t* staging, std::uint64_t* part_bits, hipEvent_t done = nullptr);
// Words of one DC group's staging slot.
std::size_t dc_group_slot_words();
// The DC chain's sizing: header blobs and chunk bits in one launch, then
// the per-group scan of chunk offsets (which adds the blobs' bits).
void launch_dc_sizes(hipStream_t s, const Geometry& geo, const Constants* k, const CodeSet* cs,
                     const std::uint32_t* fb_count, DcBlobs* blobs, std::uint32_t* blob_words, const std::int32_t* dc,
                     const std::int8_t* acs, const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                     std::uint64_t* chunk_bits, std::uint32_t* group_bytes, std::uint64_t* group_bits,
                     std::uint32_t* staging);
void launch_layout(hipStream_t s, const Geometry& geo, const Constants* k, const CodeSet* cs,
                   const std::uint64_t* ac_part_bits, const std::uint32_t* dc_group_bytes,
                   const std::uint64_t* dc_group_bits, std::uint32_t* ac_group_bytes, std::uint64_t* ac_group_bits,
                   const DcBlobs* blobs, Layout* lay, std::uint64_t* ac_base, std::uint64_t* dc_base,
                   std::uint64_t* total_out, hipEvent_t done = nullptr);
void launch_assemble(hipStream_t s, const Geometry& geo, const Constants* k, const Layout* lay,
                     const std::uint32_t* dc_staging, const std::uint32_t* ac_staging, const std::uint64_t* dc_bits,
                     const std::uint64_t* ac_bits, const std::uint64_t* ac_part_bits, const std::uint64_t* dc_base,
                     const std::uint64_t* ac_base, std::uint8_t* out, std::uint64_t cap_bytes,
                     hipEvent_t done = nullptr);
void launch_dc_emit(hipStream_t s, const Geometry& geo, const std::int32_t* dc, const std::int8_t* acs,
                    const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                    const CodeSet* cs, const DcBlobs* blobs, const std::uint32_t* blob_words,
                    const std::uint64_t* chunk_off, std::uint32_t* staging, hipEvent_t done = nullptr);

}  // namespace cjx
