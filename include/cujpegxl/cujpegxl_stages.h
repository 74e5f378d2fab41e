/* Stage harness for libcujpegxl_stages.so: the encode pipeline's internals,
 * exposed to parity tests and the benchmark.
 *
 * It plays the role of the reference's dump tools and round-trip wrappers
 * (the src/ dump tools, src/entropy_roundtrip.h:19-40, src/dc_entropy_roundtrip.h,
 * tools/benchmark/encode_bench.cc:462-527): run a stage on given data and read
 * its intermediate buffers back.  libcujpegxl_stages.so also exports the
 * public cujpegxl_* entry points; an encoder from either is usable here.
 */
#ifndef CUJPEGXL_CUJPEGXL_STAGES_H_
#define CUJPEGXL_CUJPEGXL_STAGES_H_

#include <stddef.h>
#include <stdint.h>

#include "cujpegxl/cujpegxl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Buffers readable after an encode (layouts as in the front end):
 *  0 coefficients  uint16 FP16 bits [3][blocks*64] covered-block layout
 *  1 transform map int8 [blocks] (8/16/32 first-block, 0 covered)
 *  2 quant field   int32 [blocks]
 *  3 CfL partials  float [regions][3]
 *  4 YtoX map      int8 [tiles]      5 YtoB map int8 [tiles]
 *  6 AC            int16 [3][blocks*64]
 *  7 DC            int32 [3][blocks]
 *  8 nz grid       uint8 [3][blocks] 9 nz count uint16 [3][blocks]
 * 10 histograms    uint32 [32 + 2*dc_groups][256]
 * 11 AC group bytes uint32 [ac_groups] 12 DC group bytes uint32 [dc_groups]
 * 13 code depths   uint8 [32 + 2*dc_groups][256]
 * 14 code words    uint16 [32 + 2*dc_groups][256] */
size_t cjx_stage_buffer_bytes(cujpegxl_encoder* encoder, int which);
int cjx_stage_fetch(cujpegxl_encoder* encoder, int which, void* host, size_t bytes);

/* Entropy coding and assembly of host-supplied coefficients (frame of the
 * encoder's configured size).  `ytox`/`ytob` may be NULL (all zero).
 * Returns a cujpegxl_status; *size receives the file size. */
int cjx_stage_encode_coefficients(cujpegxl_encoder* encoder, const int16_t* ac, const int32_t* dc,
                                  const int8_t* acs, const int32_t* quant_field, const int8_t* ytox,
                                  const int8_t* ytob, uint32_t global_scale, uint32_t quant_dc,
                                  uint8_t* out, size_t capacity, size_t* size);

/* Creates an encoder without the C ABI's more-than-one-AC-group rule, for
 * single-group frames (one combined TOC section). */
int cjx_stage_encoder_create_any(const cujpegxl_config* config, cujpegxl_encoder** out_encoder);

/* cujpegxl_encoder_encode without that rule (any multiple-of-8 size). */
int cjx_stage_encode_nv12(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* input,
                          cujpegxl_output* output);

/* Config-1 semantics on the device (the reference's dct.cu:43-101 forward
 * DCT8 and quant.cu:30-97 quantize_dct8, a test path there): DCT8 everywhere,
 * the uniform quant field, no colour-correlation maps, then the same entropy
 * coding and assembly; any multiple-of-8 size. */
int cjx_stage_encode_dct8_uniform(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* input,
                                  cujpegxl_output* output);

/* Per-stage device time of the following encodes (HIP events on the
 * encoder's stream): region, quantize, histograms, codes, sizes, emit,
 * readback.  `ms` receives 7 floats for the last encode.  `on`: 0 off,
 * 1 every stage, 2 the region kernel only (one event pair; the others read
 * 0).  Each timed event adds a few microseconds of idle device time. */
int cjx_stage_set_profiling(cujpegxl_encoder* encoder, int on);

/* Encoder mode bits (no counterpart in the reference), from the next encode:
 * CJX_MODE_MFMA16 runs the region kernel's DCT8/16/32 as FP16-operand
 * matrix-core products with FP32 accumulation (hi + lo split operands); 0 is
 * the plain path, whose files equal the oracle's. */
#define CJX_MODE_MFMA16 1u
/* Test modes: every region takes DCT8, DCT16 or DCT32 (wherever that size
 * fits) instead of the rate-distortion choice. */
#define CJX_MODE_FORCE8 0x10u
#define CJX_MODE_FORCE16 0x20u
#define CJX_MODE_FORCE32 0x30u
/* The transform map from the selection network wherever it is confident,
 * the rate-distortion heuristic elsewhere. */
#define CJX_MODE_LEARNED 2u

/* From the next encode, the region kernel also writes each region's 96
 * network features and the heuristic's 12 labels (108 floats per region,
 * regions in raster order; buffer 15 of cjx_stage_fetch). */
int cjx_stage_export_features(cujpegxl_encoder* encoder, int on);

/* The selection network's weights on `device` for every later encode there:
 * FP16 bit patterns of W1 [64][96] and W2 [7][64], FP32 biases b1 [64] and
 * b2 [7], the confidence threshold as a logit magnitude. */
int cjx_mlp_set_weights(int32_t device, const uint16_t* w1, const float* b1, const uint16_t* w2, const float* b2,
                        float threshold);
/* Back to the weights the library was built with. */
int cjx_mlp_default_weights(int32_t device);

/* Batches of frames of uneven sizes and distances (no counterpart in the
 * reference, whose encoder takes one frame at a time): every stage runs as
 * one launch over all the batch's frames, the batch's 14 launches are
 * captured once as a hipGraph and replayed, and the launches' shapes come
 * from an autotuner (cjx_batch_tune).  Files land in page-locked host memory
 * the library owns (cjx_batch_result gives each file's address and size,
 * valid until the next run).  `mode`: 0 or CJX_MODE_MFMA16 / CJX_MODE_LEARNED
 * bits, the same for every frame. */
typedef struct cjx_batch cjx_batch;
int cjx_batch_create(int32_t device, uint32_t count, const uint32_t* width, const uint32_t* height,
                     const float* distance, uint32_t mode, cjx_batch** out);
/* Each frame's NV12 device planes (an array of `count` inputs). */
int cjx_batch_set_inputs(cjx_batch* batch, const cujpegxl_nv12_input* inputs);
/* `count` replays of the batch's graph back to back, then a wait; the wall
 * seconds of the whole into *seconds. */
int cjx_batch_run(cjx_batch* batch, int32_t count, double* seconds);
/* `count` replays queued on the batch's own stream without a wait (another
 * batch's replays may then run beside them), and the wait for them; results
 * are valid after cjx_batch_wait. */
int cjx_batch_launch(cjx_batch* batch, int32_t count);
int cjx_batch_wait(cjx_batch* batch);
int cjx_batch_result(const cjx_batch* batch, uint32_t index, const uint8_t** data, uint64_t* size);
/* Times candidate launch shapes over `reps` replays and keeps the fastest:
 * tokenizer workgroups per frame x output words per assembly thread with one
 * workgroup per region, then the region stage as a persistent grid (one and
 * two resident rounds) with the best of those; one line per candidate, then
 * the choice, into `report`. */
int cjx_batch_tune(cjx_batch* batch, int32_t reps, char* report, size_t cap);
/* Sets a launch shape directly: tokenizer workgroups per frame (>= 1),
 * output words per assembly thread (>= 1), and region_wgs = 0 for one
 * workgroup per region or the size of a persistent region grid (a multiple
 * of 8, at most 65,536). */
int cjx_batch_set_shape(cjx_batch* batch, uint32_t token_cap, uint32_t assemble_words, uint32_t region_wgs);
void cjx_batch_destroy(cjx_batch* batch);

/* Training on the device (all pointers device memory, `stream` a hipStream_t
 * as an integer): cjx_mlp_step computes the mean loss (loss[0]) and its
 * gradient (grad, 6,663 floats in the parameter layout W1, b1, W2, b2) over
 * n regions (features x [n][96], labels [n][12]) into a partial buffer of
 * cjx_mlp_partial_floats(n) floats, and with lr > 0 takes one Adam step
 * (moments m, v; step counts from 1) on params.  cjx_mlp_forward writes the
 * outputs z [n][7]. */
size_t cjx_mlp_partial_floats(uint32_t n);
int cjx_mlp_step(const float* x, const float* labels, uint32_t n, float lambda, float* partial, float* grad,
                 float* loss, float* params, float* m, float* v, uint32_t step, float lr, uintptr_t stream);
int cjx_mlp_forward(const float* params, const float* x, uint32_t n, float* z, uintptr_t stream);
int cjx_stage_set_mode(cujpegxl_encoder* encoder, uint32_t mode);
int cjx_stage_timings(cujpegxl_encoder* encoder, float* ms, int count);

/* The region kernel's AQ terms divide and take square roots with the
 * compiler's correctly rounded sequences stripped of their operand scaling
 * and special-case steps (identities for the terms' moderate operands).
 * Counts, on the device, the results of those forms that differ from the
 * compiler's full forms over every float input of the terms' domains;
 * *mismatches = 0 is the expected answer. */
int cjx_stage_check_aq_math(int32_t device_ordinal, uint64_t* mismatches);

/* `count` synchronous cujpegxl_encoder_encode calls on the same input and
 * output, back to back, timed on the host around the whole loop (wall
 * seconds into *seconds).  No counterpart in the reference's library: the
 * native-loop counterpart of its C++ benchmark driver
 * (tools/benchmark/encode_bench.cc:462-470, 514-523), for callers that
 * time the encoder without an interpreter's per-call overhead.  Returns the
 * first failing call's status. */
int cjx_stage_encode_repeat(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* input,
                            cujpegxl_output* output, int count, double* seconds);

/* `count` synchronous encodes cycling through `n_inputs` frames (encode i
 * reads inputs[i % n_inputs]) into the same output, timed like
 * cjx_stage_encode_repeat: the reference's benchmark alternates its inputs
 * so that cache effects do not follow one frame
 * (tools/benchmark/encode_bench.cc:512-518).  Returns the first failing
 * call's status. */
int cjx_stage_encode_cycle(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* inputs, int n_inputs,
                           cujpegxl_output* output, int count, double* seconds);


/* Prefix codes of `count` (<= 96) histograms of 256 uint32 bins, built on
 * device `device_ordinal` by the encoder's code kernel (prefix_code.cc:36-397
 * semantics): per histogram the alphabet size, 256 depths, 256 LSB-first
 * codes, the header (64 words, LSB-first) and its length in bits. */
int cjx_stage_build_codes(int32_t device_ordinal, const uint32_t* hists, uint32_t count, uint32_t* alpha,
                          uint8_t* depth, uint16_t* bits, uint32_t* hdr_words, uint32_t* hdr_bits);

/* Test hook for the failure paths: the n-th device allocation an encoder
 * makes from now on (create, configure) fails as if the device were out of
 * memory; 0 cancels.  No counterpart in the reference. */
void cjx_stage_fail_allocation_after(int n);

/* Device memory for harnesses that have no other allocator. */
void* cjx_device_alloc(size_t bytes);
void cjx_device_free(void* ptr);
/* A HIP stream (as an integer, default priority) and ordered copies and
 * fills on it, for tests of the asynchronous entry point. */
uintptr_t cjx_stream_create(void);
void cjx_stream_destroy(uintptr_t stream);
int cjx_stream_synchronize(uintptr_t stream);
int cjx_copy_async(void* dst, const void* src, size_t bytes, uintptr_t stream);
int cjx_fill_async(void* dst, int value, size_t bytes, uintptr_t stream);
/* Page-locked host memory (the library DMAs straight into or out of it). */
void* cjx_host_alloc(size_t bytes);
void cjx_host_free(void* ptr);
int cjx_copy_to_device(void* dst, const void* src, size_t bytes);
int cjx_copy_to_host(void* dst, const void* src, size_t bytes);
int cjx_device_synchronize(void);
/* Device-to-device copy (a plain copy kernel of known size, used to calibrate
 * the HBM byte counters against the encoder's kernels). */
int cjx_device_copy(void* dst, const void* src, size_t bytes);

/* Constant tables as the device uses them, for table parity tests:
 * 0 dequant weights float [3*(64+256+1024)], 1 DCT bases float [64+256+1024],
 * 2 natural orders uint16 [64+256+1024].  Host-side; needs no device. */
size_t cjx_host_table(int which, void* out, size_t capacity);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CUJPEGXL_CUJPEGXL_STAGES_H_ */
