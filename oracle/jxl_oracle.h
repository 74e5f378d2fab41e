/* Test infrastructure only: the CPU oracle for the VarDCT encode path.
 *
 * A plain-C restatement of the reference encoder (m-pilia/cujpegxl) used by
 * tests/ and by bench.py's CPU-baseline leg as the checker.  The product
 * library never links, loads or calls it.  Every function names the reference
 * file:line it restates in jxl_oracle.c.
 *
 * Numeric contract shared with the HIP kernels: float arithmetic with
 * contraction off, explicit fmaf where a fused multiply-add is meant, the
 * transcendental functions jo_cbrt / jo_pow24 (table-driven on the XYB
 * conversion's domains) and jo_log2 / jo_exp2 below (built from IEEE basic
 * operations only), DCT bases computed in double and narrowed, and
 * reductions in the stated tree orders.  Integer outputs (quant field,
 * transform map, quantized coefficients, histograms, codes, bytes) are then
 * bit-exact between this file and the device path.
 */
#ifndef JXL_ORACLE_H_
#define JXL_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jo_cal {
    uint32_t global_scale;
    uint32_t quant_dc;
    uint32_t raw_quant_field;
    float global_scale_float;
    float inv_global_scale;
} jo_cal;

/* ---- numeric primitives ------------------------------------------------ */
float jo_cbrt(float x);  /* x in [2^-9, 2) */
float jo_log2(float x);
float jo_exp2(float x);
float jo_pow24(float x); /* x in [2^-4, 2) */
void jo_xyb_table(float* out); /* the 272-float table of jo_cbrt / jo_pow24 */
float jo_exp(float x);
uint16_t jo_f32_to_f16(float x);
float jo_f16_to_f32(uint16_t h);

/* ---- tables -------------------------------------------------------------- */
/* Dequant matrices [3][n*n] for n = 8, 16, 32 (libjxl 0.11.2 defaults). */
void jo_set_dequant(const float* w8, const float* w16, const float* w32);
/* Threads for the front end's per-row/tile/region/block loops (default 1);
 * outputs are the same bytes for every count. */
void jo_set_threads(int n);
void jo_calibrate(float distance, jo_cal* out);
void jo_natural_order(uint32_t n, uint32_t* out);
void jo_dct_basis(uint32_t n, float* out);
int jo_ac_cluster(uint32_t ctx);

/* ---- front end ----------------------------------------------------------- */
void jo_nv12_to_xyb(const uint8_t* luma, size_t luma_pitch, const uint8_t* chroma,
                    size_t chroma_pitch, uint32_t w, uint32_t h, float* xyb);
void jo_chroma_upsample(const uint8_t* chroma, size_t chroma_pitch, uint32_t w, uint32_t h,
                        float* cb, float* cr);
void jo_nv12_to_xyb_f64(const uint8_t* luma, size_t luma_pitch, const uint8_t* chroma,
                        size_t chroma_pitch, uint32_t w, uint32_t h, double* xyb);
void jo_gaborish(const float* xyb, uint32_t w, uint32_t h, float* out);
void jo_adaptive_quant(const float* xyb, uint32_t w, uint32_t h, float distance, int32_t* qf);
void jo_select_transforms(const float* y, uint32_t w, uint32_t h, float distance, int8_t* acs);
void jo_forward_dct(const float* planes, uint32_t w, uint32_t h, const int8_t* acs, float* coef);
void jo_forward_dct_f64(const double* planes, uint32_t w, uint32_t h, const int8_t* acs,
                        double* coef);
void jo_cfl_region_partials(const uint16_t* coef16, const int8_t* acs, uint32_t w, uint32_t h,
                            float* partials);
void jo_cfl_maps(const float* partials, uint32_t w, uint32_t h, int8_t* ytox, int8_t* ytob);
void jo_quantize(const uint16_t* coef16, const int8_t* acs, const int8_t* ytox,
                 const int8_t* ytob, const int32_t* qf, uint32_t w, uint32_t h, float distance,
                 int16_t* ac, int32_t* dc, uint8_t* nz_grid);
void jo_dc_from_llf(uint32_t m, const float* llf, float* dc);

/* Config-1 semantics: DCT8 everywhere, uniform quant field, quantize_dct8. */
void jo_quantize_dct8_uniform(const float* coef, uint32_t w, uint32_t h, float distance,
                              int32_t* ac, int32_t* dc);

/* ---- entropy and assembly ------------------------------------------------ */
typedef struct jo_frame {
    uint32_t width, height;
    uint32_t global_scale, quant_dc;
    const int32_t* dc;     /* [3][blocks] X, Y, B */
    const int32_t* ac;     /* [3][blocks*64] covered-block layout */
    const int8_t* acs;     /* [blocks] or NULL = all DCT8 */
    const int32_t* qf;     /* [blocks] per-block quant field */
    const int8_t* ytox;    /* [tiles] or NULL = 0 */
    const int8_t* ytob;
} jo_frame;

size_t jo_huffman(const uint32_t* hist, uint32_t len, uint8_t* depth, uint16_t* bits,
                  uint8_t* header, size_t cap, size_t* header_bits);
size_t jo_best_context_map(const uint8_t* map, uint32_t n, uint32_t nclusters, uint8_t* out,
                           size_t cap, size_t* nbits, int* encoding);
void jo_move_to_front(const uint8_t* in, uint32_t n, uint8_t* out);
void jo_ac_histogram(const jo_frame* f, uint32_t* hist /* 32*256 */);
void jo_dc_histograms(const jo_frame* f, uint32_t* dc_hist, uint32_t* meta_hist);
size_t jo_ac_groups(const jo_frame* f, const uint8_t* depth, const uint16_t* bits,
                    uint32_t* group_sizes, uint8_t* out, size_t cap);
size_t jo_dc_groups(const jo_frame* f, uint32_t* group_sizes, uint8_t* out, size_t cap);
size_t jo_dc_global(uint32_t gs, uint32_t qdc, uint8_t* out, size_t cap);
size_t jo_ac_global(const uint32_t* hist, uint32_t num_ac_groups, uint8_t* depth,
                    uint16_t* bits, uint8_t* out, size_t cap);
size_t jo_codestream_head(uint32_t w, uint32_t h, const uint32_t* sizes, uint32_t n,
                          uint8_t* out, size_t cap);
size_t jo_container_framing(uint64_t codestream_size, uint8_t* out, size_t cap);
size_t jo_write_codestream(const jo_frame* f, uint8_t* out, size_t cap);

/* Whole pipeline: NV12 -> .jxl file bytes (container + codestream). */
size_t jo_encode_nv12(const uint8_t* luma, size_t luma_pitch, const uint8_t* chroma,
                      size_t chroma_pitch, uint32_t w, uint32_t h, float distance, uint8_t* out,
                      size_t cap);
/* Config-1: DCT8-only, uniform quant field, no CfL maps -> .jxl bytes. */
size_t jo_encode_dct8_uniform(const uint8_t* luma, size_t luma_pitch, const uint8_t* chroma,
                              size_t chroma_pitch, uint32_t w, uint32_t h, float distance,
                              uint8_t* out, size_t cap);
size_t jo_max_output_size(uint32_t w, uint32_t h);

/* ---- subset decoder (jxl_decode.c) --------------------------------------- */
/* Frame size of a `.jxl` file (container or bare codestream); 0 or -1. */
int jd_probe(const uint8_t* file, size_t n, uint32_t* width, uint32_t* height);
/* Decodes the file back to the quantized frame: transform map (8/16/32 at
 * first blocks, 0 covered), quant field at first blocks, DC [3][blocks]
 * (X, Y, B), AC [3][blocks*64] in the covered-block layout, YtoX / YtoB
 * [tiles].  0 on success, -1 with jd_error() naming the first violation. */
int jd_decode(const uint8_t* file, size_t n, uint32_t width, uint32_t height, uint32_t* global_scale,
              uint32_t* quant_dc, int8_t* acs, int32_t* qf, int32_t* dc, int16_t* ac, int8_t* ytox,
              int8_t* ytob);
const char* jd_error(void);

/* --- the learned selection / AQ network (jxl_mlp.c) ---------------------- */
#define JO_MLP_IN 96
#define JO_MLP_HID 64
#define JO_MLP_OUT 7
#define JO_MLP_PARAMS (JO_MLP_HID * JO_MLP_IN + JO_MLP_HID + JO_MLP_OUT * JO_MLP_HID + JO_MLP_OUT)
#define JO_MLP_LABELS 12
/* Outputs z[n][7] of n feature rows, in float64. */
void jo_mlp_forward(const float* params, const float* x, uint32_t n, double* z);
/* Mean loss over n tiles and, when `grad` is not null, its gradient with
 * respect to the parameters (JO_MLP_PARAMS doubles). */
double jo_mlp_loss_grad(const float* params, const float* x, const float* labels, uint32_t n, float lambda,
                        double* grad);
/* The selection's RD cost of every 8x8 block's DCT8 (the sharpened Y plane;
 * the same function and summation order as jo_select_transforms), and when
 * not null its two sums: squared quantization errors and rate terms. */
void jo_block_costs8(const float* y, uint32_t w, uint32_t h, float distance, float* cost8, float* dist8,
                     float* rate8);

#ifdef __cplusplus
}
#endif

#endif /* JXL_ORACLE_H_ */
