/* cujpegxl C ABI, implemented for AMD Instinct MI355X (gfx950).
 *
 * Drop-in for the public interface of m-pilia/cujpegxl: the same types, enum
 * values, struct layouts and entry points (include/cujpegxl/cujpegxl.h:24-82
 * there), exported from libcujpegxl.so.1 under the version node CUJPEGXL_1.
 * The header is free of HIP types: device memory is passed as integer
 * addresses.
 */
#ifndef CUJPEGXL_CUJPEGXL_H_
#define CUJPEGXL_CUJPEGXL_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cujpegxl.h:24: the reference's ABI version, unchanged.  The host-input,
 * asynchronous, allocation-free and encoder-form entry points at the end of
 * this header are extensions, versioned on their own (CUJPEGXL_EXT_VERSION,
 * cujpegxl_ext_version). */
#define CUJPEGXL_API_VERSION 3u
#define CUJPEGXL_EXT_VERSION 2u

/* cujpegxl.h:26-34; values are ABI. */
typedef enum cujpegxl_status {
    CUJPEGXL_OK = 0,
    CUJPEGXL_NOT_IMPLEMENTED = 1,
    CUJPEGXL_INVALID_ARGUMENT = 2,
    CUJPEGXL_UNSUPPORTED_RESOLUTION = 3,
    CUJPEGXL_NO_DEVICE = 4,
    CUJPEGXL_INTERNAL_ERROR = 5,
    CUJPEGXL_BUFFER_TOO_SMALL = 6
} cujpegxl_status;

/* cujpegxl.h:36-40.  A gfx9 device with matrix cores reports FP16_TENSOR. */
typedef enum cujpegxl_backend {
    CUJPEGXL_BACKEND_UNKNOWN = 0,
    CUJPEGXL_BACKEND_FP32_SIMT = 1,
    CUJPEGXL_BACKEND_FP16_TENSOR = 2
} cujpegxl_backend;

typedef struct cujpegxl_encoder cujpegxl_encoder;

/* cujpegxl.h:44-49 */
typedef struct cujpegxl_config {
    uint32_t width;
    uint32_t height;
    float distance;
    int32_t device_ordinal;
} cujpegxl_config;

/* cujpegxl.h:53-58: NV12 (BT.709 full range) in device memory. */
typedef struct cujpegxl_nv12_input {
    uintptr_t luma;
    uintptr_t chroma;
    size_t luma_pitch;
    size_t chroma_pitch;
} cujpegxl_nv12_input;

/* cujpegxl.h:60-64: caller-owned host buffer; the library writes `size`. */
typedef struct cujpegxl_output {
    uint8_t* data;
    size_t size;
    size_t capacity;
} cujpegxl_output;

/* cujpegxl.h:66 */
uint32_t cujpegxl_api_version(void);
/* cujpegxl.h:67 */
const char* cujpegxl_status_string(cujpegxl_status status);
/* cujpegxl.h:68 */
const char* cujpegxl_backend_string(cujpegxl_backend backend);
/* cujpegxl.h:70 */
cujpegxl_status cujpegxl_query_backend(int32_t device_ordinal, cujpegxl_backend* out_backend);
/* cujpegxl.h:74: worst-case output bytes for `config`. */
cujpegxl_status cujpegxl_max_output_size(const cujpegxl_config* config, size_t* out_bytes);
/* cujpegxl.h:76-77 */
cujpegxl_status cujpegxl_encoder_create(const cujpegxl_config* config,
                                        cujpegxl_encoder** out_encoder);
/* cujpegxl.h:78-79 */
cujpegxl_status cujpegxl_encoder_configure(cujpegxl_encoder* encoder,
                                           const cujpegxl_config* config);
/* cujpegxl.h:80-81: synchronous encode of one device-resident NV12 frame into
 * a `.jxl` (ISOBMFF container) in `output`. */
cujpegxl_status cujpegxl_encoder_encode(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* input,
                                        cujpegxl_output* output);
/* cujpegxl.h:82 */
void cujpegxl_encoder_destroy(cujpegxl_encoder* encoder);

/* ---- Extensions, version 1 (no counterpart in the reference) ------------ */

/* CUJPEGXL_EXT_VERSION of the library: 1 for the three entry points below,
 * 2 with cujpegxl_encoder_set_mode.
 * Exported under the version node CUJPEGXL_2; a caller that finds this
 * symbol may use them. */
uint32_t cujpegxl_ext_version(void);

/* NV12 (BT.709 full range) in host memory: plane addresses and pitches in
 * bytes (pitches >= width). */
typedef struct cujpegxl_nv12_host_input {
    const uint8_t* luma;
    const uint8_t* chroma;
    size_t luma_pitch;
    size_t chroma_pitch;
} cujpegxl_nv12_host_input;

/* Synchronous encode of a host-resident NV12 frame: the planes are uploaded
 * into a device frame the encoder owns (through its pinned staging buffer, or
 * by direct DMA when both planes are pinned host memory), then encoded as
 * cujpegxl_encoder_encode does.  Same validation and statuses. */
cujpegxl_status cujpegxl_encoder_encode_host(cujpegxl_encoder* encoder,
                                             const cujpegxl_nv12_host_input* input,
                                             cujpegxl_output* output);

/* Asynchronous encode of a device-resident NV12 frame.  `stream` is a
 * hipStream_t as an integer (0: the null stream).  The frame's work is
 * ordered after the work already queued on `stream`, and later work queued on
 * `stream` is ordered after the encoder has finished reading the planes.
 * Returns once the work is queued; the file is delivered by
 * cujpegxl_encoder_finish.  One frame may be pending per encoder: a second
 * call before finish returns CUJPEGXL_INVALID_ARGUMENT. */
cujpegxl_status cujpegxl_encoder_encode_async(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* input,
                                              uintptr_t stream);

/* Waits for the pending frame and writes it to `output` (statuses as for
 * cujpegxl_encoder_encode, including CUJPEGXL_BUFFER_TOO_SMALL with the
 * required size in output->size; the frame is consumed either way).
 * CUJPEGXL_INVALID_ARGUMENT when no frame is pending. */
cujpegxl_status cujpegxl_encoder_finish(cujpegxl_encoder* encoder, cujpegxl_output* output);

/* ---- Allocation-free encode (no counterpart in the reference) ----------
 *
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Exported
 * under the version node CUJPEGXL_3.  Every intermediate buffer lives in one
 * device array of CUJPEGXL_NOALLOC_SCRATCH_BYTES that the library reserves on
 * each device it is loaded on, until the process ends; nothing is allocated,
 * and no stream or event is created, per call.  Frames up to
 * CUJPEGXL_NOALLOC_MAX_WIDTH x CUJPEGXL_NOALLOC_MAX_HEIGHT (and at most
 * CUJPEGXL_NOALLOC_MAX_PIXELS pixels); larger ones return
 * CUJPEGXL_UNSUPPORTED_RESOLUTION. */
#define CUJPEGXL_NOALLOC_MAX_WIDTH 4096u
#define CUJPEGXL_NOALLOC_MAX_HEIGHT 4096u
#define CUJPEGXL_NOALLOC_MAX_PIXELS (1920u * 1088u)
#define CUJPEGXL_NOALLOC_SCRATCH_BYTES (104ull << 20)

/* Device scratch bytes a frame of width x height uses (0 for shapes the
 * encoder never accepts); at most CUJPEGXL_NOALLOC_SCRATCH_BYTES for every
 * accepted frame. */
size_t cujpegxl_noalloc_workspace_bytes(uint32_t width, uint32_t height);

/* Encodes one device-resident NV12 frame, the work queued on `stream` (a
 * hipStream_t as an integer; 0: the null stream) after the work already
 * there.  The file goes into the device buffer `output` (4-byte aligned; the
 * first `capacity` bytes of the file are written) and its size into the
 * 8-byte device word at `size_out`, both once the stream's work has finished:
 * a size above `capacity` means the file was cut short.  Validation and
 * statuses as for cujpegxl_encoder_create and cujpegxl_encoder_encode, plus
 * CUJPEGXL_UNSUPPORTED_RESOLUTION above the limits.  Calls share one scratch
 * array per device, which keeps the last call's contents: calls whose work
 * overlaps on one device (on different streams) race and must be serialised
 * by the caller.  The first call on a device, and a call with a new shape or
 * distance, also waits for `stream`, as it copies the frame's constant
 * headers in order on it. */
cujpegxl_status cujpegxl_encode_noalloc(const cujpegxl_config* config, const cujpegxl_nv12_input* input,
                                        uintptr_t output, size_t capacity, uintptr_t size_out,
                                        uintptr_t stream);

/* ---- Encoder forms (no counterpart in the reference) --------------------
 *
 * Exported under the version node CUJPEGXL_4 (extension version 2).  The
 * plain path (mode 0, every encoder's start) writes the reference's files;
 * CUJPEGXL_MODE_MFMA16 computes the DCTs as FP16-operand matrix-core
 * products (coefficients within the FP16 storage error of the plain ones);
 * CUJPEGXL_MODE_LEARNED lets the compiled-in selection network choose the
 * transform sizes where it is confident (the heuristic elsewhere).  Only
 * encoder decisions change: files stay conformant JPEG XL.  Applies from the
 * next encode; not while a frame is pending. */
#define CUJPEGXL_MODE_MFMA16 1u
#define CUJPEGXL_MODE_LEARNED 2u
cujpegxl_status cujpegxl_encoder_set_mode(cujpegxl_encoder* encoder, uint32_t mode);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CUJPEGXL_CUJPEGXL_H_ */
