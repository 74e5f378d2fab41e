// The cujpegxl C ABI (cujpegxl.cc in the reference) over cjx::Encoder.
#include "cujpegxl/cujpegxl.h"

#include <cmath>
#include <new>

#include "encoder.h"
#include "handle.h"


namespace {

// cujpegxl.cc:36-48: distance first, then resolution (non-zero multiples of 8
// spanning more than one AC group).
cujpegxl_status validate_config(const cujpegxl_config& c) {
    if (!(c.distance > 0.0f) || !std::isfinite(c.distance)) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (c.width == 0 || c.height == 0 || c.width % 8 != 0 || c.height % 8 != 0) {
        return CUJPEGXL_UNSUPPORTED_RESOLUTION;
    }
    const cjx::Geometry g{cjx::Geometry::of(c.width, c.height)};
    if (g.num_ac <= 1 || g.num_dc > 32) {
        return CUJPEGXL_UNSUPPORTED_RESOLUTION;
    }
    return CUJPEGXL_OK;
}

}  // namespace

extern "C" {

uint32_t cujpegxl_api_version(void) {
    return CUJPEGXL_API_VERSION;
}

uint32_t cujpegxl_ext_version(void) {
    return CUJPEGXL_EXT_VERSION;
}

const char* cujpegxl_status_string(cujpegxl_status status) {
    switch (status) {
        case CUJPEGXL_OK: return "ok";
        case CUJPEGXL_NOT_IMPLEMENTED: return "not-implemented";
        case CUJPEGXL_INVALID_ARGUMENT: return "invalid-argument";
        case CUJPEGXL_UNSUPPORTED_RESOLUTION: return "unsupported-resolution";
        case CUJPEGXL_NO_DEVICE: return "no-device";
        case CUJPEGXL_INTERNAL_ERROR: return "internal-error";
        case CUJPEGXL_BUFFER_TOO_SMALL: return "buffer-too-small";
    }
    return "internal-error";
}

const char* cujpegxl_backend_string(cujpegxl_backend backend) {
    switch (backend) {
        case CUJPEGXL_BACKEND_FP32_SIMT: return "fp32-simt";
        case CUJPEGXL_BACKEND_FP16_TENSOR: return "fp16-tensor";
        case CUJPEGXL_BACKEND_UNKNOWN: return "unknown";
    }
    return "unknown";
}

cujpegxl_status cujpegxl_query_backend(int32_t device_ordinal, cujpegxl_backend* out_backend) {
    if (out_backend == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (!cjx::device_exists(device_ordinal)) {
        return CUJPEGXL_NO_DEVICE;
    }
    *out_backend = cjx::device_has_matrix_cores(device_ordinal) ? CUJPEGXL_BACKEND_FP16_TENSOR
                                                                 : CUJPEGXL_BACKEND_FP32_SIMT;
    return CUJPEGXL_OK;
}

cujpegxl_status cujpegxl_max_output_size(const cujpegxl_config* config, size_t* out_bytes) {
    if (config == nullptr || out_bytes == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const cujpegxl_status s{validate_config(*config)};
    if (s != CUJPEGXL_OK) {
        return s;
    }
    *out_bytes = cjx::Encoder::max_output_size(config->width, config->height);
    return CUJPEGXL_OK;
}

cujpegxl_status cujpegxl_encoder_create(const cujpegxl_config* config, cujpegxl_encoder** out_encoder) {
    if (config == nullptr || out_encoder == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    *out_encoder = nullptr;
    const cujpegxl_status s{validate_config(*config)};
    if (s != CUJPEGXL_OK) {
        return s;
    }
    if (!cjx::device_exists(config->device_ordinal)) {
        return CUJPEGXL_NO_DEVICE;
    }
    auto* e{new (std::nothrow) cujpegxl_encoder{}};
    if (e == nullptr) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    e->config = *config;
    if (!e->engine.configure(config->width, config->height, config->distance, config->device_ordinal)) {
        delete e;
        return CUJPEGXL_INTERNAL_ERROR;
    }
    *out_encoder = e;
    return CUJPEGXL_OK;
}

cujpegxl_status cujpegxl_encoder_configure(cujpegxl_encoder* encoder, const cujpegxl_config* config) {
    if (encoder == nullptr || config == nullptr || encoder->engine.pending()) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const cujpegxl_status s{validate_config(*config)};
    if (s != CUJPEGXL_OK) {
        return s;
    }
    if (!cjx::device_exists(config->device_ordinal)) {
        return CUJPEGXL_NO_DEVICE;
    }
    if (!encoder->engine.configure(config->width, config->height, config->distance, config->device_ordinal)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    encoder->config = *config;
    return CUJPEGXL_OK;
}

cujpegxl_status cujpegxl_encoder_encode(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* input,
                                        cujpegxl_output* output) {
    if (encoder == nullptr || input == nullptr || output == nullptr || output->data == nullptr ||
        input->luma == 0 || input->chroma == 0 || encoder->engine.pending()) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const cujpegxl_status s{validate_config(encoder->config)};
    if (s != CUJPEGXL_OK) {
        return s;
    }
    if (input->luma_pitch < encoder->config.width || input->chroma_pitch < encoder->config.width) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    std::size_t size{0};
    if (!encoder->engine.encode_nv12(reinterpret_cast<const std::uint8_t*>(input->luma), input->luma_pitch,
                                     reinterpret_cast<const std::uint8_t*>(input->chroma), input->chroma_pitch,
                                     output->data, output->capacity, &size)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    output->size = size;
    return size > output->capacity ? CUJPEGXL_BUFFER_TOO_SMALL : CUJPEGXL_OK;
}

void cujpegxl_encoder_destroy(cujpegxl_encoder* encoder) {
    delete encoder;
}

cujpegxl_status cujpegxl_encoder_encode_host(cujpegxl_encoder* encoder, const cujpegxl_nv12_host_input* input,
                                             cujpegxl_output* output) {
    if (encoder == nullptr || input == nullptr || output == nullptr || output->data == nullptr ||
        input->luma == nullptr || input->chroma == nullptr || encoder->engine.pending()) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const cujpegxl_status s{validate_config(encoder->config)};
    if (s != CUJPEGXL_OK) {
        return s;
    }
    if (input->luma_pitch < encoder->config.width || input->chroma_pitch < encoder->config.width) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    std::size_t size{0};
    if (!encoder->engine.encode_host(input->luma, input->luma_pitch, input->chroma, input->chroma_pitch,
                                     output->data, output->capacity, &size)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    output->size = size;
    return size > output->capacity ? CUJPEGXL_BUFFER_TOO_SMALL : CUJPEGXL_OK;
}

cujpegxl_status cujpegxl_encoder_encode_async(cujpegxl_encoder* encoder, const cujpegxl_nv12_input* input,
                                              uintptr_t stream) {
    if (encoder == nullptr || input == nullptr || input->luma == 0 || input->chroma == 0 ||
        encoder->engine.pending()) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const cujpegxl_status s{validate_config(encoder->config)};
    if (s != CUJPEGXL_OK) {
        return s;
    }
    if (input->luma_pitch < encoder->config.width || input->chroma_pitch < encoder->config.width) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (!encoder->engine.encode_async(reinterpret_cast<const std::uint8_t*>(input->luma), input->luma_pitch,
                                      reinterpret_cast<const std::uint8_t*>(input->chroma), input->chroma_pitch,
                                      reinterpret_cast<hipStream_t>(stream))) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    return CUJPEGXL_OK;
}

cujpegxl_status cujpegxl_encoder_finish(cujpegxl_encoder* encoder, cujpegxl_output* output) {
    if (encoder == nullptr || !encoder->engine.pending()) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const bool have_out{output != nullptr && output->data != nullptr};
    std::size_t size{0};
    if (!encoder->engine.finish(have_out ? output->data : nullptr, have_out ? output->capacity : 0, &size)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    if (!have_out) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    output->size = size;
    return size > output->capacity ? CUJPEGXL_BUFFER_TOO_SMALL : CUJPEGXL_OK;
}

cujpegxl_status cujpegxl_encoder_set_mode(cujpegxl_encoder* encoder, uint32_t mode) {
    static_assert(CUJPEGXL_MODE_MFMA16 == cjx::kModeMfma16 && CUJPEGXL_MODE_LEARNED == cjx::kModeLearned,
                  "public mode bits");
    if (encoder == nullptr || (mode & ~(CUJPEGXL_MODE_MFMA16 | CUJPEGXL_MODE_LEARNED)) != 0 ||
        encoder->engine.pending()) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    encoder->engine.set_mode(mode);
    return CUJPEGXL_OK;
}

}  // extern "C"
