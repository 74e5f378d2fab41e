// Stage harness (cujpegxl_stages.h).
#include "cujpegxl/cujpegxl_stages.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "common.h"
#include "batch_encoder.h"
#include "encoder.h"
#include "handle.h"
#include "kernels.h"


extern "C" {

size_t cjx_stage_buffer_bytes(cujpegxl_encoder* e, int which) {
    if (e == nullptr || which < 0 || which >= static_cast<int>(cjx::Buffer::kCount)) {
        return 0;
    }
    return e->engine.buffer_bytes(static_cast<cjx::Buffer>(which));
}

int cjx_stage_fetch(cujpegxl_encoder* e, int which, void* host, size_t bytes) {
    if (e == nullptr || host == nullptr || which < 0 || which >= static_cast<int>(cjx::Buffer::kCount)) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return e->engine.fetch(static_cast<cjx::Buffer>(which), host, bytes) ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_stage_encode_coefficients(cujpegxl_encoder* e, const int16_t* ac, const int32_t* dc, const int8_t* acs,
                                  const int32_t* qf, const int8_t* ytox, const int8_t* ytob,
                                  uint32_t global_scale, uint32_t quant_dc, uint8_t* out, size_t cap,
                                  size_t* size) {
    if (e == nullptr || ac == nullptr || dc == nullptr || acs == nullptr || qf == nullptr || size == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (!e->engine.encode_coefficients(ac, dc, acs, qf, ytox, ytob, global_scale, quant_dc, out, cap, size)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    return *size > cap ? CUJPEGXL_BUFFER_TOO_SMALL : CUJPEGXL_OK;
}

int cjx_stage_encoder_create_any(const cujpegxl_config* c, cujpegxl_encoder** out) {
    if (c == nullptr || out == nullptr || !(c->distance > 0.0f) || !std::isfinite(c->distance)) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (c->width == 0 || c->height == 0 || c->width % 8 || c->height % 8) {
        return CUJPEGXL_UNSUPPORTED_RESOLUTION;
    }
    if (!cjx::device_exists(c->device_ordinal)) {
        return CUJPEGXL_NO_DEVICE;
    }
    auto* e{new (std::nothrow) cujpegxl_encoder{}};
    if (e == nullptr) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    e->config = *c;
    if (!e->engine.configure(c->width, c->height, c->distance, c->device_ordinal)) {
        delete e;
        return CUJPEGXL_INTERNAL_ERROR;
    }
    *out = e;
    return CUJPEGXL_OK;
}

int cjx_stage_encode_nv12(cujpegxl_encoder* e, const cujpegxl_nv12_input* in, cujpegxl_output* out) {
    if (e == nullptr || in == nullptr || out == nullptr || out->data == nullptr || in->luma == 0 ||
        in->chroma == 0 || in->luma_pitch < e->config.width || in->chroma_pitch < e->config.width) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    std::size_t size{0};
    if (!e->engine.encode_nv12(reinterpret_cast<const std::uint8_t*>(in->luma), in->luma_pitch,
                               reinterpret_cast<const std::uint8_t*>(in->chroma), in->chroma_pitch, out->data,
                               out->capacity, &size)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    out->size = size;
    return size > out->capacity ? CUJPEGXL_BUFFER_TOO_SMALL : CUJPEGXL_OK;
}

int cjx_stage_encode_repeat(cujpegxl_encoder* e, const cujpegxl_nv12_input* in, cujpegxl_output* out, int count,
                            double* seconds) {
    if (count <= 0 || seconds == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const auto t0{std::chrono::steady_clock::now()};
    for (int i{0}; i < count; ++i) {
        const int s{cujpegxl_encoder_encode(e, in, out)};
        if (s != CUJPEGXL_OK) {
            return s;
        }
    }
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return CUJPEGXL_OK;
}

int cjx_stage_encode_cycle(cujpegxl_encoder* e, const cujpegxl_nv12_input* inputs, int n_inputs,
                           cujpegxl_output* out, int count, double* seconds) {
    if (inputs == nullptr || n_inputs <= 0 || count <= 0 || seconds == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const auto t0{std::chrono::steady_clock::now()};
    for (int i{0}; i < count; ++i) {
        const int s{cujpegxl_encoder_encode(e, &inputs[i % n_inputs], out)};
        if (s != CUJPEGXL_OK) {
            return s;
        }
    }
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return CUJPEGXL_OK;
}

int cjx_stage_encode_dct8_uniform(cujpegxl_encoder* e, const cujpegxl_nv12_input* in, cujpegxl_output* out) {
    if (e == nullptr || in == nullptr || out == nullptr || out->data == nullptr || in->luma == 0 ||
        in->chroma == 0 || in->luma_pitch < e->config.width || in->chroma_pitch < e->config.width) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    std::size_t size{0};
    if (!e->engine.encode_dct8_uniform(reinterpret_cast<const std::uint8_t*>(in->luma), in->luma_pitch,
                                       reinterpret_cast<const std::uint8_t*>(in->chroma), in->chroma_pitch,
                                       out->data, out->capacity, &size)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    out->size = size;
    return size > out->capacity ? CUJPEGXL_BUFFER_TOO_SMALL : CUJPEGXL_OK;
}

int cjx_stage_set_profiling(cujpegxl_encoder* e, int on) {
    if (e == nullptr || on < 0 || on > 2) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    e->engine.set_profiling(on);
    return CUJPEGXL_OK;
}

int cjx_stage_set_mode(cujpegxl_encoder* e, uint32_t mode) {
    if (e == nullptr || (mode & ~cjx::kModeMask) != 0) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    e->engine.set_mode(mode);
    return CUJPEGXL_OK;
}

int cjx_stage_export_features(cujpegxl_encoder* e, int on) {
    if (e == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return e->engine.export_features(on != 0) ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_mlp_set_weights(int32_t device, const uint16_t* w1, const float* b1, const uint16_t* w2, const float* b2,
                        float threshold) {
    if (w1 == nullptr || b1 == nullptr || w2 == nullptr || b2 == nullptr || !cjx::device_exists(device) ||
        hipSetDevice(device) != hipSuccess) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (!cjx::ensure_device_tables(device)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    return cjx::set_mlp_weights(w1, b1, w2, b2, threshold) ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_mlp_default_weights(int32_t device) {
    if (!cjx::device_exists(device) || hipSetDevice(device) != hipSuccess) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return cjx::ensure_device_tables(device) && cjx::set_default_mlp_weights() ? CUJPEGXL_OK
                                                                                : CUJPEGXL_INTERNAL_ERROR;
}

size_t cjx_mlp_partial_floats(uint32_t n) {
    return cjx::mlp_partial_floats(n);
}

int cjx_mlp_step(const float* x, const float* labels, uint32_t n, float lambda, float* partial, float* grad,
                 float* loss, float* params, float* m, float* v, uint32_t step, float lr, uintptr_t stream) {
    if (x == nullptr || labels == nullptr || n == 0 || partial == nullptr || grad == nullptr || loss == nullptr ||
        params == nullptr || (lr > 0.0f && (m == nullptr || v == nullptr || step == 0))) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    cjx::launch_mlp_step(reinterpret_cast<hipStream_t>(stream), x, labels, n, lambda, partial, grad, loss, params, m,
                         v, step, lr);
    return hipGetLastError() == hipSuccess ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_mlp_forward(const float* params, const float* x, uint32_t n, float* z, uintptr_t stream) {
    if (params == nullptr || x == nullptr || z == nullptr || n == 0) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    cjx::launch_mlp_forward(reinterpret_cast<hipStream_t>(stream), params, x, n, z);
    return hipGetLastError() == hipSuccess ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

struct cjx_batch {
    cjx::BatchEncoder engine;
};

int cjx_batch_create(int32_t device, uint32_t count, const uint32_t* width, const uint32_t* height,
                     const float* distance, uint32_t mode, cjx_batch** out) {
    if (out == nullptr || width == nullptr || height == nullptr || distance == nullptr || count == 0 ||
        !cjx::device_exists(device)) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    *out = nullptr;
    auto* b{new (std::nothrow) cjx_batch{}};
    if (b == nullptr) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    if (!b->engine.configure(static_cast<int>(count), width, height, distance, device, mode)) {
        delete b;
        return CUJPEGXL_INTERNAL_ERROR;
    }
    *out = b;
    return CUJPEGXL_OK;
}

int cjx_batch_set_inputs(cjx_batch* b, const cujpegxl_nv12_input* inputs) {
    if (b == nullptr || inputs == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const int n{b->engine.frames()};
    std::vector<const std::uint8_t*> luma(n), chroma(n);
    std::vector<std::size_t> lp(n), cp(n);
    for (int i{0}; i < n; ++i) {
        luma[i] = reinterpret_cast<const std::uint8_t*>(inputs[i].luma);
        chroma[i] = reinterpret_cast<const std::uint8_t*>(inputs[i].chroma);
        lp[i] = inputs[i].luma_pitch;
        cp[i] = inputs[i].chroma_pitch;
    }
    return b->engine.set_inputs(luma.data(), lp.data(), chroma.data(), cp.data()) ? CUJPEGXL_OK
                                                                                  : CUJPEGXL_INVALID_ARGUMENT;
}

int cjx_batch_run(cjx_batch* b, int32_t count, double* seconds) {
    if (b == nullptr || count <= 0) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return b->engine.run(count, seconds) ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_batch_launch(cjx_batch* b, int32_t count) {
    if (b == nullptr || count <= 0) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return b->engine.launch(count) ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_batch_wait(cjx_batch* b) {
    if (b == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return b->engine.wait() ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_batch_result(const cjx_batch* b, uint32_t i, const uint8_t** data, uint64_t* size) {
    if (b == nullptr || data == nullptr || size == nullptr || i >= static_cast<uint32_t>(b->engine.frames())) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    *data = b->engine.file(static_cast<int>(i));
    *size = b->engine.size(static_cast<int>(i));
    return CUJPEGXL_OK;
}

int cjx_batch_tune(cjx_batch* b, int32_t reps, char* report, size_t cap) {
    if (b == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    std::string text;
    const bool ok{b->engine.tune(reps, &text)};
    if (report != nullptr && cap > 0) {
        const std::size_t n{text.size() < cap - 1 ? text.size() : cap - 1};
        std::memcpy(report, text.data(), n);
        report[n] = 0;
    }
    return ok ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_batch_set_shape(cjx_batch* b, uint32_t token_cap, uint32_t assemble_words, uint32_t region_wgs) {
    if (b == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return b->engine.set_shape(cjx::BatchShape{token_cap, assemble_words, region_wgs}) ? CUJPEGXL_OK
                                                                                      : CUJPEGXL_INVALID_ARGUMENT;
}

void cjx_batch_destroy(cjx_batch* b) {
    delete b;
}

int cjx_stage_check_aq_math(int32_t device_ordinal, uint64_t* mismatches) {
    unsigned long long n{0};
    if (mismatches == nullptr || hipSetDevice(device_ordinal) != hipSuccess || !cjx::check_aq_math(&n)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    *mismatches = n;
    return CUJPEGXL_OK;
}

int cjx_stage_timings(cujpegxl_encoder* e, float* ms, int count) {
    if (e == nullptr || ms == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const auto& t{e->engine.stage_ms()};
    for (int i{0}; i < count && i < static_cast<int>(t.size()); ++i) {
        ms[i] = t[i];
    }
    return CUJPEGXL_OK;
}

int cjx_stage_build_codes(int32_t device_ordinal, const uint32_t* hists, uint32_t count, uint32_t* alpha,
                          uint8_t* depth, uint16_t* bits, uint32_t* hdr_words, uint32_t* hdr_bits) {
    if (hists == nullptr || alpha == nullptr || depth == nullptr || bits == nullptr || hdr_words == nullptr ||
        hdr_bits == nullptr || count == 0 || count > 96) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (!cjx::device_exists(device_ordinal) || hipSetDevice(device_ordinal) != hipSuccess) {
        return CUJPEGXL_NO_DEVICE;
    }
    return cjx::build_codes_host(hists, count, alpha, depth, bits, hdr_words, hdr_bits) ? CUJPEGXL_OK
                                                                                       : CUJPEGXL_INTERNAL_ERROR;
}

void cjx_stage_fail_allocation_after(int n) {
    cjx::fail_allocation_after(n);
}

void* cjx_device_alloc(size_t bytes) {
    void* p{nullptr};
    return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

void cjx_device_free(void* p) {
    (void)hipFree(p);
}

uintptr_t cjx_stream_create(void) {
    hipStream_t s{nullptr};
    return hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess ? reinterpret_cast<uintptr_t>(s) : 0;
}

void cjx_stream_destroy(uintptr_t stream) {
    if (stream != 0) {
        (void)hipStreamDestroy(reinterpret_cast<hipStream_t>(stream));
    }
}

int cjx_stream_synchronize(uintptr_t stream) {
    return hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)) == hipSuccess ? CUJPEGXL_OK
                                                                                      : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_copy_async(void* dst, const void* src, size_t bytes, uintptr_t stream) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, reinterpret_cast<hipStream_t>(stream)) == hipSuccess
               ? CUJPEGXL_OK
               : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_fill_async(void* dst, int value, size_t bytes, uintptr_t stream) {
    return hipMemsetAsync(dst, value, bytes, reinterpret_cast<hipStream_t>(stream)) == hipSuccess
               ? CUJPEGXL_OK
               : CUJPEGXL_INTERNAL_ERROR;
}

void* cjx_host_alloc(size_t bytes) {
    void* p{nullptr};
    return hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}

void cjx_host_free(void* p) {
    (void)hipHostFree(p);
}

int cjx_copy_to_device(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_copy_to_host(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_device_copy(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? CUJPEGXL_OK
                                                                             : CUJPEGXL_INTERNAL_ERROR;
}

int cjx_device_synchronize(void) {
    return hipDeviceSynchronize() == hipSuccess ? CUJPEGXL_OK : CUJPEGXL_INTERNAL_ERROR;
}

size_t cjx_host_table(int which, void* out, size_t cap) {
    const cjx::DeviceTables& t{cjx::host_tables()};
    const void* src{nullptr};
    size_t n{0};
    float bases[64 + 256 + 1024];
    std::uint16_t orders[64 + 256 + 1024];
    switch (which) {
        case 0:
            src = &t.dequant8[0][0];
            n = sizeof(t.dequant8) + sizeof(t.dequant16) + sizeof(t.dequant32);
            break;
        case 1:
            std::memcpy(bases, t.basis8, sizeof(t.basis8));
            std::memcpy(bases + 64, t.basis16, sizeof(t.basis16));
            std::memcpy(bases + 320, t.basis32, sizeof(t.basis32));
            src = bases;
            n = sizeof(bases);
            break;
        case 2:
            std::memcpy(orders, t.order8, sizeof(t.order8));
            std::memcpy(orders + 64, t.order16, sizeof(t.order16));
            std::memcpy(orders + 320, t.order32, sizeof(t.order32));
            src = orders;
            n = sizeof(orders);
            break;
        default:
            return 0;
    }
    if (out != nullptr && cap >= n) {
        std::memcpy(out, src, n);
    }
    return n;
}

#ifdef CJX_REGION_CUT
// Profiling builds: the region kernel alone (see cjx_set_region_cut).
int cjx_stage_region_only(cujpegxl_encoder* e, const cujpegxl_nv12_input* in) {
    if (e == nullptr || in == nullptr) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    return e->engine.region_only(reinterpret_cast<const std::uint8_t*>(in->luma), in->luma_pitch,
                                 reinterpret_cast<const std::uint8_t*>(in->chroma), in->chroma_pitch)
               ? CUJPEGXL_OK
               : CUJPEGXL_INTERNAL_ERROR;
}
#endif

}  // extern "C"
