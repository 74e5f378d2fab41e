// cujpegxl_encode_noalloc: a whole-frame encode that allocates nothing.
//
// Every intermediate buffer of the frame (coefficients, quantized values,
// token records, histograms, codes, staging, layout) lives in `scratch`, a
// file-scope device array of kNoAllocScratchBytes that the code object
// reserves on each device it is loaded on.  Host code never passes the
// array's name to a kernel or a copy: at the first call on a device, with
// that device current, hipGetSymbolAddress gives the array's address on that
// device, kept per device in g_slots, and only that address is used.
//
// The frame is queued on the caller's stream alone; the file goes into the
// caller's device buffer and its size into a caller's device word.  Every
// call on a device shares that device's array: the region kernel clears the
// accumulators each frame needs at zero, and calls whose work overlaps on one
// device (other streams, other threads' streams) race on the array.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <new>

#include "cujpegxl/cujpegxl.h"
#include "encoder.h"

namespace cjx {

// Holds the buffers of every accepted shape (at most 4096 on a side and
// 1920 x 1088 pixels; the largest need, 100 MiB, is a narrow frame crossing
// many AC groups, e.g. 520 x 4016).
constexpr std::size_t kNoAllocScratchBytes = CUJPEGXL_NOALLOC_SCRATCH_BYTES;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

}  // namespace cjx

namespace {

struct DeviceSlot {
    std::once_flag once;
    bool ok{false};
    void* ptr{nullptr};  // `scratch` on this device
    std::mutex m;        // host-side queuing of one call at a time
    cjx::Encoder* enc{nullptr};
};

DeviceSlot g_slots[64];

// Set-up of a device, once, with the device current: the array's address on
// it and the encoder state that binds its buffers there.
bool device_slot(int device, DeviceSlot*& out) {
    if (device < 0 || device >= 64) {
        return false;
    }
    DeviceSlot& d{g_slots[device]};
    std::call_once(d.once, [&d, device] {
        if (hipSetDevice(device) != hipSuccess) {
            return;
        }
        void* p{nullptr};
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(cjx::scratch)) != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            return;
        }
        d.ptr = p;
        d.enc = new (std::nothrow) cjx::Encoder;  // host state, kept until the process ends
        d.ok = d.enc != nullptr;
    });
    out = &d;
    return d.ok;
}

}  // namespace

extern "C" {

size_t cujpegxl_noalloc_workspace_bytes(uint32_t width, uint32_t height) {
    return cjx::Encoder::workspace_bytes(width, height);
}

cujpegxl_status cujpegxl_encode_noalloc(const cujpegxl_config* config, const cujpegxl_nv12_input* input,
                                        uintptr_t output, size_t capacity, uintptr_t size_out, uintptr_t stream) {
    if (config == nullptr || input == nullptr || input->luma == 0 || input->chroma == 0 || output == 0 ||
        size_out == 0 || size_out % 8 != 0 || output % 4 != 0) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (!(config->distance > 0.0f) || !std::isfinite(config->distance)) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    const std::uint32_t w{config->width}, h{config->height};
    if (w == 0 || h == 0 || w % 8 != 0 || h % 8 != 0) {
        return CUJPEGXL_UNSUPPORTED_RESOLUTION;
    }
    const cjx::Geometry g{cjx::Geometry::of(w, h)};
    if (g.num_ac <= 1 || w > CUJPEGXL_NOALLOC_MAX_WIDTH || h > CUJPEGXL_NOALLOC_MAX_HEIGHT ||
        static_cast<std::uint64_t>(w) * h > CUJPEGXL_NOALLOC_MAX_PIXELS ||
        cjx::Encoder::workspace_bytes(w, h) > cjx::kNoAllocScratchBytes) {
        return CUJPEGXL_UNSUPPORTED_RESOLUTION;
    }
    if (input->luma_pitch < w || input->chroma_pitch < w) {
        return CUJPEGXL_INVALID_ARGUMENT;
    }
    if (!cjx::device_exists(config->device_ordinal)) {
        return CUJPEGXL_NO_DEVICE;
    }
    DeviceSlot* d{nullptr};
    if (!device_slot(config->device_ordinal, d)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    std::lock_guard<std::mutex> lock{d->m};
    if (!d->enc->bind_workspace(w, h, config->distance, config->device_ordinal, d->ptr, cjx::kNoAllocScratchBytes)) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    if (!d->enc->encode_on(reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const std::uint8_t*>(input->luma),
                           input->luma_pitch, reinterpret_cast<const std::uint8_t*>(input->chroma),
                           input->chroma_pitch, reinterpret_cast<std::uint8_t*>(output), capacity,
                           reinterpret_cast<std::uint64_t*>(size_out))) {
        return CUJPEGXL_INTERNAL_ERROR;
    }
    return CUJPEGXL_OK;
}

}  // extern "C"
