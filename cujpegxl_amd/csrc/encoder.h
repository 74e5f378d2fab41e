// The frame encoder: owns one stream and every device buffer of a
// configuration, and runs the whole VarDCT pipeline for a frame.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "batch.h"
#include "common.h"

namespace cjx {

enum class Stage : int {
    kRegion = 0,   // XYB, gaborish, AQ, selection, DCT, CfL partials
    kQuantize,     // CfL maps, quantization, DC, non-zero counts
    kHistograms,   // AC cluster, DC and AcMetadata histograms
    kCodes,        // prefix codes and code headers
    kSizes,        // per-item bits, group scans, section layout
    kEmit,         // zero fill, headers, AC and DC token emission
    kReadback,     // size and bytes to the host
    kCount
};

// Internal buffers a harness can read back after an encode.
enum class Buffer : int {
    kCoeffs = 0, kAcs, kQuantField, kCflPartials, kYtoX, kYtoB, kAc, kDc, kNzGrid, kNzCount,
    kHistograms, kAcGroupBytes, kDcGroupBytes, kCodeDepth, kCodeBits, kFeatures, kCount
};

class Encoder {
public:
    Encoder() = default;
    ~Encoder();
    Encoder(const Encoder&) = delete;
    Encoder& operator=(const Encoder&) = delete;

    // Selects the device and (re)allocates for the frame size; false on error.
    bool configure(std::uint32_t width, std::uint32_t height, float distance, int device);

    // NV12 device planes -> `.jxl` bytes into `out` (capacity `cap`).
    // Returns false on a device error; `*size` is the file size either way.
    bool encode_nv12(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                     std::size_t chroma_pitch, std::uint8_t* out, std::size_t cap, std::size_t* size);

    // Asynchronous form: the frame's work is queued on the encoder's stream
    // after the work already on `caller` (which in turn waits for the region
    // pass, the only reader of the planes, before its later work); finish()
    // waits and delivers the bytes.  One frame may be pending at a time.
    bool encode_async(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                      std::size_t chroma_pitch, hipStream_t caller);
    bool finish(std::uint8_t* out, std::size_t cap, std::size_t* size);
    bool pending() const { return pending_; }
    // True once a configure has succeeded (and no later one has failed).
    bool configured() const { return configured_; }

    // NV12 host planes: uploaded into an encoder-owned device frame, then
    // encoded as encode_nv12 does.
    bool encode_host(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                     std::size_t chroma_pitch, std::uint8_t* out, std::size_t cap, std::size_t* size);

    // Config-1 semantics (the reference's dct.cu + quant.cu test path): DCT8
    // everywhere, uniform quant field, no CfL maps; same entropy coding.
    bool encode_dct8_uniform(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                             std::size_t chroma_pitch, std::uint8_t* out, std::size_t cap, std::size_t* size);

    // Entropy coding and assembly only, for coefficients supplied from the host
    // (same layouts as the front end's outputs).
    bool encode_coefficients(const std::int16_t* ac, const std::int32_t* dc, const std::int8_t* acs,
                             const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                             std::uint32_t global_scale, std::uint32_t quant_dc, std::uint8_t* out,
                             std::size_t cap, std::size_t* size);

    bool fetch(Buffer which, void* host, std::size_t bytes);
    std::size_t buffer_bytes(Buffer which) const;

    // Workspace form, for callers that cannot allocate.  Every device buffer
    // of the shape is a 256-byte aligned range of `base` (no allocation, no
    // stream, event or pinned word is created), and encode_on() queues a
    // whole frame on the caller's stream alone, the file into a device
    // output and its size into a device word.  workspace_bytes() is what
    // bind_workspace() needs for a shape.
    static std::size_t workspace_bytes(std::uint32_t width, std::uint32_t height);
    bool bind_workspace(std::uint32_t width, std::uint32_t height, float distance, int device, void* base,
                        std::size_t bytes);
    bool encode_on(hipStream_t s, const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                   std::size_t chroma_pitch, std::uint8_t* out_dev, std::size_t cap, std::uint64_t* size_dev);
#ifdef CJX_REGION_CUT
    // Profiling builds: the region kernel alone on the encoder's stream, then
    // a wait for it.
    bool region_only(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                     std::size_t chroma_pitch);
#endif
    // A workspace-bound encoder's frame as a batch job: its constant bit
    // strings uploaded (in order on `s`), its buffers, output and size word;
    // the stage grids are the caller's.
    bool fill_job(FrameJob& job, hipStream_t s, std::uint8_t* out_dev, std::size_t cap, std::uint64_t* size_dev);

    // 0: no events; 1: an event pair around every stage; 2: the region
    // kernel's pair only (each timed event on the stream costs a few
    // microseconds of idle device time between the kernels it separates).
    void set_profiling(int level) { profiling_ = level; }
    // kMode* bits (0: the plain path); applies from the next encode.
    void set_mode(std::uint32_t mode) {
        mode_ = mode;
        prm_.mode = mode;
    }
    std::uint32_t mode() const { return mode_; }
    // The region kernel also writes each region's network features and the
    // heuristic's labels (Buffer::kFeatures, 108 floats per region) from the
    // next encode; the buffer is allocated on the first call.
    bool export_features(bool on);
    // Milliseconds per stage of the last encode (profiling on; with level 2
    // only the region entry is filled).
    const std::array<float, static_cast<int>(Stage::kCount)>& stage_ms() const { return stage_ms_; }
    hipStream_t stream() const { return stream_; }
    const Geometry& geometry() const { return geo_; }

    static std::size_t max_output_size(std::uint32_t width, std::uint32_t height);

private:
    bool setup(int device);
    bool alloc();
    // Points every device buffer except the output at get(pointer, bytes).
    template <class Get>
    void carve(Get& get);
    void release();
    void null_buffers();
    // Uploads the constant bit strings on `s` (the encoder's stream when
    // null) unless they are already those of (global_scale, quant_dc).
    bool build_constants(std::uint32_t global_scale, std::uint32_t quant_dc, hipStream_t s = nullptr);
    // `fork_bound`: the quantizer's launch has bound `ev_fork_` already;
    // `listed`: it also listed the non-empty items for the tokenizer.
    bool run_entropy(std::uint8_t* out, std::size_t cap, hipEvent_t fork, bool listed, bool maps = false,
                     bool size_event = true);
    std::uint8_t* mapped_output(std::uint8_t* dst) const;
    Clears entropy_clears() const;
    // Queues a frame after the work on `caller` (null: the null stream);
    // `release_caller`: later work on `caller` waits for the region kernel.
    bool submit_nv12(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                     std::size_t chroma_pitch, hipStream_t caller, bool release_caller, std::uint8_t* dst,
                     std::size_t cap);
    bool queue_readback(std::uint8_t* direct);
    bool stage_chunk(int slot, std::size_t n);
    bool upload_nv12(const std::uint8_t* luma, std::size_t luma_pitch, const std::uint8_t* chroma,
                     std::size_t chroma_pitch);
    void mark(Stage s, hipStream_t on = nullptr);
    bool wait_done();
    static bool spin_on(hipEvent_t e);

    int device_{-1};
    bool configured_{false};
    bool workspace_{false};  // buffers bound to a caller's workspace (bind_workspace)
    float distance_{1.0f};
    Geometry geo_{};
    FrameParams prm_{};
    hipStream_t stream_{nullptr};
    hipStream_t dc_stream_{nullptr};  // DC entropy chain, joined by events
    hipEvent_t ev_fork_{nullptr}, ev_dc_done_{nullptr};  // fork onto / join from the DC stream
    hipEvent_t ev_size_{nullptr};                        // the total size has reached the host
    hipEvent_t ev_done_{nullptr};                        // the frame's last copy has finished
    hipEvent_t ev_end_{nullptr};  // the assembly kernel's completion, when it ends the frame
    bool end_bound_{false};       // ev_end_ is bound to this frame's assembly kernel
    bool size_bound_{true};  // ev_size_ is bound to this frame's layout kernel
    hipStream_t tail_{nullptr};  // layout, assembly and readback: the DC stream, after both chains
    hipEvent_t ev_caller_{nullptr}, ev_input_free_{nullptr};  // ordering against a caller's stream
    bool pending_{false};            // a frame's readback is queued, finish() not yet called
    std::size_t staged_{0};          // output bytes queued into the staging chunks
    std::size_t stage_len_[2]{0, 0};
    std::uint8_t* direct_dst_{nullptr};  // page-locked output the assembly wrote the file into
    std::uint8_t* input_{nullptr};   // device NV12 frame for host input (allocated on first use)
    int profiling_{0};
    std::uint32_t mode_{0};
    std::array<hipEvent_t, static_cast<int>(Stage::kCount) + 1> events_{};
    std::array<float, static_cast<int>(Stage::kCount)> stage_ms_{};
    std::uint32_t const_gs_{0}, const_qdc_{0};

    // Device buffers.
    std::uint16_t* coeffs_{nullptr};
    std::int8_t* acs_{nullptr};
    std::int32_t* qf_{nullptr};
    float* cfl_part_{nullptr};
    std::int8_t* ytox_{nullptr};
    std::int8_t* ytob_{nullptr};
    std::int16_t* ac_{nullptr};
    std::int32_t* dc_{nullptr};
    std::uint8_t* nz_grid_{nullptr};
    std::uint16_t* nz_count_{nullptr};
    std::uint32_t* hists_{nullptr};
    std::uint32_t* fb_count_{nullptr};
    void* codes_{nullptr};
    void* constants_{nullptr};
    void* blobs_{nullptr};
    std::uint32_t* blob_words_{nullptr};
    std::uint32_t* tokens_{nullptr};       // token records in coefficient-slot layout
    std::uint16_t* item_tokens_{nullptr};  // tokens per (block, channel) item
    std::uint32_t* item_list_{nullptr};    // first-block items (quantizer -> tokenizer), ItemList
    std::uint32_t* item_count_{nullptr};   // entries at its front and back
    std::uint32_t* hist_partial_{nullptr}; // per-workgroup AC histograms
    std::uint32_t* ac_staging_{nullptr};   // each AC group's stream at group-relative bits
    std::uint32_t* dc_staging_{nullptr};   // each DC group's section at group-relative bits
    std::uint64_t* ac_part_bits_{nullptr};   // bits of each part (row band) of each AC group
    std::uint32_t* ac_group_bytes_{nullptr};
    std::uint64_t* ac_group_bits_{nullptr};
    std::uint64_t* chunk_bits_{nullptr};
    std::uint32_t* dc_group_bytes_{nullptr};
    std::uint64_t* dc_group_bits_{nullptr};
    void* layout_{nullptr};
    std::uint64_t* ac_base_{nullptr};
    std::uint64_t* dc_base_{nullptr};
    std::uint32_t* out_{nullptr};
    std::size_t out_cap_{0};
    float* features_{nullptr};  // per-region features and labels (export on)
    bool export_{false};
    std::uint64_t* total_host_{nullptr};  // pinned; the layout kernel writes the file size here
    std::uint64_t* total_dev_{nullptr};   // the same word as a device address
    // Readback: the file is copied in chunks through two pinned slots, the
    // device-to-host copy of one chunk overlapping the host copy of the last.
    static constexpr std::size_t kStageChunk = 512 * 1024;
    std::uint8_t* stage_host_{nullptr};   // pinned, 2 * kStageChunk
    std::array<hipEvent_t, 2> stage_ready_{};
    std::vector<void*> allocations_{};
};

bool device_exists(int ordinal);
bool device_has_matrix_cores(int ordinal);
// The constant tables (and the default network weights) on `device`, once;
// the device must be current.
bool ensure_device_tables(int device);
// Test hook: the n-th device allocation from now fails (0 cancels).
void fail_allocation_after(int n);

}  // namespace cjx
