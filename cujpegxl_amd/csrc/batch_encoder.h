// A batch of frames of uneven sizes and distances encoded together: every
// stage is one launch over all the frames (batch.inl), the batch's launches
// are captured once as a hipGraph and replayed, and the launches' shapes come
// from a small autotuner.  Files land in page-locked host memory mapped for
// the device (the assembly kernels write them there), sizes likewise.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "batch.h"
#include "encoder.h"

namespace cjx {

class BatchEncoder {
public:
    static constexpr int kMaxBatch = 256;

    BatchEncoder() = default;
    ~BatchEncoder();
    BatchEncoder(const BatchEncoder&) = delete;
    BatchEncoder& operator=(const BatchEncoder&) = delete;

    // The batch's frames (sizes, distances), all in encoder mode `mode`
    // (kModeMfma16 / kModeLearned bits); buffers for all of them in one arena.
    bool configure(int n, const std::uint32_t* w, const std::uint32_t* h, const float* d, int device,
                   std::uint32_t mode);
    // Each frame's NV12 device planes.
    bool set_inputs(const std::uint8_t* const* luma, const std::size_t* luma_pitch,
                    const std::uint8_t* const* chroma, const std::size_t* chroma_pitch);
    // `count` replays of the batch's graph, then a wait; wall seconds.
    bool run(int count, double* seconds);
    // `count` replays queued on the batch's own stream, no wait: another
    // batch's replays on its own stream may run beside them.
    bool launch(int count);
    // Waits for the queued replays (files and sizes are then valid).
    bool wait();
    std::uint64_t size(int i) const;
    const std::uint8_t* file(int i) const;
    // Times every candidate launch shape over `reps` replays and keeps the
    // fastest; one line per candidate into `report`.
    bool tune(int reps, std::string* report);
    // A launch shape of the caller's (region_wgs: 0, or a multiple of 8 up to
    // 65,536); re-captures the batch's graph.
    bool set_shape(const BatchShape& shape);
    const BatchShape& shape() const { return shape_; }
    int frames() const { return static_cast<int>(jobs_.size()); }

private:
    void release();
    bool apply_shape(const BatchShape& shape);

    int device_{-1};
    std::uint32_t mode_{0};
    BatchShape shape_{1024, 16, 0};
    std::vector<std::unique_ptr<Encoder>> frames_;
    std::vector<FrameJob> jobs_;
    std::vector<std::size_t> out_off_, caps_;
    std::uint32_t total_[kBStages]{};
    void* arena_{nullptr};
    FrameJob* jobs_dev_{nullptr};
    std::uint32_t* counters_{nullptr};  // the persistent region stage's claim counters (8)
    std::uint8_t* out_host_{nullptr};
    std::uint64_t* sizes_host_{nullptr};
    hipStream_t stream_{nullptr};
    hipStream_t dc_stream_{nullptr};  // the DC chain's branch of the captured graph
    hipEvent_t fork_{nullptr}, join_{nullptr};
    hipGraph_t graph_{nullptr};
    hipGraphExec_t exec_{nullptr};
    bool inputs_set_{false};
};

}  // namespace cjx
