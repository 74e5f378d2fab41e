// Batches of frames of uneven sizes and distances, encoded by one launch per
// stage (batch.inl) captured once as a hipGraph; a project autotuner picks
// the batched launches' shapes (no counterpart in the reference, whose
// encoder runs one frame at a time, frame_encoder.cu:299-352).
#include "batch_encoder.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>

#include "kernels.h"

namespace cjx {

BatchEncoder::~BatchEncoder() {
    release();
}

void BatchEncoder::release() {
    if (device_ >= 0) {
        (void)hipSetDevice(device_);
    }
    if (stream_ != nullptr) {
        (void)hipStreamSynchronize(stream_);
    }
    if (exec_ != nullptr) {
        (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
    if (graph_ != nullptr) {
        (void)hipGraphDestroy(graph_);
        graph_ = nullptr;
    }
    frames_.clear();
    if (arena_ != nullptr) {
        (void)hipFree(arena_);
        arena_ = nullptr;
    }
    if (jobs_dev_ != nullptr) {
        (void)hipFree(jobs_dev_);
        jobs_dev_ = nullptr;
    }
    if (counters_ != nullptr) {
        (void)hipFree(counters_);
        counters_ = nullptr;
    }
    if (out_host_ != nullptr) {
        (void)hipHostFree(out_host_);
        out_host_ = nullptr;
    }
    if (sizes_host_ != nullptr) {
        (void)hipHostFree(sizes_host_);
        sizes_host_ = nullptr;
    }
    if (stream_ != nullptr) {
        (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
    }
    if (dc_stream_ != nullptr) {
        (void)hipStreamDestroy(dc_stream_);
        dc_stream_ = nullptr;
    }
    for (hipEvent_t* e : {&fork_, &join_}) {
        if (*e != nullptr) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
    }
    jobs_.clear();
    out_off_.clear();
    caps_.clear();
    inputs_set_ = false;
}

bool BatchEncoder::configure(int n, const std::uint32_t* w, const std::uint32_t* h, const float* d, int device,
                             std::uint32_t mode) {
    release();
    if (n <= 0 || n > kMaxBatch || (mode & ~kModeMask) != 0 || (mode & kModeForceMask) != 0 ||
        hipSetDevice(device) != hipSuccess || !ensure_device_tables(device)) {
        return false;
    }
    device_ = device;
    mode_ = mode;
    std::vector<std::size_t> ws(n), cap(n);
    std::size_t arena{0}, outs{0};
    for (int i{0}; i < n; ++i) {
        if (w[i] == 0 || h[i] == 0 || w[i] % 8 || h[i] % 8 || Geometry::of(w[i], h[i]).num_ac <= 1 ||
            Geometry::of(w[i], h[i]).num_dc > 32 || !(d[i] > 0.0f)) {
            return false;
        }
        ws[i] = (Encoder::workspace_bytes(w[i], h[i]) + 255) & ~static_cast<std::size_t>(255);
        cap[i] = (Encoder::max_output_size(w[i], h[i]) + 64 + 255) & ~static_cast<std::size_t>(255);
        arena += ws[i];
        outs += cap[i];
    }
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&dc_stream_, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&fork_, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&join_, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&arena_, arena) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&jobs_dev_), sizeof(FrameJob) * n) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&counters_), 8 * sizeof(std::uint32_t)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&out_host_), outs, hipHostMallocMapped) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sizes_host_), sizeof(std::uint64_t) * n, hipHostMallocMapped) !=
            hipSuccess) {
        release();
        return false;
    }
    std::uint8_t* out_dev{nullptr};
    std::uint64_t* sizes_dev{nullptr};
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&out_dev), out_host_, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&sizes_dev), sizes_host_, 0) != hipSuccess) {
        release();
        return false;
    }
    jobs_.assign(n, FrameJob{});
    std::size_t a{0}, o{0};
    for (int i{0}; i < n; ++i) {
        frames_.push_back(std::make_unique<Encoder>());
        Encoder& e{*frames_.back()};
        e.set_mode(mode & ~kModeForceMask);
        if (!e.bind_workspace(w[i], h[i], d[i], device, static_cast<std::uint8_t*>(arena_) + a, ws[i]) ||
            !e.fill_job(jobs_[i], stream_, out_dev + o, cap[i], sizes_dev + i)) {
            release();
            return false;
        }
        out_off_.push_back(o);
        caps_.push_back(cap[i]);
        a += ws[i];
        o += cap[i];
    }
    return apply_shape(shape_);
}

// Stage grids and first workgroups for `shape`, the jobs on the device, and
// a fresh graph of the batch's launches.
bool BatchEncoder::apply_shape(const BatchShape& shape) {
    if (stream_ == nullptr || hipStreamSynchronize(stream_) != hipSuccess) {
        return false;
    }
    shape_ = shape;
    std::fill(std::begin(total_), std::end(total_), 0u);
    for (FrameJob& j : jobs_) {
        batch_stage_grids(j.geo, shape_, j.grid);
        for (int s{0}; s < kBStages; ++s) {
            j.grid[s].first = total_[s];
            total_[s] += j.grid[s].nx * j.grid[s].ny;
        }
    }
    if (hipMemcpy(jobs_dev_, jobs_.data(), sizeof(FrameJob) * jobs_.size(), hipMemcpyHostToDevice) != hipSuccess) {
        return false;
    }
    if (exec_ != nullptr) {
        (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
    if (graph_ != nullptr) {
        (void)hipGraphDestroy(graph_);
        graph_ = nullptr;
    }
    if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        return false;
    }
    launch_batch(stream_, dc_stream_, fork_, join_, jobs_dev_, static_cast<int>(jobs_.size()), total_, mode_, shape_,
                 counters_);
    if (hipStreamEndCapture(stream_, &graph_) != hipSuccess ||
        hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0) != hipSuccess) {
        return false;
    }
    return true;
}

bool BatchEncoder::set_inputs(const std::uint8_t* const* luma, const std::size_t* luma_pitch,
                              const std::uint8_t* const* chroma, const std::size_t* chroma_pitch) {
    if (jobs_.empty() || hipSetDevice(device_) != hipSuccess) {
        return false;
    }
    for (std::size_t i{0}; i < jobs_.size(); ++i) {
        if (luma[i] == nullptr || chroma[i] == nullptr || luma_pitch[i] < jobs_[i].geo.width ||
            chroma_pitch[i] < jobs_[i].geo.width) {
            return false;
        }
        jobs_[i].luma = luma[i];
        jobs_[i].luma_pitch = luma_pitch[i];
        jobs_[i].chroma = chroma[i];
        jobs_[i].chroma_pitch = chroma_pitch[i];
    }
    inputs_set_ = true;
    // The graph reads the jobs from device memory: a new copy before the
    // next replay is all it takes (ordered on the batch's stream).
    return hipMemcpyAsync(jobs_dev_, jobs_.data(), sizeof(FrameJob) * jobs_.size(), hipMemcpyHostToDevice,
                          stream_) == hipSuccess &&
           hipStreamSynchronize(stream_) == hipSuccess;
}

bool BatchEncoder::launch(int count) {
    if (!inputs_set_ || exec_ == nullptr || hipSetDevice(device_) != hipSuccess) {
        return false;
    }
    for (int i{0}; i < count; ++i) {
        if (hipGraphLaunch(exec_, stream_) != hipSuccess) {
            return false;
        }
    }
    return true;
}

bool BatchEncoder::wait() {
    return stream_ != nullptr && hipSetDevice(device_) == hipSuccess && hipStreamSynchronize(stream_) == hipSuccess;
}

bool BatchEncoder::run(int count, double* seconds) {
    const auto t0{std::chrono::steady_clock::now()};
    const bool ok{launch(count) && hipStreamSynchronize(stream_) == hipSuccess};
    if (seconds != nullptr) {
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return ok;
}

std::uint64_t BatchEncoder::size(int i) const {
    return i >= 0 && i < static_cast<int>(jobs_.size()) ? sizes_host_[i] : 0;
}

const std::uint8_t* BatchEncoder::file(int i) const {
    return i >= 0 && i < static_cast<int>(jobs_.size()) ? out_host_ + out_off_[i] : nullptr;
}

bool BatchEncoder::set_shape(const BatchShape& shape) {
    if (shape.token_cap == 0 || shape.assemble_words == 0 || shape.region_wgs % 8 != 0 ||
        shape.region_wgs > 65536 || hipSetDevice(device_) != hipSuccess) {
        return false;
    }
    return apply_shape(shape);
}

// Coordinate search: the tokenizer and assembly shapes with one workgroup per
// region first, then the region stage's persistent grids with the best of
// those (one and two resident rounds of five workgroups per CU).
bool BatchEncoder::tune(int reps, std::string* report) {
    if (!inputs_set_ || reps <= 0) {
        return false;
    }
    static constexpr std::uint32_t kTokenCaps[]{256, 512, 1024, 2048};
    static constexpr std::uint32_t kAssembleWords[]{4, 16, 64};
    int cus{0};
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_) != hipSuccess || cus <= 0) {
        cus = 256;
    }
    const std::uint32_t resident{static_cast<std::uint32_t>(5 * cus) & ~7u};
    const std::uint32_t kRegionGrids[]{resident, 2 * resident};
    BatchShape best{shape_};
    double best_s{1e30};
    char line[192];
    const auto trial = [&](const BatchShape& c) {
        if (!apply_shape(c)) {
            return false;
        }
        double s{0.0};
        if (!run(2, nullptr) || !run(reps, &s)) {
            return false;
        }
        s /= reps;
        std::snprintf(line, sizeof(line), "token_cap=%u assemble_words=%u region_wgs=%u batch_ms=%.4f\n",
                      c.token_cap, c.assemble_words, c.region_wgs, 1e3 * s);
        if (report != nullptr) {
            *report += line;
        }
        if (s < best_s) {
            best_s = s;
            best = c;
        }
        return true;
    };
    for (std::uint32_t tc : kTokenCaps) {
        for (std::uint32_t aw : kAssembleWords) {
            if (!trial(BatchShape{tc, aw, 0})) {
                return false;
            }
        }
    }
    const BatchShape base{best};
    for (std::uint32_t rg : kRegionGrids) {
        if (!trial(BatchShape{base.token_cap, base.assemble_words, rg})) {
            return false;
        }
    }
    std::snprintf(line, sizeof(line), "chosen token_cap=%u assemble_words=%u region_wgs=%u batch_ms=%.4f\n",
                  best.token_cap, best.assemble_words, best.region_wgs, 1e3 * best_s);
    if (report != nullptr) {
        *report += line;
    }
    return apply_shape(best);
}

}  // namespace cjx
