// Batched launches: one kernel per stage over every frame of a batch
// (included into device.hip after the stage bodies).  A workgroup finds its
// frame by a binary search over the frames' first workgroups of the stage
// (blockIdx.x is uniform, so the search runs on scalar loads) and runs the
// stage's body with the frame's own workgroup coordinates.

namespace cjx {

// The frame of batched workgroup `b` of a stage and the workgroup's
// coordinates in that frame's own grid.
__device__ __forceinline__ const FrameJob& batch_job_at(const FrameJob* __restrict__ jobs, int count, int stage,
                                                        std::uint32_t b, Wg& wg) {
    int lo{0}, hi{count - 1};
    while (lo < hi) {
        const int mid{(lo + hi + 1) >> 1};
        if (jobs[mid].grid[stage].first <= b) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    const FrameJob& j{jobs[lo]};
    const StageGrid& g{j.grid[stage]};
    const std::uint32_t local{b - g.first};
    wg = Wg{local % g.nx, local / g.nx, g.nx, g.ny};
    return j;
}

__device__ __forceinline__ const FrameJob& batch_job(const FrameJob* __restrict__ jobs, int count, int stage,
                                                     Wg& wg) {
    return batch_job_at(jobs, count, stage, blockIdx.x, wg);
}

// The DC passes' chunk count per group and staging slot size (the values
// the per-frame launchers pass).
constexpr std::uint32_t kBatchDcChunks{(2 + 6 * 65536 + 2 * 1024 + kDcChunk - 1) / kDcChunk};
constexpr std::uint64_t kBatchDcSlotWords{static_cast<std::uint64_t>(kBatchDcChunks) * kDcChunk + 2 * kBlobWords + 64};

template <bool kMf16, bool kLearned>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void region_batch_kernel(
    const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBRegion, wg)};
    region_kernel_body<kMf16, kLearned>(wg, j.geo, j.prm, j.luma, j.luma_pitch, j.chroma, j.chroma_pitch, j.coeffs,
                                        j.acs, j.qf, j.cfl_part, j.item_count, j.clr, nullptr);
}

// The region stage as a persistent grid: workgroup i stays on XCD i mod 8
// and claims, through that XCD's counter, the batched region workgroups
// b = 8 t + (i mod 8) in turn (t = 0, 1, ...), which are the ones the
// one-workgroup-per-region launch would place on the same XCD: the same
// contiguous runs of regions per XCD.  Every workgroup leaves once its
// XCD's share is claimed; no workgroup waits for another.
template <bool kMf16, bool kLearned>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void region_persistent_batch_kernel(
    const FrameJob* __restrict__ jobs, int count, std::uint32_t total, std::uint32_t* __restrict__ counters) {
    __shared__ std::uint32_t claimed;
    const std::uint32_t x{blockIdx.x & 7u};
    for (;;) {
        __syncthreads();  // every thread is done with the previous region and with `claimed`
        if (threadIdx.x == 0) {
            claimed = atomicAdd(&counters[x], 1u);
        }
        __syncthreads();
        const std::uint32_t b{8u * claimed + x};
        if (b >= total) {
            break;
        }
        Wg wg;
        const FrameJob& j{batch_job_at(jobs, count, kBRegion, b, wg)};
        region_kernel_body<kMf16, kLearned>(wg, j.geo, j.prm, j.luma, j.luma_pitch, j.chroma, j.chroma_pitch,
                                            j.coeffs, j.acs, j.qf, j.cfl_part, j.item_count, j.clr, nullptr);
    }
}

__global__ void clear_counters_kernel(std::uint32_t* __restrict__ counters) {
    if (threadIdx.x < 8) {
        counters[threadIdx.x] = 0;
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void quantize_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBQuantize, wg)};
    tile_quantize_kernel_body(wg, j.geo, j.prm, j.coeffs, j.acs, j.qf, j.cfl_part, j.ac, j.nz_grid, j.nz_count,
                              ItemList{j.item_list, j.item_count});
}

__global__ __launch_bounds__(256) void dc_maps_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBDcMaps, wg)};
    dc_maps_kernel_body(wg, j.geo, j.prm, j.coeffs, j.acs, j.cfl_part, j.dc, j.ytox, j.ytob);
}

__global__ __launch_bounds__(256) void dc_hist_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBDcHist, wg)};
    dc_hist_kernel_body(wg, j.geo, j.dc, j.acs, j.qf, j.ytox, j.ytob, j.hists + kNumClusters * kHistStride,
                        j.hists + (kNumClusters + j.geo.num_dc) * kHistStride, j.fb_count);
}

template <BatchStage kStage>
__global__ __launch_bounds__(64 * kTrials) void huffman_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kStage, wg)};
    huffman_kernel_body(wg, j.hists, j.codes, kStage == kBHuffDc ? kNumClusters : 0);
}

__global__ __launch_bounds__(256) void dc_blobs_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBDcBlobs, wg)};
    dc_blobs_bits_kernel_body(wg, j.geo, j.k, j.codes, j.fb_count, j.blobs, j.blob_words, j.dc, j.acs, j.qf, j.ytox,
                              j.ytob, j.chunk_bits, kBatchDcChunks);
}

__global__ __launch_bounds__(1024) void dc_scan_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBDcScan, wg)};
    dc_chunk_scan_kernel_body(wg, j.geo, j.chunk_bits, kBatchDcChunks, j.blobs, j.dc_group_bytes, j.dc_group_bits,
                              j.dc_staging, kBatchDcSlotWords);
}

__global__ __launch_bounds__(256) void dc_emit_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBDcEmit, wg)};
    const std::uint32_t nd{j.geo.num_dc};
    dc_emit_kernel_body(wg, j.geo, j.dc, j.acs, j.qf, j.ytox, j.ytob, &j.codes->depth[kNumClusters][0],
                        &j.codes->code[kNumClusters][0], &j.codes->depth[kNumClusters + nd][0],
                        &j.codes->code[kNumClusters + nd][0], j.blobs, j.blob_words, j.chunk_bits,
                        kBatchDcChunks, j.dc_staging, kBatchDcSlotWords);
}

__global__ __launch_bounds__(256) void tokens_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBTokens, wg)};
    ac_tokens_kernel_body(wg, j.geo, j.ac, j.acs, j.nz_grid, j.nz_count, j.tokens, j.item_tokens, j.hist_partial,
                          j.item_list, j.item_count);
}

__global__ __launch_bounds__(256) void reduce_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBReduce, wg)};
    ac_hist_reduce_kernel_body(wg, j.hist_partial, static_cast<int>(j.grid[kBTokens].nx), j.hists);
}

__global__ __launch_bounds__(kPartThreads) void ac_emit_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBAcEmit, wg)};
    ac_group_emit_kernel_body(wg, j.geo, j.acs, j.item_tokens, j.tokens, &j.codes->depth[0][0],
                              &j.codes->code[0][0], j.ac_staging, j.ac_part_bits);
}

__global__ __launch_bounds__(256) void layout_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBLayout, wg)};
    layout_kernel_body(wg, j.geo, j.k, j.codes, j.ac_part_bits, j.dc_group_bytes, j.dc_group_bits, j.ac_group_bytes,
                       j.ac_group_bits, j.blobs, j.lay, j.ac_base, j.dc_base, j.size_out);
}

__global__ __launch_bounds__(256) void assemble_batch_kernel(const FrameJob* __restrict__ jobs, int count) {
    Wg wg;
    const FrameJob& j{batch_job(jobs, count, kBAssemble, wg)};
    const Pieces P{j.dc_staging, kBatchDcSlotWords, j.ac_staging, j.dc_group_bits, j.ac_group_bits,
                   j.ac_part_bits, j.dc_base, j.ac_base};
    assemble_kernel_body(wg, j.geo, j.k, j.lay, P, j.out, j.cap);
}

void batch_stage_grids(const Geometry& geo, const BatchShape& shape, StageGrid (&g)[kBStages]) {
    const auto set = [&](BatchStage s, std::uint32_t nx, std::uint32_t ny) {
        g[s] = StageGrid{0, nx > 0 ? nx : 1, ny > 0 ? ny : 1, 0};
    };
    const std::uint32_t cpg{dc_chunks_per_group()};
    set(kBRegion, 8 * ((geo.rw * geo.rh + 7) / 8), 1);
    set(kBQuantize, geo.cmw * geo.cmh, 1);
    set(kBDcMaps, (geo.nblocks + 255) / 256, 1);
    set(kBDcHist, cpg, geo.num_dc);
    set(kBHuffDc, 2 * geo.num_dc, 1);
    set(kBDcBlobs, cpg + 1, geo.num_dc);
    set(kBDcScan, geo.num_dc, 1);
    set(kBDcEmit, cpg, geo.num_dc);
    const std::uint32_t waves{(3 * geo.nblocks + 3) / 4};
    std::uint32_t cap{shape.token_cap < static_cast<std::uint32_t>(kTokenGrid) ? shape.token_cap
                                                                               : static_cast<std::uint32_t>(kTokenGrid)};
    set(kBTokens, waves < cap ? waves : cap, 1);
    set(kBReduce, kNumClusters * kAcSymbols / 256, 80);
    set(kBHuffAc, kNumClusters, 1);
    set(kBAcEmit, geo.num_ac * kGroupParts, 1);
    set(kBLayout, 1, 1);
    const std::uint64_t nb{geo.nblocks};
    const std::uint64_t words{(3 * nb * 384 + 48 * nb + (geo.num_ac + geo.num_dc) * 2048ull + 16384 + 3) / 4};
    const std::uint64_t per_wg{256ull * (shape.assemble_words > 0 ? shape.assemble_words : 1)};
    const std::uint64_t ag{(words + per_wg - 1) / per_wg};
    set(kBAssemble, static_cast<std::uint32_t>(ag < 2048 ? ag : 2048), 1);
}

void launch_batch(hipStream_t s, hipStream_t s_dc, hipEvent_t fork, hipEvent_t join, const FrameJob* jobs,
                  int count, const std::uint32_t (&total)[kBStages], std::uint32_t mode, const BatchShape& shape,
                  std::uint32_t* counters) {
    const bool mf{(mode & kModeMfma16) != 0}, learned{(mode & kModeLearned) != 0};
    if (shape.region_wgs == 0) {
        auto* const region{mf ? (learned ? region_batch_kernel<true, true> : region_batch_kernel<true, false>)
                              : (learned ? region_batch_kernel<false, true> : region_batch_kernel<false, false>)};
        region<<<total[kBRegion], 256, 0, s>>>(jobs, count);
    } else {
        auto* const region{mf ? (learned ? region_persistent_batch_kernel<true, true>
                                         : region_persistent_batch_kernel<true, false>)
                              : (learned ? region_persistent_batch_kernel<false, true>
                                         : region_persistent_batch_kernel<false, false>)};
        clear_counters_kernel<<<1, 64, 0, s>>>(counters);
        region<<<shape.region_wgs, 256, 0, s>>>(jobs, count, total[kBRegion], counters);
    }
    // The DC chain (DC values and colour maps, DC / AcMetadata histograms and
    // codes, header blobs, sizes, emission) reads only what the region kernel
    // wrote: it forks here and joins before the layout, beside the AC chain.
    (void)hipEventRecord(fork, s);
    (void)hipStreamWaitEvent(s_dc, fork, 0);
    dc_maps_batch_kernel<<<total[kBDcMaps], 256, 0, s_dc>>>(jobs, count);
    dc_hist_batch_kernel<<<total[kBDcHist], 256, 0, s_dc>>>(jobs, count);
    huffman_batch_kernel<kBHuffDc><<<total[kBHuffDc], 64 * kTrials, 0, s_dc>>>(jobs, count);
    dc_blobs_batch_kernel<<<total[kBDcBlobs], 256, 0, s_dc>>>(jobs, count);
    dc_scan_batch_kernel<<<total[kBDcScan], 1024, 0, s_dc>>>(jobs, count);
    dc_emit_batch_kernel<<<total[kBDcEmit], 256, 0, s_dc>>>(jobs, count);
    (void)hipEventRecord(join, s_dc);
    quantize_batch_kernel<<<total[kBQuantize], 256, 0, s>>>(jobs, count);
    tokens_batch_kernel<<<total[kBTokens], 256, 0, s>>>(jobs, count);
    reduce_batch_kernel<<<total[kBReduce], 256, 0, s>>>(jobs, count);
    huffman_batch_kernel<kBHuffAc><<<total[kBHuffAc], 64 * kTrials, 0, s>>>(jobs, count);
    ac_emit_batch_kernel<<<total[kBAcEmit], kPartThreads, 0, s>>>(jobs, count);
    (void)hipStreamWaitEvent(s, join, 0);
    layout_batch_kernel<<<total[kBLayout], 256, 0, s>>>(jobs, count);
    assemble_batch_kernel<<<total[kBAssemble], 256, 0, s>>>(jobs, count);
}

}  // namespace cjx
