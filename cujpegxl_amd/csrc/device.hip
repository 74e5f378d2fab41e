// The encoder's device code in one translation unit (the kernels share the
// __constant__ tables), plus the host-callable launchers.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstring>

#include "common.h"
#include "kernels.h"
#include "prefix.h"

namespace cjx {
__constant__ __align__(16) DeviceTables g_tables;
}  // namespace cjx

#include "frontend.inl"
#include "entropy.inl"
#include "huffman.inl"
#include "assemble.inl"
#include "mlp.inl"
#include "batch.h"
#include "batch.inl"

namespace cjx {

namespace {
#include "mlp_weights.inc"
}  // namespace

bool init_device_tables() {
    const DeviceTables& t{host_tables()};
    return hipMemcpyToSymbol(HIP_SYMBOL(g_tables), &t, sizeof(t)) == hipSuccess && set_default_mlp_weights();
}

bool set_default_mlp_weights() {
    return set_mlp_weights(kTrainedW1, kTrainedB1, kTrainedW2, kTrainedB2, kTrainedThreshold);
}

bool set_mlp_weights(const std::uint16_t* w1, const float* b1, const std::uint16_t* w2, const float* b2,
                     float threshold) {
    MlpWeights m{};
    std::memcpy(m.w1, w1, sizeof(m.w1));
    std::memcpy(m.w2, w2, sizeof(m.w2));
    std::memcpy(m.b1, b1, sizeof(m.b1));
    std::memcpy(m.b2, b2, 7 * sizeof(float));
    m.threshold = threshold;
    return hipMemcpyToSymbol(HIP_SYMBOL(g_mlp), &m, sizeof(m)) == hipSuccess;
}

std::size_t mlp_partial_floats(std::uint32_t n) {
    return static_cast<std::size_t>((n + kMlpTiles - 1) / kMlpTiles) * (kMlpParams + 1);
}

void launch_mlp_step(hipStream_t s, const float* x, const float* labels, std::uint32_t n, float lambda,
                     float* partial, float* grad, float* loss, float* params, float* m, float* v, std::uint32_t step,
                     float lr) {
    const std::uint32_t rows{(n + kMlpTiles - 1) / kMlpTiles};
    mlp_grad_kernel<<<rows, 256, 0, s>>>(params, x, labels, n, lambda, partial);
    mlp_reduce_kernel<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(partial, rows, n, grad, loss, params, m, v, step,
                                                                   lr);
}

void launch_mlp_forward(hipStream_t s, const float* params, const float* x, std::uint32_t n, float* z) {
    mlp_forward_kernel<<<(n + 255) / 256, 256, 0, s>>>(params, x, n, z);
}

// One thread per (block position, channel): non-zero AC counts of a
// first-block and the normalized grid over its covered positions, for frames
// whose coefficients are supplied directly.
__global__ void nz_kernel(Geometry geo, const std::int16_t* __restrict__ ac, const std::int8_t* __restrict__ acs,
                          std::uint8_t* __restrict__ nz_grid, std::uint16_t* __restrict__ nz_count, Clears clr) {
    clear_words(clr, Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y});
    const std::size_t i{blockIdx.x * 256ull + threadIdx.x};
    const std::size_t nblk{geo.nblocks};
    if (i >= 3 * nblk) {
        return;
    }
    const int c{static_cast<int>(i / nblk)};
    const std::size_t pos{i % nblk};
    const int n{acs[pos]};
    if (n == 0) {
        return;
    }
    const int bx{static_cast<int>(pos % geo.bw)}, by{static_cast<int>(pos / geo.bw)};
    const int m{n / 8};
    const std::int16_t* plane{ac + c * nblk * 64};
    std::uint32_t cnt{0};
    for (int e{0}; e < n * n; ++e) {
        if (e / n < m && e % n < m) {
            continue;
        }
        const int q{e >> 6};
        cnt += plane[(static_cast<std::size_t>(by + q / m) * geo.bw + bx + q % m) * 64 + (e & 63)] != 0;
    }
    nz_count[c * nblk + pos] = static_cast<std::uint16_t>(cnt);
    const std::uint32_t l2{n == 16 ? 2u : (n == 32 ? 4u : 0u)};
    for (int dy{0}; dy < m; ++dy) {
        for (int dx{0}; dx < m; ++dx) {
            nz_grid[c * nblk + static_cast<std::size_t>(by + dy) * geo.bw + bx + dx] =
                static_cast<std::uint8_t>((cnt + m * m - 1) >> l2);
        }
    }
}

__global__ void fill_u32_kernel(std::uint32_t* p, std::size_t n, std::uint32_t v) {
    for (std::size_t i{blockIdx.x * 256ull + threadIdx.x}; i < n; i += gridDim.x * 256ull) {
        p[i] = v;
    }
}

bool check_aq_math(unsigned long long* mismatches) {
    unsigned long long* dev{nullptr};
    if (hipMalloc(reinterpret_cast<void**>(&dev), sizeof(unsigned long long)) != hipSuccess) {
        return false;
    }
    bool ok{hipMemset(dev, 0, sizeof(unsigned long long)) == hipSuccess};
    if (ok) {
        // v in [0, 4): bit patterns 0 .. 0x40800000; masking_sqrt below 0.25.
        check_aq_math_kernel<<<8192, 256>>>(0u, 0x40800000u, 0x3e800000u, dev);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpy(mismatches, dev, sizeof(unsigned long long), hipMemcpyDeviceToHost) == hipSuccess;
    }
    (void)hipFree(dev);
    return ok;
}

std::size_t sizeof_constants() { return sizeof(Constants); }
std::size_t sizeof_codeset() { return sizeof(CodeSet); }
std::size_t sizeof_layout() { return sizeof(Layout); }
std::size_t sizeof_dc_blobs() { return sizeof(DcBlobs); }
std::uint32_t dc_chunks_per_group() { return (2 + 6 * 65536 + 2 * 1024 + kDcChunk - 1) / kDcChunk; }
std::uint32_t blob_words_per_group() { return 2 * kBlobWords; }

bool constants_fit(std::uint32_t head_bits, std::uint32_t ac_prefix_bits, std::uint32_t dc_pre_bits,
                   std::uint32_t dc_mid_bits) {
    return head_bits <= 32u * kHeadConstWords && ac_prefix_bits <= 32u * (kAcGlobalWords / 2) &&
           dc_pre_bits <= 32u * 64 && dc_mid_bits <= 32u * 64;
}

void fill_constants(void* dst, const std::uint32_t* head, std::uint32_t head_bits,
                    const std::uint32_t* dc_global, std::uint32_t dc_global_bits,
                    const std::uint32_t* ac_prefix, std::uint32_t ac_prefix_bits,
                    const std::uint32_t* dc_pre, std::uint32_t dc_pre_bits, const std::uint32_t* dc_mid,
                    std::uint32_t dc_mid_bits) {
    Constants* k{static_cast<Constants*>(dst)};
    std::memset(k, 0, sizeof(Constants));
    k->head_bits = head_bits;
    k->dc_global_bits = dc_global_bits;
    k->ac_prefix_bits = ac_prefix_bits;
    k->dc_pre_bits = dc_pre_bits;
    k->dc_mid_bits = dc_mid_bits;
    std::memcpy(k->head_words, head, (head_bits + 31) / 32 * 4);
    std::memcpy(k->dc_global_words, dc_global, (dc_global_bits + 31) / 32 * 4);
    std::memcpy(k->ac_prefix_words, ac_prefix, (ac_prefix_bits + 31) / 32 * 4);
    std::memcpy(k->dc_pre_words, dc_pre, (dc_pre_bits + 31) / 32 * 4);
    std::memcpy(k->dc_mid_words, dc_mid, (dc_mid_bits + 31) / 32 * 4);
}

static unsigned int grid_for(std::size_t n, std::size_t per) {
    return static_cast<unsigned int>((n + per - 1) / per);
}

void launch_region(hipStream_t s, const Geometry& geo, const FrameParams& prm, const std::uint8_t* luma,
                   std::size_t luma_pitch, const std::uint8_t* chroma, std::size_t chroma_pitch,
                   std::uint16_t* coeffs, std::int8_t* acs, std::int32_t* qf, float* cfl_part,
                   std::uint32_t* list_count, const Clears& clr, hipEvent_t start, hipEvent_t stop,
                   float* features) {
    // Eight equal runs of regions, one per XCD (the kernel maps workgroup i to
    // run i mod 8); workgroups past the last region exit at once.  The events,
    // when given, take the kernel's own start and end times (no marker
    // packets between it and its neighbours).
    const unsigned int per_xcd{(geo.rw * geo.rh + 7) / 8};
    const bool mf{(prm.mode & kModeMfma16) != 0}, learned{(prm.mode & kModeLearned) != 0};
    auto* const kernel{mf ? (learned ? region_kernel<true, true> : region_kernel<true, false>)
                          : (learned ? region_kernel<false, true> : region_kernel<false, false>)};
    hipExtLaunchKernelGGL(kernel, dim3(8 * per_xcd), dim3(256), 0, s, start, stop, 0, geo, prm, luma, luma_pitch,
                          chroma, chroma_pitch, coeffs, acs, qf, cfl_part, list_count, clr, features);
}

void launch_dct8_uniform(hipStream_t s, const Geometry& geo, const FrameParams& prm, const std::uint8_t* luma,
                         std::size_t luma_pitch, const std::uint8_t* chroma, std::size_t chroma_pitch,
                         std::int16_t* ac, std::int32_t* dc, std::int8_t* acs, std::int32_t* qf, std::int8_t* ytox,
                         std::int8_t* ytob) {
    const unsigned int per_xcd{(geo.rw * geo.rh + 7) / 8};
    dct8_uniform_kernel<<<8 * per_xcd, 256, 0, s>>>(geo, prm, luma, luma_pitch, chroma, chroma_pitch, ac, dc, acs,
                                                    qf, ytox, ytob);
}

void launch_tile_quantize(hipStream_t s, const Geometry& geo, const FrameParams& prm,
                          const std::uint16_t* coeffs, const std::int8_t* acs, const std::int32_t* qf,
                          const float* cfl_part, std::int16_t* ac, std::uint8_t* nz_grid,
                          std::uint16_t* nz_count, const ItemList& list) {
    tile_quantize_kernel<<<geo.cmw * geo.cmh, 256, 0, s>>>(geo, prm, coeffs, acs, qf, cfl_part, ac, nz_grid,
                                                           nz_count, list);
}

void launch_dc_maps(hipStream_t s, const Geometry& geo, const FrameParams& prm, const std::uint16_t* coeffs,
                    const std::int8_t* acs, const float* cfl_part, std::int32_t* dc, std::int8_t* ytox,
                    std::int8_t* ytob) {
    dc_maps_kernel<<<grid_for(geo.nblocks, 256), 256, 0, s>>>(geo, prm, coeffs, acs, cfl_part, dc, ytox, ytob);
}

void launch_nz(hipStream_t s, const Geometry& geo, const std::int16_t* ac, const std::int8_t* acs,
               std::uint8_t* nz_grid, std::uint16_t* nz_count, const Clears& clr) {
    nz_kernel<<<grid_for(3ull * geo.nblocks, 256), 256, 0, s>>>(geo, ac, acs, nz_grid, nz_count, clr);
}

void launch_fill_u32(hipStream_t s, std::uint32_t* p, std::size_t n, std::uint32_t v) {
    const unsigned int g{grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024};
    fill_u32_kernel<<<g > 0 ? g : 1, 256, 0, s>>>(p, n, v);
}


void launch_ac_tokens(hipStream_t s, const Geometry& geo, const std::int16_t* ac, const std::int8_t* acs,
                      const std::uint8_t* nz_grid, const std::uint16_t* nz_count, std::uint32_t* tokens,
                      std::uint16_t* item_tokens, std::uint32_t* partial, std::uint32_t* hist,
                      const ItemList& list) {
    // One resident round of workgroups (occupancy x CUs of the current
    // device, at most kTokenGrid): a second partial round would double the
    // kernel's time for a few waves.
    static const unsigned int resident{[] {
        int per_cu{0}, dev{0}, cus{0};
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ac_tokens_kernel, 256, 0) != hipSuccess ||
            hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            per_cu <= 0 || cus <= 0) {
            return static_cast<unsigned int>(kTokenGrid);
        }
        const int n{per_cu * cus};
        return static_cast<unsigned int>(n < kTokenGrid ? n : kTokenGrid);
    }()};
    const unsigned int waves{(3 * geo.nblocks + 3) / 4};
    const unsigned int grid{waves < resident ? waves : resident};
    ac_tokens_kernel<<<grid, 256, 0, s>>>(geo, ac, acs, nz_grid, nz_count, tokens, item_tokens, partial, list.items,
                                          list.count);
    ac_hist_reduce_kernel<<<dim3(kNumClusters * kAcSymbols / 256, 80), 256, 0, s>>>(partial, static_cast<int>(grid),
                                                                                    hist);
}

std::size_t ac_partial_words() {
    return static_cast<std::size_t>(kTokenGrid) * kNumClusters * kAcSymbols;
}

void launch_dc_hist(hipStream_t s, const Geometry& geo, const std::int32_t* dc, const std::int8_t* acs,
                    const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                    std::uint32_t* dc_hist, std::uint32_t* meta_hist, std::uint32_t* fb_count) {
    const dim3 grid{dc_chunks_per_group(), geo.num_dc};
    dc_hist_kernel<<<grid, 256, 0, s>>>(geo, dc, acs, qf, ytox, ytob, dc_hist, meta_hist, fb_count);
}

void launch_huffman(hipStream_t s, int first, int count, const std::uint32_t* hists, CodeSet* cs) {
    huffman_kernel<<<count, 64 * kTrials, 0, s>>>(hists, cs, first);
}

bool build_codes_host(const std::uint32_t* hists, std::uint32_t count, std::uint32_t* alpha, std::uint8_t* depth,
                      std::uint16_t* bits, std::uint32_t* hdr_words, std::uint32_t* hdr_bits) {
    if (count == 0 || count > 96) {
        return false;
    }
    std::uint32_t* dh{nullptr};
    CodeSet* dcs{nullptr};
    const std::size_t hbytes{static_cast<std::size_t>(count) * kHistStride * sizeof(std::uint32_t)};
    bool ok{hipMalloc(&dh, hbytes) == hipSuccess && hipMalloc(&dcs, sizeof(CodeSet)) == hipSuccess &&
            hipMemcpy(dh, hists, hbytes, hipMemcpyHostToDevice) == hipSuccess};
    auto* host{new CodeSet};
    if (ok) {
        huffman_kernel<<<count, 64 * kTrials, 0, nullptr>>>(dh, dcs, 0);
        ok = hipGetLastError() == hipSuccess && hipMemcpy(host, dcs, sizeof(CodeSet), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (ok) {
        for (std::uint32_t h{0}; h < count; ++h) {
            alpha[h] = host->alpha[h];
            hdr_bits[h] = host->hdr_bits[h];
            std::memcpy(depth + h * kHistStride, host->depth[h], kHistStride);
            std::memcpy(bits + h * kHistStride, host->code[h], kHistStride * sizeof(std::uint16_t));
            std::memcpy(hdr_words + h * kHdrWords, host->hdr[h], kHdrWords * sizeof(std::uint32_t));
        }
    }
    delete host;
    (void)hipFree(dh);
    (void)hipFree(dcs);
    return ok;
}


std::size_t ac_group_slot_words() {
    return static_cast<std::size_t>(kGroupParts) * kPartSlotWords;
}

std::size_t ac_group_parts() {
    return kGroupParts;
}

void launch_ac_group_emit(hipStream_t s, const Geometry& geo, const std::int8_t* acs,
                          const std::uint16_t* item_tokens, const std::uint32_t* tokens, const CodeSet* cs,
                          std::uint32_t* staging, std::uint64_t* part_bits, hipEvent_t done) {
    hipExtLaunchKernelGGL(ac_group_emit_kernel, dim3(geo.num_ac * kGroupParts), dim3(kPartThreads), 0, s, nullptr,
                          done, 0, geo, acs, item_tokens, tokens, &cs->depth[0][0], &cs->code[0][0], staging,
                          part_bits);
}

std::size_t dc_group_slot_words() {
    // Every DC item is at most 32 bits; plus the two header blobs.
    return static_cast<std::size_t>(dc_chunks_per_group()) * kDcChunk + 2 * kBlobWords + 64;
}

void launch_dc_sizes(hipStream_t s, const Geometry& geo, const Constants* k, const CodeSet* cs,
                     const std::uint32_t* fb_count, DcBlobs* blobs, std::uint32_t* blob_words, const std::int32_t* dc,
                     const std::int8_t* acs, const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                     std::uint64_t* chunk_bits, std::uint32_t* group_bytes, std::uint64_t* group_bits,
                     std::uint32_t* staging) {
    const std::uint32_t cpg{dc_chunks_per_group()};
    dc_blobs_bits_kernel<<<dim3(cpg + 1, geo.num_dc), 256, 0, s>>>(geo, k, cs, fb_count, blobs, blob_words, dc, acs,
                                                                   qf, ytox, ytob, chunk_bits, cpg);
    dc_chunk_scan_kernel<<<geo.num_dc, 1024, 0, s>>>(geo, chunk_bits, cpg, blobs, group_bytes, group_bits, staging,
                                                     dc_group_slot_words());
}

void launch_layout(hipStream_t s, const Geometry& geo, const Constants* k, const CodeSet* cs,
                   const std::uint64_t* ac_part_bits, const std::uint32_t* dc_group_bytes,
                   const std::uint64_t* dc_group_bits, std::uint32_t* ac_group_bytes, std::uint64_t* ac_group_bits,
                   const DcBlobs* blobs, Layout* lay, std::uint64_t* ac_base, std::uint64_t* dc_base,
                   std::uint64_t* total_out, hipEvent_t done) {
    hipExtLaunchKernelGGL(layout_kernel, dim3(1), dim3(256), 0, s, nullptr, done, 0, geo, k, cs, ac_part_bits,
                          dc_group_bytes, dc_group_bits, ac_group_bytes, ac_group_bits, blobs, lay, ac_base, dc_base,
                          total_out);
}

void launch_assemble(hipStream_t s, const Geometry& geo, const Constants* k, const Layout* lay,
                     const std::uint32_t* dc_staging, const std::uint32_t* ac_staging, const std::uint64_t* dc_bits,
                     const std::uint64_t* ac_bits, const std::uint64_t* ac_part_bits, const std::uint64_t* dc_base,
                     const std::uint64_t* ac_base, std::uint8_t* out, std::uint64_t cap_bytes, hipEvent_t done) {
    const Pieces P{dc_staging, dc_group_slot_words(), ac_staging, dc_bits, ac_bits, ac_part_bits, dc_base, ac_base};
    hipExtLaunchKernelGGL(assemble_kernel, dim3(2048), dim3(256), 0, s, nullptr, done, 0, geo, k, lay, P,
                          reinterpret_cast<std::uint32_t*>(out), cap_bytes);
}

void launch_dc_emit(hipStream_t s, const Geometry& geo, const std::int32_t* dc, const std::int8_t* acs,
                    const std::int32_t* qf, const std::int8_t* ytox, const std::int8_t* ytob,
                    const CodeSet* cs, const DcBlobs* blobs, const std::uint32_t* blob_words,
                    const std::uint64_t* chunk_off, std::uint32_t* staging, hipEvent_t done) {
    const std::uint32_t cpg{dc_chunks_per_group()};
    hipExtLaunchKernelGGL(dc_emit_kernel, dim3(cpg, geo.num_dc), dim3(256), 0, s, nullptr, done, 0, geo, dc, acs, qf,
                          ytox, ytob, &cs->depth[kNumClusters][0], &cs->code[kNumClusters][0],
                          &cs->depth[kNumClusters + geo.num_dc][0], &cs->code[kNumClusters + geo.num_dc][0], blobs,
                          blob_words, chunk_off, cpg, staging, dc_group_slot_words());
}

}  // namespace cjx

#ifdef CJX_REGION_CUT
// Profiling builds: the region kernel's cut step (0 = none).
extern "C" int cjx_set_region_cut(int k) {
    return hipMemcpyToSymbol(HIP_SYMBOL(cjx::g_region_cut), &k, sizeof(int)) == hipSuccess ? 0 : 1;
}
#endif
