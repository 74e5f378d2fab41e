// Codestream assembly kernels (included into device.hip).
//
// After the histograms, everything up to the final bytes stays on the device:
// one workgroup per histogram builds its prefix code and header; small
// single-workgroup kernels stitch the variable parts (alphabet sizes, code
// headers, first-block counts, TOC) onto constant bit strings the host built
// once per configuration, lay out the sections, and write the codestream head.

namespace cjx {

constexpr int kHdrWords = 64;      // 2048 bits: bound on one code header
constexpr int kBlobWords = 128;    // per DC-group header blob
constexpr int kAcGlobalWords = 8192;
constexpr int kHeadWords = 1024;

struct Constants {
    // Bit strings precomputed on the host (see encoder.cc).
    std::uint32_t head_bits, dc_global_bits, ac_prefix_bits, dc_pre_bits, dc_mid_bits;
    std::uint32_t head_words[kHeadWords / 4];
    std::uint32_t dc_global_words[4];
    std::uint32_t ac_prefix_words[kAcGlobalWords / 2];
    std::uint32_t dc_pre_words[64];
    std::uint32_t dc_mid_words[64];
};

struct CodeSet {
    // Per histogram (32 AC clusters, then DC of each group, then AcMetadata of
    // each group).
    std::uint8_t depth[96][kHistStride];
    std::uint16_t code[96][kHistStride];
    std::uint32_t hdr[96][kHdrWords];
    std::uint32_t hdr_bits[96];
    std::uint32_t alpha[96];
};

struct Layout {
    std::uint64_t total_bytes;
    std::uint64_t codestream_bytes;
    std::uint32_t framing_bytes;
    std::uint32_t head_bytes;
    std::uint64_t dc_global_bit;   // absolute bit positions in the file
    std::uint64_t ac_global_bit;
    std::uint32_t ac_global_bits;
    std::uint32_t overflow;
    std::uint32_t ac_global_words[kAcGlobalWords];
    std::uint32_t head_words[kHeadWords];
    std::uint32_t head_total_bits;
};

// One wave per histogram: alphabet size, depths, codes and header.
__global__ __launch_bounds__(64) void huffman_kernel(const std::uint32_t* __restrict__ hists, CodeSet* __restrict__ cs) {
    __shared__ WaveCodeScratch scratch;
    __shared__ std::uint32_t hist[kHistStride];
    __shared__ std::uint8_t depth[kHistStride];
    __shared__ std::uint16_t code[kHistStride];
    __shared__ std::uint32_t hdr[kHdrWords];
    const int h{static_cast<int>(blockIdx.x)};
    const int t{static_cast<int>(threadIdx.x)};
    std::uint32_t top{0};
    for (int i{t}; i < kHistStride; i += 64) {
        const std::uint32_t v{hists[h * kHistStride + i]};
        hist[i] = v;
        depth[i] = 0;
        code[i] = 0;
        if (v != 0) {
            top = static_cast<std::uint32_t>(i) + 1u;
        }
    }
    for (int i{t}; i < kHdrWords; i += 64) {
        hdr[i] = 0;
    }
    const int alpha{static_cast<int>(wave_max_u32(top > 1u ? top : 1u))};
    __syncthreads();
    std::uint32_t nbits{0};
    if (alpha > 1) {
        nbits = alpha <= 64 ? wave_build_code<1>(hist, alpha, depth, code, hdr, kHdrWords, scratch)
                            : wave_build_code<4>(hist, alpha, depth, code, hdr, kHdrWords, scratch);
    }
    __syncthreads();
    for (int i{t}; i < kHistStride; i += 64) {
        cs->depth[h][i] = depth[i];
        cs->code[h][i] = code[i];
    }
    for (int i{t}; i < kHdrWords; i += 64) {
        cs->hdr[h][i] = hdr[i];
    }
    if (t == 0) {
        cs->hdr_bits[h] = nbits;
        cs->alpha[h] = static_cast<std::uint32_t>(alpha);
    }
}

// Per DC group: blob_pre = constant prefix + data alphabet + code header;
// blob_mid = (count - 1) + constant prefix + alphabet + header.
__global__ __launch_bounds__(64) void dc_blobs_kernel(Geometry geo, const Constants* __restrict__ k,
                                                      const CodeSet* __restrict__ cs,
                                                      const std::uint32_t* __restrict__ fb_count,
                                                      DcBlobs* __restrict__ blobs,
                                                      std::uint32_t* __restrict__ blob_words) {
    const int g{static_cast<int>(blockIdx.x)};
    const int t{static_cast<int>(threadIdx.x)};
    std::uint32_t* pre{blob_words + static_cast<std::size_t>(2 * g) * kBlobWords};
    std::uint32_t* mid{pre + kBlobWords};
    for (int i{t}; i < kBlobWords; i += 64) {
        pre[i] = 0;
        mid[i] = 0;
    }
    __syncthreads();
    if (t == 0) {
        const DcGroupGeo d{dc_group_geo(geo, g)};
        const int hd{kNumClusters + g}, hm{kNumClusters + static_cast<int>(geo.num_dc) + g};
        BitSink w{pre, 0, kBlobWords * 32};
        copy_bits(w, k->dc_pre_words, k->dc_pre_bits);
        write_var_len_u16(w, cs->alpha[hd] - 1);
        copy_bits(w, cs->hdr[hd], cs->hdr_bits[hd]);
        blobs->pre_bits[g] = w.ok() ? w.pos : 0u;
        if (!w.ok()) {
            atomicOr(&blobs->overflow, 1u);
        }
        blobs->pre_off[g] = static_cast<std::uint32_t>(2 * g) * kBlobWords;
        BitSink m{mid, 0, kBlobWords * 32};
        m.write(ceil_log2(static_cast<std::uint32_t>(d.n)), fb_count[g] - 1);
        copy_bits(m, k->dc_mid_words, k->dc_mid_bits);
        write_var_len_u16(m, cs->alpha[hm] - 1);
        copy_bits(m, cs->hdr[hm], cs->hdr_bits[hm]);
        blobs->mid_bits[g] = m.ok() ? m.pos : 0u;
        if (!m.ok()) {
            atomicOr(&blobs->overflow, 1u);
        }
        blobs->mid_off[g] = static_cast<std::uint32_t>(2 * g + 1) * kBlobWords;
    }
}

// AcGlobal, section sizes, TOC, head, framing and every section's position.
// `single` frames (one AC group) concatenate DcGlobal, the DC group, AcGlobal
// and the AC tokens into one section without padding in between.
__global__ __launch_bounds__(64) void layout_kernel(Geometry geo, const Constants* __restrict__ k,
                                                    const CodeSet* __restrict__ cs,
                                                    const std::uint32_t* __restrict__ ac_group_bytes,
                                                    const std::uint32_t* __restrict__ dc_group_bytes,
                                                    const std::uint64_t* __restrict__ ac_single_bits,
                                                    const std::uint64_t* __restrict__ dc_single_bits,
                                                    const DcBlobs* __restrict__ blobs,
                                                    Layout* __restrict__ lay, std::uint64_t* __restrict__ ac_base,
                                                    std::uint64_t* __restrict__ dc_base) {
    const int t{static_cast<int>(threadIdx.x)};
    for (int i{t}; i < kAcGlobalWords; i += 64) {
        lay->ac_global_words[i] = 0;
    }
    for (int i{t}; i < kHeadWords; i += 64) {
        lay->head_words[i] = 0;
    }
    __syncthreads();
    if (t != 0) {
        return;
    }
    BitSink a{lay->ac_global_words, 0, kAcGlobalWords * 32};
    copy_bits(a, k->ac_prefix_words, k->ac_prefix_bits);
    for (int c{0}; c < kNumClusters; ++c) {
        write_var_len_u16(a, cs->alpha[c] - 1);
    }
    for (int c{0}; c < kNumClusters; ++c) {
        copy_bits(a, cs->hdr[c], cs->hdr_bits[c]);
    }
    lay->ac_global_bits = a.pos;
    const bool single{geo.num_ac == 1};
    std::uint32_t overflow{a.ok() ? 0u : 1u};
    overflow |= blobs->overflow;

    // Section sizes in codestream order: DcGlobal, DcGroups, AcGlobal, AcGroups.
    const Distr toc[4]{{false, 0, 10}, {false, 1024, 14}, {false, 17408, 22}, {false, 4211712, 30}};
    BitSink hw{lay->head_words, 0, kHeadWords * 32};
    copy_bits(hw, k->head_words, k->head_bits);
    hw.write(1, 0);                 // no permutation
    hw.pos = (hw.pos + 7u) & ~7u;   // byte boundary
    std::uint64_t body{0};
    if (single) {
        const std::uint64_t bits{k->dc_global_bits + dc_single_bits[0] + a.pos + ac_single_bits[0]};
        const std::uint32_t bytes{static_cast<std::uint32_t>((bits + 7) / 8)};
        write_u32(hw, toc, bytes);
        body = bytes;
    } else {
        const std::uint32_t dcg{(k->dc_global_bits + 7) / 8};
        write_u32(hw, toc, dcg);
        body += dcg;
        for (std::uint32_t g{0}; g < geo.num_dc; ++g) {
            write_u32(hw, toc, dc_group_bytes[g]);
            body += dc_group_bytes[g];
        }
        const std::uint32_t acg{(a.pos + 7) / 8};
        write_u32(hw, toc, acg);
        body += acg;
        for (std::uint32_t g{0}; g < geo.num_ac; ++g) {
            write_u32(hw, toc, ac_group_bytes[g]);
            body += ac_group_bytes[g];
        }
    }
    hw.pos = (hw.pos + 7u) & ~7u;
    overflow |= hw.ok() ? 0u : 1u;
    const std::uint32_t head_bytes{hw.pos / 8};
    const std::uint64_t cs_bytes{head_bytes + body};
    const std::uint32_t framing{cs_bytes >= 0x100000000ull - 8 ? 48u : 40u};
    lay->head_total_bits = hw.pos;
    lay->head_bytes = head_bytes;
    lay->framing_bytes = framing;
    lay->codestream_bytes = cs_bytes;
    lay->total_bytes = overflow ? ~0ull : framing + cs_bytes;
    lay->overflow = overflow;
    std::uint64_t at{(static_cast<std::uint64_t>(framing) + head_bytes) * 8};
    lay->dc_global_bit = at;
    if (single) {
        at += k->dc_global_bits;
        dc_base[0] = at;
        at += dc_single_bits[0];
        lay->ac_global_bit = at;
        at += a.pos;
        ac_base[0] = at;
        return;
    }
    at += ((k->dc_global_bits + 7) / 8) * 8;
    for (std::uint32_t g{0}; g < geo.num_dc; ++g) {
        dc_base[g] = at;
        at += static_cast<std::uint64_t>(dc_group_bytes[g]) * 8;
    }
    lay->ac_global_bit = at;
    at += ((a.pos + 7) / 8) * 8;
    for (std::uint32_t g{0}; g < geo.num_ac; ++g) {
        ac_base[g] = at;
        at += static_cast<std::uint64_t>(ac_group_bytes[g]) * 8;
    }
}

__global__ void zero_kernel(const Layout* __restrict__ lay, std::uint32_t* __restrict__ out_words,
                            std::uint64_t cap_words) {
    std::uint64_t words{(lay->total_bytes + 3) / 4};
    words = words < cap_words ? words : cap_words;
    for (std::uint64_t i{blockIdx.x * 256ull + threadIdx.x}; i < words; i += gridDim.x * 256ull) {
        out_words[i] = 0;
    }
}

// Container framing (container.cc:35-54), head, DcGlobal, AcGlobal.
__global__ __launch_bounds__(64) void headers_kernel(const Constants* __restrict__ k,
                                                     const Layout* __restrict__ lay,
                                                     std::uint32_t* __restrict__ out_words,
                                                     std::uint64_t cap_words) {
    const int t{static_cast<int>(threadIdx.x)};
    if (t == 0) {
        constexpr std::uint8_t hdr[32]{0,   0,   0, 0x0c, 'J', 'X', 'L', ' ', 0x0d, 0x0a, 0x87,
                                       0x0a, 0,  0, 0,    0x14, 'f', 't', 'y', 'p', 'j',  'x',
                                       'l', ' ', 0, 0,    0,    0,   'j', 'x', 'l', ' '};
        std::uint8_t f[48];
        for (int i{0}; i < 32; ++i) {
            f[i] = hdr[i];
        }
        const bool large{lay->framing_bytes == 48};
        const std::uint64_t box{lay->codestream_bytes + (large ? 16 : 8)};
        const std::uint32_t field{large ? 1u : static_cast<std::uint32_t>(box)};
        for (int i{0}; i < 4; ++i) {
            f[32 + i] = static_cast<std::uint8_t>(field >> (8 * (3 - i)));
        }
        f[36] = 'j';
        f[37] = 'x';
        f[38] = 'l';
        f[39] = 'c';
        for (int i{0}; i < 8; ++i) {
            f[40 + i] = static_cast<std::uint8_t>(box >> (8 * (7 - i)));
        }
        for (std::uint32_t i{0}; i < lay->framing_bytes; ++i) {
            or_bits(out_words, cap_words, static_cast<std::uint64_t>(i) * 8, f[i], 8);
        }
    }
    const std::uint64_t head_bit{static_cast<std::uint64_t>(lay->framing_bytes) * 8};
    for (std::uint32_t i{static_cast<std::uint32_t>(t) * 32}; i < lay->head_total_bits; i += 64 * 32) {
        const std::uint32_t n{lay->head_total_bits - i < 32 ? lay->head_total_bits - i : 32};
        or_bits(out_words, cap_words, head_bit + i, lay->head_words[i >> 5], n);
    }
    for (std::uint32_t i{static_cast<std::uint32_t>(t) * 32}; i < k->dc_global_bits; i += 64 * 32) {
        const std::uint32_t n{k->dc_global_bits - i < 32 ? k->dc_global_bits - i : 32};
        or_bits(out_words, cap_words, lay->dc_global_bit + i, k->dc_global_words[i >> 5], n);
    }
    for (std::uint32_t i{static_cast<std::uint32_t>(t) * 32}; i < lay->ac_global_bits; i += 64 * 32) {
        const std::uint32_t n{lay->ac_global_bits - i < 32 ? lay->ac_global_bits - i : 32};
        or_bits(out_words, cap_words, lay->ac_global_bit + i, lay->ac_global_words[i >> 5], n);
    }
}

}  // namespace cjx
