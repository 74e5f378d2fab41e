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
constexpr int kMaxAcGroups = 32 * 64;  // at most 32 DC groups of 64 AC groups
// Codestream head: the constant headers (under 32 words) and the TOC, whose
// 2 + 32 + 2048 entries take at most 32 bits each at the largest accepted
// frame (2 + 30 for sections of 4,211,712 bytes or more).
constexpr int kHeadConstWords = 256;
constexpr int kHeadWords = ((kHeadConstWords * 32 + (2 + 32 + kMaxAcGroups) * 32 + 8 + 8191) / 8192) * 256;
static_assert(kHeadWords % 256 == 0 && kHeadWords * 32 >= kHeadConstWords * 32 + (2 + 32 + kMaxAcGroups) * 32 + 8,
              "the head holds the worst-case TOC");

struct Constants {
    // Bit strings precomputed on the host (see encoder.cc).
    std::uint32_t head_bits, dc_global_bits, ac_prefix_bits, dc_pre_bits, dc_mid_bits;
    std::uint32_t head_words[kHeadConstWords];
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
    std::uint32_t framing_words[12];  // container framing, 40 or 48 bytes
};

// One 8-wave workgroup per histogram (histograms first, first + 1, ...):
// create_huffman_tree's count floors 1, 2, 4, ... are independent trials, so
// in round r wave w tries floor 2^(8r + w), and the smallest floor whose tree
// fits 15 bits wins (the serial builder stops at that same floor); floors up
// to 2^7 take one round.  Floors above 2^15 continue on wave 0.  Wave 0 then
// builds the codes and the header.
constexpr int kTrials = 8;
constexpr int kTrialRounds = 2;

struct HuffmanShared {
    WaveHuffScratch trial[kTrials];
    std::uint8_t trial_depth[kTrials][kHistStride];
    std::uint32_t trial_max[kTrials];
    WaveCodeScratch code_scratch;
    std::uint32_t hist[kHistStride];
    std::uint8_t depth[kHistStride];
    std::uint16_t code[kHistStride];
    std::uint32_t hdr[kHdrWords];
};

template <int R>
__device__ void huffman_depths(HuffmanShared& S, int alpha) {
    const int w{static_cast<int>(threadIdx.x) >> 6};
    const int lane{static_cast<int>(threadIdx.x) & 63};
    const Leaves<R> L{wave_leaves<R>(S.hist, alpha)};
    if (L.n < 2) {
        return;  // one symbol or none: the simple header, all depths stay 0
    }
    int win{-1};
    for (int round{0}; round < kTrialRounds && win < 0; ++round) {
        if (round > 0) {
            for (int i{lane}; i < kHistStride; i += 64) {
                S.trial_depth[w][i] = 0;
            }
            wave_sync();
        }
        const std::uint32_t maxd{tree_trial<R>(L, 1u << (round * kTrials + w), S.trial_depth[w], S.trial[w])};
        if (lane == 0) {
            S.trial_max[w] = maxd;
        }
        __syncthreads();
        for (int i{0}; i < kTrials; ++i) {
            if (S.trial_max[i] <= 15u) {
                win = i;
                break;
            }
        }
        __syncthreads();
    }
    if (win < 0 && w == 0) {
        for (std::uint32_t count_limit{1u << (kTrials * kTrialRounds)};; count_limit *= 2) {
            for (int i{static_cast<int>(threadIdx.x)}; i < kHistStride; i += 64) {
                S.trial_depth[0][i] = 0;
            }
            wave_sync();
            if (tree_trial<R>(L, count_limit, S.trial_depth[0], S.trial[0]) <= 15u) {
                break;
            }
        }
    }
    __syncthreads();
    const int src{win < 0 ? 0 : win};
    for (int i{static_cast<int>(threadIdx.x)}; i < kHistStride; i += 64 * kTrials) {
        S.depth[i] = S.trial_depth[src][i];
    }
}

__device__ __attribute__((always_inline)) inline void huffman_kernel_body(const Wg& wg, const std::uint32_t* __restrict__ hists,
                                                               CodeSet* __restrict__ cs, int first) {
    __shared__ HuffmanShared S;
    const int h{first + static_cast<int>(wg.x)};
    const int t{static_cast<int>(threadIdx.x)};
    const int w{t >> 6};
    for (int i{t}; i < kHistStride; i += 64 * kTrials) {
        S.hist[i] = hists[h * kHistStride + i];
        S.depth[i] = 0;
        S.code[i] = 0;
    }
    for (int i{t}; i < kTrials * kHistStride; i += 64 * kTrials) {
        S.trial_depth[i / kHistStride][i % kHistStride] = 0;
    }
    if (t < kHdrWords) {
        S.hdr[t] = 0;
    }
    __syncthreads();
    std::uint32_t top{0};
    for (int i{t & 63}; i < kHistStride; i += 64) {
        if (S.hist[i] != 0) {
            top = static_cast<std::uint32_t>(i) + 1u;
        }
    }
    const int alpha{static_cast<int>(wave_max_u32(top > 1u ? top : 1u))};
    if (alpha > 1) {
        if (alpha <= 64) {
            huffman_depths<1>(S, alpha);
        } else {
            huffman_depths<4>(S, alpha);
        }
    }
    __syncthreads();
    std::uint32_t nbits{0};
    if (w == 0 && alpha > 1) {
        nbits = alpha <= 64 ? wave_finish_code<1>(S.hist, alpha, S.depth, S.code, S.hdr, kHdrWords, S.code_scratch)
                            : wave_finish_code<4>(S.hist, alpha, S.depth, S.code, S.hdr, kHdrWords, S.code_scratch);
    }
    __syncthreads();
    for (int i{t}; i < kHistStride; i += 64 * kTrials) {
        cs->depth[h][i] = S.depth[i];
        cs->code[h][i] = S.code[i];
    }
    if (t < kHdrWords) {
        cs->hdr[h][t] = S.hdr[t];
    }
    if (t == 0) {
        cs->hdr_bits[h] = nbits;
        cs->alpha[h] = static_cast<std::uint32_t>(alpha);
    }
}

__global__ __launch_bounds__(64 * kTrials) void huffman_kernel(const std::uint32_t* __restrict__ hists,
                                                               CodeSet* __restrict__ cs, int first) {
    huffman_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, hists, cs, first);
}

// Per DC group: blob_pre = constant prefix + data alphabet + code header;
// blob_mid = (count - 1) + constant prefix + alphabet + header.
// One DC group's header blobs (group g, threads of the workgroup take the
// clearing; thread 0 writes the bits).
__device__ void dc_blobs_body(const Geometry& geo, const Constants* __restrict__ k, const CodeSet* __restrict__ cs,
                              const std::uint32_t* __restrict__ fb_count, DcBlobs* __restrict__ blobs,
                              std::uint32_t* __restrict__ blob_words, int g) {
    const int t{static_cast<int>(threadIdx.x)};
    std::uint32_t* pre{blob_words + static_cast<std::size_t>(2 * g) * kBlobWords};
    std::uint32_t* mid{pre + kBlobWords};
    for (int i{t}; i < kBlobWords; i += static_cast<int>(blockDim.x)) {
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

// The DC chain's sizing pass: workgroup columns 0 .. chunks_per_group - 1 of
// group row g size the group's item chunks, column chunks_per_group builds
// its two header blobs (independent work, one launch).
__device__ __attribute__((always_inline)) inline void dc_blobs_bits_kernel_body(const Wg& wg, 
    Geometry geo, const Constants* __restrict__ k, const CodeSet* __restrict__ cs,
    const std::uint32_t* __restrict__ fb_count, DcBlobs* __restrict__ blobs, std::uint32_t* __restrict__ blob_words,
    const std::int32_t* __restrict__ dc, const std::int8_t* __restrict__ acs, const std::int32_t* __restrict__ qf,
    const std::int8_t* __restrict__ ytox, const std::int8_t* __restrict__ ytob,
    std::uint64_t* __restrict__ chunk_bits, std::uint32_t chunks_per_group) {
    if (wg.x == chunks_per_group) {
        dc_blobs_body(geo, k, cs, fb_count, blobs, blob_words, static_cast<int>(wg.y));
    } else {
        dc_chunk_bits_body(wg, geo, dc, acs, qf, ytox, ytob, &cs->depth[kNumClusters][0],
                           &cs->depth[kNumClusters + geo.num_dc][0], chunk_bits, chunks_per_group);
    }
}

__global__ __launch_bounds__(256) void dc_blobs_bits_kernel(
    Geometry geo, const Constants* __restrict__ k, const CodeSet* __restrict__ cs,
    const std::uint32_t* __restrict__ fb_count, DcBlobs* __restrict__ blobs, std::uint32_t* __restrict__ blob_words,
    const std::int32_t* __restrict__ dc, const std::int8_t* __restrict__ acs, const std::int32_t* __restrict__ qf,
    const std::int8_t* __restrict__ ytox, const std::int8_t* __restrict__ ytob,
    std::uint64_t* __restrict__ chunk_bits, std::uint32_t chunks_per_group) {
    dc_blobs_bits_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, k, cs, fb_count, blobs, blob_words, dc, acs, qf, ytox, ytob, chunk_bits, chunks_per_group);
}

// U32 field coder choice (field_coder.h:46-107) as one code word: selector
// in the low 2 bits, then the offset bits.
__device__ __forceinline__ void u32_code(const Distr* d, std::uint32_t value, std::uint32_t& code,
                                         std::uint32_t& len) {
    std::uint32_t best{0}, best_bits{64};
    for (std::uint32_t s{0}; s < 4; ++s) {
        if (d[s].direct) {
            if (d[s].v == value) {
                best = s;
                best_bits = 2;
                break;
            }
            continue;
        }
        if (value < d[s].v ||
            static_cast<std::uint64_t>(value) >= static_cast<std::uint64_t>(d[s].v) + (1ull << d[s].bits)) {
            continue;
        }
        if (2 + d[s].bits < best_bits) {
            best = s;
            best_bits = 2 + d[s].bits;
        }
    }
    if (d[best].direct) {
        code = best;
        len = 2;
        return;
    }
    const std::uint32_t nb{d[best].bits};
    const std::uint32_t off{nb < 32 ? (value - d[best].v) & ((1u << nb) - 1u) : value - d[best].v};
    code = best | (off << 2);
    len = 2 + nb;
}

// AcGlobal, section sizes, TOC, head, framing and every section's position,
// by one 256-thread workgroup.  AcGlobal (constant prefix, 32 alphabet sizes,
// 32 code headers) and the head (constant bits, TOC) are assembled in LDS:
// every piece's bit offset comes from a block scan of the piece lengths, and
// the section bases from scans of the section sizes.  `single` frames (one AC
// group) concatenate DcGlobal, the DC group, AcGlobal and the AC tokens into
// one section without padding in between.
__device__ __attribute__((always_inline)) inline void layout_kernel_body(const Wg& wg, Geometry geo, const Constants* __restrict__ k,
                                                     const CodeSet* __restrict__ cs,
                                                     const std::uint64_t* __restrict__ ac_part_bits,
                                                     const std::uint32_t* __restrict__ dc_group_bytes,
                                                     const std::uint64_t* __restrict__ dc_single_bits,
                                                     std::uint32_t* __restrict__ ac_group_bytes,
                                                     std::uint64_t* __restrict__ ac_group_bits,
                                                     const DcBlobs* __restrict__ blobs,
                                                     Layout* __restrict__ lay, std::uint64_t* __restrict__ ac_base,
                                                     std::uint64_t* __restrict__ dc_base,
                                                     std::uint64_t* __restrict__ total_out) {
    __shared__ std::uint32_t agw[kAcGlobalWords];
    __shared__ std::uint32_t hw[kHeadWords];
    __shared__ std::uint64_t agbits[kMaxAcGroups];
    __shared__ std::uint64_t sh[17];
    const int t{static_cast<int>(threadIdx.x)};
    // The reads that depend on nothing computed here, issued together before
    // the first wait: each thread's first word of the AcGlobal prefix and of
    // the head (the head's constant words are one per thread), its code
    // header's alphabet size or bit count, its DC group sizes for the TOC
    // (entry t: group t - 1) and for the section positions (group t).
    static_assert(kHeadConstWords == 256, "one head word per thread");
    const std::uint32_t overflow0{blobs->overflow};
    const std::uint32_t pre_bits{k->ac_prefix_bits};
    const std::uint32_t hb{k->head_bits};
    const std::uint32_t dc_global_bits{k->dc_global_bits};
    const std::uint32_t prefix_w0{k->ac_prefix_words[t]};
    const std::uint32_t head_w0{k->head_words[t]};
    const std::uint32_t alpha_t{t < kNumClusters ? cs->alpha[t] : 0u};
    const std::uint32_t hdr_bits_t{t >= kNumClusters && t < 2 * kNumClusters ? cs->hdr_bits[t - kNumClusters] : 0u};
    const std::uint32_t dc_bytes_t{static_cast<std::uint32_t>(t) < geo.num_dc ? dc_group_bytes[t] : 0u};
    const std::uint32_t dc_bytes_prev{t >= 1 && static_cast<std::uint32_t>(t - 1) < geo.num_dc ? dc_group_bytes[t - 1]
                                                                                                  : 0u};
    // AC group sizes: the sums of their parts' bits.
    for (int g{t}; g < static_cast<int>(geo.num_ac); g += 256) {
        std::uint64_t sum{0};
        for (int p{0}; p < kGroupParts; ++p) {
            sum += ac_part_bits[g * kGroupParts + p];
        }
        agbits[g] = sum;
        ac_group_bits[g] = sum;
        ac_group_bytes[g] = static_cast<std::uint32_t>((sum + 7) / 8);
    }
    for (int i{t}; i < kAcGlobalWords; i += 256) {
        agw[i] = 0;
    }
    for (int i{t}; i < kHeadWords; i += 256) {
        hw[i] = 0;
    }
    __syncthreads();
    std::uint32_t overflow{overflow0};

    // AcGlobal.
    for (std::uint32_t i{static_cast<std::uint32_t>(t)}; i < (pre_bits + 31) / 32; i += 256) {
        const std::uint32_t rem{pre_bits - 32 * i};
        const std::uint32_t w{i == static_cast<std::uint32_t>(t) ? prefix_w0 : k->ac_prefix_words[i]};
        agw[i] = rem >= 32 ? w : (w & ((1u << rem) - 1u));
    }
    std::uint32_t len{0}, val{0};
    if (t < kNumClusters) {
        const std::uint32_t n{alpha_t - 1};
        if (n == 0) {
            len = 1;
        } else {
            const std::uint32_t nb{31u - static_cast<std::uint32_t>(__builtin_clz(n))};
            len = 5 + nb;
            val = 1u | (nb << 1) | ((n - (1u << nb)) << 5);
        }
    } else if (t < 2 * kNumClusters) {
        len = hdr_bits_t;
    }
    std::uint64_t pieces{0};
    const std::uint64_t off{pre_bits + block_excl_scan(len, sh, pieces)};
    const std::uint64_t abits{pre_bits + pieces};
    if (abits > static_cast<std::uint64_t>(kAcGlobalWords) * 32) {
        overflow = 1;
    } else if (t < kNumClusters) {
        lds_or_bits(agw, kAcGlobalWords, static_cast<std::uint32_t>(off), val, len);
    } else if (t < 2 * kNumClusters) {
        const std::uint32_t* src{cs->hdr[t - kNumClusters]};
        for (std::uint32_t i{0}; i < len; i += 32) {
            lds_or_bits(agw, kAcGlobalWords, static_cast<std::uint32_t>(off) + i, src[i >> 5],
                        len - i < 32 ? len - i : 32u);
        }
    }

    // Head: constant bits, no permutation (one 0 bit), byte-aligned TOC.
    for (std::uint32_t i{static_cast<std::uint32_t>(t)}; i < (hb + 31) / 32; i += 256) {
        const std::uint32_t rem{hb - 32 * i};
        const std::uint32_t w{head_w0};  // i == t: the head has at most 256 words
        hw[i] = rem >= 32 ? w : (w & ((1u << rem) - 1u));
    }
    const Distr toc[4]{{false, 0, 10}, {false, 1024, 14}, {false, 17408, 22}, {false, 4211712, 30}};
    const bool single{geo.num_ac == 1};
    const std::uint32_t dcg{(dc_global_bits + 7) / 8};
    const std::uint32_t acg{static_cast<std::uint32_t>((abits + 7) / 8)};
    const std::uint32_t entries{single ? 1u : 2u + geo.num_dc + geo.num_ac};
    std::uint64_t at{(hb + 1u + 7u) & ~7u};
    std::uint64_t body{0};
    for (std::uint32_t e0{0}; e0 < entries; e0 += 256) {
        const std::uint32_t e{e0 + static_cast<std::uint32_t>(t)};
        std::uint32_t size{0}, code{0}, clen{0};
        if (e < entries) {
            if (single) {
                const std::uint64_t bits{dc_global_bits + dc_single_bits[0] + abits + agbits[0]};
                size = static_cast<std::uint32_t>((bits + 7) / 8);
            } else if (e == 0) {
                size = dcg;
            } else if (e <= geo.num_dc) {
                size = dc_bytes_prev;  // e = t: the TOC's DC entries all lie in its first 256
            } else if (e == geo.num_dc + 1) {
                size = acg;
            } else {
                size = static_cast<std::uint32_t>((agbits[e - geo.num_dc - 2] + 7) / 8);
            }
            u32_code(toc, size, code, clen);
        }
        std::uint64_t chunk_bits{0}, chunk_bytes{0};
        const std::uint64_t p{at + block_excl_scan(clen, sh, chunk_bits)};
        (void)block_excl_scan(size, sh, chunk_bytes);
        if (clen != 0 && p + clen <= static_cast<std::uint64_t>(kHeadWords) * 32) {
            lds_or_bits(hw, kHeadWords, static_cast<std::uint32_t>(p), code, clen);
        }
        at += chunk_bits;
        body += chunk_bytes;
    }
    at = (at + 7) & ~7ull;
    if (at > static_cast<std::uint64_t>(kHeadWords) * 32) {
        overflow = 1;
    }
    const std::uint32_t head_bytes{static_cast<std::uint32_t>(at / 8)};
    const std::uint64_t cs_bytes{head_bytes + body};
    const std::uint32_t framing{cs_bytes >= 0x100000000ull - 8 ? 48u : 40u};

    // Section positions (absolute bits in the file).
    const std::uint64_t dcg_bit{(static_cast<std::uint64_t>(framing) + head_bytes) * 8};
    std::uint64_t ac_global_bit{0};
    if (single) {
        if (t == 0) {
            dc_base[0] = dcg_bit + dc_global_bits;
            ac_global_bit = dcg_bit + dc_global_bits + dc_single_bits[0];
            ac_base[0] = ac_global_bit + abits;
        }
    } else {
        const std::uint64_t first{dcg_bit + static_cast<std::uint64_t>(dcg) * 8};
        std::uint64_t dc_total{0};
        const std::uint32_t g{static_cast<std::uint32_t>(t)};
        const std::uint64_t dsz{g < geo.num_dc ? static_cast<std::uint64_t>(dc_bytes_t) * 8 : 0ull};
        const std::uint64_t dpre{block_excl_scan(dsz, sh, dc_total)};
        if (g < geo.num_dc) {
            dc_base[g] = first + dpre;
        }
        ac_global_bit = first + dc_total;
        std::uint64_t run{ac_global_bit + static_cast<std::uint64_t>(acg) * 8};
        for (std::uint32_t g0{0}; g0 < geo.num_ac; g0 += 256) {
            const std::uint32_t ga{g0 + static_cast<std::uint32_t>(t)};
            const std::uint64_t asz{ga < geo.num_ac ? (agbits[ga] + 7) / 8 * 8 : 0ull};
            std::uint64_t chunk{0};
            const std::uint64_t apre{block_excl_scan(asz, sh, chunk)};
            if (ga < geo.num_ac) {
                ac_base[ga] = run + apre;
            }
            run += chunk;
        }
    }
    // Container framing (container.cc:35-54): the 32-byte signature and ftyp
    // box, then the jxlc box header (8 bytes, or 16 with a 64-bit size).
    // Lane i forms bytes 4i..4i+3.
    if (t < static_cast<int>(framing / 4)) {
        constexpr std::uint8_t sig[32]{0,   0,   0, 0x0c, 'J', 'X', 'L', ' ', 0x0d, 0x0a, 0x87,
                                       0x0a, 0,  0, 0,    0x14, 'f', 't', 'y', 'p', 'j',  'x',
                                       'l', ' ', 0, 0,    0,    0,   'j', 'x', 'l', ' '};
        const bool large{framing == 48};
        const std::uint64_t box{cs_bytes + (large ? 16 : 8)};
        const std::uint32_t field{large ? 1u : static_cast<std::uint32_t>(box)};
        std::uint32_t word{0};
        for (int j{0}; j < 4; ++j) {
            const int i{4 * t + j};
            std::uint8_t b;
            if (i < 32) {
                b = sig[i];
            } else if (i < 36) {
                b = static_cast<std::uint8_t>(field >> (8 * (35 - i)));
            } else if (i < 40) {
                b = static_cast<std::uint8_t>("jxlc"[i - 36]);
            } else {
                b = static_cast<std::uint8_t>(box >> (8 * (47 - i)));
            }
            word |= static_cast<std::uint32_t>(b) << (8 * j);
        }
        lay->framing_words[t] = word;
    }
    __syncthreads();
    for (int i{t}; i < kAcGlobalWords; i += 256) {
        lay->ac_global_words[i] = agw[i];
    }
    for (int i{t}; i < kHeadWords; i += 256) {
        lay->head_words[i] = hw[i];
    }
    if (t == 0) {
        lay->ac_global_bits = static_cast<std::uint32_t>(abits);
        lay->head_total_bits = static_cast<std::uint32_t>(at);
        lay->head_bytes = head_bytes;
        lay->framing_bytes = framing;
        lay->codestream_bytes = cs_bytes;
        lay->total_bytes = overflow ? ~0ull : framing + cs_bytes;
        *total_out = lay->total_bytes;
        lay->overflow = overflow;
        lay->dc_global_bit = dcg_bit;
        lay->ac_global_bit = ac_global_bit;
    }
}

__global__ __launch_bounds__(256) void layout_kernel(Geometry geo, const Constants* __restrict__ k,
                                                     const CodeSet* __restrict__ cs,
                                                     const std::uint64_t* __restrict__ ac_part_bits,
                                                     const std::uint32_t* __restrict__ dc_group_bytes,
                                                     const std::uint64_t* __restrict__ dc_single_bits,
                                                     std::uint32_t* __restrict__ ac_group_bytes,
                                                     std::uint64_t* __restrict__ ac_group_bits,
                                                     const DcBlobs* __restrict__ blobs,
                                                     Layout* __restrict__ lay, std::uint64_t* __restrict__ ac_base,
                                                     std::uint64_t* __restrict__ dc_base,
                                                     std::uint64_t* __restrict__ total_out) {
    layout_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, k, cs, ac_part_bits, dc_group_bytes, dc_single_bits, ac_group_bytes, ac_group_bits, blobs, lay, ac_base, dc_base, total_out);
}

// The file is a sequence of bit strings ("pieces") laid end to end, each at
// the position the layout gave it: framing, head (headers and TOC), DcGlobal,
// the DC groups, AcGlobal, the AC groups.  Gaps between pieces (section
// padding) are zero bits.  Each output word is built once from the pieces
// that overlap it and stored whole, so the output needs no clearing and no
// atomics.
struct Pieces {
    const std::uint32_t* dc_staging;
    std::uint64_t dc_slot_words;
    const std::uint32_t* ac_staging;  // kGroupParts slots of kPartSlotWords per group
    const std::uint64_t* dc_bits;
    const std::uint64_t* ac_bits;
    const std::uint64_t* ac_part_bits;
    const std::uint64_t* dc_base;
    const std::uint64_t* ac_base;
};

constexpr int kMaxPieces = 4 + 32 + kMaxAcGroups;

// A piece's bits are one run at `src`, or, for an AC group (`parts` set),
// the concatenation of its parts' runs.
struct Piece {
    std::uint64_t dst;
    std::uint64_t nbits;
    const std::uint32_t* src;
    const std::uint64_t* parts;
};

// `len` (1..32) bits of a run from bit `o`.
__device__ __forceinline__ std::uint32_t run_bits(const std::uint32_t* src, std::uint64_t o, std::uint32_t len) {
    const std::uint32_t sh{static_cast<std::uint32_t>(o & 31u)};
    std::uint32_t x{src[o >> 5] >> sh};
    if (sh != 0 && len > 32 - sh) {
        x |= src[(o >> 5) + 1] << (32 - sh);
    }
    return len < 32 ? x & ((1u << len) - 1u) : x;
}

__device__ __forceinline__ std::uint32_t piece_bits(const Piece& pc, std::uint64_t o, std::uint32_t len) {
    if (pc.parts == nullptr) {
        return run_bits(pc.src, o, len);
    }
    std::uint32_t v{0}, got{0};
    std::uint64_t start{0};
    for (int p{0}; p < kGroupParts && got < len; ++p) {
        const std::uint64_t pb{pc.parts[p]};
        if (o < start + pb) {
            const std::uint64_t off{o - start};
            const std::uint32_t take{static_cast<std::uint32_t>(
                pb - off < static_cast<std::uint64_t>(len - got) ? pb - off : static_cast<std::uint64_t>(len - got))};
            v |= run_bits(pc.src + static_cast<std::size_t>(p) * kPartSlotWords, off, take) << got;
            got += take;
            o += take;
        }
        start += pb;
    }
    return v;
}

__device__ __forceinline__ Piece piece_at(const Geometry& geo, const Constants* k, const Layout* lay,
                                          const Pieces& P, int i) {
    const int nd{static_cast<int>(geo.num_dc)};
    if (i == 0) {
        return {0, static_cast<std::uint64_t>(lay->framing_bytes) * 8, lay->framing_words, nullptr};
    }
    if (i == 1) {
        return {static_cast<std::uint64_t>(lay->framing_bytes) * 8, lay->head_total_bits, lay->head_words, nullptr};
    }
    if (i == 2) {
        return {lay->dc_global_bit, k->dc_global_bits, k->dc_global_words, nullptr};
    }
    if (i < 3 + nd) {
        const int g{i - 3};
        return {P.dc_base[g], P.dc_bits[g], P.dc_staging + static_cast<std::size_t>(g) * P.dc_slot_words, nullptr};
    }
    if (i == 3 + nd) {
        return {lay->ac_global_bit, lay->ac_global_bits, lay->ac_global_words, nullptr};
    }
    const int g{i - 4 - nd};
    return {P.ac_base[g], P.ac_bits[g], P.ac_staging + static_cast<std::size_t>(g) * kGroupParts * kPartSlotWords,
            P.ac_part_bits + static_cast<std::size_t>(g) * kGroupParts};
}

__device__ __attribute__((always_inline)) inline void assemble_kernel_body(const Wg& wg, Geometry geo, const Constants* __restrict__ k,
                                                       const Layout* __restrict__ lay, Pieces P,
                                                       std::uint32_t* __restrict__ out_words,
                                                       std::uint64_t cap_bytes) {
    __shared__ std::uint64_t end_bit[kMaxPieces];
    if (lay->overflow != 0) {
        return;
    }
    const int n{4 + static_cast<int>(geo.num_dc) + static_cast<int>(geo.num_ac)};
    for (int i{static_cast<int>(threadIdx.x)}; i < n; i += 256) {
        const Piece pc{piece_at(geo, k, lay, P, i)};
        end_bit[i] = pc.dst + pc.nbits;
    }
    __syncthreads();
    // Whole words up to the capacity, then the bytes of a partial last word
    // that still fit (the output may be the caller's mapped host buffer).
    std::uint64_t words{(lay->total_bytes + 3) / 4};
    const std::uint64_t cap_words{(cap_bytes + 3) / 4};
    words = words < cap_words ? words : cap_words;
    for (std::uint64_t w{wg.x * 256ull + threadIdx.x}; w < words; w += wg.nx * 256ull) {
        const std::uint64_t b0{w * 32};
        // First piece ending after b0 (pieces are ordered and disjoint).
        int lo{0}, hi{n};
        while (lo < hi) {
            const int mid{(lo + hi) >> 1};
            if (end_bit[mid] > b0) {
                hi = mid;
            } else {
                lo = mid + 1;
            }
        }
        std::uint32_t v{0};
        for (int i{lo}; i < n; ++i) {
            const Piece pc{piece_at(geo, k, lay, P, i)};
            if (pc.dst >= b0 + 32) {
                break;
            }
            const std::uint64_t s{pc.dst > b0 ? pc.dst : b0};
            const std::uint64_t e{pc.dst + pc.nbits < b0 + 32 ? pc.dst + pc.nbits : b0 + 32};
            if (s >= e) {
                continue;
            }
            v |= piece_bits(pc, s - pc.dst, static_cast<std::uint32_t>(e - s)) << static_cast<std::uint32_t>(s - b0);
        }
        if ((w + 1) * 4 <= cap_bytes) {
            out_words[w] = v;
        } else {
            auto* ob{reinterpret_cast<std::uint8_t*>(out_words)};
            for (std::uint64_t b{w * 4}; b < cap_bytes; ++b) {
                ob[b] = static_cast<std::uint8_t>(v >> (8 * (b - w * 4)));
            }
        }
    }
}

__global__ __launch_bounds__(256) void assemble_kernel(Geometry geo, const Constants* __restrict__ k,
                                                       const Layout* __restrict__ lay, Pieces P,
                                                       std::uint32_t* __restrict__ out_words,
                                                       std::uint64_t cap_bytes) {
    assemble_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, k, lay, P, out_words, cap_bytes);
}

}  // namespace cjx
