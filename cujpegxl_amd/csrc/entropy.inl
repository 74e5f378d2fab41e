// Entropy-coding kernels (included into device.hip).
//
// AC: one wave tokenizes one block-channel (an "item": first-block x channel
// in decoder order Y, X, B).  Each 64-coefficient chunk of the block's scan
// order is classified with one ballot; a lane's remaining-non-zero count and
// previous-coefficient flag -- the two serial dependencies of the AC context
// model (entropy.cu:648-681, ac_context.h:78-84) -- come from a popcount of
// the ballot below it, so every lane forms its token in parallel.  The same
// tokenizer drives three passes: cluster histograms, per-item bit counts, and
// bit emission at scanned offsets.
//
// DC: a DcGroup section is a list of items (header blobs, DC residual tokens,
// AcMetadata tokens), each with a known bit length once the codes exist;
// chunked scans give every item its bit offset and the emitter ORs its bits in.

namespace cjx {

__constant__ std::uint8_t c_ctx_map[39]{0, 1, 2, 2, 3,  3,  4,  5,  6,  6,  6,  6,  6,
                                        7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14,
                                        7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14};
__constant__ std::uint16_t c_freq_ctx[64]{
    0,  0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 15, 16, 16, 17, 17,
    18, 18, 19, 19, 20, 20, 21, 21, 22, 22, 23, 23, 23, 23, 24, 24, 24, 24, 25, 25, 25, 25,
    26, 26, 26, 26, 27, 27, 27, 27, 28, 28, 28, 28, 29, 29, 29, 29, 30, 30, 30, 30};
__constant__ std::uint16_t c_nz_ctx[64]{
    0,   0,   31,  62,  62,  93,  93,  93,  93,  123, 123, 123, 123, 152, 152, 152,
    152, 152, 152, 152, 152, 180, 180, 180, 180, 180, 180, 180, 180, 180, 180, 180,
    180, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
    206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206};

// ac_context.h:108-131
__device__ __forceinline__ int ac_cluster(std::uint32_t ctx) {
    constexpr std::uint32_t count_contexts{15u * 37u};
    if (ctx < count_contexts) {
        const std::uint32_t bucket{ctx / 15u};
        const int group{static_cast<int>(ctx % 15u) < 7 ? 0 : 1};
        const int band{bucket < 4 ? 0 : (bucket < 12 ? 1 : (bucket < 24 ? 2 : 3))};
        return group * 4 + band;
    }
    const std::uint32_t rel{ctx - count_contexts};
    const int group{static_cast<int>(rel / 458u) < 7 ? 0 : 1};
    const std::uint32_t z{rel % 458u};
    const int band{z < 32    ? 0
                   : z < 64  ? 1
                   : z < 96  ? 2
                   : z < 128 ? 3
                   : z < 160 ? 4
                   : z < 200 ? 5
                   : z < 240 ? 6
                   : z < 280 ? 7
                   : z < 328 ? 8
                   : z < 376 ? 9
                   : z < 424 ? 10
                             : 11};
    return 8 + group * 12 + band;
}


// Inclusive prefix sums over the wave by DPP moves (no LDS crossbar): within
// each row of 16 lanes by shifts of 1, 2, 4 and 8 (lanes shifted in from
// outside the row read 0), then the last lane of rows 0 and 2 into rows 1
// and 3 and the last lane of row 1 into rows 2 and 3 (row broadcasts; rows
// the mask leaves out add 0).
template <int kCtrl, int kRowMask>
__device__ __forceinline__ std::uint32_t dpp_u32(std::uint32_t x) {
    return static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(x), kCtrl, kRowMask, 0xf, kCtrl < 0x140));
}

template <int kCtrl, int kRowMask>
__device__ __forceinline__ std::uint64_t dpp_u64(std::uint64_t x) {
    const std::uint32_t lo{dpp_u32<kCtrl, kRowMask>(static_cast<std::uint32_t>(x))};
    const std::uint32_t hi{dpp_u32<kCtrl, kRowMask>(static_cast<std::uint32_t>(x >> 32))};
    return (static_cast<std::uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ std::uint32_t wave_incl_scan_u32(std::uint32_t x) {
    x += dpp_u32<0x111, 0xf>(x);
    x += dpp_u32<0x112, 0xf>(x);
    x += dpp_u32<0x114, 0xf>(x);
    x += dpp_u32<0x118, 0xf>(x);
    x += dpp_u32<0x142, 0xa>(x);
    return x + dpp_u32<0x143, 0xc>(x);
}

__device__ __forceinline__ std::uint64_t wave_incl_scan_u64(std::uint64_t x) {
    x += dpp_u64<0x111, 0xf>(x);
    x += dpp_u64<0x112, 0xf>(x);
    x += dpp_u64<0x114, 0xf>(x);
    x += dpp_u64<0x118, 0xf>(x);
    x += dpp_u64<0x142, 0xa>(x);
    return x + dpp_u64<0x143, 0xc>(x);
}

// Sum over the wave, in every lane (the scan's last lane, read as a scalar).
__device__ __forceinline__ int wave_sum_int(int v) {
    return __builtin_amdgcn_readlane(static_cast<int>(wave_incl_scan_u32(static_cast<std::uint32_t>(v))), 63);
}

// Exclusive prefix over the wave.
__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
    (void)lane;
    return static_cast<int>(wave_incl_scan_u32(static_cast<std::uint32_t>(v)) - static_cast<std::uint32_t>(v));
}

// ORs `len` (<= 32) bits of `value` at absolute bit position `pos`; words at
// or beyond `cap_words` are never touched (the layout makes that unreachable;
// the check keeps an inconsistent layout from writing out of bounds).
__device__ __forceinline__ void or_bits(std::uint32_t* words, std::uint64_t cap_words, std::uint64_t pos,
                                        std::uint32_t value, std::uint32_t len) {
    if (len == 0) {
        return;
    }
    if (len < 32) {
        value &= (1u << len) - 1u;
    }
    const std::uint64_t wi{pos >> 5};
    const std::uint32_t b{static_cast<std::uint32_t>(pos & 31u)};
    if (wi < cap_words) {
        atomicOr(&words[wi], value << b);
    }
    if (b + len > 32 && wi + 1 < cap_words) {
        atomicOr(&words[wi + 1], value >> (32 - b));
    }
}

// ORs `len` (<= 32) bits at bit `pos` of an LDS word buffer of `cap` words.
__device__ __forceinline__ void lds_or_bits(std::uint32_t* words, std::uint32_t cap, std::uint32_t pos,
                                            std::uint32_t value, std::uint32_t len) {
    if (len == 0) {
        return;
    }
    if (len < 32) {
        value &= (1u << len) - 1u;
    }
    const std::uint32_t wi{pos >> 5}, b{pos & 31u};
    if (wi < cap) {
        atomicOr(&words[wi], value << b);
    }
    if (b + len > 32 && wi + 1 < cap) {
        atomicOr(&words[wi + 1], value >> (32 - b));
    }
}

// Orders one wave's LDS accesses: writes before the barrier are seen by
// every lane's reads after it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

struct AcItem {
    int n;             // transform side, 0 if covered
    int c;             // XYB channel
    int bx, by;
    std::uint32_t nz;  // non-zero AC count
    std::uint32_t nz_ctx;
    int bctx;
};

__device__ __forceinline__ AcItem ac_item(const Geometry& geo, const std::int8_t* acs,
                                          const std::uint8_t* nz_grid, const std::uint16_t* nz_count,
                                          std::uint32_t item) {
    AcItem it{};
    const std::uint32_t pos{item / 3};
    const int ci{static_cast<int>(item % 3)};
    it.c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
    it.bx = static_cast<int>(pos % geo.bw);
    it.by = static_cast<int>(pos / geo.bw);
    it.n = acs[pos];
    if (it.n == 0) {
        return it;
    }
    const std::size_t nblk{geo.nblocks};
    it.nz = nz_count[it.c * nblk + pos];
    const int ord{it.n == 16 ? 2 : (it.n == 32 ? 3 : 0)};
    it.bctx = c_ctx_map[(it.c < 2 ? (it.c ^ 1) : 2) * 13 + ord];
    const bool hl{(it.bx % kAcGroupBlocks) > 0}, ht{(it.by % kAcGroupBlocks) > 0};
    const int top{ht ? nz_grid[it.c * nblk + pos - geo.bw] : 0};
    const int left{hl ? nz_grid[it.c * nblk + pos - 1] : 0};
    std::uint32_t pred{!hl ? (ht ? static_cast<std::uint32_t>(top) : 32u)
                           : (!ht ? static_cast<std::uint32_t>(left)
                                  : static_cast<std::uint32_t>((top + left + 1) / 2))};
    pred = pred >= 64 ? 64 : pred;
    it.nz_ctx = (pred < 8 ? pred : 4 + pred / 2) * 15u + static_cast<std::uint32_t>(it.bctx);
    return it;
}

// Tokenizes two items per wave, one per 32-lane half (item `it` of this
// lane's half, its coefficients staged in `coef` block after block, so that
// a scan-order entry (block << 6 | position) indexes it directly; `has`
// false for a half without an item).  `emit(valid, k,
// cluster, sym, nbits, extra)` is called once per lane per chunk, k being the
// token's index in the item: 0 for the non-zero-count token (lane 0 of the
// half, first call), then one token per scan position from the first
// non-LLF position up to the last non-zero coefficient, so the valid lanes of
// a chunk are a prefix of the half.  A lane's remaining-non-zero count and
// previous-coefficient flag (entropy.cu:648-681, ac_context.h:78-84) come
// from popcounts of the half's ballot bits below it.
// The tokenizer's lookup tables, staged in LDS once per workgroup (the
// lookups are per-lane gathers).
// Zero-density contexts per block context: (nz_ctx[rem] + freq_ctx[k]) * 2 +
// prev stays below 458 because rem + k <= 64 (ac_context.h:78-84).
constexpr int kZeroDensityContexts = 458;

struct TokenTables {
    std::uint16_t order8[64];
    std::uint16_t order16[256];
    std::uint16_t order32[1024];
    std::uint16_t nz_ctx[64];
    std::uint16_t freq_ctx[64];
    std::uint8_t zero_band[kZeroDensityContexts];  // ac_cluster's band of each zero-density context
};

__device__ __forceinline__ void load_token_tables(TokenTables& T) {
    for (int i{static_cast<int>(threadIdx.x)}; i < 1024; i += static_cast<int>(blockDim.x)) {
        T.order32[i] = g_tables.order32[i];
        if (i < 256) {
            T.order16[i] = g_tables.order16[i];
        }
        if (i < 64) {
            T.order8[i] = g_tables.order8[i];
            T.nz_ctx[i] = c_nz_ctx[i];
            T.freq_ctx[i] = c_freq_ctx[i];
        }
        if (i < kZeroDensityContexts) {
            T.zero_band[i] = static_cast<std::uint8_t>(ac_cluster(15u * 37u + static_cast<std::uint32_t>(i)) - 8);
        }
    }
}

template <typename Emit>
__device__ __forceinline__ void tokenize_item_pair(const TokenTables& T, const std::int16_t* coef, const AcItem& it,
                                                   bool has, int lane, Emit&& emit) {
    const int half{lane >> 5}, sub{lane & 31};
    {
        std::uint32_t sym{0}, nb{0}, bits{0};
        hybrid420(it.nz, sym, nb, bits);
        emit(has && sub == 0, 0u, has ? ac_cluster(it.nz_ctx) : 0, sym, nb, bits);
    }
    const int n{has ? it.n : 8};
    const int side{n / 8};
    const std::uint32_t cb{static_cast<std::uint32_t>(side * side)};
    const std::uint32_t size{static_cast<std::uint32_t>(n * n)};
    const std::uint32_t l2{n == 16 ? 2u : (n == 32 ? 4u : 0u)};
    const int zgroup{it.bctx < 7 ? 8 : 20};  // 8 + 12 * (block context >= 7)
    const std::uint16_t* order{n == 32 ? T.order32 : (n == 16 ? T.order16 : T.order8)};
    std::uint32_t seen{0};
    std::uint32_t prev_carry{it.nz > size / 16 ? 0u : 1u};
    const std::uint32_t below{(1u << sub) - 1u};
    for (std::uint32_t base{cb};; base += 32) {
        const bool active{has && base < size && seen < it.nz};
        if (__ballot(active) == 0) {
            break;
        }
        const std::uint32_t k{base + static_cast<std::uint32_t>(sub)};
        int v{0};
        if (active && k < size) {
            v = coef[order[k]];
        }
        const std::uint32_t gm{static_cast<std::uint32_t>(__ballot(v != 0) >> (32 * half))};
        const std::uint32_t before{seen + static_cast<std::uint32_t>(__popc(gm & below))};
        const std::uint32_t prev{sub == 0 ? prev_carry : ((gm >> (sub - 1)) & 1u)};
        const bool valid{active && k < size && before < it.nz};
        std::uint32_t cl{0}, sym{0}, nb{0}, bits{0};
        if (valid) {
            const std::uint32_t rem{((it.nz - before) + cb - 1) >> l2};
            const std::uint32_t kk{k >> l2};
            cl = static_cast<std::uint32_t>(zgroup + T.zero_band[(T.nz_ctx[rem] + T.freq_ctx[kk]) * 2u + prev]);
            hybrid420(pack_signed(v), sym, nb, bits);
        }
        emit(valid, 1u + (k - cb), static_cast<int>(cl), sym, nb, bits);
        seen += static_cast<std::uint32_t>(__popc(gm));
        prev_carry = gm >> 31;
    }
}

// Token records: cluster | symbol << 5 | extra-bit count << 13 | extra << 18
// (hybrid-uint extra bits of 16-bit magnitudes fit in 14 bits).  Token k of
// an item lives in the k-th coefficient slot of the item's covered blocks
// (an item has at most 1 + n*n - (n/8)^2 <= n*n tokens).
__device__ __forceinline__ std::uint32_t token_record(int cl, std::uint32_t sym, std::uint32_t nb,
                                                      std::uint32_t extra) {
    return static_cast<std::uint32_t>(cl) | (sym << 5) | (nb << 13) | (extra << 18);
}

__device__ __forceinline__ std::size_t token_slot(const Geometry& geo, int c, int bx, int by, int m,
                                                  std::uint32_t k) {
    const int q{static_cast<int>(k >> 6)};
    const int l2m{m >> 1};  // m in {1, 2, 4}
    return static_cast<std::size_t>(c) * geo.nblocks * 64 +
           (static_cast<std::size_t>(by + (q >> l2m)) * geo.bw + static_cast<std::size_t>(bx + (q & (m - 1)))) * 64 +
           (k & 63u);
}

// Pass 1: tokenize every item once: token records, tokens per item and the
// per-cluster histograms.  AC symbols are below 64 (hybrid-uint tokens of
// 16-bit magnitudes and of counts <= 1024), so each workgroup counts into a
// 32 x 64 LDS histogram and writes it out whole as one partial; a second
// kernel sums the partials.  Grid-stride over item pairs, one wave per pair:
// the descriptors of 32 pairs are loaded at once (one item per lane), and
// each pair's coefficient blocks are read as 16-byte vectors while the
// previous pair is tokenized from its LDS copy.
constexpr int kAcSymbols = 64;
// LDS row stride of a cluster's histogram: one word more than the symbols,
// so equal symbols of different clusters fall in different banks.
constexpr int kHistRow = kAcSymbols + 1;
// Two copies of the histogram, taken by even and odd lanes (neighbouring
// lanes often carry the same cluster and symbol); the second starts 16
// banks after the first, so one key's two words sit in different banks.
constexpr int kHistCopy = kNumClusters * kHistRow + 16;
constexpr int kTokenGrid = 2048;
constexpr int kTokenBatch = 32;

__device__ __forceinline__ AcItem shfl_item(const AcItem& d, int src) {
    AcItem o;
    o.n = __shfl(d.n, src, 64);
    o.c = __shfl(d.c, src, 64);
    o.bx = __shfl(d.bx, src, 64);
    o.by = __shfl(d.by, src, 64);
    o.nz = static_cast<std::uint32_t>(__shfl(static_cast<int>(d.nz), src, 64));
    o.nz_ctx = static_cast<std::uint32_t>(__shfl(static_cast<int>(d.nz_ctx), src, 64));
    o.bctx = __shfl(d.bctx, src, 64);
    return o;
}

// The m x m coefficient blocks of an item (m = n / 8; 8 to 128 vectors of
// 16 bytes), lane `sub` of the half taking vectors sub, sub + 32, sub + 64
// and sub + 96 (four named registers, so the set never goes to scratch).
struct ItemVecs {
    uint4 r0, r1, r2, r3;
};

__device__ __forceinline__ int item_vecs(int n) {
    return n == 32 ? 128 : (n == 16 ? 32 : (n == 8 ? 8 : 0));
}

__device__ __forceinline__ uint4 load_item_vec(const std::int16_t* plane, const Geometry& geo, const AcItem& it,
                                               int l2s, int u) {
    const int q{u >> 3};
    const std::size_t blk{static_cast<std::size_t>(it.by + (q >> l2s)) * geo.bw +
                          static_cast<std::size_t>(it.bx + (q & ((1 << l2s) - 1)))};
    return *reinterpret_cast<const uint4*>(plane + blk * 64 + (u & 7) * 8);
}

__device__ __forceinline__ void load_item_coefs(const Geometry& geo, const std::int16_t* ac, const AcItem& it,
                                                int sub, ItemVecs& v) {
    const int l2s{it.n == 32 ? 2 : (it.n == 16 ? 1 : 0)};
    const int vecs{item_vecs(it.n)};
    const std::int16_t* plane{ac + static_cast<std::size_t>(it.c) * geo.nblocks * 64};
    if (sub < vecs) {
        v.r0 = load_item_vec(plane, geo, it, l2s, sub);
    }
    if (sub + 32 < vecs) {
        v.r1 = load_item_vec(plane, geo, it, l2s, sub + 32);
        v.r2 = load_item_vec(plane, geo, it, l2s, sub + 64);
        v.r3 = load_item_vec(plane, geo, it, l2s, sub + 96);
    }
}

__device__ __forceinline__ void store_item_coefs(std::int16_t* coef, const AcItem& it, int sub, const ItemVecs& v) {
    const int vecs{item_vecs(it.n)};
    if (sub < vecs) {
        *reinterpret_cast<uint4*>(coef + sub * 8) = v.r0;
    }
    if (sub + 32 < vecs) {
        *reinterpret_cast<uint4*>(coef + (sub + 32) * 8) = v.r1;
        *reinterpret_cast<uint4*>(coef + (sub + 64) * 8) = v.r2;
        *reinterpret_cast<uint4*>(coef + (sub + 96) * 8) = v.r3;
    }
}

__device__ __attribute__((always_inline)) inline void ac_tokens_kernel_body(const Wg& wg, Geometry geo, const std::int16_t* __restrict__ ac,
                                                        const std::int8_t* __restrict__ acs,
                                                        const std::uint8_t* __restrict__ nz_grid,
                                                        const std::uint16_t* __restrict__ nz_count,
                                                        std::uint32_t* __restrict__ tokens,
                                                        std::uint16_t* __restrict__ item_tokens,
                                                        std::uint32_t* __restrict__ partial,
                                                        const std::uint32_t* __restrict__ list,
                                                        const std::uint32_t* __restrict__ list_count) {
    __shared__ std::uint32_t h2[2 * kHistCopy];
    __shared__ TokenTables T;
    __shared__ __align__(16) std::int16_t stage[4][2][1024];
    for (int i{static_cast<int>(threadIdx.x)}; i < 2 * kHistCopy; i += 256) {
        h2[i] = 0;
    }
    load_token_tables(T);
    __syncthreads();
    const int lane{static_cast<int>(threadIdx.x) & 63};
    const int half{lane >> 5}, sub{lane & 31};
    std::uint32_t* const h{h2 + (lane & 1) * kHistCopy};
    std::int16_t* const coef{stage[threadIdx.x >> 6][half]};
    const std::uint32_t wave{wg.x * 4 + (threadIdx.x >> 6)};
    const std::uint32_t stride{2 * wg.nx * 4};
    // Walk the quantizer's list of items with non-zeros when there is one
    // (the others follow below; covered blocks' items already have 0
    // tokens), else every item.
    const std::uint32_t items{list != nullptr ? list_count[0] : geo.nblocks * 3};
    for (std::uint32_t base{2 * wave}; base < items; base += stride * kTokenBatch) {
        const std::uint32_t li{base + stride * static_cast<std::uint32_t>(lane >> 1) +
                               static_cast<std::uint32_t>(lane & 1)};
        const AcItem d{li < items ? ac_item(geo, acs, nz_grid, nz_count, list != nullptr ? list[li] : li) : AcItem{}};
        const int npairs{static_cast<int>(min(static_cast<std::uint32_t>(kTokenBatch),
                                              (items - base + stride - 1) / stride))};
        AcItem cur{shfl_item(d, half)};
        ItemVecs v{};
        load_item_coefs(geo, ac, cur, sub, v);
        for (int t{0}; t < npairs; ++t) {
            const std::uint32_t item{base + stride * static_cast<std::uint32_t>(t) + static_cast<std::uint32_t>(half)};
            wave_sync();
            store_item_coefs(coef, cur, sub, v);
            wave_sync();
            AcItem nx{};
            if (t + 1 < npairs) {
                nx = shfl_item(d, 2 * (t + 1) + half);
                load_item_coefs(geo, ac, nx, sub, v);
            }
            const bool has{cur.n != 0};
            std::uint32_t count{0};
            if (__ballot(has) != 0) {
                const AcItem it{cur};
                const std::uint32_t m{has ? static_cast<std::uint32_t>(it.n / 8) : 1u};
                const std::uint32_t l2m{m >> 1};
                // token_slot in 32 bits (3 * blocks * 64 < 2^32 for the
                // at most 32 DC groups a frame may have).
                const std::uint32_t slot0{static_cast<std::uint32_t>(token_slot(geo, has ? it.c : 0, it.bx, it.by, 1, 0))};
                const std::uint32_t row{geo.bw * 64};
                tokenize_item_pair(T, coef, it, has, lane,
                                   [&](bool valid, std::uint32_t k, int cl, std::uint32_t sym, std::uint32_t nb,
                                       std::uint32_t extra) {
                                       if (valid) {
                                           atomicAdd(&h[cl * kHistRow + (sym & (kAcSymbols - 1))], 1u);
                                           const std::uint32_t q{k >> 6};
                                           tokens[slot0 + (q >> l2m) * row + ((q & (m - 1)) << 6) + (k & 63u)] =
                                               token_record(cl, sym, nb, extra);
                                       }
                                       count += static_cast<std::uint32_t>(
                                           __popc(static_cast<std::uint32_t>(__ballot(valid) >> (32 * half))));
                                   });
            }
            if (sub == 0 && item < items) {
                // The item's own index: block * 3 + (0: Y, 1: X, 2: B).
                const std::uint32_t own{(static_cast<std::uint32_t>(cur.by) * geo.bw + static_cast<std::uint32_t>(cur.bx)) * 3u +
                                        (cur.c == 1 ? 0u : (cur.c == 0 ? 1u : 2u))};
                item_tokens[own] = static_cast<std::uint16_t>(count);
            }
            cur = nx;
        }
    }
    // First blocks without non-zero coefficients, one per thread: their one
    // token is the zero count, in the count's context.
    if (list != nullptr) {
        const std::uint32_t zeros{list_count[1]};
        const std::uint32_t last{3u * geo.nblocks - 1u};
        for (std::uint32_t i{wg.x * 256u + threadIdx.x}; i < zeros; i += wg.nx * 256u) {
            const AcItem it{ac_item(geo, acs, nz_grid, nz_count, list[last - i])};
            std::uint32_t sym{0}, nb{0}, bits{0};
            hybrid420(0u, sym, nb, bits);
            const int cl{ac_cluster(it.nz_ctx)};
            atomicAdd(&h[cl * kHistRow + (sym & (kAcSymbols - 1))], 1u);
            tokens[token_slot(geo, it.c, it.bx, it.by, 1, 0)] = token_record(cl, sym, nb, bits);
            item_tokens[(static_cast<std::uint32_t>(it.by) * geo.bw + static_cast<std::uint32_t>(it.bx)) * 3u +
                        (it.c == 1 ? 0u : (it.c == 0 ? 1u : 2u))] = 1;
        }
    }
    __syncthreads();
    std::uint32_t* mine{partial + static_cast<std::size_t>(wg.x) * kNumClusters * kAcSymbols};
    for (int i{static_cast<int>(threadIdx.x)}; i < kNumClusters * kAcSymbols; i += 256) {
        const int r{(i >> 6) * kHistRow + (i & (kAcSymbols - 1))};
        mine[i] = h2[r] + h2[kHistCopy + r];
    }
}

__global__ __launch_bounds__(256) void ac_tokens_kernel(Geometry geo, const std::int16_t* __restrict__ ac,
                                                        const std::int8_t* __restrict__ acs,
                                                        const std::uint8_t* __restrict__ nz_grid,
                                                        const std::uint16_t* __restrict__ nz_count,
                                                        std::uint32_t* __restrict__ tokens,
                                                        std::uint16_t* __restrict__ item_tokens,
                                                        std::uint32_t* __restrict__ partial,
                                                        const std::uint32_t* __restrict__ list,
                                                        const std::uint32_t* __restrict__ list_count) {
    ac_tokens_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, ac, acs, nz_grid, nz_count, tokens, item_tokens, partial, list, list_count);
}

// Sums `parts` partial histograms of 32 x 64 bins into the 32 x 256 cluster
// histograms: grid (bins / 256, slices); each thread adds its slice.
__device__ __attribute__((always_inline)) inline void ac_hist_reduce_kernel_body(const Wg& wg, const std::uint32_t* __restrict__ partial, int parts,
                                                             std::uint32_t* __restrict__ hist) {
    const int bin{static_cast<int>(wg.x) * 256 + static_cast<int>(threadIdx.x)};
    const int per{(parts + static_cast<int>(wg.ny) - 1) / static_cast<int>(wg.ny)};
    const int p0{static_cast<int>(wg.y) * per};
    const int p1{min(parts, p0 + per)};
    std::uint32_t sum{0};
    for (int p{p0}; p < p1; ++p) {
        sum += partial[static_cast<std::size_t>(p) * kNumClusters * kAcSymbols + bin];
    }
    if (sum != 0) {
        atomicAdd(&hist[(bin / kAcSymbols) * kHistStride + bin % kAcSymbols], sum);
    }
}

__global__ __launch_bounds__(256) void ac_hist_reduce_kernel(const std::uint32_t* __restrict__ partial, int parts,
                                                             std::uint32_t* __restrict__ hist) {
    ac_hist_reduce_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, partial, parts, hist);
}

// Block-wide exclusive scan of one value per thread (blockDim.x multiple of 64).
__device__ std::uint64_t block_excl_scan(std::uint64_t v, std::uint64_t* sh, std::uint64_t& total) {
    const int tid{static_cast<int>(threadIdx.x)};
    const int lane{tid & 63}, wv{tid >> 6}, nw{static_cast<int>(blockDim.x) >> 6};
    const std::uint64_t x{wave_incl_scan_u64(v)};
    if (lane == 63) {
        sh[wv] = x;
    }
    __syncthreads();
    if (tid == 0) {
        std::uint64_t run{0};
        for (int i{0}; i < nw; ++i) {
            const std::uint64_t t{sh[i]};
            sh[i] = run;
            run += t;
        }
        sh[nw] = run;
    }
    __syncthreads();
    const std::uint64_t r{sh[wv] + x - v};
    total = sh[nw];
    __syncthreads();
    return r;
}

// ---- AC groups emitted in parts -----------------------------------------------
//
// Each AC group is cut into kGroupParts parts of kPartRows block rows (each a
// contiguous run of the group's stream), and one 256-thread workgroup writes
// a part's token stream at part-relative bit positions into the part's
// staging slot: the part's items in decoder order (block rows, blocks left to
// right, channels Y, X, B) and their token counts give a flat token numbering
// (token f belongs to the last item whose prefix is <= f); block scans of the
// token lengths give the positions; the stream is assembled in LDS and stored
// once.  The part's bit count falls out of the same pass (a group's size is
// the sum of its parts').  A stream longer than the LDS copy is emitted again
// with atomicOr into the zeroed slot.
constexpr int kGroupParts = 8;
constexpr int kPartRows = kAcGroupBlocks / kGroupParts;
constexpr int kPartThreads = 256;
constexpr int kPartItems = 3 * kAcGroupBlocks * kPartRows;
constexpr int kPartWords = 2048;
constexpr int kGroupPerThread = 8;
// Bound on one part's stream: 28 bits (15-bit code + 13 extra bits) per
// coefficient slot of its blocks in three channels.
constexpr std::size_t kPartSlotWords = 3ull * kAcGroupBlocks * kPartRows * 64 * 28 / 32 + 64;

struct GroupShared {
    std::uint32_t lut[kNumClusters * kAcSymbols];  // depth | code << 16
    std::uint32_t prefix[kPartItems + 1];
    std::uint32_t slot0[kPartItems];  // token slot of the item's first token
    std::int8_t m[kPartItems];
    std::uint32_t buf[kPartWords];
    std::uint64_t sh[17];
    std::uint32_t overflow;
};

__device__ __attribute__((always_inline)) inline void ac_group_emit_kernel_body(const Wg& wg, 
    Geometry geo, const std::int8_t* __restrict__ acs, const std::uint16_t* __restrict__ item_tokens,
    const std::uint32_t* __restrict__ tokens, const std::uint8_t* __restrict__ depth,
    const std::uint16_t* __restrict__ codes, std::uint32_t* __restrict__ staging,
    std::uint64_t* __restrict__ part_bits) {
    __shared__ GroupShared S;
    const int tid{static_cast<int>(threadIdx.x)};
    const int g{static_cast<int>(wg.x) / kGroupParts};
    const int part{static_cast<int>(wg.x) % kGroupParts};
    const int bx0{(g % static_cast<int>(geo.xg)) * kAcGroupBlocks};
    const int by0{(g / static_cast<int>(geo.xg)) * kAcGroupBlocks + part * kPartRows};
    const int gbw{min(kAcGroupBlocks, static_cast<int>(geo.bw) - bx0)};
    const int gbh{min(kPartRows, static_cast<int>(geo.bh) - by0)};
    if (gbh <= 0) {
        if (tid == 0) {
            part_bits[wg.x] = 0;
        }
        return;
    }
    const int row_items{3 * gbw};
    const int items{row_items * gbh};
    std::uint32_t* const slot{staging + static_cast<std::size_t>(wg.x) * kPartSlotWords};

    // 1. Token prefix over the part's items (at most two per thread).  Every
    //    read of this step (the items' transforms and token counts, the code
    //    table) is issued before the first wait; a covered block's count is
    //    read too and then taken as 0.
    static_assert(kPartItems <= 2 * kPartThreads, "two items per thread");
    static_assert((kNumClusters * kAcSymbols) % kPartThreads == 0, "code table entries per thread");
    constexpr int kLutPerThread{kNumClusters * kAcSymbols / kPartThreads};
    int item_n[2]{0, 0};
    std::uint32_t item_cnt[2]{0u, 0u};
#pragma unroll
    for (int k{0}; k < 2; ++k) {
        const int i{k * kPartThreads + tid};
        if (i < items) {
            const int row{i / row_items}, rem{i % row_items};
            const std::size_t pos{static_cast<std::size_t>(by0 + row) * geo.bw + bx0 + rem / 3};
            item_n[k] = acs[pos];
            item_cnt[k] = item_tokens[pos * 3 + rem % 3];
        }
    }
    std::uint32_t lut[kLutPerThread];
#pragma unroll
    for (int k{0}; k < kLutPerThread; ++k) {
        const int i{k * kPartThreads + tid};
        const int at{(i / kAcSymbols) * kHistStride + i % kAcSymbols};
        lut[k] = static_cast<std::uint32_t>(depth[at]) | (static_cast<std::uint32_t>(codes[at]) << 16);
    }
    std::uint32_t carry{0};
#pragma unroll
    for (int k{0}; k < 2; ++k) {
        const int base{k * kPartThreads};
        if (base >= items) {
            break;
        }
        const int i{base + tid};
        std::uint32_t cnt{0};
        if (i < items) {
            const int row{i / row_items}, rem{i % row_items};
            const int n{item_n[k]};
            const int ci{rem % 3};
            S.m[i] = static_cast<std::int8_t>(n / 8);
            S.slot0[i] = static_cast<std::uint32_t>(
                token_slot(geo, ci == 0 ? 1 : (ci == 1 ? 0 : 2), bx0 + rem / 3, by0 + row, 1, 0));
            cnt = n != 0 ? item_cnt[k] : 0u;
        }
        std::uint64_t total{0};
        const std::uint64_t p{block_excl_scan(cnt, S.sh, total)};
        if (i < items) {
            S.prefix[i] = carry + static_cast<std::uint32_t>(p);
        }
        carry += static_cast<std::uint32_t>(total);
    }
    if (tid == 0) {
        S.prefix[items] = carry;
        S.overflow = 0;
    }
#pragma unroll
    for (int k{0}; k < kLutPerThread; ++k) {
        S.lut[k * kPartThreads + tid] = lut[k];
    }
    for (int i{tid}; i < kPartWords; i += kPartThreads) {
        S.buf[i] = 0;
    }
    __syncthreads();

    // 2. Token lengths, positions and bits; twice if the LDS copy overflows.
    // Each thread takes kGroupPerThread consecutive tokens: one binary search
    // for the first, then a forward walk over the item prefix.
    std::uint64_t run{0};
    for (int pass{0}; pass < 2; ++pass) {
        const bool global{pass == 1};
        if (global) {
            if (S.overflow == 0) {
                break;
            }
            const std::uint64_t nw{(run + 31) / 32};
            for (std::uint64_t i{static_cast<std::uint64_t>(tid)}; i < nw && i < kPartSlotWords; i += kPartThreads) {
                slot[i] = 0;
            }
            __syncthreads();
            run = 0;
        }
        for (std::uint32_t f0{0}; f0 < carry; f0 += kPartThreads * kGroupPerThread) {
            const std::uint32_t fb{f0 + static_cast<std::uint32_t>(tid) * kGroupPerThread};
            std::uint32_t rec[kGroupPerThread];
            int lo{0};
            if (fb < carry) {
                int hi{items - 1};
                while (lo < hi) {
                    const int mid{(lo + hi + 1) >> 1};
                    if (S.prefix[mid] <= fb) {
                        lo = mid;
                    } else {
                        hi = mid - 1;
                    }
                }
            }
#pragma unroll
            for (int j{0}; j < kGroupPerThread; ++j) {
                const std::uint32_t f{fb + static_cast<std::uint32_t>(j)};
                rec[j] = 0xffffffffu;
                if (f < carry) {
                    while (S.prefix[lo + 1] <= f) {
                        ++lo;
                    }
                    const std::uint32_t k{f - S.prefix[lo]};
                    const std::uint32_t m{static_cast<std::uint32_t>(S.m[lo])};
                    const std::uint32_t q{k >> 6};
                    rec[j] = tokens[S.slot0[lo] + (q >> (m >> 1)) * (geo.bw * 64) + ((q & (m - 1)) << 6) + (k & 63u)];
                }
            }
            std::uint32_t len[kGroupPerThread], val[kGroupPerThread];
            std::uint32_t sum{0};
#pragma unroll
            for (int j{0}; j < kGroupPerThread; ++j) {
                len[j] = 0;
                val[j] = 0;
                const std::uint32_t r{rec[j]};
                if (r != 0xffffffffu) {
                    const std::uint32_t e{S.lut[(r & 31u) * kAcSymbols + ((r >> 5) & (kAcSymbols - 1))]};
                    const std::uint32_t d{e & 0xffffu};
                    len[j] = d + ((r >> 13) & 31u);
                    val[j] = ((e >> 16) | ((r >> 18) << d)) & ((1u << len[j]) - 1u);
                }
                sum += len[j];
            }
            std::uint64_t chunk{0};
            std::uint64_t at{run + block_excl_scan(sum, S.sh, chunk)};
#pragma unroll
            for (int j{0}; j < kGroupPerThread; ++j) {
                if (len[j] != 0) {
                    const std::uint64_t wi{at >> 5};
                    const std::uint32_t b{static_cast<std::uint32_t>(at & 31u)};
                    const bool two{b + len[j] > 32};
                    if (global) {
                        if (wi < kPartSlotWords) {
                            atomicOr(&slot[wi], val[j] << b);
                        }
                        if (two && wi + 1 < kPartSlotWords) {
                            atomicOr(&slot[wi + 1], val[j] >> (32 - b));
                        }
                    } else if (wi + (two ? 1 : 0) < static_cast<std::uint64_t>(kPartWords)) {
                        atomicOr(&S.buf[wi], val[j] << b);
                        if (two) {
                            atomicOr(&S.buf[wi + 1], val[j] >> (32 - b));
                        }
                    } else {
                        S.overflow = 1;
                    }
                }
                at += len[j];
            }
            run += chunk;
        }
        __syncthreads();
    }
    if (S.overflow == 0) {
        const std::uint64_t nw{(run + 31) / 32};
        for (std::uint64_t i{static_cast<std::uint64_t>(tid)}; i < nw; i += kPartThreads) {
            slot[i] = S.buf[i];
        }
    }
    if (tid == 0) {
        part_bits[wg.x] = run;
    }
}

__global__ __launch_bounds__(kPartThreads) void ac_group_emit_kernel(
    Geometry geo, const std::int8_t* __restrict__ acs, const std::uint16_t* __restrict__ item_tokens,
    const std::uint32_t* __restrict__ tokens, const std::uint8_t* __restrict__ depth,
    const std::uint16_t* __restrict__ codes, std::uint32_t* __restrict__ staging,
    std::uint64_t* __restrict__ part_bits) {
    ac_group_emit_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, acs, item_tokens, tokens, depth, codes, staging, part_bits);
}

// ---- DC groups -------------------------------------------------------------

// Item kinds of a DcGroup section, in order (vardct_frame.cc:543-554):
//   0: blob_pre | [0,3n): DC residuals, channels Y, X, B, group raster |
//   blob_mid | meta: YtoX (cw*ch), YtoB (cw*ch), strategy per block (first
//   blocks only), quant field per block (first blocks only), EPF zeros (n).
struct DcGroupGeo {
    int bx0, by0, dgw, dgh, n, cw, ch;
    std::uint32_t items;
};

__device__ __forceinline__ DcGroupGeo dc_group_geo(const Geometry& geo, int g) {
    DcGroupGeo d{};
    d.bx0 = (g % static_cast<int>(geo.xdg)) * 256;
    d.by0 = (g / static_cast<int>(geo.xdg)) * 256;
    d.dgw = min(256, static_cast<int>(geo.bw) - d.bx0);
    d.dgh = min(256, static_cast<int>(geo.bh) - d.by0);
    d.n = d.dgw * d.dgh;
    d.cw = (d.dgw + 7) / 8;
    d.ch = (d.dgh + 7) / 8;
    d.items = static_cast<std::uint32_t>(2 + 6 * d.n + 2 * d.cw * d.ch);
    return d;
}

// dc_predict.h:22-49
__device__ __forceinline__ std::int32_t gradient_predict(bool hl, bool ht, std::int32_t wv,
                                                         std::int32_t nv, std::int32_t nwv) {
    const std::int32_t left{hl ? wv : (ht ? nv : 0)};
    const std::int32_t top{ht ? nv : left};
    const std::int32_t tl{(hl && ht) ? nwv : left};
    const std::int32_t lo{left < top ? left : top}, hi{left < top ? top : left};
    if (tl < lo) {
        return hi;
    }
    if (tl > hi) {
        return lo;
    }
    return static_cast<std::int32_t>(static_cast<std::int64_t>(left) + top - tl);
}

__device__ __forceinline__ std::uint32_t codes_or_zero(const std::uint16_t* codes, std::uint32_t idx) {
    return codes != nullptr ? static_cast<std::uint32_t>(codes[idx]) : 0u;
}

// Walks the items of a DcGroup section in order.  Sections: 0 blob_pre,
// 1 DC residuals (channel ci, block x, y), 2 blob_mid, 3 YtoX and 4 YtoB
// (tile x, y), 5 strategy and 6 quant field (block x, y), 7 EPF zeros.
// Positioning at item k divides once; advancing is incremental.
struct DcCursor {
    int sec, ci, x, y;
};

__device__ __forceinline__ DcCursor dc_cursor_at(const DcGroupGeo& d, std::uint32_t k) {
    DcCursor c{};
    const std::uint32_t n{static_cast<std::uint32_t>(d.n)};
    const std::uint32_t cwch{static_cast<std::uint32_t>(d.cw * d.ch)};
    if (k == 0) {
        c.sec = 0;
    } else if (k < 1 + 3 * n) {
        const std::uint32_t i{k - 1};
        const std::uint32_t p{i % n};
        c.sec = 1;
        c.ci = static_cast<int>(i / n);
        c.x = static_cast<int>(p % static_cast<std::uint32_t>(d.dgw));
        c.y = static_cast<int>(p / static_cast<std::uint32_t>(d.dgw));
    } else if (k == 1 + 3 * n) {
        c.sec = 2;
    } else {
        const std::uint32_t j{k - (2 + 3 * n)};
        if (j < 2 * cwch) {
            const std::uint32_t jj{j < cwch ? j : j - cwch};
            c.sec = j < cwch ? 3 : 4;
            c.x = static_cast<int>(jj % static_cast<std::uint32_t>(d.cw));
            c.y = static_cast<int>(jj / static_cast<std::uint32_t>(d.cw));
        } else if (j < 2 * cwch + 2 * n) {
            const std::uint32_t jj{(j - 2 * cwch) % n};
            c.sec = j < 2 * cwch + n ? 5 : 6;
            c.x = static_cast<int>(jj % static_cast<std::uint32_t>(d.dgw));
            c.y = static_cast<int>(jj / static_cast<std::uint32_t>(d.dgw));
        } else {
            c.sec = 7;
        }
    }
    return c;
}

__device__ __forceinline__ void dc_cursor_next(const DcGroupGeo& d, DcCursor& c) {
    switch (c.sec) {
        case 0:
        case 2:
            ++c.sec;
            c.ci = 0;
            c.x = 0;
            c.y = 0;
            break;
        case 1:
            if (++c.x == d.dgw) {
                c.x = 0;
                if (++c.y == d.dgh) {
                    c.y = 0;
                    if (++c.ci == 3) {
                        c.sec = 2;
                    }
                }
            }
            break;
        case 3:
        case 4:
            if (++c.x == d.cw) {
                c.x = 0;
                if (++c.y == d.ch) {
                    c.y = 0;
                    ++c.sec;
                }
            }
            break;
        case 5:
        case 6:
            if (++c.x == d.dgw) {
                c.x = 0;
                if (++c.y == d.dgh) {
                    c.y = 0;
                    ++c.sec;
                }
            }
            break;
        default:
            break;
    }
}

// Token of the cursor's item (sections 1 and 3-7).  Returns 0 for a DC code,
// 1 for an AcMetadata code, 2 for an absent item (a covered block's
// strategy / quant-field slot).
__device__ __forceinline__ int dc_cursor_token(const Geometry& geo, const DcGroupGeo& d, const DcCursor& c,
                                               const std::int32_t* dc, const std::int8_t* acs,
                                               const std::int32_t* qf, const std::int8_t* ytox,
                                               const std::int8_t* ytob, std::uint32_t& sym, std::uint32_t& nb,
                                               std::uint32_t& bits) {
    if (c.sec == 1) {
        const int ch{c.ci == 0 ? 1 : (c.ci == 1 ? 0 : 2)};
        const std::int32_t* pl{dc + static_cast<std::size_t>(ch) * geo.nblocks};
        const std::size_t b{static_cast<std::size_t>(d.by0 + c.y) * geo.bw + d.bx0 + c.x};
        const bool hl{c.x > 0}, ht{c.y > 0};
        const std::int32_t pred{gradient_predict(hl, ht, hl ? pl[b - 1] : 0, ht ? pl[b - geo.bw] : 0,
                                                 (hl && ht) ? pl[b - geo.bw - 1] : 0)};
        hybrid420(pack_signed(pl[b] - pred), sym, nb, bits);
        return 0;
    }
    std::int32_t v{0};
    if (c.sec == 3 || c.sec == 4) {
        const std::size_t t{static_cast<std::size_t>(d.by0 / 8 + c.y) * geo.cmw + d.bx0 / 8 + c.x};
        v = c.sec == 3 ? ytox[t] : ytob[t];
    } else if (c.sec == 5 || c.sec == 6) {
        const std::size_t b{static_cast<std::size_t>(d.by0 + c.y) * geo.bw + d.bx0 + c.x};
        const int side{acs[b]};
        if (side == 0) {
            return 2;
        }
        v = c.sec == 5 ? (side == 16 ? 4 : (side == 32 ? 5 : 0)) : qf[b] - 1;
    }
    hybrid420(pack_signed(v), sym, nb, bits);
    return 1;
}

constexpr int kDcChunk = 1024;               // items per workgroup in the DC passes
constexpr int kDcPerThread = kDcChunk / 256;  // consecutive items per thread

// DC and AcMetadata token histograms per group, plus the first-block count.
__device__ __attribute__((always_inline)) inline void dc_hist_kernel_body(const Wg& wg, Geometry geo, const std::int32_t* __restrict__ dc,
                                                      const std::int8_t* __restrict__ acs,
                                                      const std::int32_t* __restrict__ qf,
                                                      const std::int8_t* __restrict__ ytox,
                                                      const std::int8_t* __restrict__ ytob,
                                                      std::uint32_t* __restrict__ dc_hist,
                                                      std::uint32_t* __restrict__ meta_hist,
                                                      std::uint32_t* __restrict__ fb_count) {
    __shared__ std::uint32_t h[2 * kHistStride];
    __shared__ std::uint32_t cnt;
    const int g{static_cast<int>(wg.y)};
    const DcGroupGeo d{dc_group_geo(geo, g)};
    for (int i{static_cast<int>(threadIdx.x)}; i < 2 * kHistStride; i += 256) {
        h[i] = 0;
    }
    if (threadIdx.x == 0) {
        cnt = 0;
    }
    __syncthreads();
    // Lane-strided items (neighbouring lanes read neighbouring blocks); the
    // histogram does not depend on the order.
    const std::uint32_t k0{wg.x * kDcChunk};
    const std::uint32_t k1{min(k0 + static_cast<std::uint32_t>(kDcChunk), d.items)};
    std::uint32_t fb{0};
    for (std::uint32_t k{k0 + threadIdx.x}; k < k1; k += 256) {
        const DcCursor c{dc_cursor_at(d, k)};
        if (c.sec == 0 || c.sec == 2) {
            continue;
        }
        std::uint32_t sym, nb, bits;
        const int kind{dc_cursor_token(geo, d, c, dc, acs, qf, ytox, ytob, sym, nb, bits)};
        if (kind != 2) {
            atomicAdd(&h[(kind == 0 ? 0 : kHistStride) + static_cast<int>(sym)], 1u);
            fb += (kind == 1 && c.sec == 5) ? 1u : 0u;
        }
    }
    if (fb != 0) {
        atomicAdd(&cnt, fb);
    }
    __syncthreads();
    for (int i{static_cast<int>(threadIdx.x)}; i < kHistStride; i += 256) {
        if (h[i]) {
            atomicAdd(&dc_hist[g * kHistStride + i], h[i]);
        }
        if (h[kHistStride + i]) {
            atomicAdd(&meta_hist[g * kHistStride + i], h[kHistStride + i]);
        }
    }
    if (threadIdx.x == 0 && cnt) {
        atomicAdd(&fb_count[g], cnt);
    }
}

__global__ __launch_bounds__(256) void dc_hist_kernel(Geometry geo, const std::int32_t* __restrict__ dc,
                                                      const std::int8_t* __restrict__ acs,
                                                      const std::int32_t* __restrict__ qf,
                                                      const std::int8_t* __restrict__ ytox,
                                                      const std::int8_t* __restrict__ ytob,
                                                      std::uint32_t* __restrict__ dc_hist,
                                                      std::uint32_t* __restrict__ meta_hist,
                                                      std::uint32_t* __restrict__ fb_count) {
    dc_hist_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, dc, acs, qf, ytox, ytob, dc_hist, meta_hist, fb_count);
}

struct DcBlobs {
    // Per group: blob bit lengths and word offsets into `words`.
    std::uint32_t pre_bits[64], mid_bits[64];
    std::uint32_t pre_off[64], mid_off[64];
    std::uint32_t overflow;  // a header blob exceeded its slot
};

// Code length and code of the cursor's item (0 bits for an absent item).
__device__ __forceinline__ std::uint32_t dc_cursor_code(const Geometry& geo, const DcGroupGeo& d,
                                                        const DcCursor& c, int g, const std::int32_t* dc,
                                                        const std::int8_t* acs, const std::int32_t* qf,
                                                        const std::int8_t* ytox, const std::int8_t* ytob,
                                                        const std::uint8_t* dc_depth, const std::uint16_t* dc_codes,
                                                        const std::uint8_t* meta_depth,
                                                        const std::uint16_t* meta_codes, std::uint32_t& val) {
    std::uint32_t sym, nb, bits;
    const int kind{dc_cursor_token(geo, d, c, dc, acs, qf, ytox, ytob, sym, nb, bits)};
    if (kind == 2) {
        val = 0;
        return 0;
    }
    const std::uint32_t idx{static_cast<std::uint32_t>(g) * kHistStride + sym};
    const std::uint32_t dd{(kind == 0 ? dc_depth : meta_depth)[idx]};
    val = codes_or_zero(kind == 0 ? dc_codes : meta_codes, idx) | (bits << dd);
    return dd + nb;
}

// Chunk bit totals of every DC group (grid: chunks x groups).
// Bits of one chunk of a DC group's items, the two header blobs not
// counted (the scan adds them: their sizes are built beside this pass).
__device__ void dc_chunk_bits_body(
    const Wg& wg, const Geometry& geo, const std::int32_t* __restrict__ dc, const std::int8_t* __restrict__ acs,
    const std::int32_t* __restrict__ qf, const std::int8_t* __restrict__ ytox,
    const std::int8_t* __restrict__ ytob, const std::uint8_t* __restrict__ dc_depth,
    const std::uint8_t* __restrict__ meta_depth, std::uint64_t* __restrict__ chunk_bits,
    std::uint32_t chunks_per_group) {
    __shared__ std::uint64_t sh[5];
    const int g{static_cast<int>(wg.y)};
    const DcGroupGeo d{dc_group_geo(geo, g)};
    const std::uint32_t k0{wg.x * kDcChunk + threadIdx.x * kDcPerThread};
    const std::uint32_t k1{min(k0 + kDcPerThread, d.items)};
    std::uint64_t run{0};
    if (k0 < k1) {
        DcCursor c{dc_cursor_at(d, k0)};
        for (std::uint32_t k{k0}; k < k1; ++k, dc_cursor_next(d, c)) {
            if (c.sec == 0 || c.sec == 2) {
                continue;
            }
            std::uint32_t val;
            run += dc_cursor_code(geo, d, c, g, dc, acs, qf, ytox, ytob, dc_depth, nullptr, meta_depth, nullptr,
                                  val);
        }
    }
    std::uint64_t total{0};
    block_excl_scan(run, sh, total);
    if (threadIdx.x == 0) {
        chunk_bits[static_cast<std::size_t>(g) * chunks_per_group + wg.x] = total;
    }
}

// Exclusive scan of chunk totals per group (one workgroup per group).
// Also clears the words of the group's staging slot that its emission will
// OR into.
__device__ __attribute__((always_inline)) inline void dc_chunk_scan_kernel_body(const Wg& wg, Geometry geo, std::uint64_t* __restrict__ chunk_bits,
                                                             std::uint32_t chunks_per_group,
                                                             const DcBlobs* __restrict__ blobs,
                                                             std::uint32_t* __restrict__ group_bytes,
                                                             std::uint64_t* __restrict__ group_bits,
                                                             std::uint32_t* __restrict__ staging,
                                                             std::uint64_t slot_words) {
    __shared__ std::uint64_t sh[17];
    const int g{static_cast<int>(wg.x)};
    const DcGroupGeo d{dc_group_geo(geo, g)};
    const std::uint32_t used{(d.items + kDcChunk - 1) / kDcChunk};
    const std::uint32_t t{threadIdx.x};
    std::uint64_t v{0};
    if (t < used) {
        v = chunk_bits[static_cast<std::size_t>(g) * chunks_per_group + t];
        // The blobs are items 0 and 1 + 3 n of the group.
        v += t == 0 ? blobs->pre_bits[g] : 0u;
        v += t == (1u + 3u * static_cast<std::uint32_t>(d.n)) / kDcChunk ? blobs->mid_bits[g] : 0u;
    }
    std::uint64_t total{0};
    const std::uint64_t ex{block_excl_scan(v, sh, total)};
    if (t < used) {
        chunk_bits[static_cast<std::size_t>(g) * chunks_per_group + t] = ex;
    }
    if (t == 0) {
        group_bytes[g] = static_cast<std::uint32_t>((total + 7) / 8);
        group_bits[g] = total;
    }
    std::uint32_t* const slot{staging + static_cast<std::size_t>(g) * slot_words};
    const std::uint64_t nw{(total + 31) / 32 < slot_words ? (total + 31) / 32 : slot_words};
    for (std::uint64_t i{t}; i < nw; i += 1024) {
        slot[i] = 0;
    }
}

__global__ __launch_bounds__(1024) void dc_chunk_scan_kernel(Geometry geo, std::uint64_t* __restrict__ chunk_bits,
                                                             std::uint32_t chunks_per_group,
                                                             const DcBlobs* __restrict__ blobs,
                                                             std::uint32_t* __restrict__ group_bytes,
                                                             std::uint64_t* __restrict__ group_bits,
                                                             std::uint32_t* __restrict__ staging,
                                                             std::uint64_t slot_words) {
    dc_chunk_scan_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, chunk_bits, chunks_per_group, blobs, group_bytes, group_bits, staging, slot_words);
}

// Emission of one chunk of a DC group into the group's staging slot, at
// group-relative bit positions: the chunk's items occupy one contiguous bit
// range, assembled in LDS (interior words plain-stored, the two boundary
// words ORed into the cleared slot); a range longer than the LDS copy is
// emitted with global atomics instead.
constexpr int kDcEmitWords = 2048;

__device__ __attribute__((always_inline)) inline void dc_emit_kernel_body(const Wg& wg, 
    Geometry geo, const std::int32_t* __restrict__ dc, const std::int8_t* __restrict__ acs,
    const std::int32_t* __restrict__ qf, const std::int8_t* __restrict__ ytox,
    const std::int8_t* __restrict__ ytob, const std::uint8_t* __restrict__ dc_depth,
    const std::uint16_t* __restrict__ dc_codes, const std::uint8_t* __restrict__ meta_depth,
    const std::uint16_t* __restrict__ meta_codes, const DcBlobs* __restrict__ blobs,
    const std::uint32_t* __restrict__ blob_words, const std::uint64_t* __restrict__ chunk_off,
    std::uint32_t chunks_per_group, std::uint32_t* __restrict__ staging, std::uint64_t slot_words) {
    __shared__ std::uint64_t sh[5];
    __shared__ std::uint32_t buf[kDcEmitWords];
    const int g{static_cast<int>(wg.y)};
    const int tid{static_cast<int>(threadIdx.x)};
    const DcGroupGeo d{dc_group_geo(geo, g)};
    if (wg.x * kDcChunk >= d.items) {
        return;
    }
    const std::uint32_t k0{wg.x * kDcChunk + threadIdx.x * kDcPerThread};
    const std::uint32_t k1{min(k0 + kDcPerThread, d.items)};
    std::uint64_t run{0};
    if (k0 < k1) {
        DcCursor c{dc_cursor_at(d, k0)};
        for (std::uint32_t k{k0}; k < k1; ++k, dc_cursor_next(d, c)) {
            if (c.sec == 0 || c.sec == 2) {
                run += c.sec == 0 ? blobs->pre_bits[g] : blobs->mid_bits[g];
                continue;
            }
            std::uint32_t val;
            run += dc_cursor_code(geo, d, c, g, dc, acs, qf, ytox, ytob, dc_depth, nullptr, meta_depth, nullptr,
                                  val);
        }
    }
    std::uint64_t total{0};
    const std::uint64_t rel0{block_excl_scan(run, sh, total)};
    std::uint32_t* const out_words{staging + static_cast<std::size_t>(g) * slot_words};
    const std::uint64_t cap_words{slot_words};
    const std::uint64_t start{chunk_off[static_cast<std::size_t>(g) * chunks_per_group + wg.x]};
    const std::uint64_t w0{start >> 5};
    const std::uint64_t nw{((start + total + 31) >> 5) - w0};
    const bool staged{nw <= static_cast<std::uint64_t>(kDcEmitWords)};
    if (staged) {
        for (int i{tid}; i < static_cast<int>(nw); i += 256) {
            buf[i] = 0;
        }
    }
    __syncthreads();
    const std::uint64_t lbase{w0 << 5};
    auto put = [&](std::uint64_t at, std::uint32_t v, std::uint32_t len) {
        if (staged) {
            lds_or_bits(buf, kDcEmitWords, static_cast<std::uint32_t>(at - lbase), v, len);
        } else {
            or_bits(out_words, cap_words, at, v, len);
        }
    };
    std::uint64_t pos{start + rel0};
    if (k0 < k1) {
        DcCursor c{dc_cursor_at(d, k0)};
        for (std::uint32_t k{k0}; k < k1; ++k, dc_cursor_next(d, c)) {
            if (c.sec == 0 || c.sec == 2) {
                const std::uint32_t nbits{c.sec == 0 ? blobs->pre_bits[g] : blobs->mid_bits[g]};
                const std::uint32_t* src{blob_words + (c.sec == 0 ? blobs->pre_off[g] : blobs->mid_off[g])};
                for (std::uint32_t i{0}; i < nbits; i += 32) {
                    put(pos + i, src[i >> 5], nbits - i < 32 ? nbits - i : 32u);
                }
                pos += nbits;
                continue;
            }
            std::uint32_t val;
            const std::uint32_t len{dc_cursor_code(geo, d, c, g, dc, acs, qf, ytox, ytob, dc_depth, dc_codes,
                                                   meta_depth, meta_codes, val)};
            put(pos, val, len);
            pos += len;
        }
    }
    if (!staged) {
        return;
    }
    __syncthreads();
    for (int i{tid}; i < static_cast<int>(nw); i += 256) {
        const std::uint64_t gi{w0 + static_cast<std::uint64_t>(i)};
        if (gi >= cap_words) {
            continue;
        }
        const std::uint32_t v{buf[i]};
        if (i == 0 || i == static_cast<int>(nw) - 1) {
            if (v != 0) {
                atomicOr(&out_words[gi], v);
            }
        } else {
            out_words[gi] = v;
        }
    }
}

__global__ __launch_bounds__(256) void dc_emit_kernel(
    Geometry geo, const std::int32_t* __restrict__ dc, const std::int8_t* __restrict__ acs,
    const std::int32_t* __restrict__ qf, const std::int8_t* __restrict__ ytox,
    const std::int8_t* __restrict__ ytob, const std::uint8_t* __restrict__ dc_depth,
    const std::uint16_t* __restrict__ dc_codes, const std::uint8_t* __restrict__ meta_depth,
    const std::uint16_t* __restrict__ meta_codes, const DcBlobs* __restrict__ blobs,
    const std::uint32_t* __restrict__ blob_words, const std::uint64_t* __restrict__ chunk_off,
    std::uint32_t chunks_per_group, std::uint32_t* __restrict__ staging, std::uint64_t slot_words) {
    dc_emit_kernel_body(Wg{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}, geo, dc, acs, qf, ytox, ytob, dc_depth, dc_codes, meta_depth, meta_codes, blobs, blob_words, chunk_off, chunks_per_group, staging, slot_words);
}

}  // namespace cjx
