// Length-limited prefix codes built by one wave (included into device.hip).
//
// Same results as the serial builder in prefix.h (prefix_code.cc:36-397):
// leaves in descending symbol order, stable sort by count, two-queue merge,
// count floor doubled until the depth limit holds, canonical LSB-first codes,
// simple / code-length-code headers.  The parallel parts:
//   - leaf positions from ballots, the stable sort as a rank computation over
//     (count, leaf position);
//   - the two-queue merge and the top-down depth walk stay sequential, but
//     their queues live in VGPRs (one value per lane, read with readlane and
//     written with a lane select at a wave-uniform index), so no step waits on
//     memory;
//   - canonical codes from per-depth ballot ranks;
//   - the code-length-coded header as a wave scan of per-symbol bit lengths.
// Every lane runs the same control flow; loop bounds and queue indices are
// wave-uniform.

namespace cjx {

constexpr std::uint32_t kQueueEnd = 0xffffffffu;

// Orders one wave's LDS accesses (its LDS instructions execute in order; this
// keeps the compiler from moving them across the point).

// R values per lane: element idx lives in register idx / 64 of lane idx % 64.
template <int R>
struct LaneArray {
    std::uint32_t v[R];

    __device__ __forceinline__ std::uint32_t get(int idx) const {
        const int l{idx & 63};
        if constexpr (R == 1) {
            return static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v[0]), l));
        } else {
            const int r{idx >> 6};
            const int x{r == 0   ? __builtin_amdgcn_readlane(static_cast<int>(v[0]), l)
                        : r == 1 ? __builtin_amdgcn_readlane(static_cast<int>(v[1]), l)
                        : r == 2 ? __builtin_amdgcn_readlane(static_cast<int>(v[2]), l)
                                 : __builtin_amdgcn_readlane(static_cast<int>(v[3]), l)};
            return static_cast<std::uint32_t>(x);
        }
    }

    // Lane idx % 64 of register idx / 64 takes x (a compare and a select).
    __device__ __forceinline__ void set(int idx, std::uint32_t x) {
        const bool mine{(static_cast<int>(threadIdx.x) & 63) == (idx & 63)};
        if constexpr (R == 1) {
            v[0] = mine ? x : v[0];
        } else {
            const int r{idx >> 6};
#pragma unroll
            for (int k{0}; k < R; ++k) {
                v[k] = (mine && r == k) ? x : v[k];
            }
        }
    }
};

__device__ __forceinline__ std::uint32_t wave_max_u32(std::uint32_t v) {
#pragma unroll
    for (int off{32}; off > 0; off >>= 1) {
        const std::uint32_t o{static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), off, 64))};
        v = o > v ? o : v;
    }
    return v;
}

// LDS scratch of one wave's build.
struct WaveHuffScratch {
    std::uint32_t leaf_total[256];  // input order
    std::uint32_t sorted_total[256];
    std::uint32_t sorted_symbol[256];
};

// Value of x in lane `src` (wave-uniform) of the wave.
__device__ __forceinline__ std::uint32_t lane_value(std::uint32_t x, int src) {
    return static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), src));
}

// The merge and depths below for n <= 64 leaves.  Each queue is one VGPR
// (lane = queue index) and a merge step has no branch: the two picks are
// selects on wave-uniform values.  The depths come from the parent links by
// pointer jumping: every node holds (distance d, ancestor p), initially
// (1, parent) and (0, root) at the root, and six rounds of d += d[p],
// p = p[p] leave each node's depth in d (depths are below 64).  Ancestors
// are internal nodes, so every lookup reads the internal nodes' registers.
__device__ std::uint32_t merge_depths_small(WaveHuffScratch& s, int n, std::uint8_t* depth) {
    const int lane{static_cast<int>(threadIdx.x) & 63};
    const std::uint32_t leaf{lane < n ? s.sorted_total[lane] : kQueueEnd};
    const std::uint32_t sym{lane < n ? s.sorted_symbol[lane] : 0u};
    // Two-queue merge: leaves (tree index i < n) and internal nodes
    // (internal k = tree index n + 1 + k), smaller total first, leaf on ties;
    // internal nodes not yet made read as kQueueEnd.
    std::uint32_t inner{kQueueEnd}, kid_l{0}, kid_r{0};
    int i{0}, j{0};
    for (int k{0}; k < n - 1; ++k) {
        const std::uint32_t a0{i < n ? lane_value(leaf, i) : kQueueEnd};
        const std::uint32_t b0{lane_value(inner, j)};
        const bool leaf0{a0 <= b0};
        const std::uint32_t t0{leaf0 ? a0 : b0};
        const int c0{leaf0 ? i : n + 1 + j};
        i += leaf0 ? 1 : 0;
        j += leaf0 ? 0 : 1;
        const std::uint32_t a1{i < n ? lane_value(leaf, i) : kQueueEnd};
        const std::uint32_t b1{lane_value(inner, j)};
        const bool leaf1{a1 <= b1};
        const std::uint32_t t1{leaf1 ? a1 : b1};
        const int c1{leaf1 ? i : n + 1 + j};
        i += leaf1 ? 1 : 0;
        j += leaf1 ? 0 : 1;
        if (lane == k) {
            inner = t0 + t1;
            kid_l = static_cast<std::uint32_t>(c0);
            kid_r = static_cast<std::uint32_t>(c1);
        }
    }
    // Parent links in LDS (leaf_total is dead after the sort): node ids are
    // leaf i -> i, internal k -> 64 + k; the root (internal n - 2) is its own
    // parent.
    std::uint32_t* const par{s.leaf_total};
    const std::uint32_t root{64u + static_cast<std::uint32_t>(n - 2)};
    if (lane < n - 1) {
        const std::uint32_t me{64u + static_cast<std::uint32_t>(lane)};
        const int nn{n};
        par[kid_l < static_cast<std::uint32_t>(nn) ? kid_l : kid_l - static_cast<std::uint32_t>(nn) - 1u + 64u] = me;
        par[kid_r < static_cast<std::uint32_t>(nn) ? kid_r : kid_r - static_cast<std::uint32_t>(nn) - 1u + 64u] = me;
    }
    if (lane == 0) {
        par[root] = root;
    }
    wave_sync();
    std::uint32_t pa{lane < n ? par[lane] : root};
    std::uint32_t pb{lane < n - 1 ? par[64 + lane] : root};
    std::uint32_t da{lane < n ? 1u : 0u};
    std::uint32_t db{lane < n - 2 ? 1u : 0u};
    wave_sync();
#pragma unroll
    for (int r{0}; r < 6; ++r) {
        const int ia{static_cast<int>(pa - 64u)}, ib{static_cast<int>(pb - 64u)};
        const std::uint32_t d_a{static_cast<std::uint32_t>(__shfl(static_cast<int>(db), ia, 64))};
        const std::uint32_t p_a{static_cast<std::uint32_t>(__shfl(static_cast<int>(pb), ia, 64))};
        const std::uint32_t d_b{static_cast<std::uint32_t>(__shfl(static_cast<int>(db), ib, 64))};
        const std::uint32_t p_b{static_cast<std::uint32_t>(__shfl(static_cast<int>(pb), ib, 64))};
        da += d_a;
        pa = p_a;
        db += d_b;
        pb = p_b;
    }
    if (lane < n) {
        depth[sym] = static_cast<std::uint8_t>(da);
    }
    return lane < n ? da : 0u;
}

// Two-queue merge and top-down depths of the n (>= 2) sorted leaves in
// s.sorted_total / s.sorted_symbol; writes depth[symbol] and returns this
// lane's largest leaf depth.  Q * 64 >= n.
template <int Q>
__device__ std::uint32_t merge_depths(WaveHuffScratch& s, int n, std::uint8_t* depth) {
    if constexpr (Q == 1) {
        return merge_depths_small(s, n, depth);
    }
    const int lane{static_cast<int>(threadIdx.x) & 63};
    LaneArray<Q> leaf{}, sym{};
#pragma unroll
    for (int r{0}; r < Q; ++r) {
        const int t{64 * r + lane};
        leaf.v[r] = t < n ? s.sorted_total[t] : kQueueEnd;
        sym.v[r] = t < n ? s.sorted_symbol[t] : 0u;
    }
    // Two-queue merge: leaves (tree index i < n) and internal nodes
    // (internal k = tree index n + 1 + k), smaller total first, leaf on ties.
    LaneArray<Q> inner{}, kid_l{}, kid_r{};
    int i{0}, j{0};
    for (int k{0}; k < n - 1; ++k) {
        std::uint32_t a{i < n ? leaf.get(i) : kQueueEnd};
        std::uint32_t b{j < k ? inner.get(j) : kQueueEnd};
        int left, right;
        std::uint32_t lt, rt;
        if (a <= b) {
            left = i;
            lt = a;
            ++i;
            a = i < n ? leaf.get(i) : kQueueEnd;
        } else {
            left = n + 1 + j;
            lt = b;
            ++j;
            b = j < k ? inner.get(j) : kQueueEnd;
        }
        if (a <= b) {
            right = i;
            rt = a;
            ++i;
        } else {
            right = n + 1 + j;
            rt = b;
            ++j;
        }
        inner.set(k, lt + rt);
        kid_l.set(k, static_cast<std::uint32_t>(left));
        kid_r.set(k, static_cast<std::uint32_t>(right));
    }
    // Depths top-down from the root (internal n - 2): every child was
    // created before its parent.
    LaneArray<Q> node_depth{}, leaf_depth{};
    node_depth.set(n - 2, 0u);
    for (int k{n - 2}; k >= 0; --k) {
        const std::uint32_t d{node_depth.get(k) + 1u};
        const int cl{static_cast<int>(kid_l.get(k))}, cr{static_cast<int>(kid_r.get(k))};
        if (cl > n) {
            node_depth.set(cl - n - 1, d);
        } else {
            leaf_depth.set(cl, d);
        }
        if (cr > n) {
            node_depth.set(cr - n - 1, d);
        } else {
            leaf_depth.set(cr, d);
        }
    }
    std::uint32_t maxd{0};
#pragma unroll
    for (int r{0}; r < Q; ++r) {
        if (64 * r + lane < n) {
            depth[sym.v[r]] = static_cast<std::uint8_t>(leaf_depth.v[r]);
            maxd = leaf_depth.v[r] > maxd ? leaf_depth.v[r] : maxd;
        }
    }
    return maxd;
}

// Leaves of data[0..length) in create_huffman_tree's input order (descending
// symbol): per lane, the counts of symbols 64 r + lane, whether present, and
// the input position of each present one.
template <int R>
struct Leaves {
    std::uint32_t cnt[R];
    int pos[R];
    int n;
};

template <int R>
__device__ __forceinline__ Leaves<R> wave_leaves(const std::uint32_t* data, int length) {
    const int lane{static_cast<int>(threadIdx.x) & 63};
    Leaves<R> L{};
    std::uint64_t present[R];
#pragma unroll
    for (int r{0}; r < R; ++r) {
        const int sym{64 * r + lane};
        L.cnt[r] = sym < length ? data[sym] : 0u;
        present[r] = __ballot(L.cnt[r] != 0);
        L.n += __popcll(present[r]);
    }
    const std::uint64_t above{lane == 63 ? 0ull : (~0ull << (lane + 1))};
    int higher{0};
#pragma unroll
    for (int r{R - 1}; r >= 0; --r) {
        L.pos[r] = higher + __popcll(present[r] & above);
        higher += __popcll(present[r]);
    }
    return L;
}

// One pass of create_huffman_tree's loop (prefix_code.cc:36-96) with count
// floor `count_limit` for n >= 2 leaves: writes depth[symbol] of every present
// symbol and returns the largest depth (wave-uniform).
template <int R>
__device__ std::uint32_t tree_trial(const Leaves<R>& L, std::uint32_t count_limit, std::uint8_t* depth,
                                    WaveHuffScratch& s) {
    const int lane{static_cast<int>(threadIdx.x) & 63};
    const int n{L.n};
    std::uint32_t tot[R];
#pragma unroll
    for (int r{0}; r < R; ++r) {
        tot[r] = L.cnt[r] > count_limit - 1 ? L.cnt[r] : count_limit - 1;
        if (L.cnt[r] != 0) {
            s.leaf_total[L.pos[r]] = tot[r];
        }
    }
    wave_sync();
    // Stable ascending sort: rank = smaller totals + equal totals earlier.
    int rank[R];
#pragma unroll
    for (int r{0}; r < R; ++r) {
        rank[r] = 0;
    }
    for (int b{0}; b < n; ++b) {
        const std::uint32_t t{s.leaf_total[b]};
#pragma unroll
        for (int r{0}; r < R; ++r) {
            rank[r] += (t < tot[r] || (t == tot[r] && b < L.pos[r])) ? 1 : 0;
        }
    }
#pragma unroll
    for (int r{0}; r < R; ++r) {
        if (L.cnt[r] != 0) {
            s.sorted_total[rank[r]] = tot[r];
            s.sorted_symbol[rank[r]] = static_cast<std::uint32_t>(64 * r + lane);
        }
    }
    wave_sync();
    const std::uint32_t maxd{n <= 64 ? merge_depths<1>(s, n, depth) : merge_depths<4>(s, n, depth)};
    wave_sync();
    return wave_max_u32(maxd);
}

// create_huffman_tree (prefix_code.cc:36-96) by one wave, floors tried in
// order.  data[0..length) and depth[0..length) are in LDS; depth is zero on
// entry for absent symbols.  R * 64 >= length.
template <int R>
__device__ void wave_create_tree(const std::uint32_t* data, int length, int limit, std::uint8_t* depth,
                                 WaveHuffScratch& s) {
    const int lane{static_cast<int>(threadIdx.x) & 63};
    const Leaves<R> L{wave_leaves<R>(data, length)};
    if (L.n == 0) {
        return;
    }
    if (L.n == 1) {
#pragma unroll
        for (int r{0}; r < R; ++r) {
            if (L.cnt[r] != 0) {
                depth[64 * r + lane] = 1;
            }
        }
        return;
    }
    for (std::uint32_t count_limit{1};; count_limit *= 2) {
        if (static_cast<int>(tree_trial<R>(L, count_limit, depth, s)) <= limit) {
            return;
        }
    }
}

// Canonical codes (prefix_code.cc:111-130): bits[s] for depth[s] > 0.
template <int R>
__device__ void wave_depths_to_codes(const std::uint8_t* depth, int len, std::uint16_t* bits) {
    const int lane{static_cast<int>(threadIdx.x) & 63};
    const std::uint64_t below{(1ull << lane) - 1ull};
    int dep[R];
#pragma unroll
    for (int r{0}; r < R; ++r) {
        const int sym{64 * r + lane};
        dep[r] = sym < len ? depth[sym] : 0;
    }
    std::uint32_t next[16];
    {
        std::uint32_t code{0}, prev{0};
        next[0] = 0;
        for (int d{1}; d < 16; ++d) {
            code = (code + prev) << 1;
            next[d] = code;
            std::uint32_t c{0};
#pragma unroll
            for (int r{0}; r < R; ++r) {
                c += static_cast<std::uint32_t>(__popcll(__ballot(dep[r] == d)));
            }
            prev = c;
        }
    }
    for (int d{1}; d < 16; ++d) {
        std::uint32_t base{next[d]};
#pragma unroll
        for (int r{0}; r < R; ++r) {
            const std::uint64_t m{__ballot(dep[r] == d)};
            if (dep[r] == d) {
                const std::uint32_t c{base + static_cast<std::uint32_t>(__popcll(m & below))};
                bits[64 * r + lane] = static_cast<std::uint16_t>(__builtin_bitreverse32(c) >> (32 - d));
            }
            base += static_cast<std::uint32_t>(__popcll(m));
        }
    }
}

struct WaveCodeScratch {
    WaveHuffScratch tree;
    std::uint32_t cl_hist[32];
    std::uint8_t cl_depth[32];
    std::uint16_t cl_bits[32];
    std::uint32_t pos;
};

// store_huffman_tree (prefix_code.cc:211-354) for depth[0..num): returns the
// header length in bits written into hdr (cap words, zeroed).
template <int R>
__device__ std::uint32_t wave_store_tree(const std::uint8_t* depth, int num, std::uint32_t* hdr,
                                         std::uint32_t cap, WaveCodeScratch& s) {
    const int lane{static_cast<int>(threadIdx.x) & 63};
    int dep[R];
    std::uint32_t last{0};
#pragma unroll
    for (int r{0}; r < R; ++r) {
        const int sym{64 * r + lane};
        dep[r] = sym < num ? depth[sym] : 0;
        if (dep[r] != 0) {
            last = static_cast<std::uint32_t>(sym) + 1u;
        }
    }
    const int new_len{static_cast<int>(wave_max_u32(last))};
    // Histogram of depth values over [0, new_len).
    for (int v{0}; v < 18; ++v) {
        std::uint32_t c{0};
#pragma unroll
        for (int r{0}; r < R; ++r) {
            c += static_cast<std::uint32_t>(__popcll(__ballot(64 * r + lane < new_len && dep[r] == v)));
        }
        if (lane == 0) {
            s.cl_hist[v] = c;
        }
    }
    if (lane < 18) {
        s.cl_depth[lane] = 0;
        s.cl_bits[lane] = 0;
    }
    wave_sync();
    int num_codes{0}, code{0};
    for (int i{0}; i < 18; ++i) {
        if (s.cl_hist[i]) {
            if (num_codes == 0) {
                code = i;
                num_codes = 1;
            } else if (num_codes == 1) {
                num_codes = 2;
                break;
            }
        }
    }
    wave_create_tree<1>(s.cl_hist, 18, 5, s.cl_depth, s.tree);
    wave_sync();
    wave_depths_to_codes<1>(s.cl_depth, 18, s.cl_bits);
    wave_sync();
    {
        // The code-length code's depths in the format's order (1, 2, 3, 4,
        // 0, 5, 17, 6, 16, 7, ..., 15), lane i taking entry i: the stored run
        // [skip, to_store) goes out as one prefix scan of the entries'
        // lengths (the fixed code of depths 0..5: lengths 2,4,3,2,2,4,
        // values 0,7,3,2,1,15, packed in nibbles).
        const int ord{lane < 4    ? lane + 1
                      : lane == 4 ? 0
                      : lane == 5 ? 5
                      : lane == 6 ? 17
                      : lane == 7 ? 6
                      : lane == 8 ? 16
                                  : lane - 2};
        const std::uint32_t l{lane < 18 ? static_cast<std::uint32_t>(s.cl_depth[ord]) : 0u};
        const std::uint64_t nonzero{__ballot(lane < 18 && l != 0)};
        const int to_store{num_codes > 1 ? (nonzero != 0 ? 64 - static_cast<int>(__builtin_clzll(nonzero)) : 0) : 18};
        const int skip{(nonzero & 3ull) != 0 ? 0 : ((nonzero & 4ull) != 0 ? 2 : 3)};
        const bool stored{lane >= skip && lane < to_store};
        const std::uint32_t nb{stored ? (0x422342u >> (4 * l)) & 15u : 0u};
        const std::uint32_t val{stored ? (0xf12370u >> (4 * l)) & 15u : 0u};
        const std::uint32_t pre{static_cast<std::uint32_t>(wave_excl_scan(static_cast<int>(nb), lane))};
        const std::uint32_t total{2u + static_cast<std::uint32_t>(wave_sum_int(static_cast<int>(nb)))};
        wave_sync();  // every lane has read cl_depth
        if (total <= cap * 32) {
            if (lane == 0) {
                lds_or_bits(hdr, cap, 0, static_cast<std::uint32_t>(skip), 2);
            }
            lds_or_bits(hdr, cap, 2 + pre, val, nb);
        }
        if (lane == 0) {
            if (num_codes == 1) {
                s.cl_depth[code] = 0;
            }
            s.pos = total <= cap * 32 ? total : cap * 32 + 1;
        }
    }
    wave_sync();
    std::uint32_t at{s.pos};
#pragma unroll
    for (int r{0}; r < R; ++r) {
        const bool in{64 * r + lane < new_len};
        const std::uint32_t l{in ? static_cast<std::uint32_t>(s.cl_depth[dep[r]]) : 0u};
        const std::uint32_t v{in ? static_cast<std::uint32_t>(s.cl_bits[dep[r]]) : 0u};
        const std::uint32_t pre{static_cast<std::uint32_t>(wave_excl_scan(static_cast<int>(l), lane))};
        if (at + pre + l <= cap * 32) {
            lds_or_bits(hdr, cap, at + pre, v, l);
        }
        at += static_cast<std::uint32_t>(wave_sum_int(static_cast<int>(l)));
    }
    return at;
}

// build_and_store_huffman_tree (prefix_code.cc:358-397) for hist[0..alpha),
// alpha >= 2, with depth[] already holding create_huffman_tree's result when
// more than one symbol is present (all zero otherwise); bits are zeroed by the
// caller.  One wave.  Returns the header length in bits, or 0 if it would not
// fit `cap` words.
template <int R>
__device__ std::uint32_t wave_finish_code(const std::uint32_t* hist, int alpha, const std::uint8_t* depth,
                                          std::uint16_t* bits, std::uint32_t* hdr, std::uint32_t cap,
                                          WaveCodeScratch& s) {
    const int lane{static_cast<int>(threadIdx.x) & 63};
    std::uint64_t present[R];
    int count{0};
#pragma unroll
    for (int r{0}; r < R; ++r) {
        const int sym{64 * r + lane};
        present[r] = __ballot(sym < alpha && hist[sym] != 0);
        count += __popcll(present[r]);
    }
    // The first four present symbols, ascending.
    int s4[4]{0, 0, 0, 0};
    {
        int got{0};
#pragma unroll
        for (int r{0}; r < R; ++r) {
            std::uint64_t m{present[r]};
            while (m != 0 && got < 4) {
                s4[got++] = 64 * r + static_cast<int>(__builtin_ctzll(m));
                m &= m - 1;
            }
        }
    }
    int max_bits{0};
    for (int c{alpha - 1}; c; c >>= 1) {
        ++max_bits;
    }
    if (count <= 1) {
        std::uint32_t pos{0};
        if (lane == 0) {
            BitSink w{hdr, 0, cap * 32};
            w.write(4, 1);
            w.write(static_cast<std::uint32_t>(max_bits), static_cast<std::uint32_t>(s4[0]));
            pos = w.ok() ? w.pos : 0u;
        }
        return static_cast<std::uint32_t>(__shfl(static_cast<int>(pos), 0, 64));
    }
    wave_depths_to_codes<R>(depth, alpha, bits);
    wave_sync();
    if (count <= 4) {
        std::uint32_t pos{0};
        if (lane == 0) {
            BitSink w{hdr, 0, cap * 32};
            w.write(2, 1);
            w.write(2, static_cast<std::uint32_t>(count - 1));
            for (int i{0}; i < count; ++i) {
                for (int j{i + 1}; j < count; ++j) {
                    if (depth[s4[j]] < depth[s4[i]]) {
                        const int t{s4[j]};
                        s4[j] = s4[i];
                        s4[i] = t;
                    }
                }
            }
            for (int i{0}; i < count; ++i) {
                w.write(static_cast<std::uint32_t>(max_bits), static_cast<std::uint32_t>(s4[i]));
            }
            if (count == 4) {
                w.write(1, depth[s4[0]] == 1 ? 1u : 0u);
            }
            pos = w.ok() ? w.pos : 0u;
        }
        return static_cast<std::uint32_t>(__shfl(static_cast<int>(pos), 0, 64));
    }
    const std::uint32_t nb{wave_store_tree<R>(depth, alpha, hdr, cap, s)};
    return nb <= cap * 32 ? nb : 0u;
}

}  // namespace cjx
