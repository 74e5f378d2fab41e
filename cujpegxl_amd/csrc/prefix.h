// Length-limited prefix codes and the JPEG XL histogram headers, usable on
// the host (constant blobs built once per configuration) and on the device
// (the per-frame codes, one workgroup per histogram).
//
// Follows the reference's prefix_code.cc:36-397 (libjxl enc_huffman_tree /
// enc_huffman): leaves entered in descending symbol order, stable sort by
// count, two-queue merge, count floor doubled until the depth limit holds,
// canonical LSB-first codes, simple form for <= 4 symbols, otherwise the
// code-length-code form without run-length symbols.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define CJX_HDI __host__ __device__ inline

namespace cjx {

// LSB-first bit sink over a zeroed word buffer (single writer).
struct BitSink {
    std::uint32_t* words;
    std::uint32_t pos;  // bits written
    std::uint32_t cap_bits;

    CJX_HDI void write(std::uint32_t n, std::uint32_t v) {
        if (n == 0) {
            return;
        }
        if (pos + n > cap_bits) {
            pos = cap_bits + 1;  // overflow marker, checked by the caller
            return;
        }
        if (n < 32) {
            v &= (1u << n) - 1u;
        }
        const std::uint32_t w{pos >> 5}, b{pos & 31u};
        words[w] |= v << b;
        if (b + n > 32) {
            words[w + 1] |= v >> (32 - b);
        }
        pos += n;
    }
    CJX_HDI bool ok() const { return pos <= cap_bits; }
};

// Copies `nbits` bits from an LSB-first word buffer into a sink.
template <typename Sink>
CJX_HDI void copy_bits(Sink& w, const std::uint32_t* src, std::uint32_t nbits) {
    std::uint32_t i{0};
    for (; i + 32 <= nbits; i += 32) {
        w.write(32, src[i >> 5]);
    }
    if (i < nbits) {
        w.write(nbits - i, src[i >> 5]);
    }
}

struct HuffNode {
    std::uint32_t total;
    std::int16_t left;
    std::int16_t right_or_value;
};

// Scratch for one build: nodes for up to 256 leaves plus internal nodes,
// sort buffer and node depths.
struct HuffScratch {
    HuffNode tree[2 * 256 + 4];
    HuffNode tmp[256];
    std::uint8_t node_depth[2 * 256 + 4];
};

CJX_HDI void stable_sort_by_total(HuffNode* a, HuffNode* tmp, int n) {
    // Bottom-up merge sort; ties keep their input order.
    for (int width{1}; width < n; width *= 2) {
        for (int lo{0}; lo < n; lo += 2 * width) {
            const int mid{lo + width < n ? lo + width : n};
            const int hi{lo + 2 * width < n ? lo + 2 * width : n};
            int i{lo}, j{mid}, k{lo};
            while (i < mid && j < hi) {
                tmp[k++] = a[j].total < a[i].total ? a[j++] : a[i++];
            }
            while (i < mid) {
                tmp[k++] = a[i++];
            }
            while (j < hi) {
                tmp[k++] = a[j++];
            }
        }
        for (int i{0}; i < n; ++i) {
            a[i] = tmp[i];
        }
    }
}

// create_huffman_tree (prefix_code.cc:36-96) into `depth` (length entries,
// zero-initialized by the caller; entries of absent symbols are left as is).
CJX_HDI void create_huffman_tree(const std::uint32_t* data, int length, int limit,
                                 std::uint8_t* depth, HuffScratch& s) {
    for (std::uint32_t count_limit{1};; count_limit *= 2) {
        int n{0};
        for (int i{length - 1}; i >= 0; --i) {
            if (data[i]) {
                const std::uint32_t c{data[i] > count_limit - 1 ? data[i] : count_limit - 1};
                s.tree[n].total = c;
                s.tree[n].left = -1;
                s.tree[n].right_or_value = static_cast<std::int16_t>(i);
                ++n;
            }
        }
        if (n == 1) {
            depth[s.tree[0].right_or_value] = 1;
            return;
        }
        stable_sort_by_total(s.tree, s.tmp, n);
        const HuffNode sentinel{0xffffffffu, -1, -1};
        int size{n};
        s.tree[size++] = sentinel;
        s.tree[size++] = sentinel;
        int i{0}, j{n + 1};
        for (int k{n - 1}; k != 0; --k) {
            int left, right;
            if (s.tree[i].total <= s.tree[j].total) {
                left = i++;
            } else {
                left = j++;
            }
            if (s.tree[i].total <= s.tree[j].total) {
                right = i++;
            } else {
                right = j++;
            }
            const int j_end{size - 1};
            s.tree[j_end].total = s.tree[left].total + s.tree[right].total;
            s.tree[j_end].left = static_cast<std::int16_t>(left);
            s.tree[j_end].right_or_value = static_cast<std::int16_t>(right);
            s.tree[size++] = sentinel;
        }
        // Depths top-down: every child is created before its parent.
        const int root{2 * n - 1};
        s.node_depth[root] = 0;
        int maxd{0};
        for (int p{root}; p > n; --p) {
            const std::uint8_t d{static_cast<std::uint8_t>(s.node_depth[p] + 1)};
            s.node_depth[s.tree[p].left] = d;
            s.node_depth[s.tree[p].right_or_value] = d;
        }
        for (int t{0}; t < n; ++t) {
            depth[s.tree[t].right_or_value] = s.node_depth[t];
        }
        for (int t{0}; t < length; ++t) {
            maxd = depth[t] > maxd ? depth[t] : maxd;
        }
        if (maxd <= limit) {
            return;
        }
    }
}

CJX_HDI std::uint16_t reverse_bits(int num_bits, std::uint16_t bits) {
    std::uint16_t r{0};
    for (int i{0}; i < num_bits; ++i) {
        r = static_cast<std::uint16_t>((r << 1) | ((bits >> i) & 1u));
    }
    return r;
}

// prefix_code.cc:111-130
CJX_HDI void depths_to_codes(const std::uint8_t* depth, int len, std::uint16_t* bits) {
    std::uint16_t bl_count[16]{};
    std::uint16_t next_code[16]{};
    for (int i{0}; i < len; ++i) {
        ++bl_count[depth[i]];
    }
    bl_count[0] = 0;
    int code{0};
    for (int i{1}; i < 16; ++i) {
        code = (code + bl_count[i - 1]) << 1;
        next_code[i] = static_cast<std::uint16_t>(code);
    }
    for (int i{0}; i < len; ++i) {
        if (depth[i]) {
            bits[i] = reverse_bits(depth[i], next_code[depth[i]]++);
        }
    }
}

// prefix_code.cc:211-354 (no run-length symbols for these alphabets).
template <typename Sink>
CJX_HDI void store_huffman_tree(const std::uint8_t* depths, int num, Sink& w, HuffScratch& s) {
    int new_len{num};
    while (new_len > 0 && depths[new_len - 1] == 0) {
        --new_len;
    }
    std::uint32_t hist[18]{};
    for (int i{0}; i < new_len; ++i) {
        ++hist[depths[i]];
    }
    int num_codes{0}, code{0};
    for (int i{0}; i < 18; ++i) {
        if (hist[i]) {
            if (num_codes == 0) {
                code = i;
                num_codes = 1;
            } else if (num_codes == 1) {
                num_codes = 2;
                break;
            }
        }
    }
    std::uint8_t cl_depth[18]{};
    std::uint16_t cl_bits[18]{};
    create_huffman_tree(hist, 18, 5, cl_depth, s);
    depths_to_codes(cl_depth, 18, cl_bits);
    constexpr std::uint8_t order[18]{1, 2, 3, 4, 0, 5, 17, 6, 16, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    constexpr std::uint8_t sym[6]{0, 7, 3, 2, 1, 15};
    constexpr std::uint8_t len[6]{2, 4, 3, 2, 2, 4};
    int to_store{18};
    if (num_codes > 1) {
        for (; to_store > 0; --to_store) {
            if (cl_depth[order[to_store - 1]] != 0) {
                break;
            }
        }
    }
    int skip{0};
    if (cl_depth[order[0]] == 0 && cl_depth[order[1]] == 0) {
        skip = 2;
        if (cl_depth[order[2]] == 0) {
            skip = 3;
        }
    }
    w.write(2, static_cast<std::uint32_t>(skip));
    for (int i{skip}; i < to_store; ++i) {
        const std::uint8_t l{cl_depth[order[i]]};
        w.write(len[l], sym[l]);
    }
    if (num_codes == 1) {
        cl_depth[code] = 0;
    }
    for (int i{0}; i < new_len; ++i) {
        w.write(cl_depth[depths[i]], cl_bits[depths[i]]);
    }
}

// build_and_store_huffman_tree (prefix_code.cc:358-397). depth/bits must hold
// `length` entries; they are cleared here.
template <typename Sink>
CJX_HDI void build_and_store_huffman(const std::uint32_t* hist, int length, std::uint8_t* depth,
                                     std::uint16_t* bits, Sink& w, HuffScratch& s) {
    for (int i{0}; i < length; ++i) {
        depth[i] = 0;
        bits[i] = 0;
    }
    int count{0};
    int s4[4]{0, 0, 0, 0};
    for (int i{0}; i < length; ++i) {
        if (hist[i]) {
            if (count < 4) {
                s4[count] = i;
            } else if (count > 4) {
                break;
            }
            ++count;
        }
    }
    int max_bits{0};
    for (int c{length - 1}; c; c >>= 1) {
        ++max_bits;
    }
    if (count <= 1) {
        w.write(4, 1);
        w.write(static_cast<std::uint32_t>(max_bits), static_cast<std::uint32_t>(s4[0]));
        return;
    }
    create_huffman_tree(hist, length, 15, depth, s);
    depths_to_codes(depth, length, bits);
    if (count <= 4) {
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
    } else {
        store_huffman_tree(depth, length, w, s);
    }
}

// ---- field coders (field_coder.h, histogram_writer.cc:17-57) --------------

struct Distr {
    bool direct;
    std::uint32_t v;
    std::uint32_t bits;
};

template <typename Sink>
CJX_HDI void write_u32(Sink& w, const Distr* d, std::uint32_t value) {
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
    w.write(2, best);
    if (!d[best].direct) {
        w.write(d[best].bits, value - d[best].v);
    }
}

CJX_HDI std::uint32_t ceil_log2(std::uint32_t x) {
    if (x <= 1) {
        return 0;
    }
    std::uint32_t b{0};
    for (std::uint32_t v{x - 1}; v; v >>= 1) {
        ++b;
    }
    return b;
}

template <typename Sink>
CJX_HDI void write_var_len_u16(Sink& w, std::uint32_t n) {
    if (n == 0) {
        w.write(1, 0);
        return;
    }
    w.write(1, 1);
    std::uint32_t nb{0};
    for (std::uint32_t v{n}; v > 1; v >>= 1) {
        ++nb;
    }
    w.write(4, nb);
    w.write(nb, n - (1u << nb));
}

template <typename Sink>
CJX_HDI void write_uint_config(Sink& w, std::uint32_t se, std::uint32_t msb, std::uint32_t lsb) {
    w.write(ceil_log2(15 + 1), se);
    if (se == 15) {
        return;
    }
    w.write(ceil_log2(se + 1), msb);
    w.write(ceil_log2(se - msb + 1), lsb);
}

// hybrid_uint.h:23-40
CJX_HDI void hybrid_uint(std::uint32_t se, std::uint32_t msb, std::uint32_t lsb, std::uint32_t value,
                         std::uint32_t& tok, std::uint32_t& nb, std::uint32_t& bits) {
    const std::uint32_t split{1u << se};
    if (value < split) {
        tok = value;
        nb = 0;
        bits = 0;
        return;
    }
    std::uint32_t n{31u - static_cast<std::uint32_t>(__builtin_clz(value))};
    const std::uint32_t m{value - (1u << n)};
    tok = split + ((n - se) << (msb + lsb)) + ((m >> (n - msb)) << lsb) + (m & ((1u << lsb) - 1));
    nb = n - msb - lsb;
    bits = (value >> lsb) & ((1u << nb) - 1);
}

CJX_HDI void hybrid420(std::uint32_t value, std::uint32_t& tok, std::uint32_t& nb, std::uint32_t& bits) {
    hybrid_uint(4, 2, 0, value, tok, nb, bits);
}

CJX_HDI std::uint32_t pack_signed(std::int32_t v) {
    return (static_cast<std::uint32_t>(v) << 1) ^ static_cast<std::uint32_t>(v >> 31);
}

}  // namespace cjx
