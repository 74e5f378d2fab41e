/* Test infrastructure only: a decoder for the JPEG XL subset the encoder
 * writes, back to the quantized frame data.
 *
 * It reads what a JPEG XL decoder reads, in the order ISO/IEC 18181-1 gives,
 * and refuses anything outside the subset with an error naming it: the
 * container (signature, ftyp, jxlc box) or a bare codestream; SizeHeader;
 * all-default ImageMetadata and CustomTransformData; a regular VarDCT frame
 * header (one pass, no crop, replace blending, last frame, default loop
 * filter); an unpermuted TOC; DcGlobal (default DC dequantization, block
 * context map and colour correlation, no global tree); DC groups (modular DC
 * with a single-leaf tree per image, AcMetadata); AcGlobal (default
 * dequantization matrices, natural coefficient orders, clustered prefix
 * codes) and the AC groups.  Entropy codes: prefix codes only (simple and
 * code-length-coded, with repeat codes), simple or entropy-coded context maps
 * with move-to-front, hybrid-uint configurations; no LZ77, no ANS.
 *
 * The reference's wire format for these sections is
 * tools/bitstream/vardct_frame.h:53-58 and vardct_frame.cc:414-554; the
 * header fields follow src/bitstream/codestream.cc:20-98.  Each section must
 * end exactly at its TOC size (zero padding to the byte), so the TOC, every
 * code and every token count are checked by the decode itself.
 *
 * Nothing of the encoder is shared except the natural coefficient orders
 * (jo_natural_order, pinned against the reference's coeff_order fixtures).
 */
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "jxl_oracle.h"

/* ---- errors ---------------------------------------------------------------- */

static char g_err[256];

const char* jd_error(void) {
    return g_err;
}

/* Records the message; returns 0 (false) for the readers' `ok` results. */
static int fail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return 0;
}

/* ---- bit reader (LSB first) ------------------------------------------------ */

typedef struct br {
    const uint8_t* p;
    uint64_t pos, end; /* bits */
    int err;
} br;

static uint32_t bits(br* b, unsigned n) {
    if (n == 0) {
        return 0;
    }
    if (b->pos + n > b->end) {
        b->err = 1;
        b->pos = b->end;
        return 0;
    }
    uint64_t w = 0;
    const uint64_t byte = b->pos >> 3;
    const uint64_t last = (b->end + 7) >> 3;
    for (unsigned i = 0; i < 8 && byte + i < last; ++i) {
        w |= (uint64_t)b->p[byte + i] << (8 * i);
    }
    const uint32_t v = (uint32_t)((w >> (b->pos & 7)) & (n == 32 ? 0xffffffffull : ((1ull << n) - 1)));
    b->pos += n;
    return v;
}

/* Zero bits up to the next byte boundary; false if any is set. */
static int pad_to_byte(br* b) {
    const unsigned n = (unsigned)((8 - (b->pos & 7)) & 7);
    return bits(b, n) == 0;
}

typedef struct dist {
    int val;          /* 1: the value itself, 0: offset + `n` bits */
    uint32_t v;
    unsigned n;
} dist;

#define VAL(x) {1, (x), 0}
#define BITS(n) {0, 0, (n)}
#define BOFF(n, o) {0, (o), (n)}

static uint32_t u32(br* b, const dist d[4]) {
    const dist* s = &d[bits(b, 2)];
    return s->val ? s->v : s->v + bits(b, s->n);
}

static uint64_t u64(br* b) {
    switch (bits(b, 2)) {
        case 0: return 0;
        case 1: return 1 + bits(b, 4);
        case 2: return 17 + bits(b, 8);
        default: {
            uint64_t v = bits(b, 12);
            unsigned shift = 12;
            while (bits(b, 1)) {
                if (shift == 60) {
                    v |= (uint64_t)bits(b, 4) << 60;
                    break;
                }
                v |= (uint64_t)bits(b, 8) << shift;
                shift += 8;
            }
            return v;
        }
    }
}

static unsigned ceil_log2(uint64_t x) {
    unsigned n = 0;
    while ((1ull << n) < x) {
        ++n;
    }
    return n;
}

static int32_t unpack_signed(uint32_t u) {
    return (u & 1) ? -(int32_t)((u + 1) >> 1) : (int32_t)(u >> 1);
}

/* ---- prefix codes (Brotli format, canonical, up to 15 bits) ---------------- */

typedef struct pcode {
    int single;          /* one symbol, coded with zero bits */
    uint32_t sym0;
    uint16_t count[16];  /* symbols per code length */
    uint16_t* sorted;    /* symbols by (length, value) */
} pcode;

static void pcode_free(pcode* c) {
    free(c->sorted);
    c->sorted = NULL;
}

/* Canonical code from lengths; false if the code is over-subscribed or
 * incomplete (a single used symbol is coded with zero bits). */
static int pcode_build(pcode* c, const uint8_t* len, uint32_t alpha) {
    memset(c, 0, sizeof(*c));
    uint32_t used = 0, last = 0;
    for (uint32_t i = 0; i < alpha; ++i) {
        if (len[i] > 15) {
            return 0;
        }
        if (len[i]) {
            ++c->count[len[i]];
            ++used;
            last = i;
        }
    }
    if (used <= 1) {
        c->single = 1;
        c->sym0 = last;
        return 1;
    }
    int64_t space = 1 << 15;
    for (int l = 1; l <= 15; ++l) {
        space -= (int64_t)c->count[l] << (15 - l);
    }
    if (space != 0) {
        return 0;
    }
    uint32_t off[16] = {0};
    for (int l = 2; l <= 15; ++l) {
        off[l] = off[l - 1] + c->count[l - 1];
    }
    c->sorted = (uint16_t*)malloc(sizeof(uint16_t) * used);
    for (uint32_t i = 0; i < alpha; ++i) {
        if (len[i]) {
            c->sorted[off[len[i]]++] = (uint16_t)i;
        }
    }
    return 1;
}

/* Code words are stored LSB first with their bits reversed, so the bits come
 * out most significant first. */
static uint32_t pcode_read(br* b, const pcode* c) {
    if (c->single) {
        return c->sym0;
    }
    int code = 0, first = 0, index = 0;
    for (int l = 1; l <= 15; ++l) {
        code |= (int)bits(b, 1);
        const int n = c->count[l];
        if (code - first < n) {
            return c->sorted[index + code - first];
        }
        index += n;
        first = (first + n) << 1;
        code <<= 1;
    }
    b->err = 2;
    return 0;
}

static int read_prefix_code(br* b, uint32_t alpha, pcode* out) {
    uint8_t* len = (uint8_t*)calloc(alpha, 1);
    int ok = 1;
    const uint32_t hskip = bits(b, 2);
    if (hskip == 1) {
        const unsigned max_bits = ceil_log2(alpha);
        const uint32_t nsym = bits(b, 2) + 1;
        uint32_t s[4];
        for (uint32_t i = 0; i < nsym; ++i) {
            s[i] = bits(b, max_bits);
            if (s[i] >= alpha) {
                ok = 0;
            }
            for (uint32_t j = 0; j < i; ++j) {
                if (s[j] == s[i]) {
                    ok = 0;
                }
            }
        }
        if (ok) {
            static const uint8_t l3[3] = {1, 2, 2}, l4a[4] = {2, 2, 2, 2}, l4b[4] = {1, 2, 3, 3};
            if (nsym == 1) {
                len[s[0]] = 1; /* a single symbol: zero bits */
            } else if (nsym == 2) {
                len[s[0]] = len[s[1]] = 1;
            } else if (nsym == 3) {
                for (int i = 0; i < 3; ++i) {
                    len[s[i]] = l3[i];
                }
            } else {
                const uint8_t* l = bits(b, 1) ? l4b : l4a;
                for (int i = 0; i < 4; ++i) {
                    len[s[i]] = l[i];
                }
            }
        }
    } else {
        static const uint8_t order[18] = {1, 2, 3, 4, 0, 5, 17, 6, 16, 7, 8, 9, 10, 11, 12, 13, 14, 15};
        uint8_t cl[18] = {0};
        int space = 32, num = 0;
        for (uint32_t i = hskip; i < 18 && space > 0; ++i) {
            /* Static code of the code-length code lengths: 00->0, 10->4 (bits
             * read 0,1), 01->3, 11 then 0->2, 1110->1, 1111->5. */
            uint32_t v = bits(b, 2), l;
            if (v == 0) {
                l = 0;
            } else if (v == 1) {
                l = 4;
            } else if (v == 2) {
                l = 3;
            } else if (!bits(b, 1)) {
                l = 2;
            } else {
                l = bits(b, 1) ? 5 : 1;
            }
            cl[order[i]] = (uint8_t)l;
            if (l) {
                space -= 32 >> l;
                ++num;
            }
        }
        pcode clc;
        if ((num != 1 && space != 0) || !pcode_build(&clc, cl, 18)) {
            ok = 0;
        } else {
            uint32_t i = 0;
            int prev = 8, repeat = 0, repeat_len = 0;
            int32_t left = 1 << 15;
            while (i < alpha && left > 0 && !b->err) {
                const uint32_t sym = pcode_read(b, &clc);
                if (sym < 16) {
                    len[i++] = (uint8_t)sym;
                    repeat = 0;
                    if (sym) {
                        prev = (int)sym;
                        left -= (1 << 15) >> sym;
                    }
                } else {
                    const unsigned extra = sym == 16 ? 2 : 3;
                    const int nl = sym == 16 ? prev : 0;
                    if (repeat_len != nl) {
                        repeat = 0;
                        repeat_len = nl;
                    }
                    const int old = repeat;
                    if (repeat > 0) {
                        repeat = (repeat - 2) << extra;
                    }
                    repeat += (int)bits(b, extra) + 3;
                    const int delta = repeat - old;
                    if (i + (uint32_t)delta > alpha) {
                        ok = 0;
                        break;
                    }
                    for (int k = 0; k < delta; ++k) {
                        len[i++] = (uint8_t)nl;
                    }
                    if (nl) {
                        left -= delta * ((1 << 15) >> nl);
                    }
                }
            }
            if (left != 0) {
                ok = 0;
            }
            pcode_free(&clc);
        }
    }
    if (ok && !pcode_build(out, len, alpha)) {
        ok = 0;
    }
    free(len);
    return ok && !b->err;
}

/* ---- entropy codes ---------------------------------------------------------- */

typedef struct ecode {
    uint32_t nctx, ncl;
    uint8_t* cmap;
    uint32_t se[256], msb[256], lsb[256];
    pcode code[256];
} ecode;

static void ecode_free(ecode* e) {
    for (uint32_t i = 0; i < e->ncl; ++i) {
        pcode_free(&e->code[i]);
    }
    free(e->cmap);
    memset(e, 0, sizeof(*e));
}

static int read_ecode(br* b, uint32_t nctx, ecode* e);

static uint32_t read_value(br* b, const ecode* e, uint32_t ctx) {
    const uint32_t c = e->cmap[ctx];
    const uint32_t tok = pcode_read(b, &e->code[c]);
    const uint32_t se = e->se[c], msb = e->msb[c], lsb = e->lsb[c];
    const uint32_t split = 1u << se;
    if (tok < split) {
        return tok;
    }
    const uint32_t nb = se - (msb + lsb) + ((tok - split) >> (msb + lsb));
    if (nb > 31 - msb - lsb) {
        b->err = 3;
        return 0;
    }
    const uint32_t low = tok & ((1u << lsb) - 1);
    const uint32_t hi = ((tok >> lsb) & ((1u << msb) - 1)) | (1u << msb);
    return (((hi << nb) | bits(b, nb)) << lsb) | low;
}

static int read_cmap(br* b, uint32_t nctx, uint8_t* map) {
    if (bits(b, 1)) {
        const unsigned bpe = bits(b, 2);
        for (uint32_t i = 0; i < nctx; ++i) {
            map[i] = (uint8_t)bits(b, bpe);
        }
        return !b->err;
    }
    const int mtf = (int)bits(b, 1);
    ecode sub;
    if (!read_ecode(b, 1, &sub)) {
        return 0;
    }
    int ok = 1;
    for (uint32_t i = 0; i < nctx && ok; ++i) {
        const uint32_t v = read_value(b, &sub, 0);
        if (v > 255) {
            ok = 0;
        }
        map[i] = (uint8_t)v;
    }
    ecode_free(&sub);
    if (ok && mtf) {
        uint8_t s[256];
        for (int i = 0; i < 256; ++i) {
            s[i] = (uint8_t)i;
        }
        for (uint32_t i = 0; i < nctx; ++i) {
            const uint8_t r = map[i], v = s[r];
            map[i] = v;
            memmove(s + 1, s, r);
            s[0] = v;
        }
    }
    return ok && !b->err;
}

static int read_ecode(br* b, uint32_t nctx, ecode* e) {
    memset(e, 0, sizeof(*e));
    e->nctx = nctx;
    e->cmap = (uint8_t*)calloc(nctx, 1);
    if (bits(b, 1)) {
        return fail("LZ77 is outside the subset");
    }
    if (nctx > 1 && !read_cmap(b, nctx, e->cmap)) {
        return fail("bad context map at bit %llu", (unsigned long long)b->pos);
    }
    uint32_t ncl = 0;
    for (uint32_t i = 0; i < nctx; ++i) {
        ncl = e->cmap[i] + 1u > ncl ? e->cmap[i] + 1u : ncl;
    }
    e->ncl = ncl;
    if (!bits(b, 1)) {
        return fail("ANS codes are outside the subset");
    }
    for (uint32_t c = 0; c < ncl; ++c) {
        e->se[c] = bits(b, 4);
        if (e->se[c] != 15) {
            e->msb[c] = bits(b, ceil_log2(e->se[c] + 1));
            if (e->msb[c] > e->se[c]) {
                return fail("bad hybrid-uint config");
            }
            e->lsb[c] = bits(b, ceil_log2(e->se[c] - e->msb[c] + 1));
            if (e->msb[c] + e->lsb[c] > e->se[c]) {
                return fail("bad hybrid-uint config");
            }
        }
    }
    uint32_t alpha[256];
    for (uint32_t c = 0; c < ncl; ++c) {
        if (bits(b, 1)) {
            const unsigned n = bits(b, 4);
            alpha[c] = 1 + (1u << n) + bits(b, n);
        } else {
            alpha[c] = 1;
        }
        if (alpha[c] > (1u << 15)) {
            return fail("alphabet too large");
        }
    }
    for (uint32_t c = 0; c < ncl; ++c) {
        if (alpha[c] <= 1) {
            e->code[c].single = 1;
            e->code[c].sym0 = 0;
        } else if (!read_prefix_code(b, alpha[c], &e->code[c])) {
            return fail("bad prefix code %u at bit %llu", c, (unsigned long long)b->pos);
        }
    }
    return !b->err;
}

/* ---- modular images with a single-leaf tree --------------------------------- */

typedef struct leaf {
    uint32_t predictor;
    int32_t offset;
    uint32_t multiplier;
    ecode data;
} leaf;

static int read_modular_header(br* b, leaf* lf) {
    static const dist ntr[4] = {VAL(0), VAL(1), BOFF(4, 2), BOFF(8, 18)};
    if (bits(b, 1)) {
        return fail("global MA tree is outside the subset");
    }
    if (!bits(b, 1)) {
        return fail("custom weighted-predictor parameters are outside the subset");
    }
    if (u32(b, ntr) != 0) {
        return fail("modular transforms are outside the subset");
    }
    ecode tree;
    if (!read_ecode(b, 6, &tree)) {
        return 0;
    }
    if (read_value(b, &tree, 1) != 0) {
        ecode_free(&tree);
        return fail("MA trees with decision nodes are outside the subset");
    }
    lf->predictor = read_value(b, &tree, 2);
    lf->offset = unpack_signed(read_value(b, &tree, 3));
    const uint32_t mul_log = read_value(b, &tree, 4);
    const uint32_t mul_bits = read_value(b, &tree, 5);
    ecode_free(&tree);
    if (mul_log > 30 || mul_bits > 30) {
        return fail("bad leaf multiplier");
    }
    lf->multiplier = (mul_bits + 1) << mul_log;
    if (lf->predictor != 0 && lf->predictor != 5) {
        return fail("predictor %u is outside the subset", lf->predictor);
    }
    return read_ecode(b, 1, &lf->data);
}

static int read_channel(br* b, const leaf* lf, int32_t* ch, uint32_t w, uint32_t h) {
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const int64_t v = (int64_t)unpack_signed(read_value(b, &lf->data, 0)) * lf->multiplier + lf->offset;
            int64_t pred = 0;
            if (lf->predictor == 5) {
                const int64_t W = x > 0 ? ch[y * w + x - 1] : (y > 0 ? ch[(y - 1) * w + x] : 0);
                const int64_t N = y > 0 ? ch[(y - 1) * w + x] : W;
                const int64_t NW = (x > 0 && y > 0) ? ch[(y - 1) * w + x - 1] : W;
                const int64_t lo = W < N ? W : N, hi = W < N ? N : W;
                const int64_t g = W + N - NW;
                pred = g < lo ? lo : (g > hi ? hi : g);
            }
            const int64_t r = v + pred;
            if (r < INT32_MIN || r > INT32_MAX) {
                return fail("modular value out of range");
            }
            ch[y * w + x] = (int32_t)r;
        }
        if (b->err) {
            return fail("modular data runs past its section");
        }
    }
    return 1;
}

/* ---- frame ---------------------------------------------------------------- */

typedef struct frame {
    uint32_t w, h, bw, bh, xg, yg, nac, xdg, ydg, ndc, cmw, cmh;
    uint32_t gs, qdc;
    int8_t* acs;
    int32_t *qf, *dc;
    int16_t* ac;
    int8_t *ytox, *ytob;
    uint32_t nhist;
    ecode acode;
    uint32_t order8[64], order16[256], order32[1024];
} frame;

static int read_dc_global(br* b, frame* f) {
    static const dist gs[4] = {BOFF(11, 1), BOFF(11, 2049), BOFF(12, 4097), BOFF(16, 8193)};
    static const dist qdc[4] = {VAL(16), BOFF(5, 1), BOFF(8, 1), BOFF(16, 1)};
    if (!bits(b, 1)) {
        return fail("custom DC dequantization is outside the subset");
    }
    f->gs = u32(b, gs);
    f->qdc = u32(b, qdc);
    if (!bits(b, 1)) {
        return fail("custom block context map is outside the subset");
    }
    if (!bits(b, 1)) {
        return fail("custom colour correlation is outside the subset");
    }
    if (bits(b, 1)) {
        return fail("a global MA tree is outside the subset");
    }
    return !b->err;
}

static int read_dc_group(br* b, frame* f, uint32_t g) {
    const uint32_t bx0 = (g % f->xdg) * 256, by0 = (g / f->xdg) * 256;
    const uint32_t dgw = f->bw - bx0 < 256 ? f->bw - bx0 : 256;
    const uint32_t dgh = f->bh - by0 < 256 ? f->bh - by0 : 256;
    const size_t n = (size_t)dgw * dgh, nblk = (size_t)f->bw * f->bh;
    bits(b, 2); /* extra precision */
    leaf lf;
    if (!read_modular_header(b, &lf)) {
        return 0;
    }
    int32_t* ch = (int32_t*)malloc(sizeof(int32_t) * n);
    static const int plane[3] = {1, 0, 2}; /* modular channels Y, X, B */
    int ok = 1;
    for (int c = 0; c < 3 && ok; ++c) {
        ok = read_channel(b, &lf, ch, dgw, dgh);
        for (uint32_t y = 0; y < dgh && ok; ++y) {
            for (uint32_t x = 0; x < dgw; ++x) {
                f->dc[plane[c] * nblk + (size_t)(by0 + y) * f->bw + bx0 + x] = ch[y * dgw + x];
            }
        }
    }
    ecode_free(&lf.data);
    free(ch);
    if (!ok) {
        return 0;
    }
    /* AcMetadata: first-block count, then YtoX, YtoB, strategy + quant field
     * rows, EPF sharpness. */
    const uint32_t count = 1 + bits(b, ceil_log2(n));
    const uint32_t cw = (dgw + 7) / 8, chh = (dgh + 7) / 8;
    if (!read_modular_header(b, &lf)) {
        return 0;
    }
    int32_t* m = (int32_t*)malloc(sizeof(int32_t) * (2 * (size_t)cw * chh + 2 * (size_t)count + n));
    int32_t* tx = m;
    int32_t* tb = tx + (size_t)cw * chh;
    int32_t* sq = tb + (size_t)cw * chh;
    int32_t* epf = sq + 2 * (size_t)count;
    ok = read_channel(b, &lf, tx, cw, chh) && read_channel(b, &lf, tb, cw, chh) &&
         read_channel(b, &lf, sq, count, 2) && read_channel(b, &lf, epf, dgw, dgh);
    ecode_free(&lf.data);
    for (uint32_t y = 0; y < chh && ok; ++y) {
        for (uint32_t x = 0; x < cw; ++x) {
            const int32_t a = tx[y * cw + x], c = tb[y * cw + x];
            if (a < -128 || a > 127 || c < -128 || c > 127) {
                ok = fail("colour correlation factor out of range");
                break;
            }
            f->ytox[(size_t)(by0 / 8 + y) * f->cmw + bx0 / 8 + x] = (int8_t)a;
            f->ytob[(size_t)(by0 / 8 + y) * f->cmw + bx0 / 8 + x] = (int8_t)c;
        }
    }
    for (size_t i = 0; i < n && ok; ++i) {
        if (epf[i] != 0) {
            ok = fail("non-zero EPF sharpness is outside the subset");
        }
    }
    /* Place the varblocks: raster order over the group, each strategy at the
     * first position its blocks do not already cover. */
    uint32_t next = 0;
    for (uint32_t y = 0; y < dgh && ok; ++y) {
        for (uint32_t x = 0; x < dgw && ok; ++x) {
            const size_t pos = (size_t)(by0 + y) * f->bw + bx0 + x;
            if (f->acs[pos] != -1) {
                continue;
            }
            if (next >= count) {
                ok = fail("fewer strategies than blocks in DC group %u", g);
                break;
            }
            const int32_t s = sq[next], q = sq[count + next];
            ++next;
            const int side = s == 0 ? 1 : (s == 4 ? 2 : (s == 5 ? 4 : 0));
            if (side == 0) {
                ok = fail("strategy %d is outside the subset", s);
                break;
            }
            if (x + side > dgw || y + side > dgh || (x % side) != 0 || (y % side) != 0) {
                ok = fail("misplaced varblock at block (%u, %u)", bx0 + x, by0 + y);
                break;
            }
            if (q < 0 || q > 255) {
                ok = fail("quant field %d out of range", q + 1);
                break;
            }
            for (int dy = 0; dy < side; ++dy) {
                for (int dx = 0; dx < side; ++dx) {
                    const size_t p = pos + (size_t)dy * f->bw + dx;
                    if (f->acs[p] != -1) {
                        ok = fail("overlapping varblocks at block (%u, %u)", bx0 + x, by0 + y);
                    }
                    f->acs[p] = 0;
                }
            }
            f->acs[pos] = (int8_t)(8 * side);
            f->qf[pos] = q + 1;
        }
    }
    if (ok && next != count) {
        ok = fail("DC group %u lists %u strategies for %u varblocks", g, count, next);
    }
    free(m);
    return ok && !b->err;
}

static int read_ac_global(br* b, frame* f) {
    static const dist used[4] = {VAL(0x5F), VAL(0x13), VAL(0), BOFF(13, 0)};
    if (!bits(b, 1)) {
        return fail("custom dequantization matrices are outside the subset");
    }
    f->nhist = 1 + bits(b, ceil_log2(f->nac));
    if (u32(b, used) != 0) {
        return fail("custom coefficient orders are outside the subset");
    }
    return read_ecode(b, 15u * 495u * f->nhist, &f->acode);
}

static const uint8_t CTX_MAP[39] = {0, 1, 2, 2, 3,  3,  4,  5,  6,  6,  6,  6,  6,
                                    7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14,
                                    7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14};
static const uint8_t FREQ_CTX[64] = {
    0,  0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 15, 16, 16, 17, 17,
    18, 18, 19, 19, 20, 20, 21, 21, 22, 22, 23, 23, 23, 23, 24, 24, 24, 24, 25, 25, 25, 25,
    26, 26, 26, 26, 27, 27, 27, 27, 28, 28, 28, 28, 29, 29, 29, 29, 30, 30, 30, 30};
static const uint8_t NZ_CTX[64] = {
    0,   0,   31,  62,  62,  93,  93,  93,  93,  123, 123, 123, 123, 152, 152, 152,
    152, 152, 152, 152, 152, 180, 180, 180, 180, 180, 180, 180, 180, 180, 180, 180,
    180, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
    206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206};

static int read_ac_group(br* b, frame* f, uint32_t g) {
    const uint32_t bx0 = (g % f->xg) * 32, by0 = (g / f->xg) * 32;
    const uint32_t gbw = f->bw - bx0 < 32 ? f->bw - bx0 : 32;
    const uint32_t gbh = f->bh - by0 < 32 ? f->bh - by0 : 32;
    const size_t nblk = (size_t)f->bw * f->bh;
    const uint32_t sel = bits(b, ceil_log2(f->nhist));
    if (sel >= f->nhist) {
        return fail("histogram selector out of range");
    }
    const uint32_t ctx_base = sel * 15u * 495u;
    int32_t* grid = (int32_t*)calloc((size_t)3 * gbw * gbh, sizeof(int32_t));
    static const int corder[3] = {1, 0, 2};
    int ok = 1;
    for (uint32_t sy = 0; sy < gbh && ok; ++sy) {
        for (uint32_t sx = 0; sx < gbw && ok; ++sx) {
            const size_t pos = (size_t)(by0 + sy) * f->bw + bx0 + sx;
            const int n = f->acs[pos];
            if (n == 0) {
                continue;
            }
            const uint32_t side = (uint32_t)n / 8, cb = side * side, size = (uint32_t)(n * n);
            const uint32_t l2 = n == 16 ? 2 : (n == 32 ? 4 : 0);
            const int ord = n == 16 ? 2 : (n == 32 ? 3 : 0);
            const uint32_t* order = n == 32 ? f->order32 : (n == 16 ? f->order16 : f->order8);
            for (int ci = 0; ci < 3 && ok; ++ci) {
                const int c = corder[ci];
                int32_t* gr = grid + (size_t)c * gbw * gbh;
                const int bctx = CTX_MAP[(c < 2 ? (c ^ 1) : 2) * 13 + ord];
                const int hl = sx > 0, ht = sy > 0;
                const int32_t top = ht ? gr[(sy - 1) * gbw + sx] : 0, left = hl ? gr[sy * gbw + sx - 1] : 0;
                uint32_t pred = !hl ? (ht ? (uint32_t)top : 32u) : (!ht ? (uint32_t)left : (uint32_t)((top + left + 1) / 2));
                pred = pred > 64 ? 64 : pred;
                const uint32_t nzctx = (pred < 8 ? pred : 4 + pred / 2) * 15u + (uint32_t)bctx;
                const uint32_t nz = read_value(b, &f->acode, ctx_base + nzctx);
                if (nz > size - cb) {
                    ok = fail("non-zero count %u exceeds the varblock at (%u, %u)", nz, bx0 + sx, by0 + sy);
                    break;
                }
                const int32_t norm = (int32_t)((nz + cb - 1) >> l2);
                for (uint32_t dy = 0; dy < side; ++dy) {
                    for (uint32_t dx = 0; dx < side; ++dx) {
                        gr[(sy + dy) * gbw + sx + dx] = norm;
                    }
                }
                int16_t* pl = f->ac + (size_t)c * nblk * 64;
                uint32_t rem = nz, prev = nz > size / 16 ? 0u : 1u;
                const uint32_t off = 15u * 37u + 458u * (uint32_t)bctx;
                for (uint32_t k = cb; k < size && rem > 0; ++k) {
                    const uint32_t zl = (rem + cb - 1) >> l2, kk = k >> l2;
                    const uint32_t ctx = off + (NZ_CTX[zl] + FREQ_CTX[kk]) * 2u + prev;
                    const int32_t v = unpack_signed(read_value(b, &f->acode, ctx_base + ctx));
                    if (v < -32768 || v > 32767) {
                        ok = fail("coefficient out of the int16 range");
                        break;
                    }
                    const uint32_t e = order[k], q = e >> 6;
                    pl[((size_t)(by0 + sy + q / side) * f->bw + bx0 + sx + q % side) * 64 + (e & 63)] = (int16_t)v;
                    prev = v != 0;
                    rem -= prev;
                }
                if (ok && rem != 0) {
                    ok = fail("varblock at (%u, %u) ends with %u non-zeros left", bx0 + sx, by0 + sy, rem);
                }
                if (b->err) {
                    ok = fail("AC group %u runs past its section", g);
                }
            }
        }
    }
    free(grid);
    return ok;
}

/* Section reader: exactly `bytes` long, zero padding after the last bit. */
static int section_done(br* b, const char* what, uint32_t i) {
    if (b->err) {
        return fail("%s %u runs past its TOC size", what, i);
    }
    if (!pad_to_byte(b) || b->pos != b->end) {
        return fail("%s %u ends at bit %llu of %llu", what, i, (unsigned long long)b->pos,
                    (unsigned long long)b->end);
    }
    return 1;
}

static uint32_t box_u32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

/* The codestream inside the container (or the whole file when bare). */
static int find_codestream(const uint8_t* file, size_t n, const uint8_t** cs, size_t* csn) {
    static const uint8_t sig[12] = {0, 0, 0, 0x0c, 'J', 'X', 'L', ' ', 0x0d, 0x0a, 0x87, 0x0a};
    if (n >= 2 && file[0] == 0xff && file[1] == 0x0a) {
        *cs = file;
        *csn = n;
        return 1;
    }
    if (n < 12 || memcmp(file, sig, 12) != 0) {
        return fail("not a JPEG XL file");
    }
    size_t at = 12;
    while (at + 8 <= n) {
        uint64_t size = box_u32(file + at);
        size_t hdr = 8;
        if (size == 1) {
            if (at + 16 > n) {
                return fail("truncated box header");
            }
            size = (uint64_t)box_u32(file + at + 8) << 32 | box_u32(file + at + 12);
            hdr = 16;
        } else if (size == 0) {
            size = n - at;
        }
        if (size < hdr || at + size > n) {
            return fail("box overruns the file");
        }
        if (memcmp(file + at + 4, "jxlc", 4) == 0) {
            *cs = file + at + hdr;
            *csn = (size_t)size - hdr;
            return 1;
        }
        at += (size_t)size;
    }
    return fail("no jxlc box");
}

static int read_headers(br* b, uint32_t* w, uint32_t* h) {
    static const dist dim[4] = {BOFF(9, 1), BOFF(13, 1), BOFF(18, 1), BOFF(30, 1)};
    if (bits(b, 8) != 0xff || bits(b, 8) != 0x0a) {
        return fail("bad codestream signature");
    }
    const int small = (int)bits(b, 1);
    *h = small ? 8 * (bits(b, 5) + 1) : u32(b, dim);
    const uint32_t ratio = bits(b, 3);
    if (ratio != 0) {
        static const uint32_t num[8] = {0, 1, 12, 4, 3, 16, 5, 2}, den[8] = {1, 1, 10, 3, 2, 9, 4, 1};
        *w = (uint32_t)((uint64_t)*h * num[ratio] / den[ratio]);
    } else {
        *w = small ? 8 * (bits(b, 5) + 1) : u32(b, dim);
    }
    if (!bits(b, 1)) {
        return fail("non-default image metadata is outside the subset");
    }
    if (!bits(b, 1)) {
        return fail("custom transform data is outside the subset");
    }
    if (!pad_to_byte(b)) {
        return fail("non-zero padding after the image header");
    }
    return !b->err;
}

int jd_probe(const uint8_t* file, size_t n, uint32_t* w, uint32_t* h) {
    const uint8_t* cs;
    size_t csn;
    if (!find_codestream(file, n, &cs, &csn)) {
        return -1;
    }
    br b = {cs, 0, (uint64_t)csn * 8, 0};
    return read_headers(&b, w, h) ? 0 : -1;
}

static int read_frame_header(br* b) {
    static const dist ft[4] = {VAL(0), VAL(1), VAL(2), VAL(3)};
    static const dist ups[4] = {VAL(1), VAL(2), VAL(4), VAL(8)};
    static const dist passes[4] = {VAL(1), VAL(2), VAL(3), BOFF(3, 4)};
    static const dist blend[4] = {VAL(0), VAL(1), VAL(2), BOFF(2, 3)};
    static const dist name[4] = {VAL(0), BITS(4), BOFF(5, 16), BOFF(10, 48)};
    if (bits(b, 1)) {
        return fail("all-default frame header (modular) is outside the subset");
    }
    if (u32(b, ft) != 0) {
        return fail("only regular frames are in the subset");
    }
    if (bits(b, 1)) {
        return fail("modular frames are outside the subset");
    }
    if (u64(b) != 0) {
        return fail("frame flags are outside the subset");
    }
    if (u32(b, ups) != 1) {
        return fail("upsampling is outside the subset");
    }
    bits(b, 3); /* x_qm_scale */
    bits(b, 3); /* b_qm_scale */
    if (u32(b, passes) != 1) {
        return fail("progressive passes are outside the subset");
    }
    if (bits(b, 1)) {
        return fail("cropped frames are outside the subset");
    }
    if (u32(b, blend) != 0) {
        return fail("blending is outside the subset");
    }
    if (!bits(b, 1)) {
        return fail("multi-frame images are outside the subset");
    }
    const uint32_t nlen = u32(b, name);
    for (uint32_t i = 0; i < nlen; ++i) {
        bits(b, 8);
    }
    if (!bits(b, 1)) {
        return fail("custom loop filter is outside the subset");
    }
    if (u64(b) != 0) {
        return fail("frame header extensions are outside the subset");
    }
    return !b->err;
}

int jd_decode(const uint8_t* file, size_t n, uint32_t width, uint32_t height, uint32_t* global_scale,
              uint32_t* quant_dc, int8_t* acs, int32_t* qf, int32_t* dc, int16_t* ac, int8_t* ytox,
              int8_t* ytob) {
    const uint8_t* cs;
    size_t csn;
    if (!find_codestream(file, n, &cs, &csn)) {
        return -1;
    }
    br b = {cs, 0, (uint64_t)csn * 8, 0};
    uint32_t w, h;
    if (!read_headers(&b, &w, &h)) {
        return -1;
    }
    if (w != width || h != height || w % 8 || h % 8) {
        fail("frame is %ux%u, expected %ux%u (multiples of 8)", w, h, width, height);
        return -1;
    }
    if (!read_frame_header(&b)) {
        return -1;
    }
    frame f;
    memset(&f, 0, sizeof(f));
    f.w = w;
    f.h = h;
    f.bw = w / 8;
    f.bh = h / 8;
    f.xg = (f.bw + 31) / 32;
    f.yg = (f.bh + 31) / 32;
    f.nac = f.xg * f.yg;
    f.xdg = (f.bw + 255) / 256;
    f.ydg = (f.bh + 255) / 256;
    f.ndc = f.xdg * f.ydg;
    f.cmw = (f.bw + 7) / 8;
    f.cmh = (f.bh + 7) / 8;
    const size_t nblk = (size_t)f.bw * f.bh;
    f.acs = acs;
    f.qf = qf;
    f.dc = dc;
    f.ac = ac;
    f.ytox = ytox;
    f.ytob = ytob;
    memset(acs, 0xff, nblk); /* -1: not yet covered */
    memset(qf, 0, nblk * sizeof(int32_t));
    memset(dc, 0, 3 * nblk * sizeof(int32_t));
    memset(ac, 0, 3 * nblk * 64 * sizeof(int16_t));
    jo_natural_order(8, f.order8);
    jo_natural_order(16, f.order16);
    jo_natural_order(32, f.order32);

    static const dist toc_dist[4] = {BITS(10), BOFF(14, 1024), BOFF(22, 17408), BOFF(30, 4211712)};
    if (bits(&b, 1)) {
        fail("permuted TOC is outside the subset");
        return -1;
    }
    if (!pad_to_byte(&b)) {
        fail("non-zero padding before the TOC");
        return -1;
    }
    const int single = f.nac == 1;
    const uint32_t entries = single ? 1 : 2 + f.ndc + f.nac;
    uint32_t* sizes = (uint32_t*)malloc(sizeof(uint32_t) * entries);
    for (uint32_t i = 0; i < entries; ++i) {
        sizes[i] = u32(&b, toc_dist);
    }
    if (!pad_to_byte(&b) || b.err) {
        free(sizes);
        fail("bad TOC");
        return -1;
    }
    uint64_t at = b.pos;
    int ok = 1;
    for (uint32_t i = 0; i < entries && ok; ++i) {
        if (at + (uint64_t)sizes[i] * 8 > (uint64_t)csn * 8) {
            ok = fail("section %u overruns the codestream", i);
            break;
        }
        br s = {cs, at, at + (uint64_t)sizes[i] * 8, 0};
        if (single) {
            ok = read_dc_global(&s, &f) && read_dc_group(&s, &f, 0) && read_ac_global(&s, &f) &&
                 read_ac_group(&s, &f, 0) && section_done(&s, "section", 0);
        } else if (i == 0) {
            ok = read_dc_global(&s, &f) && section_done(&s, "DcGlobal", 0);
        } else if (i <= f.ndc) {
            ok = read_dc_group(&s, &f, i - 1) && section_done(&s, "DcGroup", i - 1);
        } else if (i == f.ndc + 1) {
            ok = read_ac_global(&s, &f) && section_done(&s, "AcGlobal", 0);
        } else {
            ok = read_ac_group(&s, &f, i - f.ndc - 2) && section_done(&s, "AcGroup", i - f.ndc - 2);
        }
        at += (uint64_t)sizes[i] * 8;
    }
    if (ok && at != (uint64_t)csn * 8) {
        ok = fail("%llu bytes after the last section", (unsigned long long)(csn - at / 8));
    }
    free(sizes);
    ecode_free(&f.acode);
    if (!ok) {
        return -1;
    }
    *global_scale = f.gs;
    *quant_dc = f.qdc;
    return 0;
}
