/*
 * gpzip_oracle.c -- CPU restatement of gpzip's GPU gzip compress path.
 *
 * TEST INFRASTRUCTURE ONLY: the expected outputs for the HIP library's
 * parity tests and the CPU baseline that bench.py reports.  The product never links or calls it.
 *
 * The reference is Rust + WGSL on wgpu and cannot be built here, so this
 * file restates its algorithm in plain C, one function per reference
 * function, each citing the file:line it follows (paths relative to the
 * reference repository with the crates/gpzip- prefix dropped).
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   orc_match_find       - against orc_match_find_brute (independent,
 *                          table-free statement of the candidate rule),
 *                          hand-derived known answers (q x 64, abc x 8,
 *                          < 3 bytes, 32 distinct bytes, the candidate order
 *                          and window cases, and hash3 collisions that
 *                          shadow a real candidate) and invariants
 *                          (every match verifies, 3<=len<=258, dist<=32768,
 *                          reconstruct(walk(tok)) == input).
 *   orc_hash3            - the collision known answers above: a build with
 *                          another multiplier (ORC_HASH_MUL) or byte order
 *                          fails them (tests/test_oracle.py builds such a
 *                          mutant and checks that it does).
 *   orc_greedy_walk      - closed-form cases + the round trip.
 *   orc_lookahead_walk   - (the ratio mode's rule, not the reference's)
 *                          hand-derived cases + the round trip.
 *   orc_strict_walk      - the reference's exact-equality walk cases
 *                          (walk.rs:575-695) as fixtures.
 *   orc_build_code_lengths - reference property tests (huffman.rs:158-232),
 *                          brute-force optimality on small alphabets, Kraft,
 *                          hand-derived halving cases (limits 7 and 4), and
 *                          an independent heap model on random vectors; a
 *                          build that rounds the halving up (ORC_HALVE) fails
 *                          the halving cases.
 *   orc_encode_block_fast - zlib inflate (independent decoder), the
 *                          hand-derived empty-stream KAT, gzip -t.
 *   orc_crc32            - zlib.crc32.
 */
#include "gpzip_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* hash3 -- lz77_hash_build.wgsl:36-43 (byte order: lz77_hash_build.wgsl:30-34).
 * ORC_HASH_MUL and ORC_HASH_REVERSED exist so that a test can build mutants
 * and show that the known answers reject them. */
#ifndef ORC_HASH_MUL
#define ORC_HASH_MUL 0x9E3779B1u
#endif
uint32_t orc_hash3(const uint8_t *in, size_t p) {
#ifdef ORC_HASH_REVERSED /* mutant: bytes in the other order */
    uint32_t x = ((uint32_t)in[p + 2] << 16) | ((uint32_t)in[p + 1] << 8) | (uint32_t)in[p];
#else
    uint32_t x = ((uint32_t)in[p] << 16) | ((uint32_t)in[p + 1] << 8) | (uint32_t)in[p + 2];
#endif
    return (uint32_t)(x * (uint32_t)ORC_HASH_MUL) >> (32u - ORC_HASH_BITS);
}

/* try_candidate -- lz77_hash_lookup.wgsl:45-67.  Returns the accepted length
 * (0 = rejected) and writes the distance. */
static uint32_t try_candidate(const uint8_t *in, size_t n, uint32_t raw, size_t p,
                              uint32_t window, uint32_t cur_best, uint32_t *dist_out) {
    if (raw == 0u || raw == 0xFFFFFFFFu) return 0;
    size_t cand = (size_t)raw - 1u;
    if (cand >= p) return 0;
    size_t dist = p - cand;
    if (dist > window) return 0;
    if (in[cand] != in[p] || in[cand + 1] != in[p + 1] || in[cand + 2] != in[p + 2]) return 0;
    uint32_t len = 3;
    for (;;) {
        if (len >= ORC_MAX_MATCH) break;
        if (p + len >= n) break;
        if (in[cand + len] != in[p + len]) break;
        len++;
    }
    if (len > cur_best) {
        *dist_out = (uint32_t)dist;
        return len;
    }
    return 0;
}

/* The per-position loop shared by both match finders:
 * lz77_hash_lookup.wgsl:69-119.  newest(seg)/oldest(seg) give the raw table
 * values (p+1, or the sentinels 0 / 0xFFFFFFFF). */
typedef uint32_t (*cand_fn)(void *ctx, uint32_t h, size_t seg);

static uint32_t lookup_one(const uint8_t *in, size_t n, size_t p, uint32_t window,
                           cand_fn newest, cand_fn oldest, void *ctx) {
    if (p + 2 >= n) return in[p];
    uint32_t h = orc_hash3(in, p);
    size_t p_seg = p >> ORC_SEG_LOG2;
    uint32_t best_len = 0, best_dist = 0;
    for (size_t off = 0; off <= p_seg; off++) {
        size_t seg = p_seg - off;
        uint32_t d = 0;
        uint32_t l = try_candidate(in, n, newest(ctx, h, seg), p, window, best_len, &d);
        if (l) {
            best_len = l;
            best_dist = d;
        }
        if (best_len >= ORC_MAX_MATCH) break;
        if (best_len < ORC_OLDEST_SKIP) {
            l = try_candidate(in, n, oldest(ctx, h, seg), p, window, best_len, &d);
            if (l) {
                best_len = l;
                best_dist = d;
            }
            if (best_len >= ORC_MAX_MATCH) break;
        }
    }
    return best_len == 0 ? (uint32_t)in[p] : ((best_len << 16) | best_dist);
}

/* Dense [hash][segment] tables, lz77_hash.rs:120-181 (idx = h*num_segs+seg). */
typedef struct {
    uint32_t *oldest, *newest;
    size_t num_segs;
} dense_tab;

static uint32_t dense_newest(void *c, uint32_t h, size_t seg) {
    dense_tab *t = (dense_tab *)c;
    return t->newest[(size_t)h * t->num_segs + seg];
}
static uint32_t dense_oldest(void *c, uint32_t h, size_t seg) {
    dense_tab *t = (dense_tab *)c;
    return t->oldest[(size_t)h * t->num_segs + seg];
}

int orc_match_find(const uint8_t *in, size_t n, uint32_t window, uint32_t *tok) {
    if (n == 0) return 0; /* lz77_hash.rs:308-311 */
    size_t seg_size = (size_t)1 << ORC_SEG_LOG2;
    dense_tab t;
    t.num_segs = (n + seg_size - 1) / seg_size;
    if (t.num_segs == 0) t.num_segs = 1;
    size_t entries = ((size_t)1 << ORC_HASH_BITS) * t.num_segs;
    t.oldest = (uint32_t *)malloc(entries * sizeof(uint32_t));
    t.newest = (uint32_t *)calloc(entries, sizeof(uint32_t)); /* clear_buffer, lz77_hash.rs:587 */
    if (!t.oldest || !t.newest) {
        free(t.oldest);
        free(t.newest);
        return -1;
    }
    memset(t.oldest, 0xFF, entries * sizeof(uint32_t)); /* reset blob, lz77_hash.rs:586 */
    /* build pass: lz77_hash_build.wgsl:45-54 */
    for (size_t p = 0; p + 2 < n; p++) {
        uint32_t h = orc_hash3(in, p);
        size_t idx = (size_t)h * t.num_segs + (p >> ORC_SEG_LOG2);
        uint32_t v = (uint32_t)p + 1u;
        if (v < t.oldest[idx]) t.oldest[idx] = v;
        if (v > t.newest[idx]) t.newest[idx] = v;
    }
    /* lookup pass */
    for (size_t p = 0; p < n; p++)
        tok[p] = lookup_one(in, n, p, window, dense_newest, dense_oldest, &t);
    free(t.oldest);
    free(t.newest);
    return 0;
}

/* Table-free statement: newest/oldest of hash h in segment seg found by a
 * linear scan of that segment's positions q with q + 2 < n. */
typedef struct {
    const uint8_t *in;
    size_t n;
} brute_ctx;

static uint32_t brute_scan(void *c, uint32_t h, size_t seg, int want_newest) {
    brute_ctx *b = (brute_ctx *)c;
    size_t lo = seg << ORC_SEG_LOG2, hi = (seg + 1) << ORC_SEG_LOG2;
    uint32_t best = want_newest ? 0u : 0xFFFFFFFFu;
    for (size_t q = lo; q < hi && q + 2 < b->n; q++) {
        if (orc_hash3(b->in, q) != h) continue;
        uint32_t v = (uint32_t)q + 1u;
        if (want_newest) {
            if (v > best) best = v;
        } else if (v < best) {
            best = v;
        }
    }
    return best;
}
static uint32_t brute_newest(void *c, uint32_t h, size_t seg) { return brute_scan(c, h, seg, 1); }
static uint32_t brute_oldest(void *c, uint32_t h, size_t seg) { return brute_scan(c, h, seg, 0); }

int orc_match_find_brute(const uint8_t *in, size_t n, uint32_t window, uint32_t *tok) {
    brute_ctx b = {in, n};
    for (size_t p = 0; p < n; p++) tok[p] = lookup_one(in, n, p, window, brute_newest, brute_oldest, &b);
    return 0;
}

/* ------------------------------------------------------------------ */
/* greedy_walk (lazy) -- codec-gpu/src/gpu/lz77.rs:260-284 */
size_t orc_greedy_walk(const uint32_t *pp, const uint8_t *in, size_t n, uint32_t *out) {
    size_t p = 0, k = 0;
    while (p < n) {
        uint32_t t = pp[p];
        uint32_t len = t >> 16;
        if (len == 0) {
            out[k++] = t;
            p += 1;
            continue;
        }
        if (p + 1 < n) {
            uint32_t nl = pp[p + 1] >> 16;
            if (nl != 0 && nl > len) {
                out[k++] = (uint32_t)in[p]; /* Token::literal(input[p]) */
                p += 1;
                continue;
            }
        }
        out[k++] = t;
        p += len;
    }
    return k;
}

/* The opt-in ratio mode's walk (this library's own rule, no reference
 * counterpart): greedy_walk's demotion, extended to a match at p+2 that is
 * longer than the one at p by two or more.  Positions past n count as
 * literals.  Pinned by hand-derived cases and the round trip. */
size_t orc_lookahead_walk(const uint32_t *pp, const uint8_t *in, size_t n, uint32_t *out) {
    size_t p = 0, k = 0;
    while (p < n) {
        uint32_t t = pp[p];
        uint32_t len = t >> 16;
        if (len == 0) {
            out[k++] = t;
            p += 1;
            continue;
        }
        uint32_t n1 = p + 1 < n ? pp[p + 1] >> 16 : 0;
        uint32_t n2 = p + 2 < n ? pp[p + 2] >> 16 : 0;
        if (n1 > len || n2 > len + 1) {
            out[k++] = (uint32_t)in[p];
            p += 1;
            continue;
        }
        out[k++] = t;
        p += len;
    }
    return k;
}

/* host_strict_greedy -- codec-gpu/src/gpu/walk.rs:512-525 */
size_t orc_strict_walk(const uint32_t *pp, size_t n, uint32_t *out) {
    size_t p = 0, k = 0;
    while (p < n) {
        uint32_t t = pp[p];
        out[k++] = t;
        uint32_t len = t >> 16;
        p += len == 0 ? 1 : len;
    }
    return k;
}

/* reconstruct -- codec-gpu/src/gpu/lz77.rs:288-304 */
size_t orc_reconstruct(const uint32_t *tok, size_t ntok, uint8_t *out, size_t out_cap) {
    size_t o = 0;
    for (size_t i = 0; i < ntok; i++) {
        uint32_t len = tok[i] >> 16, d = tok[i] & 0xFFFFu;
        if (len == 0) {
            if (o >= out_cap) return (size_t)-1;
            out[o++] = (uint8_t)d;
        } else {
            if (d == 0 || d > o || o + len > out_cap) return (size_t)-1;
            size_t start = o - d;
            for (uint32_t j = 0; j < len; j++) out[o++] = out[start + j];
        }
    }
    return o;
}

/* ------------------------------------------------------------------ */
/* Huffman -- codec-gpu/src/gpu/huffman.rs */

typedef struct {
    uint64_t f;
    size_t id;
} hnode;

static int hless(hnode a, hnode b) { return a.f < b.f || (a.f == b.f && a.id < b.id); }

static void heap_push(hnode *h, size_t *sz, hnode v) {
    size_t i = (*sz)++;
    h[i] = v;
    while (i > 0) {
        size_t par = (i - 1) / 2;
        if (!hless(h[i], h[par])) break;
        hnode tmp = h[i];
        h[i] = h[par];
        h[par] = tmp;
        i = par;
    }
}

static hnode heap_pop(hnode *h, size_t *sz) {
    hnode top = h[0];
    h[0] = h[--(*sz)];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < *sz && hless(h[l], h[m])) m = l;
        if (r < *sz && hless(h[r], h[m])) m = r;
        if (m == i) break;
        hnode tmp = h[i];
        h[i] = h[m];
        h[m] = tmp;
        i = m;
    }
    return top;
}

/* standard_huffman -- huffman.rs:63-100 (BinaryHeap<Reverse<(freq, id)>>:
 * pops the smallest (freq, id)); walk -- huffman.rs:102-117. */
static void standard_huffman(const uint32_t *freq, size_t n, uint8_t *lens) {
    hnode *heap = (hnode *)malloc(sizeof(hnode) * (2 * n + 1));
    int *left = (int *)malloc(sizeof(int) * 2 * n);
    int *right = (int *)malloc(sizeof(int) * 2 * n);
    size_t *stack = (size_t *)malloc(sizeof(size_t) * 4 * n);
    uint32_t *depth = (uint32_t *)malloc(sizeof(uint32_t) * 4 * n);
    size_t sz = 0;
    for (size_t i = 0; i < 2 * n; i++) left[i] = right[i] = -1;
    for (size_t i = 0; i < n; i++)
        if (freq[i] > 0) heap_push(heap, &sz, (hnode){freq[i], i});
    size_t next_id = n;
    while (sz > 1) {
        hnode a = heap_pop(heap, &sz);
        hnode b = heap_pop(heap, &sz);
        size_t id = next_id++;
        left[id] = (int)a.id;
        right[id] = (int)b.id;
        heap_push(heap, &sz, (hnode){a.f + b.f, id});
    }
    size_t root = heap_pop(heap, &sz).id;
    memset(lens, 0, n);
    size_t sp = 0;
    stack[sp] = root;
    depth[sp++] = 0;
    while (sp > 0) {
        sp--;
        size_t node = stack[sp];
        uint32_t d = depth[sp];
        if (node < n) {
            lens[node] = (uint8_t)(d > 0 ? d : 1);
            continue;
        }
        if (right[node] >= 0) {
            stack[sp] = (size_t)right[node];
            depth[sp++] = d + 1;
        }
        if (left[node] >= 0) {
            stack[sp] = (size_t)left[node];
            depth[sp++] = d + 1;
        }
    }
    free(heap);
    free(left);
    free(right);
    free(stack);
    free(depth);
}

/* try_build_code_lengths -- huffman.rs:26-60.  ORC_HALVE is the halving
 * step (huffman.rs:52-56, rounding down); a test builds a mutant that rounds
 * up and shows that the known answers reject it. */
#ifndef ORC_HALVE
#define ORC_HALVE(f) ((f) / 2u)
#endif
int orc_build_code_lengths(const uint32_t *freq, size_t n, uint32_t max_bits, uint8_t *lens) {
    size_t used = 0, only = 0;
    for (size_t i = 0; i < n; i++)
        if (freq[i] > 0) {
            used++;
            only = i;
        }
    memset(lens, 0, n);
    if (used == 0) return 1;
    if (used == 1) {
        lens[only] = 1;
        return 1;
    }
    uint32_t required = 0; /* ceil(log2(used)) */
    while (((size_t)1 << required) < used) required++;
    if (required > max_bits) return 0;
    uint32_t *work = (uint32_t *)malloc(sizeof(uint32_t) * n);
    memcpy(work, freq, sizeof(uint32_t) * n);
    for (;;) {
        standard_huffman(work, n, lens);
        uint32_t mx = 0;
        for (size_t i = 0; i < n; i++)
            if (lens[i] > mx) mx = lens[i];
        if (mx <= max_bits) break;
        for (size_t i = 0; i < n; i++)
            if (work[i] > 1) work[i] = ORC_HALVE(work[i]);
    }
    free(work);
    return 1;
}

/* canonical_codes -- huffman.rs:122-152 (RFC 1951 3.2.2) */
void orc_canonical_codes(const uint8_t *lens, size_t n, uint32_t *codes) {
    uint32_t max_bits = 0;
    for (size_t i = 0; i < n; i++)
        if (lens[i] > max_bits) max_bits = lens[i];
    memset(codes, 0, sizeof(uint32_t) * n);
    if (max_bits == 0) return;
    uint32_t bl_count[33] = {0}, next_code[33] = {0};
    for (size_t i = 0; i < n; i++)
        if (lens[i]) bl_count[lens[i]]++;
    uint32_t code = 0;
    for (uint32_t b = 1; b <= max_bits; b++) {
        code = (code + bl_count[b - 1]) << 1;
        next_code[b] = code;
    }
    for (size_t i = 0; i < n; i++)
        if (lens[i]) codes[i] = next_code[lens[i]]++;
}

/* ------------------------------------------------------------------ */
/* DEFLATE -- codec-gpu/src/gpu/deflate.rs */

/* RFC 1951 3.2.5 length/distance rows (deflate.rs:502-563, 575-681). */
static const uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
static const uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                      2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
static const uint16_t DIST_BASE[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                       33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                       1025, 1537, 2049, 3073, 4097, 6145,  8193,  12289, 16385, 24577};
static const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                       6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
/* META_ORDER -- deflate.rs:17-19 */
static const uint8_t META_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

static int len_sym(uint32_t len) { /* index into LEN_BASE; symbol = 257 + idx */
    int i = 28;
    while (i > 0 && LEN_BASE[i] > len) i--;
    return i;
}
static int dist_sym(uint32_t d) {
    int i = 29;
    while (i > 0 && DIST_BASE[i] > d) i--;
    return i;
}

/* reverse_bits -- deflate.rs:686-692 */
static uint32_t rev_bits(uint32_t code, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++)
        if (code >> i & 1u) r |= 1u << (bits - 1 - i);
    return r;
}

/* VecBitWriter -- deflate.rs:697-745: LSB-first packing, final byte
 * zero-padded. */
typedef struct {
    uint8_t *out;
    size_t cap, len;
    uint64_t acc;
    uint32_t nbits;
    int overflow;
} bitw;

static void bw_put(bitw *w, uint32_t value, uint32_t count) {
    w->acc |= (uint64_t)value << w->nbits;
    w->nbits += count;
    while (w->nbits >= 8) {
        if (w->len < w->cap)
            w->out[w->len] = (uint8_t)w->acc;
        else
            w->overflow = 1;
        w->len++;
        w->acc >>= 8;
        w->nbits -= 8;
    }
}

static void bw_finish(bitw *w) {
    if (w->nbits > 0) {
        if (w->len < w->cap)
            w->out[w->len] = (uint8_t)w->acc;
        else
            w->overflow = 1;
        w->len++;
        w->acc = 0;
        w->nbits = 0;
    }
}

/* rle_encode_lengths -- deflate.rs:254-326.  Each entry: code | extra_bits<<8 |
 * extra_value<<16. */
static size_t rle_encode(const uint8_t *lens, size_t n, uint32_t *out) {
    size_t k = 0, i = 0;
    while (i < n) {
        uint8_t cur = lens[i];
        size_t run = 1;
        while (i + run < n && lens[i + run] == cur) run++;
        if (cur == 0) {
            while (run >= 11) {
                size_t m = run < 138 ? run : 138;
                out[k++] = 18u | (7u << 8) | ((uint32_t)(m - 11) << 16);
                run -= m;
                i += m;
            }
            while (run >= 3) {
                size_t m = run < 10 ? run : 10;
                out[k++] = 17u | (3u << 8) | ((uint32_t)(m - 3) << 16);
                run -= m;
                i += m;
            }
            while (run > 0) {
                out[k++] = 0;
                run--;
                i++;
            }
        } else {
            out[k++] = cur;
            i++;
            run--;
            while (run >= 3) {
                size_t m = run < 6 ? run : 6;
                out[k++] = 16u | (2u << 8) | ((uint32_t)(m - 3) << 16);
                run -= m;
                i += m;
            }
            while (run > 0) {
                out[k++] = cur;
                run--;
                i++;
            }
        }
    }
    return k;
}

/* trim_to_min -- deflate.rs:329-335 */
static size_t trim_to_min(const uint8_t *lens, size_t n, size_t min) {
    while (n > min && lens[n - 1] == 0) n--;
    return n;
}

/* write_fixed_block_fast -- deflate.rs:763-797 (unreachable for these
 * alphabets, restated for completeness). */
static void write_fixed_block(const uint32_t *tok, size_t ntok, bitw *w) {
    bw_put(w, 3u, 3);
    for (size_t i = 0; i <= ntok; i++) {
        uint32_t sym, code, bits;
        uint32_t len = i < ntok ? tok[i] >> 16 : 0;
        sym = i == ntok ? 256u : (len == 0 ? (tok[i] & 0xFFFFu) : 257u + (uint32_t)len_sym(len));
        if (sym <= 143) {
            code = 0x30 + sym;
            bits = 8;
        } else if (sym <= 255) {
            code = 0x190 + (sym - 144);
            bits = 9;
        } else if (sym <= 279) {
            code = sym - 256;
            bits = 7;
        } else {
            code = 0xC0 + (sym - 280);
            bits = 8;
        }
        bw_put(w, rev_bits(code, bits), bits);
        if (i < ntok && len != 0) {
            int li = len_sym(len);
            if (LEN_EXTRA[li]) bw_put(w, len - LEN_BASE[li], LEN_EXTRA[li]);
            uint32_t d = tok[i] & 0xFFFFu;
            int di = dist_sym(d);
            bw_put(w, rev_bits((uint32_t)di, 5), 5);
            if (DIST_EXTRA[di]) bw_put(w, d - DIST_BASE[di], DIST_EXTRA[di]);
        }
    }
}

/* try_write_dynamic_block_fast -- deflate.rs:830-959 */
static int write_dynamic_block(const uint32_t *tok, size_t ntok, bitw *w) {
    uint32_t ll_freq[286] = {0}, d_freq[30] = {0};
    for (size_t i = 0; i < ntok; i++) {
        uint32_t len = tok[i] >> 16;
        if (len == 0) {
            ll_freq[tok[i] & 0xFFFFu]++;
        } else {
            ll_freq[257 + len_sym(len)]++;
            d_freq[dist_sym(tok[i] & 0xFFFFu)]++;
        }
    }
    ll_freq[256]++;

    uint8_t ll_lens[286], d_lens[30];
    if (!orc_build_code_lengths(ll_freq, 286, 15, ll_lens)) return 0;
    if (!orc_build_code_lengths(d_freq, 30, 15, d_lens)) return 0;

    size_t used_d = 0, only = 0;
    for (size_t i = 0; i < 30; i++)
        if (d_lens[i]) {
            if (!used_d) only = i;
            used_d++;
        }
    if (used_d == 0) {
        d_lens[0] = 1;
        d_lens[1] = 1;
    } else if (used_d == 1) {
        d_lens[only] = 1;
        d_lens[1 - (only < 1 ? only : 1)] = 1;
    }
    size_t used_ll = 0;
    for (size_t i = 0; i < 286; i++) used_ll += ll_lens[i] > 0;
    if (used_ll < 2) {
        ll_lens[0] = 1;
        ll_lens[256] = 1;
    }

    uint32_t ll_codes[286], d_codes[30], ll_pre[286], d_pre[30];
    orc_canonical_codes(ll_lens, 286, ll_codes);
    orc_canonical_codes(d_lens, 30, d_codes);
    for (size_t i = 0; i < 286; i++) ll_pre[i] = ll_lens[i] ? rev_bits(ll_codes[i], ll_lens[i]) : 0;
    for (size_t i = 0; i < 30; i++) d_pre[i] = d_lens[i] ? rev_bits(d_codes[i], d_lens[i]) : 0;

    size_t nlit = trim_to_min(ll_lens, 286, 257);
    size_t ndist = trim_to_min(d_lens, 30, 1);
    uint8_t combined[316];
    memcpy(combined, ll_lens, nlit);
    memcpy(combined + nlit, d_lens, ndist);
    uint32_t rle[316];
    size_t nrle = rle_encode(combined, nlit + ndist, rle);

    uint32_t m_freq[19] = {0};
    for (size_t i = 0; i < nrle; i++) m_freq[rle[i] & 0xFFu]++;
    uint8_t m_lens[19];
    if (!orc_build_code_lengths(m_freq, 19, 7, m_lens)) return 0;
    uint32_t m_codes[19], m_pre[19];
    orc_canonical_codes(m_lens, 19, m_codes);
    for (size_t i = 0; i < 19; i++) m_pre[i] = m_lens[i] ? rev_bits(m_codes[i], m_lens[i]) : 0;
    size_t hclen_count = 19;
    while (hclen_count > 4 && m_lens[META_ORDER[hclen_count - 1]] == 0) hclen_count--;

    bw_put(w, 5u, 3); /* BFINAL=1, BTYPE=10 */
    bw_put(w, (uint32_t)(nlit - 257), 5);
    bw_put(w, (uint32_t)(ndist - 1), 5);
    bw_put(w, (uint32_t)(hclen_count - 4), 4);
    for (size_t i = 0; i < hclen_count; i++) bw_put(w, m_lens[META_ORDER[i]], 3);
    for (size_t i = 0; i < nrle; i++) {
        uint32_t c = rle[i] & 0xFFu, eb = (rle[i] >> 8) & 0xFFu, ev = rle[i] >> 16;
        bw_put(w, m_pre[c], m_lens[c]);
        if (eb) bw_put(w, ev, eb);
    }
    for (size_t i = 0; i < ntok; i++) {
        uint32_t len = tok[i] >> 16;
        if (len == 0) {
            uint32_t s = tok[i] & 0xFFFFu;
            bw_put(w, ll_pre[s], ll_lens[s]);
        } else {
            int li = len_sym(len);
            uint32_t s = 257u + (uint32_t)li;
            bw_put(w, ll_pre[s], ll_lens[s]);
            if (LEN_EXTRA[li]) bw_put(w, len - LEN_BASE[li], LEN_EXTRA[li]);
            uint32_t d = tok[i] & 0xFFFFu;
            int di = dist_sym(d);
            bw_put(w, d_pre[di], d_lens[di]);
            if (DIST_EXTRA[di]) bw_put(w, d - DIST_BASE[di], DIST_EXTRA[di]);
        }
    }
    bw_put(w, ll_pre[256], ll_lens[256]);
    return 1;
}

size_t orc_encode_block_fast(const uint32_t *tok, size_t ntok, uint8_t *out, size_t out_cap) {
    for (size_t i = 0; i < ntok; i++) { /* the alphabet the reference's LUTs cover */
        uint32_t len = tok[i] >> 16, v = tok[i] & 0xFFFFu;
        if (len == 0 ? v > 255 : (len < 3 || len > 258 || v < 1 || v > 32768)) return (size_t)-1;
    }
    bitw w = {out, out_cap, 0, 0, 0, 0};
    if (!write_dynamic_block(tok, ntok, &w)) {
        bitw f = {out, out_cap, 0, 0, 0, 0};
        write_fixed_block(tok, ntok, &f);
        w = f;
    }
    bw_finish(&w);
    return w.overflow ? (size_t)-1 : w.len;
}

/* ------------------------------------------------------------------ */
/* CRC-32 (bitwise definition of the reflected IEEE polynomial). */
uint32_t orc_crc32(const uint8_t *in, size_t n) {
    static uint32_t table[256];
    static int init = 0;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
            table[i] = c;
        }
        init = 1;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ in[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

/* gzip_wrap -- deflate.rs:962-980 */
size_t orc_gzip_wrap(const uint8_t *deflate, size_t dlen, const uint8_t *orig, size_t n,
                     uint8_t *out, size_t out_cap) {
    static const uint8_t hdr[10] = {0x1f, 0x8b, 0x08, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0xff};
    if (dlen + 18 > out_cap) return (size_t)-1;
    memcpy(out, hdr, 10);
    memmove(out + 10, deflate, dlen);
    uint32_t crc = orc_crc32(orig, n), isz = (uint32_t)n;
    for (int i = 0; i < 4; i++) out[10 + dlen + i] = (uint8_t)(crc >> (8 * i));
    for (int i = 0; i < 4; i++) out[14 + dlen + i] = (uint8_t)(isz >> (8 * i));
    return dlen + 18;
}

static size_t member_bound(size_t n) { return 2 * n + 1024 + 18; }

size_t orc_chunk_bound(size_t n, size_t sub_chunk) {
    if (n <= sub_chunk) return member_bound(n);
    size_t members = (n + sub_chunk - 1) / sub_chunk;
    return 2 * n + members * (1024 + 18);
}

/* encode_one -- codec-gpu/src/gpu.rs:271-287 (host encoder path); rule 2
 * walks with orc_lookahead_walk (the ratio mode) instead of greedy_walk. */
size_t orc_gzip_member_rule(const uint8_t *in, size_t n, int rule, uint8_t *out, size_t out_cap) {
    uint32_t *raw = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *walked = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    size_t dcap = member_bound(n);
    uint8_t *defl = (uint8_t *)malloc(dcap);
    size_t res = (size_t)-1;
    if (raw && walked && defl && orc_match_find(in, n, ORC_WINDOW, raw) == 0) {
        size_t nw = rule == 2 ? orc_lookahead_walk(raw, in, n, walked) : orc_greedy_walk(raw, in, n, walked);
        size_t dl = orc_encode_block_fast(walked, nw, defl, dcap);
        if (dl != (size_t)-1) res = orc_gzip_wrap(defl, dl, in, n, out, out_cap);
    }
    free(raw);
    free(walked);
    free(defl);
    return res;
}

size_t orc_gzip_member(const uint8_t *in, size_t n, uint8_t *out, size_t out_cap) {
    return orc_gzip_member_rule(in, n, 0, out, out_cap);
}

/* gzip_chunk_fn -- codec-gpu/src/gpu.rs:193-208 */
size_t orc_gzip_chunk_rule(const uint8_t *in, size_t n, size_t sub_chunk, int rule, uint8_t *out, size_t out_cap) {
    if (sub_chunk == 0) return (size_t)-1;
    if (n <= sub_chunk) return orc_gzip_member_rule(in, n, rule, out, out_cap);
    size_t o = 0;
    for (size_t off = 0; off < n; off += sub_chunk) {
        size_t len = n - off < sub_chunk ? n - off : sub_chunk;
        size_t m = orc_gzip_member_rule(in + off, len, rule, out + o, out_cap - o);
        if (m == (size_t)-1) return m;
        o += m;
    }
    return o;
}

size_t orc_gzip_chunk(const uint8_t *in, size_t n, size_t sub_chunk, uint8_t *out, size_t out_cap) {
    return orc_gzip_chunk_rule(in, n, sub_chunk, 0, out, out_cap);
}
