/*
 * orot_oracle.c -- CPU restatement of orot's parallel DEFLATE compress path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * How each part is pinned (see tests/test_oracle.py):
 *   crc32 / adler32          orot KATs crc32("123456789") = 0xCBF43926,
 *                            adler32("Wikipedia") = 0x11E60398
 *                            (tests/unit/test_deflate_formats.cpp:74-88) and
 *                            Python's zlib.crc32 / zlib.adler32 on long inputs.
 *   *_combine                zlib on split inputs (closed form of the CRC /
 *                            Adler recurrences).
 *   RFC 1951 tables          orot's table KATs (test_deflate_huffman.cpp:21-76).
 *   lz77 L1..L9 tokens       golden token digests written by orot itself
 *                            (tests/golden/, made by oracle/ref_driver.cpp
 *                            from the orot sources), plus a brute-force
 *                            re-derivation of the L1 candidate rule.
 *   L0..L3 output bytes      golden orot payloads / sizes / digests, zlib
 *                            inflate, Appendix-style closed forms (empty input,
 *                            "a").
 *   limited Huffman lengths  brute-force optimum on tiny alphabets and
 *                            Kraft equality.
 *   lazy_parse (L4..L9)      the same golden token digests: orot's standard
 *                            path runs through lazy_parse with its own match
 *                            functions.
 *   L6-class M(x)            brute force of its definition, with and without
 *                            a dictionary, and the 32 KiB window edge.
 *   L6-class tokens          lazy_parse over M(x), and a Python restatement of
 *                            orot's lazy rule over the same M(x).
 *   L4..L12 block encoder    forced stored / fixed / dynamic emissions: each
 *                            inflates and the choice is the smallest; zlib
 *                            and orot's inflater on every kind of output; the
 *                            size within 3% of orot's at every level 4-12 (the
 *                            encoder is the HIP path's own design, not orot's,
 *                            so orot's sizes are the ratio reference only).
 */
#include "orot_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ======================================================================
 * Checksums -- src/simd/simd_detect.cpp:33-70 (scalar definitions; the x86
 * SSE4.2 build computes CRC-32C instead, src/simd/x86/match_sse42.cpp:76-99,
 * which gzip readers reject -- the IEEE polynomial is restated here).
 * ====================================================================== */

static uint32_t crc_tab[8][256];
static pthread_once_t crc_once = PTHREAD_ONCE_INIT;

static void crc_init(void) {
    for (uint32_t i = 0; i < 256; ++i) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
        crc_tab[0][i] = c;
    }
    for (uint32_t i = 0; i < 256; ++i)
        for (int t = 1; t < 8; ++t)
            crc_tab[t][i] = (crc_tab[t - 1][i] >> 8) ^ crc_tab[0][crc_tab[t - 1][i] & 0xFF];
}

uint32_t oracle_crc32(uint32_t crc, const uint8_t* p, size_t n) {
    pthread_once(&crc_once, crc_init);
    uint32_t c = ~crc;
    /* slicing-by-8 (same function as the byte-at-a-time Sarwate loop) */
    while (n >= 8) {
        uint32_t lo = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
        uint32_t hi = (uint32_t)p[4] | ((uint32_t)p[5] << 8) | ((uint32_t)p[6] << 16) | ((uint32_t)p[7] << 24);
        lo ^= c;
        c = crc_tab[7][lo & 0xFF] ^ crc_tab[6][(lo >> 8) & 0xFF] ^ crc_tab[5][(lo >> 16) & 0xFF] ^
            crc_tab[4][lo >> 24] ^ crc_tab[3][hi & 0xFF] ^ crc_tab[2][(hi >> 8) & 0xFF] ^
            crc_tab[1][(hi >> 16) & 0xFF] ^ crc_tab[0][hi >> 24];
        p += 8;
        n -= 8;
    }
    while (n--) c = (c >> 8) ^ crc_tab[0][(c ^ *p++) & 0xFF];
    return ~c;
}

uint32_t oracle_adler32(uint32_t adler, const uint8_t* p, size_t n) {
    uint32_t s1 = adler & 0xFFFF, s2 = adler >> 16;
    while (n > 0) {
        size_t k = n < 5552 ? n : 5552; /* NMAX: no 32-bit overflow before the reduction */
        for (size_t i = 0; i < k; ++i) {
            s1 += p[i];
            s2 += s1;
        }
        s1 %= 65521u;
        s2 %= 65521u;
        p += k;
        n -= k;
    }
    return (s2 << 16) | s1;
}

/* Reflected GF(2) product a*b mod P (bit 31 = x^0). */
static uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
    uint32_t prod = 0;
    for (int i = 0; i < 32; ++i) {
        if (a & 0x80000000u) prod ^= b;
        a <<= 1;
        b = (b >> 1) ^ (0xEDB88320u & (0u - (b & 1u)));
    }
    return prod;
}

/* x^(8*len) mod P by square-and-multiply. */
static uint32_t gf2_x8n(uint64_t len) {
    uint32_t result = 0x80000000u; /* x^0 */
    uint32_t sq = 0x00800000u;     /* x^8 */
    while (len) {
        if (len & 1) result = gf2_mulmod(result, sq);
        sq = gf2_mulmod(sq, sq);
        len >>= 1;
    }
    return result;
}

uint32_t oracle_crc32_combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b) {
    return gf2_mulmod(gf2_x8n(len_b), crc_a) ^ crc_b;
}

uint32_t oracle_adler32_combine(uint32_t adler_a, uint32_t adler_b, uint64_t len_b) {
    const uint32_t M = 65521u;
    uint32_t a1 = adler_a & 0xFFFF, a2 = adler_a >> 16;
    uint32_t b1 = adler_b & 0xFFFF, b2 = adler_b >> 16;
    uint32_t rem = (uint32_t)(len_b % M);
    uint32_t s1 = (a1 + b1 + M - 1) % M;
    uint32_t s2 = (uint32_t)(((uint64_t)rem * a1 + a2 + b2 + M - rem) % M);
    return (s2 << 16) | s1;
}

/* ======================================================================
 * Bounds -- src/formats/raw_deflate.cpp:9-12, src/api/deflate_api.cpp:81-98
 * ====================================================================== */

size_t oracle_compress_bound(size_t n, int format) {
    size_t b = n + (n >> 12) + (n >> 14) + (n >> 25) + 13;
    if (format == 1) b += 6;
    if (format == 2) b += 18;
    return b;
}

/* ======================================================================
 * RFC 1951 tables -- src/deflate/huffman.cpp:15-53, 79-119, 125-159
 * ====================================================================== */

static const uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                      2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
static const uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
static const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                                       6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
static const uint16_t DIST_BASE[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,    49,    65,    97,    129,
                                       193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};

static int len_code(int len) { /* 3..258 -> 0..28 */
    int c = 0;
    while (c < 28 && LEN_BASE[c + 1] <= len) ++c;
    return c;
}
static int dist_code(int d) { /* 1..32768 -> 0..29 */
    int c = 0;
    while (c < 29 && DIST_BASE[c + 1] <= d) ++c;
    return c;
}

static uint8_t fixed_ll_len(int sym) {
    if (sym < 144) return 8;
    if (sym < 256) return 9;
    if (sym < 280) return 7;
    return 8;
}

static uint16_t rev_bits(uint32_t code, int len) {
    uint32_t r = 0;
    for (int i = 0; i < len; ++i) r |= ((code >> i) & 1u) << (len - 1 - i);
    return (uint16_t)r;
}

/* Canonical codes, RFC 1951 section 3.2.2, bit-reversed for LSB-first output
 * (src/deflate/huffman.cpp:374-397). */
static void canon_codes(const uint8_t* lens, int n, uint16_t* codes) {
    int bl_count[16] = {0};
    uint32_t next[17] = {0};
    for (int i = 0; i < n; ++i)
        if (lens[i]) bl_count[lens[i]]++;
    uint32_t code = 0;
    for (int b = 1; b <= 15; ++b) {
        code = (code + (uint32_t)bl_count[b - 1]) << 1;
        next[b] = code;
    }
    for (int i = 0; i < n; ++i) codes[i] = lens[i] ? rev_bits(next[lens[i]]++, lens[i]) : 0;
}

/* ======================================================================
 * Bit buffer (src/deflate/bit_writer.hpp:27-85): LSB-first concatenation.
 * The flush granularity of orot's writer does not change the bytes, so the
 * restatement ORs bits into a zero-filled byte buffer.
 * ====================================================================== */

typedef struct {
    uint8_t* buf;
    uint64_t pos; /* bit position */
} bitbuf;

static void put_bits(bitbuf* b, uint32_t v, int n) {
    for (int i = 0; i < n; ++i) {
        if ((v >> i) & 1u) b->buf[(b->pos + (uint64_t)i) >> 3] |= (uint8_t)(1u << ((b->pos + (uint64_t)i) & 7));
    }
    b->pos += (uint64_t)n;
}

static void align_byte(bitbuf* b) { b->pos = (b->pos + 7) & ~(uint64_t)7; }

/* ======================================================================
 * LZ77 -- src/deflate/lz77.cpp (level table :18-32, fast path :163-191 and
 * :378-406, standard path :66-156, :199-207, :435-496).
 * ====================================================================== */

typedef struct {
    int max_chain, nice_len, lazy_depth, bt4, fast_path, hash_bits, miss_limit;
} lz_cfg;

static const lz_cfg LEVELS[13] = {
    {0, 0, 0, 0, 0, 16, 8},       {1, 32, 0, 0, 1, 12, 8},      {16, 64, 0, 0, 1, 14, 8},
    {32, 128, 0, 0, 1, 14, 8},    {64, 128, 1, 0, 0, 16, 5},    {128, 128, 1, 0, 0, 16, 5},
    {128, 258, 1, 0, 0, 16, 5},   {256, 258, 2, 0, 0, 16, 6},   {512, 258, 2, 0, 0, 16, 6},
    {768, 258, 3, 0, 0, 16, 6},   {1024, 258, 4, 1, 0, 16, 8},  {2048, 258, 4, 1, 0, 16, 8},
    {4096, 258, 4, 1, 0, 16, 8},
};

static int clamp_level(int level) { return level < 0 ? 0 : (level > 12 ? 12 : level); }

static uint32_t le32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

static uint32_t hash4(const uint8_t* p, int bits) { return (le32(p) * 0x9E3779B1u) >> (32 - bits); }

/* Common-prefix length capped at max_len (src/simd/simd_detect.cpp:12-31). */
static int lcp(const uint8_t* a, const uint8_t* b, int max_len) {
    int l = 0;
    while (l < max_len && a[l] == b[l]) ++l;
    return l;
}

#define WIN_MASK 0x7FFF

static size_t lz77_fast(const uint8_t* src, int n, int bits, uint32_t* head, uint32_t* out) {
    size_t nt = 0;
    int pos = 0;
    while (pos < n) {
        int len = 0, dist = 0;
        if (pos + 3 <= n && n - pos >= 4) {
            const uint32_t h = hash4(src + pos, bits);
            const uint32_t cur = head[h];
            head[h] = (uint32_t)pos; /* the probed position is inserted at once */
            if (cur != 0) {
                const int d = (pos - (int)cur) & WIN_MASK;
                /* first-byte check reads src[cur], not src[pos - d] (lz77.cpp:184) */
                if (d != 0 && src[cur] == src[pos]) {
                    const int l = lcp(src + pos, src + pos - d, n - pos < 258 ? n - pos : 258);
                    if (l >= 3) {
                        len = l;
                        dist = d;
                    }
                }
            }
        }
        if (len < 3) {
            out[nt++] = src[pos++];
        } else {
            out[nt++] = ORACLE_TOK_MATCH | ((uint32_t)len << 16) | (uint32_t)dist;
            for (int i = pos + 1; i < pos + len && i + 4 <= n; ++i) head[hash4(src + i, bits)] = (uint32_t)i;
            pos += len;
        }
    }
    return nt;
}

/* The fast path over a chunk that follows `halo` bytes of its buffer (the
 * L1-class dictionary mode of the HIP path; orot's own dictionary insertion,
 * lz77.cpp:43-55, hashes with 16 bits into tables the fast path reads with
 * hash_bits and stores dictionary offsets, so orot has no usable L1-class
 * dictionary to follow).  Positions count from the dictionary's first byte,
 * which is the empty sentinel as position 0 is; every dictionary position is
 * in the table before the chunk's first probe, exactly as every earlier chunk
 * position is; tokens cover the chunk only. */
static size_t lz77_fast_h(const uint8_t* src, int halo, int n, int bits, uint32_t* head, uint32_t* out) {
    const int nx = halo + n;
    for (int q = 0; q < halo && q + 4 <= nx; ++q) head[hash4(src + q, bits)] = (uint32_t)q;
    size_t nt = 0;
    int pos = halo;
    while (pos < nx) {
        int len = 0, dist = 0;
        if (nx - pos >= 4) {
            const uint32_t h = hash4(src + pos, bits);
            const uint32_t cur = head[h];
            head[h] = (uint32_t)pos;
            if (cur != 0) {
                const int d = (pos - (int)cur) & WIN_MASK;
                if (d != 0 && src[cur] == src[pos]) {
                    const int l = lcp(src + pos, src + pos - d, nx - pos < 258 ? nx - pos : 258);
                    if (l >= 3) {
                        len = l;
                        dist = d;
                    }
                }
            }
        }
        if (len < 3) {
            out[nt++] = src[pos++];
        } else {
            out[nt++] = ORACLE_TOK_MATCH | ((uint32_t)len << 16) | (uint32_t)dist;
            for (int i = pos + 1; i < pos + len && i + 4 <= nx; ++i) head[hash4(src + i, bits)] = (uint32_t)i;
            pos += len;
        }
    }
    return nt;
}

size_t oracle_l1_lz77_h(const uint8_t* src, size_t halo, size_t n, int level, uint32_t* tokens) {
    const lz_cfg* c = &LEVELS[clamp_level(level)];
    if (!c->fast_path || n == 0) return 0;
    uint32_t* head = (uint32_t*)calloc((size_t)1 << c->hash_bits, sizeof(uint32_t));
    if (!head) return 0;
    const size_t r = lz77_fast_h(src, (int)halo, (int)n, c->hash_bits, head, tokens);
    free(head);
    return r;
}

/* Chain walk of the standard path (lz77.cpp:66-156).  The 8-byte and
 * byte-at-best_len rejects of the original only skip candidates that cannot
 * beat best_len, so they are omitted without changing the result. */
static int match_find(const uint8_t* src, int pos, int n, const uint32_t* head, const uint32_t* prev, int chain,
                      int nice, int miss_limit, int* best_dist) {
    const int max_match = n - pos < 258 ? n - pos : 258;
    *best_dist = 0;
    if (max_match < 3 || n - pos < 4) return 0;
    int best = 2;
    uint32_t cur = head[hash4(src + pos, 16)];
    int misses = 0;
    while (chain-- > 0 && cur != 0) {
        const int d = (pos - (int)cur) & WIN_MASK;
        if (d == 0) break;
        const uint8_t* cand = src + pos - d;
        if (le32(cand) != le32(src + pos)) {
            if (++misses >= miss_limit) break;
            cur = prev[cur & WIN_MASK];
            continue;
        }
        misses = 0;
        const int l = lcp(src + pos, cand, max_match);
        if (l > best) {
            best = l;
            *best_dist = d;
            if (l >= nice) break;
        }
        cur = prev[cur & WIN_MASK];
    }
    return best;
}

static void insert16(const uint8_t* src, int pos, int n, uint32_t* head, uint32_t* prev) {
    if (pos + 4 > n) return;
    const uint32_t h = hash4(src + pos, 16);
    prev[pos & WIN_MASK] = head[h];
    head[h] = (uint32_t)pos;
}

/* The lazy parse of the standard path (lz77.cpp:435-496), over two match
 * functions of the position alone (A.3: before the probe at p every q < p
 * with q + 4 <= n has been inserted, in increasing order, whatever the parse
 * did):
 *   main(p)  = the best match at p with all q < p inserted;
 *   probe(p) = the lazy probe at p + 1 with that same table (p itself not
 *              yet inserted).
 * A length below 3 means no match.  orot's standard path and the HIP L6-class
 * matcher differ only in these two functions, so the orot token digests of
 * levels 4-9 pin this loop for both. */
typedef struct {
    void* ctx;
    int (*main)(void* ctx, int p, int* dist);
    int (*probe)(void* ctx, int p, int* dist);
} lz_matches;

static size_t lazy_parse(const uint8_t* src, int n, int lazy_depth, const lz_matches* m, uint32_t* out) {
    size_t nt = 0;
    int pos = 0;
    while (pos < n) {
        if (pos + 3 > n) {
            out[nt++] = src[pos++];
            continue;
        }
        int bd = 0;
        int bl = m->main(m->ctx, pos, &bd);
        if (bl < 3) {
            out[nt++] = src[pos++];
            continue;
        }
        /* a literal is emitted while the probe at pos + 1 beats the match by
         * more than one byte, at most lazy_depth times (lz77.cpp:460-480) */
        int k = lazy_depth;
        while (k > 0 && pos + bl + 1 < n && pos + 1 + 3 <= n) {
            int nd = 0;
            const int nl = m->probe(m->ctx, pos, &nd);
            if (nl <= bl + 1) break;
            out[nt++] = src[pos++];
            bl = nl;
            bd = nd;
            --k;
        }
        out[nt++] = ORACLE_TOK_MATCH | ((uint32_t)bl << 16) | (uint32_t)bd;
        pos += bl;
    }
    return nt;
}

/* orot's match functions: the hash chains are brought up to date (every
 * q < p inserted) before each probe. */
typedef struct {
    const uint8_t* src;
    int n;
    const lz_cfg* c;
    uint32_t* head;
    uint32_t* prev;
    int inserted; /* positions [0, inserted) are in the chains */
} std_ctx;

static void std_sync(std_ctx* s, int p) {
    for (; s->inserted < p; ++s->inserted) insert16(s->src, s->inserted, s->n, s->head, s->prev);
}

static int std_main(void* ctx, int p, int* dist) {
    std_ctx* s = (std_ctx*)ctx;
    std_sync(s, p);
    return match_find(s->src, p, s->n, s->head, s->prev, s->c->max_chain, s->c->nice_len, s->c->miss_limit, dist);
}

static int std_probe(void* ctx, int p, int* dist) {
    std_ctx* s = (std_ctx*)ctx;
    std_sync(s, p);
    const int lazy_chain = (s->c->max_chain >> 1) > 1 ? (s->c->max_chain >> 1) : 1;
    return match_find(s->src, p + 1, s->n, s->head, s->prev, lazy_chain, s->c->nice_len, s->c->miss_limit, dist);
}

static size_t lz77_std(const uint8_t* src, int n, const lz_cfg* c, uint32_t* head, uint32_t* prev, uint32_t* out) {
    std_ctx s = {src, n, c, head, prev, 0};
    const lz_matches m = {&s, std_main, std_probe};
    return lazy_parse(src, n, c->lazy_depth, &m, out);
}

size_t oracle_lz77(const uint8_t* src, size_t n, int level, uint32_t* tokens) {
    const lz_cfg* c = &LEVELS[clamp_level(level)];
    if (n == 0) return 0;
    if (c->max_chain == 0) {
        for (size_t i = 0; i < n; ++i) tokens[i] = src[i];
        return n;
    }
    if (c->bt4) return (size_t)-1;
    if (c->fast_path) {
        uint32_t* head = (uint32_t*)calloc((size_t)1 << c->hash_bits, sizeof(uint32_t));
        if (!head) return (size_t)-1;
        size_t r = lz77_fast(src, (int)n, c->hash_bits, head, tokens);
        free(head);
        return r;
    }
    uint32_t* head = (uint32_t*)calloc(65536 + 32768, sizeof(uint32_t));
    if (!head) return (size_t)-1;
    size_t r = lz77_std(src, (int)n, c, head, head + 65536, tokens);
    free(head);
    return r;
}

void oracle_l1_candidates(const uint8_t* src, size_t n, int hash_bits, uint16_t* d16) {
    uint32_t* head = (uint32_t*)calloc((size_t)1 << hash_bits, sizeof(uint32_t));
    for (size_t p = 0; p < n; ++p) {
        d16[p] = 0;
        if (p + 4 > n) continue;
        const uint32_t h = hash4(src + p, hash_bits);
        const uint32_t cur = head[h];
        head[h] = (uint32_t)p;
        if (cur == 0) continue;
        const uint32_t d = ((uint32_t)p - cur) & WIN_MASK;
        if (d != 0 && src[cur] == src[p]) d16[p] = (uint16_t)d;
    }
    free(head);
}

/* ======================================================================
 * HIP L6-class matcher (levels 4..12).  Not orot's parse: a GPU-friendly
 * definition with the same shape as orot's standard path (lz77.cpp:435-496):
 *   C(p)     latest q < p, q + 4 <= n, with h12(q) == h12(p), 0 if none
 *            (every earlier position is in the table);
 *   M(x)     longest match among the chain C(x), C(C(x)), ... of at most
 *            max_chain candidates with x - cand <= 32768; a candidate counts
 *            only if its first 4 bytes equal x's; ties keep the nearer one;
 *            the walk stops at nice_len;
 *   parse    lazy, lazy_depth steps, the probe at pos + 1 being M(pos + 1):
 *            a literal is emitted when M(pos + 1) > best + 1 (orot's rule).
 * Cross-chunk dictionary: a chunk that is not the first of its buffer sees
 * the up to 32 KiB before it (the "halo") as positions -H..-1 -- they are in
 * the chains and matches may start there -- while tokens are made for the
 * chunk's own positions only.  orot's parallel path starts every chunk
 * empty (src/parallel/parallel_compressor.cpp:56-70).
 * ====================================================================== */

typedef struct {
    int max_chain, nice_len, lazy_depth;
} l6_cfg;

static const l6_cfg L6LEVELS[13] = {
    {0, 0, 0},    {0, 0, 0},    {0, 0, 0},    {0, 0, 0},     {16, 128, 1},  {32, 128, 1},   {32, 258, 1},
    {64, 258, 2}, {128, 258, 2}, {256, 258, 3}, {512, 258, 3}, {1024, 258, 3}, {2048, 258, 3},
};

#define L6_HASH_BITS 12

#define L6_HALO 32768

/* src points at the halo's first byte; mlen / mdist get n entries, for
 * positions halo .. halo + n - 1. */
void oracle_l6_match_h(const uint8_t* src, size_t halo, size_t n, int level, uint32_t* mlen, uint32_t* mdist) {
    const l6_cfg* c = &L6LEVELS[clamp_level(level)];
    const size_t nx = halo + n;
    uint32_t* head = (uint32_t*)calloc((size_t)1 << L6_HASH_BITS, sizeof(uint32_t));
    uint32_t* C = (uint32_t*)calloc(nx ? nx : 1, sizeof(uint32_t));
    for (size_t p = 0; p < nx; ++p) {
        C[p] = 0;
        if (p + 4 > nx) continue;
        const uint32_t h = hash4(src + p, L6_HASH_BITS);
        C[p] = head[h];
        head[h] = (uint32_t)p + 1;
    }
    for (size_t i = 0; i < n; ++i) {
        const size_t x = halo + i;
        uint32_t best = 0, bd = 0;
        if (x + 4 <= nx) {
            const int cap = nx - x < 258 ? (int)(nx - x) : 258;
            uint32_t cur = C[x];
            for (int k = 0; k < c->max_chain && cur; ++k) {
                const uint32_t q = cur - 1;
                if (x - q > 32768) break;
                if (le32(src + q) == le32(src + x)) {
                    const uint32_t l = (uint32_t)lcp(src + x, src + q, cap);
                    if (l > best) {
                        best = l;
                        bd = (uint32_t)(x - q);
                        if ((int)l >= c->nice_len) break;
                    }
                }
                cur = C[q];
            }
        }
        mlen[i] = best >= 3 ? best : 0;
        mdist[i] = best >= 3 ? bd : 0;
    }
    free(C);
    free(head);
}

void oracle_l6_match(const uint8_t* src, size_t n, int level, uint32_t* mlen, uint32_t* mdist) {
    oracle_l6_match_h(src, 0, n, level, mlen, mdist);
}

/* The L6-class match functions: M(x) from a table, the probe at p + 1 being
 * M(p + 1). */
typedef struct {
    const uint32_t* ml;
    const uint32_t* md;
    int n;
} l6_ctx;

static int l6_main(void* ctx, int p, int* dist) {
    const l6_ctx* s = (const l6_ctx*)ctx;
    *dist = (int)s->md[p];
    return (int)s->ml[p];
}

static int l6_probe(void* ctx, int p, int* dist) {
    const l6_ctx* s = (const l6_ctx*)ctx;
    if (p + 1 >= s->n) {
        *dist = 0;
        return 0;
    }
    *dist = (int)s->md[p + 1];
    return (int)s->ml[p + 1];
}

size_t oracle_l6_lz77_h(const uint8_t* src, size_t halo, size_t n, int level, uint32_t* tokens) {
    const l6_cfg* c = &L6LEVELS[clamp_level(level)];
    if (n == 0) return 0;
    uint32_t* ml = (uint32_t*)calloc(2 * n, sizeof(uint32_t));
    uint32_t* md = ml + n;
    oracle_l6_match_h(src, halo, n, level, ml, md);
    l6_ctx s = {ml, md, (int)n};
    const lz_matches m = {&s, l6_main, l6_probe};
    const size_t nt = lazy_parse(src + halo, (int)n, c->lazy_depth, &m, tokens);
    free(ml);
    return nt;
}

size_t oracle_l6_lz77(const uint8_t* src, size_t n, int level, uint32_t* tokens) {
    return oracle_l6_lz77_h(src, 0, n, level, tokens);
}

/* ======================================================================
 * Block statistics and orot's Fixed/Stored emission
 * (src/deflate/deflate_block.cpp:21-70, 107-183, 389-474).
 * ====================================================================== */

typedef struct {
    uint32_t lit[288];
    uint32_t dist[32];
    uint64_t raw;
} blk_stats;

static void block_stats(const uint32_t* t, size_t nt, blk_stats* s) {
    memset(s, 0, sizeof(*s));
    for (size_t i = 0; i < nt; ++i) {
        if (t[i] & ORACLE_TOK_MATCH) {
            const int len = (int)((t[i] >> 16) & 0x1FF), d = (int)(t[i] & 0xFFFF);
            s->lit[257 + len_code(len)]++;
            s->dist[dist_code(d)]++;
            s->raw += (uint64_t)len;
        } else {
            s->lit[t[i] & 0xFF]++;
            s->raw++;
        }
    }
    s->lit[256]++;
}

static uint64_t fixed_bits(const blk_stats* s) {
    uint64_t b = 3;
    for (int i = 0; i < 288; ++i) b += (uint64_t)fixed_ll_len(i) * s->lit[i];
    for (int i = 0; i < 32; ++i) b += 5ull * s->dist[i];
    for (int i = 0; i < 29; ++i) b += (uint64_t)LEN_EXTRA[i] * s->lit[257 + i];
    for (int i = 0; i < 30; ++i) b += (uint64_t)DIST_EXTRA[i] * s->dist[i];
    return b;
}

static void emit_stored(bitbuf* b, const uint8_t* raw, size_t len, int is_last) {
    size_t off = 0;
    do {
        const size_t piece = len - off < 65535 ? len - off : 65535;
        const int last = is_last && off + piece == len;
        put_bits(b, (uint32_t)last, 1);
        put_bits(b, 0, 2);
        align_byte(b);
        put_bits(b, (uint32_t)piece & 0xFFFF, 16);
        put_bits(b, (~(uint32_t)piece) & 0xFFFF, 16);
        memcpy(b->buf + (b->pos >> 3), raw + off, piece);
        b->pos += (uint64_t)piece * 8;
        off += piece;
        if (piece == 0) break;
    } while (off < len);
}

static void emit_tokens(bitbuf* b, const uint32_t* t, size_t nt, const uint8_t* ll, const uint16_t* llc,
                        const uint8_t* dl, const uint16_t* dlc) {
    for (size_t i = 0; i < nt; ++i) {
        if (t[i] & ORACLE_TOK_MATCH) {
            const int len = (int)((t[i] >> 16) & 0x1FF), d = (int)(t[i] & 0xFFFF);
            const int lc = len_code(len), dc = dist_code(d);
            put_bits(b, llc[257 + lc], ll[257 + lc]);
            if (LEN_EXTRA[lc]) put_bits(b, (uint32_t)(len - LEN_BASE[lc]), LEN_EXTRA[lc]);
            put_bits(b, dlc[dc], dl[dc]);
            if (DIST_EXTRA[dc]) put_bits(b, (uint32_t)(d - DIST_BASE[dc]), DIST_EXTRA[dc]);
        } else {
            put_bits(b, llc[t[i] & 0xFF], ll[t[i] & 0xFF]);
        }
    }
    put_bits(b, llc[256], ll[256]);
}

static void fixed_tables(uint8_t* ll, uint16_t* llc, uint8_t* dl, uint16_t* dlc) {
    for (int i = 0; i < 288; ++i) ll[i] = fixed_ll_len(i);
    for (int i = 0; i < 32; ++i) dl[i] = 5;
    canon_codes(ll, 288, llc);
    canon_codes(dl, 32, dlc);
}

int oracle_encode_block_fixed(const uint32_t* tokens, size_t ntok, const uint8_t* raw, size_t raw_len, int is_last,
                              uint8_t* dst, uint64_t* bitpos) {
    bitbuf b = {dst, *bitpos};
    blk_stats s;
    block_stats(tokens, ntok, &s);
    const uint64_t stored = ((uint64_t)raw_len + 5) * 8;
    int is_stored = stored <= fixed_bits(&s);
    if (is_stored) {
        emit_stored(&b, raw, raw_len, is_last);
    } else {
        uint8_t ll[288], dl[32];
        uint16_t llc[288], dlc[32];
        fixed_tables(ll, llc, dl, dlc);
        put_bits(&b, (uint32_t)(is_last ? 1 : 0), 1);
        put_bits(&b, 1, 2);
        emit_tokens(&b, tokens, ntok, ll, llc, dl, dlc);
    }
    *bitpos = b.pos;
    return is_stored;
}

/* ======================================================================
 * Length-limited Huffman by package-merge (Larmore & Hirschberg 1990).
 * Symbols are ordered by (freq, index); on equal weight a leaf precedes a
 * package.  Replaces orot's clamped heap Huffman (src/deflate/huffman.cpp:
 * 180-360), which can emit over-subscribed codes.
 * ====================================================================== */

int oracle_limited_lengths(const uint32_t* freq, int nsym, int max_bits, uint8_t* lens) {
    int idx[288];
    int n = 0;
    memset(lens, 0, (size_t)nsym);
    for (int i = 0; i < nsym; ++i)
        if (freq[i]) idx[n++] = i;
    if (n == 0) return 0;
    if (n == 1) {
        lens[idx[0]] = 1;
        return 1;
    }
    /* insertion sort by (freq, index): stable, deterministic */
    for (int i = 1; i < n; ++i) {
        int v = idx[i], j = i - 1;
        while (j >= 0 && freq[idx[j]] > freq[v]) {
            idx[j + 1] = idx[j];
            --j;
        }
        idx[j + 1] = v;
    }
    static __thread uint64_t wt[16][576];
    static __thread uint8_t leaf[16][576];
    int cnt[16];
    /* level max_bits: leaves only */
    for (int i = 0; i < n; ++i) {
        wt[max_bits][i] = freq[idx[i]];
        leaf[max_bits][i] = 1;
    }
    cnt[max_bits] = n;
    for (int l = max_bits - 1; l >= 1; --l) {
        const int np = cnt[l + 1] / 2;
        int a = 0, p = 0, k = 0;
        while (a < n || p < np) {
            const uint64_t pw = p < np ? wt[l + 1][2 * p] + wt[l + 1][2 * p + 1] : UINT64_MAX;
            if (a < n && (uint64_t)freq[idx[a]] <= pw) {
                wt[l][k] = freq[idx[a++]];
                leaf[l][k++] = 1;
            } else {
                wt[l][k] = pw;
                leaf[l][k++] = 0;
                ++p;
            }
        }
        cnt[l] = k;
    }
    int take = 2 * n - 2;
    for (int l = 1; l <= max_bits && take > 0; ++l) {
        int leaves = 0;
        for (int i = 0; i < take; ++i) leaves += leaf[l][i];
        for (int i = 0; i < leaves; ++i) lens[idx[i]]++;
        take = 2 * (take - leaves);
    }
    return n;
}

/* ======================================================================
 * Dynamic block (HIP L6-class encoder).  Code-length run-length rules follow
 * orot's encode_code_lengths (src/deflate/deflate_block.cpp:196-290):
 * runs capped at 138; zero runs >= 3 -> 17 / 18; non-zero runs >= 4 -> the
 * symbol then 16 x min(run-1, 6), the remainder re-scanned.
 * ====================================================================== */

static const uint8_t CL_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

typedef struct {
    int hlit, hdist, hclen, nrle;
    uint8_t rle_sym[320];
    uint8_t rle_x[320];
    uint8_t cl[19];
    uint16_t clc[19];
} dyn_header;

static void build_header(const uint8_t* ll_in, const uint8_t* dl_in, uint8_t* ll, uint8_t* dl, dyn_header* h) {
    memcpy(ll, ll_in, 288);
    memcpy(dl, dl_in, 32);
    if (!ll[256]) ll[256] = 1;
    int any = 0;
    for (int i = 0; i < 32; ++i) any |= dl[i];
    if (!any) dl[0] = 1;
    int hlit = 286;
    while (hlit > 257 && !ll[hlit - 1]) --hlit;
    int hdist = 30;
    while (hdist > 1 && !dl[hdist - 1]) --hdist;
    uint8_t comb[320];
    memcpy(comb, ll, (size_t)hlit);
    memcpy(comb + hlit, dl, (size_t)hdist);
    const int total = hlit + hdist;
    int nr = 0;
    for (int i = 0; i < total;) {
        const uint8_t s = comb[i];
        int run = 1;
        while (i + run < total && comb[i + run] == s && run < 138) ++run;
        if (s == 0 && run >= 3) {
            h->rle_sym[nr] = run <= 10 ? 17 : 18;
            h->rle_x[nr++] = (uint8_t)(run <= 10 ? run - 3 : run - 11);
            i += run;
        } else if (s != 0 && run >= 4) {
            h->rle_sym[nr] = s;
            h->rle_x[nr++] = 0xFF;
            const int rep = run - 1 < 6 ? run - 1 : 6;
            h->rle_sym[nr] = 16;
            h->rle_x[nr++] = (uint8_t)(rep - 3);
            i += 1 + rep;
        } else {
            h->rle_sym[nr] = s;
            h->rle_x[nr++] = 0xFF;
            ++i;
        }
    }
    uint32_t clf[19] = {0};
    for (int i = 0; i < nr; ++i) clf[h->rle_sym[i]]++;
    const int act = oracle_limited_lengths(clf, 19, 7, h->cl);
    if (act == 1) { /* a one-code code-length alphabet must still be complete */
        for (int i = 0; i < 19; ++i)
            if (!h->cl[i]) {
                h->cl[i] = 1;
                break;
            }
    }
    canon_codes(h->cl, 19, h->clc);
    int hclen = 19;
    while (hclen > 4 && !h->cl[CL_ORDER[hclen - 1]]) --hclen;
    h->hlit = hlit;
    h->hdist = hdist;
    h->hclen = hclen;
    h->nrle = nr;
}

static uint64_t header_bits(const dyn_header* h) {
    uint64_t b = 3 + 5 + 5 + 4 + 3ull * (uint64_t)h->hclen;
    for (int i = 0; i < h->nrle; ++i) {
        const int s = h->rle_sym[i];
        b += h->cl[s];
        if (s == 16) b += 2;
        if (s == 17) b += 3;
        if (s == 18) b += 7;
    }
    return b;
}

static uint64_t data_bits(const blk_stats* s, const uint8_t* ll, const uint8_t* dl) {
    uint64_t b = 0;
    for (int i = 0; i < 288; ++i) b += (uint64_t)ll[i] * s->lit[i];
    for (int i = 0; i < 32; ++i) b += (uint64_t)dl[i] * s->dist[i];
    for (int i = 0; i < 29; ++i) b += (uint64_t)LEN_EXTRA[i] * s->lit[257 + i];
    for (int i = 0; i < 30; ++i) b += (uint64_t)DIST_EXTRA[i] * s->dist[i];
    return b;
}

uint64_t oracle_dynamic_block_bits(const uint32_t* lit_freq, const uint32_t* dist_freq, const uint8_t* ll_in,
                                   const uint8_t* dl_in) {
    blk_stats s;
    memcpy(s.lit, lit_freq, sizeof(s.lit));
    memcpy(s.dist, dist_freq, sizeof(s.dist));
    uint8_t ll[288], dl[32];
    dyn_header h;
    build_header(ll_in, dl_in, ll, dl, &h);
    return header_bits(&h) + data_bits(&s, ll, dl);
}

static void emit_dynamic(bitbuf* b, const uint32_t* t, size_t nt, const uint8_t* ll_in, const uint8_t* dl_in,
                         int is_last) {
    uint8_t ll[288], dl[32];
    uint16_t llc[288], dlc[32];
    dyn_header h;
    build_header(ll_in, dl_in, ll, dl, &h);
    put_bits(b, (uint32_t)(is_last ? 1 : 0), 1);
    put_bits(b, 2, 2);
    put_bits(b, (uint32_t)(h.hlit - 257), 5);
    put_bits(b, (uint32_t)(h.hdist - 1), 5);
    put_bits(b, (uint32_t)(h.hclen - 4), 4);
    for (int i = 0; i < h.hclen; ++i) put_bits(b, h.cl[CL_ORDER[i]], 3);
    for (int i = 0; i < h.nrle; ++i) {
        const int s = h.rle_sym[i];
        put_bits(b, h.clc[s], h.cl[s]);
        if (s == 16) put_bits(b, h.rle_x[i], 2);
        if (s == 17) put_bits(b, h.rle_x[i], 3);
        if (s == 18) put_bits(b, h.rle_x[i], 7);
    }
    canon_codes(ll, 288, llc);
    canon_codes(dl, 32, dlc);
    emit_tokens(b, t, nt, ll, llc, dl, dlc);
}

void oracle_encode_block_forced(const uint32_t* tokens, size_t ntok, const uint8_t* raw, int is_last, int btype,
                                uint8_t* dst, uint64_t* bitpos) {
    bitbuf b = {dst, *bitpos};
    blk_stats s;
    block_stats(tokens, ntok, &s);
    if (btype == 0) {
        emit_stored(&b, raw, (size_t)s.raw, is_last);
    } else if (btype == 1) {
        uint8_t fl[288], fd[32];
        uint16_t flc[288], fdc[32];
        fixed_tables(fl, flc, fd, fdc);
        put_bits(&b, (uint32_t)(is_last ? 1 : 0), 1);
        put_bits(&b, 1, 2);
        emit_tokens(&b, tokens, ntok, fl, flc, fd, fdc);
    } else {
        uint8_t ll[288], dl[32];
        oracle_limited_lengths(s.lit, 288, 15, ll);
        oracle_limited_lengths(s.dist, 32, 15, dl);
        emit_dynamic(&b, tokens, ntok, ll, dl, is_last);
    }
    *bitpos = b.pos;
}

void oracle_encode_blocks_dyn(const uint32_t* tokens, size_t ntok, const uint8_t* raw, size_t raw_len, int is_last,
                              int max_syms, uint8_t* dst, uint64_t* bitpos) {
    bitbuf b = {dst, *bitpos};
    size_t t0 = 0, r0 = 0;
    if (max_syms <= 0) max_syms = 1 << 30;
    do {
        const size_t nt = ntok - t0 < (size_t)max_syms ? ntok - t0 : (size_t)max_syms;
        const int last = is_last && t0 + nt == ntok;
        blk_stats s;
        block_stats(tokens + t0, nt, &s);
        uint8_t ll[288], dl[32];
        oracle_limited_lengths(s.lit, 288, 15, ll);
        oracle_limited_lengths(s.dist, 32, 15, dl);
        const uint64_t fb = fixed_bits(&s);
        const uint64_t db = oracle_dynamic_block_bits(s.lit, s.dist, ll, dl);
        const uint64_t sb = (s.raw + 5) * 8;
        if (sb <= fb && sb <= db) {
            emit_stored(&b, raw + r0, (size_t)s.raw, last);
        } else if (fb <= db) {
            uint8_t fl[288], fd[32];
            uint16_t flc[288], fdc[32];
            fixed_tables(fl, flc, fd, fdc);
            put_bits(&b, (uint32_t)(last ? 1 : 0), 1);
            put_bits(&b, 1, 2);
            emit_tokens(&b, tokens + t0, nt, fl, flc, fd, fdc);
        } else {
            emit_dynamic(&b, tokens + t0, nt, ll, dl, last);
        }
        t0 += nt;
        r0 += (size_t)s.raw;
    } while (t0 < ntok);
    (void)raw_len;
    *bitpos = b.pos;
}

/* ======================================================================
 * Parallel driver -- src/parallel/parallel_compressor.cpp:16-150 with the
 * thread pool of src/parallel/thread_pool.hpp:17-88 reduced to a static
 * partition of chunks over pthreads.  Phase structure is kept: chunk set-up
 * (allocation and zero fill) on the caller, parallel LZ77, serial emission,
 * serial checksum.
 * ====================================================================== */

#define DYN_MAX_SYMS 16384

typedef struct {
    const uint8_t* src;
    size_t len;
    size_t halo;  /* L6 class: bytes before src that are the chunk's dictionary */
    uint32_t* tokens;
    size_t ntok;
    int level;
    int l6gpu;
} chunk_job;

typedef struct {
    chunk_job* jobs;
    size_t njobs;
    size_t next;
    pthread_mutex_t mu;
} job_queue;

static void* worker(void* arg) {
    job_queue* q = (job_queue*)arg;
    for (;;) {
        pthread_mutex_lock(&q->mu);
        size_t i = q->next++;
        pthread_mutex_unlock(&q->mu);
        if (i >= q->njobs) break;
        chunk_job* j = &q->jobs[i];
        if (j->l6gpu) {
            j->ntok = oracle_l6_lz77_h(j->src - j->halo, j->halo, j->len, j->level, j->tokens);
        } else if (j->halo) {
            const int bits = LEVELS[clamp_level(j->level)].hash_bits;
            uint32_t* head = (uint32_t*)calloc((size_t)1 << bits, sizeof(uint32_t));
            j->ntok = head ? lz77_fast_h(j->src - j->halo, (int)j->halo, (int)j->len, bits, head, j->tokens) : 0;
            free(head);
        } else {
            j->ntok = oracle_lz77(j->src, j->len, j->level, j->tokens);
        }
    }
    return NULL;
}

static size_t default_block_size(int level) {
    if (level <= 3) return 131072;
    if (level <= 6) return 524288;
    return 1048576;
}

static size_t parallel_compress_impl(int level, int format, int threads, size_t block_size, const uint8_t* src,
                                     size_t n, uint8_t* dst, size_t cap, int l6gpu, int whole, int l1dict);

size_t oracle_parallel_compress(int level, int format, int threads, size_t block_size, const uint8_t* src, size_t n,
                                uint8_t* dst, size_t cap) {
    return parallel_compress_impl(level, format, threads, block_size, src, n, dst, cap, 0, 0, 0);
}

size_t oracle_l1_dict_compress(int level, int format, int threads, size_t block_size, const uint8_t* src, size_t n,
                               uint8_t* dst, size_t cap) {
    const int lv = clamp_level(level);
    if (lv < 1 || lv > 3) return oracle_gpu_parallel_compress(level, format, threads, block_size, src, n, dst, cap);
    return parallel_compress_impl(level, format, threads, block_size, src, n, dst, cap, 0, 0, 1);
}

/* The L6 class takes chunks of at most 1 MiB: a larger block runs as 1 MiB
 * pieces, each with the 32 KiB before it as its dictionary. */
#define L6_MAX_CHUNK ((size_t)1 << 20)

size_t oracle_gpu_parallel_compress(int level, int format, int threads, size_t block_size, const uint8_t* src,
                                    size_t n, uint8_t* dst, size_t cap) {
    const int l6 = clamp_level(level) >= 4;
    if (l6 && block_size > L6_MAX_CHUNK) block_size = L6_MAX_CHUNK;
    return parallel_compress_impl(level, format, threads, block_size, src, n, dst, cap, l6, 0, 0);
}

size_t oracle_whole_compress(int level, int format, const uint8_t* src, size_t n, uint8_t* dst, size_t cap) {
    const int lv = clamp_level(level);
    if (cap < oracle_compress_bound(n, format)) return 0; /* raw_deflate.cpp:20, zlib/gzip wrappers */
    size_t r;
    if (lv >= 4) r = oracle_gpu_parallel_compress(lv, format, 1, 0, src, n, dst, cap);
    else r = parallel_compress_impl(lv, format, 1, n ? n : 1, src, n, dst, cap, 0, 1, 0);
    if (r && format == 2) dst[8] = (uint8_t)(lv >= 9 ? 2 : (lv <= 2 ? 4 : 0)); /* gzip_wrapper.cpp:31 */
    return r;
}

static size_t parallel_compress_impl(int level, int format, int threads, size_t block_size, const uint8_t* src,
                                     size_t n, uint8_t* dst, size_t cap, int l6gpu, int whole, int l1dict) {
    const int lv = clamp_level(level);
    if (LEVELS[lv].bt4 && !l6gpu) return 0;
    if (block_size == 0) block_size = default_block_size(level);
    const size_t nb = n ? (n + block_size - 1) / block_size : 1;
    chunk_job* jobs = (chunk_job*)calloc(nb, sizeof(chunk_job));
    if (!jobs) return 0;
    int ok = 1;
    for (size_t i = 0; i < nb; ++i) {
        jobs[i].src = src + i * block_size;
        jobs[i].len = n ? (n - i * block_size < block_size ? n - i * block_size : block_size) : 0;
        jobs[i].level = level;
        jobs[i].l6gpu = l6gpu;
        jobs[i].halo = (l6gpu || l1dict) ? (i * block_size < L6_HALO ? i * block_size : L6_HALO) : 0;
        jobs[i].tokens = (uint32_t*)calloc(jobs[i].len ? jobs[i].len : 1, sizeof(uint32_t));
        if (!jobs[i].tokens) ok = 0;
    }
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if ((size_t)threads > nb) threads = (int)nb;
    job_queue q = {jobs, nb, 0, PTHREAD_MUTEX_INITIALIZER};
    if (ok) {
        pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
        for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, worker, &q);
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
        free(th);
    }
    size_t written = 0;
    if (ok) {
        /* header (parallel_compressor.cpp:86-101) */
        if (format == 1) {
            if (cap < 6) ok = 0;
            else {
                dst[0] = 0x78;
                dst[1] = 0x9C;
                written = 2;
            }
        } else if (format == 2) {
            static const uint8_t gz[10] = {0x1F, 0x8B, 8, 0, 0, 0, 0, 0, 0, 255};
            if (cap < 10) ok = 0;
            else {
                memcpy(dst, gz, 10);
                written = 10;
            }
        }
    }
    const size_t trailer = format == 1 ? 4 : (format == 2 ? 8 : 0);
    if (ok && written + trailer > cap) ok = 0;
    if (ok) {
        /* serial emission into a zero-filled scratch sized at the bound */
        const size_t scratch = oracle_compress_bound(n, 0) + 64 + nb * 8;
        uint8_t* tmp = (uint8_t*)calloc(scratch, 1);
        if (!tmp) ok = 0;
        else {
            uint64_t bp = 0;
            for (size_t i = 0; i < nb; ++i) {
                const int last = i == nb - 1;
                if (lv == 0) {
                    bitbuf b = {tmp, bp};
                    emit_stored(&b, jobs[i].src, jobs[i].len, last);
                    bp = b.pos;
                } else if (whole && jobs[i].ntok == jobs[i].len) {
                    /* whole buffer, no match at all: stored (block_compressor.cpp:60-61) */
                    bitbuf b = {tmp, bp};
                    emit_stored(&b, jobs[i].src, jobs[i].len, last);
                    bp = b.pos;
                } else if (LEVELS[lv].fast_path) {
                    oracle_encode_block_fixed(jobs[i].tokens, jobs[i].ntok, jobs[i].src, jobs[i].len, last, tmp, &bp);
                } else {
                    oracle_encode_blocks_dyn(jobs[i].tokens, jobs[i].ntok, jobs[i].src, jobs[i].len, last,
                                             DYN_MAX_SYMS, tmp, &bp);
                }
            }
            const size_t raw = (size_t)((bp + 7) >> 3);
            if (written + raw + trailer > cap) ok = 0;
            else {
                memcpy(dst + written, tmp, raw);
                written += raw;
            }
            free(tmp);
        }
    }
    if (ok && format == 1) {
        const uint32_t a = oracle_adler32(1, src, n);
        dst[written++] = (uint8_t)(a >> 24);
        dst[written++] = (uint8_t)(a >> 16);
        dst[written++] = (uint8_t)(a >> 8);
        dst[written++] = (uint8_t)a;
    } else if (ok && format == 2) {
        const uint32_t c = oracle_crc32(0, src, n);
        const uint32_t isz = (uint32_t)n;
        for (int k = 0; k < 4; ++k) dst[written++] = (uint8_t)(c >> (8 * k));
        for (int k = 0; k < 4; ++k) dst[written++] = (uint8_t)(isz >> (8 * k));
    }
    for (size_t i = 0; i < nb; ++i) free(jobs[i].tokens);
    free(jobs);
    return ok ? written : 0;
}
