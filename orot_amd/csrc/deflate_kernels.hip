// deflate_kernels.hip -- gfx950 kernels for orot's parallel DEFLATE compress
// path (src/parallel/parallel_compressor.cpp:36-150).
//
// Pipeline for one input buffer split into chunks of `chunk` bytes (one
// DEFLATE block per chunk, fresh match tables per chunk, as orot does):
//
//   checksum_kernel   one workgroup per chunk: CRC-32 register or Adler-32
//                     partial sums of the chunk (lane slices + GF(2) shifts).
//   l1_lz77_kernel    one workgroup (8 waves) per chunk, levels 1-3:
//                     A) candidate distance for every position (orot's
//                        single-probe head table, lz77.cpp:163-191), built
//                        by 8 waves over 8 sub-ranges with private LDS
//                        tables and a prefix fix-up;
//                     B) greedy parse (lz77.cpp:378-406) as 512 speculative
//                        lane segments plus a serial resolve of the segment
//                        entries on wave 0 with wave-wide match lengths;
//                     C) fixed-Huffman bit counts, block-wide exclusive scan,
//                        stored/fixed choice (deflate_block.cpp:414-421) and
//                        emission into a per-chunk staging bitstream.
//   finalize_kernel   bit offset of every chunk (stored blocks depend on the
//                     incoming offset mod 8), checksum combine, framing.
//   assemble_kernel   one workgroup per chunk: shifts the chunk's bits to its
//                     offset in the output (stored chunks copy the input).
#include "deflate_gpu.h"

namespace orot_gpu {
namespace {

constexpr uint32_t CRC_POLY = 0xEDB88320u;
constexpr uint32_t ADLER_MOD = 65521u;
constexpr int L1_THREADS = 512;
constexpr int L1_WAVES = L1_THREADS / 64;
constexpr uint32_t MAX_CHUNK_L1 = 256u * 1024u;  // 12-bit tables: 8 sub-ranges of 32 KiB

// ---------------------------------------------------------------------------
// GF(2) arithmetic for CRC-32 combination (reflected: bit 31 = x^0).
// ---------------------------------------------------------------------------
constexpr uint32_t gf2_mul_ce(uint32_t a, uint32_t b) {
    uint32_t p = 0;
    for (int i = 0; i < 32; ++i) {
        if (a & 0x80000000u) p ^= b;
        a <<= 1;
        b = (b >> 1) ^ (CRC_POLY & (0u - (b & 1u)));
    }
    return p;
}

struct PowTable {
    uint32_t v[64];  // v[k] = x^(8 * 2^k) mod P
};
constexpr PowTable make_pow_table() {
    PowTable t{};
    uint32_t x = 0x00800000u;  // x^8
    for (int k = 0; k < 64; ++k) {
        t.v[k] = x;
        x = gf2_mul_ce(x, x);
    }
    return t;
}
struct CrcTable {
    uint32_t t[4][256];
};
constexpr CrcTable make_crc_table() {
    CrcTable c{};
    for (uint32_t i = 0; i < 256; ++i) {
        uint32_t v = i;
        for (int k = 0; k < 8; ++k) v = (v >> 1) ^ (CRC_POLY & (0u - (v & 1u)));
        c.t[0][i] = v;
    }
    for (uint32_t i = 0; i < 256; ++i)
        for (int s = 1; s < 4; ++s) c.t[s][i] = (c.t[s - 1][i] >> 8) ^ c.t[0][c.t[s - 1][i] & 0xFF];
    return c;
}

__constant__ PowTable c_pow = make_pow_table();
__constant__ CrcTable c_crc = make_crc_table();

__device__ __forceinline__ uint32_t gf2_mul(uint32_t a, uint32_t b) {
    uint32_t p = 0;
#pragma unroll 8
    for (int i = 0; i < 32; ++i) {
        if (a & 0x80000000u) p ^= b;
        a <<= 1;
        b = (b >> 1) ^ (CRC_POLY & (0u - (b & 1u)));
    }
    return p;
}

// x^(8*len) mod P
__device__ uint32_t gf2_x8n(uint64_t len) {
    uint32_t r = 0x80000000u;
    int k = 0;
    while (len) {
        if (len & 1) r = gf2_mul(c_pow.v[k], r);
        len >>= 1;
        ++k;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ld32u(const uint8_t* p) {
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}
__device__ __forceinline__ uint64_t ld64u(const uint8_t* p) {
    uint64_t v;
    __builtin_memcpy(&v, p, 8);
    return v;
}
__device__ __forceinline__ uint64_t ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }

template <int HB>
__device__ __forceinline__ uint32_t hash4(uint32_t v) {
    return (v * 0x9E3779B1u) >> (32 - HB);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Fixed Huffman symbols (RFC 1951 3.2.6), LSB-first (bit-reversed) codes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lit_bits(uint32_t b) { return b < 144 ? 8u : 9u; }

__device__ __forceinline__ void fixed_lit(uint32_t b, uint32_t& v, uint32_t& len) {
    const uint32_t code = b < 144 ? 0x30u + b : 0x190u + (b - 144u);
    len = lit_bits(b);
    v = __builtin_bitreverse32(code) >> (32 - len);
}

struct LenSym {
    uint32_t code, ebits, extra;
};
// match length 3..258 -> (code index 0..28, extra bit count, extra value)
__device__ __forceinline__ LenSym len_sym(uint32_t L) {
    if (L <= 10) return {L - 3, 0, 0};
    if (L == 258) return {28, 0, 0};
    const uint32_t x = L - 3;
    const uint32_t e = 31 - __builtin_clz(x) - 2;
    return {4 * e + (x >> e), e, x & ((1u << e) - 1u)};
}
// distance 1..32768 -> (code 0..29, extra bit count, extra value)
__device__ __forceinline__ LenSym dist_sym(uint32_t d) {
    if (d <= 4) return {d - 1, 0, 0};
    const uint32_t x = d - 1;
    const uint32_t e = 31 - __builtin_clz(x) - 1;
    return {2 * e + 2 + ((x >> e) & 1u), e, x & ((1u << e) - 1u)};
}

__device__ __forceinline__ uint32_t match_bits(uint32_t L, uint32_t d) {
    const LenSym ls = len_sym(L), ds = dist_sym(d);
    return (ls.code < 23 ? 7u : 8u) + ls.ebits + 5u + ds.ebits;  // symbols 257..279 are 7-bit
}

__device__ __forceinline__ void fixed_match(uint32_t L, uint32_t d, uint32_t& v, uint32_t& len) {
    const LenSym ls = len_sym(L), ds = dist_sym(d);
    const uint32_t sym = 257 + ls.code;
    const uint32_t cl = sym < 280 ? 7u : 8u;
    const uint32_t code = sym < 280 ? sym - 256u : 0xC0u + (sym - 280u);
    const uint32_t rc = __builtin_bitreverse32(code) >> (32 - cl);
    const uint32_t rd = __builtin_bitreverse32(ds.code) >> 27;
    v = rc | (ls.extra << cl) | (rd << (cl + ls.ebits)) | (ds.extra << (cl + ls.ebits + 5));
    len = cl + ls.ebits + 5 + ds.ebits;
}

// ---------------------------------------------------------------------------
// Match length helpers (src/simd/simd_detect.cpp:12-31 semantics: exact
// common-prefix length capped at `cap`).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lcp_lane(const uint8_t* a, const uint8_t* b, uint32_t cap) {
    uint32_t l = 0;
    while (l + 8 <= cap) {
        const uint64_t x = ld64u(a + l) ^ ld64u(b + l);
        if (x) return l + (uint32_t)(__builtin_ctzll(x) >> 3);
        l += 8;
    }
    while (l < cap && a[l] == b[l]) ++l;
    return l;
}

// Whole wave compares 256 bytes per round; every lane returns the same value.
__device__ uint32_t lcp_wave(const uint8_t* a, const uint8_t* b, uint32_t cap, int lane) {
    for (uint32_t r = 0; r < cap; r += 256) {
        const uint32_t off = r + 4u * (uint32_t)lane;
        uint32_t x = 0;
        if (off + 4 <= cap) {
            x = ld32u(a + off) ^ ld32u(b + off);
        } else if (off < cap) {
            for (uint32_t k = 0; k < cap - off; ++k) x |= (uint32_t)(a[off + k] ^ b[off + k]) << (8 * k);
        }
        const uint64_t bm = ballot(x != 0);
        if (bm) {
            const int f = __builtin_ctzll(bm);
            const uint32_t xf = __shfl(x, f, 64);
            return r + 4u * (uint32_t)f + (uint32_t)(__builtin_ctz(xf) >> 3);
        }
    }
    return cap;
}

// Greedy match length at a token start (orot match_find_fast, lz77.cpp:163-191
// with the candidate precomputed in D).
__device__ __forceinline__ uint32_t m1_lane(const uint8_t* s, uint32_t n, const uint16_t* D, uint32_t pos) {
    if (pos + 4 > n) return 0;
    const uint32_t d = D[pos];
    if (!d) return 0;
    const uint32_t cap = min(258u, n - pos);
    const uint32_t l = lcp_lane(s + pos, s + pos - d, cap);
    return l >= 3 ? l : 0;
}

__device__ __forceinline__ uint32_t m1_wave(const uint8_t* s, uint32_t n, const uint16_t* D, uint32_t pos, int lane) {
    if (pos + 4 > n) return 0;
    const uint32_t d = D[pos];
    if (!d) return 0;
    const uint32_t cap = min(258u, n - pos);
    const uint32_t l = lcp_wave(s + pos, s + pos - d, cap, lane);
    return l >= 3 ? l : 0;
}

__device__ __forceinline__ void clear_bits(uint32_t* bm, uint32_t x, uint32_t y, int lane) {
    if (x >= y) return;
    const uint32_t w0 = x >> 5, w1 = (y - 1) >> 5;
    for (uint32_t w = w0 + (uint32_t)lane; w <= w1; w += 64) {
        const uint32_t lo = (w == w0) ? (x & 31u) : 0u;
        const uint32_t hi = (w == w1) ? ((y - 1) & 31u) : 31u;
        const uint32_t mask = (hi == 31u ? 0xFFFFFFFFu : ((2u << hi) - 1u)) & ~((1u << lo) - 1u);
        bm[w] &= ~mask;
    }
}

// ---------------------------------------------------------------------------
// Checksum partials per chunk.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void checksum_kernel(const uint8_t* __restrict__ src, uint64_t n_total,
                                                       uint32_t chunk, int kind, ChunkInfo* __restrict__ info) {
    __shared__ uint32_t T[4][256];
    __shared__ uint32_t r0[256];
    __shared__ uint32_t r1[256];
    const int tid = threadIdx.x;
    const uint32_t c = blockIdx.x;
    const uint64_t base = (uint64_t)c * chunk;
    const uint32_t len = (uint32_t)min<uint64_t>(chunk, n_total - base);
    const uint8_t* s = src + base;
    const uint32_t seg = (((len + 255u) / 256u) + 15u) & ~15u;
    const uint32_t a = min(len, (uint32_t)tid * seg), b = min(len, a + seg);
    const uint32_t after = len - b;
    if (kind == 0) {
        for (int i = tid; i < 1024; i += 256) T[i >> 8][i & 255] = c_crc.t[i >> 8][i & 255];
        __syncthreads();
        uint32_t crc = 0;
        uint32_t p = a;
        for (; p + 4 <= b; p += 4) {
            crc ^= ld32u(s + p);
            crc = T[3][crc & 0xFF] ^ T[2][(crc >> 8) & 0xFF] ^ T[1][(crc >> 16) & 0xFF] ^ T[0][crc >> 24];
        }
        for (; p < b; ++p) crc = (crc >> 8) ^ T[0][(crc ^ s[p]) & 0xFF];
        r0[tid] = gf2_mul(gf2_x8n(after), crc);
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) r0[tid] ^= r0[tid + w];
            __syncthreads();
        }
        if (tid == 0) info[c].crc = r0[0];
    } else {
        uint32_t s1 = 0, s2 = 0;
        uint32_t p = a;
        while (p < b) {
            const uint32_t e = min(b, p + 4096u);
            for (; p + 4 <= e; p += 4) {
                const uint32_t w = ld32u(s + p);
                s1 += w & 0xFF;
                s2 += s1;
                s1 += (w >> 8) & 0xFF;
                s2 += s1;
                s1 += (w >> 16) & 0xFF;
                s2 += s1;
                s1 += w >> 24;
                s2 += s1;
            }
            for (; p < e; ++p) {
                s1 += s[p];
                s2 += s1;
            }
            s1 %= ADLER_MOD;
            s2 %= ADLER_MOD;
        }
        r0[tid] = s1;
        r1[tid] = (uint32_t)(((uint64_t)s1 * (after % ADLER_MOD) + s2) % ADLER_MOD);
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) {
                r0[tid] = (r0[tid] + r0[tid + w]) % ADLER_MOD;
                r1[tid] = (r1[tid] + r1[tid + w]) % ADLER_MOD;
            }
            __syncthreads();
        }
        if (tid == 0) {
            info[c].adler_s1 = r0[0];
            info[c].adler_s2 = r1[0];
        }
    }
}

// ---------------------------------------------------------------------------
// L1-class LZ77 + fixed-Huffman emission, one workgroup per chunk.
// ---------------------------------------------------------------------------
template <int HB>
struct L1Cfg {
    static constexpr int W = HB == 12 ? 8 : 4;  // sub-ranges (one wave each)
    static constexpr uint32_t TAB = 1u << HB;
    static constexpr uint32_t TAB_BYTES = W * TAB * 2u;
};

constexpr uint32_t L1_BM_WORDS = MAX_CHUNK_L1 / 32;
constexpr uint32_t L1_PHASE_B_BYTES = (L1_BM_WORDS + L1_THREADS + 16) * 4;

template <int HB>
__global__ __launch_bounds__(L1_THREADS) void l1_lz77_kernel(const uint8_t* __restrict__ src, uint64_t n_total,
                                                             uint32_t chunk, uint32_t nchunks,
                                                             uint16_t* __restrict__ d16, uint16_t* __restrict__ m1g,
                                                             uint32_t* __restrict__ stage, uint64_t stage_words,
                                                             ChunkInfo* __restrict__ info) {
    extern __shared__ uint32_t smem[];
    constexpr int W = L1Cfg<HB>::W;
    constexpr uint32_t TAB = L1Cfg<HB>::TAB;
    const uint32_t c = blockIdx.x;
    const uint64_t base = (uint64_t)c * chunk;
    const uint32_t n = (uint32_t)min<uint64_t>(chunk, n_total - base);
    const uint8_t* s = src + base;
    uint16_t* D = d16 + base;
    uint16_t* M = m1g + base;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    // ---- Phase A1: latest earlier position with the same hash, within each
    // sub-range.  D holds (local index + 1) of that position, 0 if none.
    const uint32_t nh = n >= 4 ? n - 3 : 0;  // positions p with p + 4 <= n are hashed
    const uint32_t SR = (((nh + W - 1) / W) + 63u) & ~63u;
    uint16_t* tabs = reinterpret_cast<uint16_t*>(smem);
    if (wv < W) {
        uint16_t* tab = tabs + (uint32_t)wv * TAB;
        for (uint32_t i = lane; i < TAB / 2; i += 64) reinterpret_cast<uint32_t*>(tab)[i] = 0;
        const uint32_t a = (uint32_t)wv * SR;
        const uint32_t b = min(a + SR, nh);
        for (uint32_t p0 = a; p0 < b; p0 += 64) {
            const uint32_t p = p0 + (uint32_t)lane;
            const bool valid = p < b;
            const uint32_t h = valid ? hash4<HB>(ld32u(s + p)) : 0u;
            uint64_t m = ballot(valid);
#pragma unroll
            for (int k = 0; k < HB; ++k) {
                const bool bit = (h >> k) & 1u;
                const uint64_t bk = ballot(bit);
                m &= bit ? bk : ~bk;
            }
            if (valid) {
                const uint64_t below = m & ((1ull << lane) - 1ull);
                uint32_t res;
                if (below) res = p0 + (uint32_t)(63 - __builtin_clzll(below)) - a + 1u;
                else res = tab[h];
                if ((m >> lane) == 1ull) tab[h] = (uint16_t)(p - a + 1u);
                D[p] = (uint16_t)res;
            }
        }
    }
    __syncthreads();

    // ---- Phase A2: resolve first occurrences against earlier sub-ranges,
    // then apply orot's candidate rules: C == 0 is "empty", distance masked
    // to 15 bits, first byte compared at C (lz77.cpp:178-184).
    for (uint32_t p = tid; p < n; p += L1_THREADS) {
        uint32_t out = 0;
        if (p < nh) {
            const uint32_t w = p / SR;
            const uint32_t t = D[p];
            uint32_t C = 0;
            bool have = false;
            if (t) {
                C = w * SR + t - 1u;
                have = true;
            } else {
                const uint32_t h = hash4<HB>(ld32u(s + p));
                for (int w2 = (int)w - 1; w2 >= 0; --w2) {
                    const uint32_t u = tabs[(uint32_t)w2 * TAB + h];
                    if (u) {
                        C = (uint32_t)w2 * SR + u - 1u;
                        have = true;
                        break;
                    }
                }
            }
            if (have && C != 0) {
                const uint32_t d = (p - C) & 0x7FFFu;
                if (d && s[C] == s[p]) out = d;
            }
        }
        D[p] = (uint16_t)out;
    }
    __syncthreads();

    // ---- Phase B: speculative greedy parse per lane segment.
    uint32_t* bm = smem;
    uint32_t* exit_spec = smem + L1_BM_WORDS;
    uint32_t* scan = exit_spec + L1_THREADS;
    const uint32_t S = max(32u, (((n + L1_THREADS - 1) / L1_THREADS) + 31u) & ~31u);
    const uint32_t NS = (n + S - 1) / S;
    const uint32_t nbw = (n + 31) >> 5;
    for (uint32_t i = tid; i < nbw; i += L1_THREADS) bm[i] = 0;
    __syncthreads();
    if ((uint32_t)tid < NS) {
        uint32_t pos = (uint32_t)tid * S;
        const uint32_t end = min(pos + S, n);
        while (pos < end) {
            bm[pos >> 5] |= 1u << (pos & 31);
            const uint32_t l = m1_lane(s, n, D, pos);
            M[pos] = (uint16_t)l;
            pos += l ? l : 1u;
        }
        exit_spec[tid] = pos;
    }
    __syncthreads();

    // ---- Serial resolve of segment entries (wave 0, lockstep).
    if (wv == 0 && NS > 1) {
        uint32_t e = exit_spec[0];
        uint32_t si = 1;
        while (si < NS) {
            if (e == si * S) {
                const uint32_t j = si + (uint32_t)lane;
                const bool mm = (j >= NS) || (exit_spec[j] != (j + 1) * S);
                const uint64_t bmask = ballot(mm);
                if (bmask == 0) {
                    si += 64;
                    e = si * S;
                    continue;
                }
                const uint32_t jj = si + (uint32_t)__builtin_ctzll(bmask);
                if (jj >= NS) break;
                e = exit_spec[jj];
                si = jj + 1;
                continue;
            }
            const uint32_t st = si * S, en = min(st + S, n);
            if (e >= en) {
                clear_bits(bm, st, en, lane);
                ++si;
                continue;
            }
            clear_bits(bm, st, e, lane);
            uint32_t pos = e, ne;
            for (;;) {
                if (bm[pos >> 5] & (1u << (pos & 31))) {
                    ne = exit_spec[si];
                    break;
                }
                if (lane == 0) bm[pos >> 5] |= 1u << (pos & 31);
                const uint32_t l = m1_wave(s, n, D, pos, lane);
                if (lane == 0) M[pos] = (uint16_t)l;
                const uint32_t np = pos + (l ? l : 1u);
                clear_bits(bm, pos + 1, min(np, en), lane);
                if (np >= en) {
                    ne = np;
                    break;
                }
                pos = np;
            }
            e = ne;
            ++si;
        }
    }
    __syncthreads();

    // ---- Phase C: bit counts, scan, block choice, emission.
    const uint32_t st = (uint32_t)tid * S;
    const uint32_t en = (uint32_t)tid < NS ? min(st + S, n) : st;
    uint32_t mybits = 0;
    for (uint32_t w = st >> 5; w < ((en + 31) >> 5); ++w) {
        uint32_t x = bm[w];
        while (x) {
            const uint32_t p = (w << 5) + (uint32_t)__builtin_ctz(x);
            x &= x - 1;
            const uint32_t l = M[p];
            mybits += l ? match_bits(l, D[p]) : lit_bits(s[p]);
        }
    }
    const uint32_t incl = wave_incl_scan(mybits, lane);
    if (lane == 63) scan[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < L1_WAVES; ++i) {
            const uint32_t t = scan[i];
            scan[i] = acc;
            acc += t;
        }
        scan[L1_WAVES] = acc;
    }
    __syncthreads();
    const uint32_t excl = scan[wv] + incl - mybits;
    const uint64_t fixed = 3ull + scan[L1_WAVES] + 7ull;
    const bool is_last = c == nchunks - 1;
    const bool stored = (uint64_t)(n + 5) * 8ull <= fixed;
    if (tid == 0) {
        info[c].fixed_bits = fixed;
        info[c].raw_len = n;
        info[c].is_stored = stored ? 1u : 0u;
    }
    if (stored) return;
    uint32_t* stg = stage + (uint64_t)c * stage_words;
    const uint32_t nw = (uint32_t)((fixed + 31) >> 5) + 1u;
    for (uint32_t i = tid; i < nw; i += L1_THREADS) stg[i] = 0;
    __syncthreads();
    if (tid == 0) atomicOr(&stg[0], (is_last ? 1u : 0u) | 2u);  // BFINAL, BTYPE = 01
    if (mybits) {
        const uint32_t bitpos = 3u + excl;
        uint32_t wi = bitpos >> 5;
        uint32_t nacc = bitpos & 31u;
        bool partial = nacc != 0;
        uint64_t acc = 0;
        for (uint32_t w = st >> 5; w < ((en + 31) >> 5); ++w) {
            uint32_t x = bm[w];
            while (x) {
                const uint32_t p = (w << 5) + (uint32_t)__builtin_ctz(x);
                x &= x - 1;
                const uint32_t l = M[p];
                uint32_t v, len;
                if (l) fixed_match(l, D[p], v, len);
                else fixed_lit(s[p], v, len);
                acc |= (uint64_t)v << nacc;
                nacc += len;
                if (nacc >= 32) {
                    if (partial) {
                        atomicOr(&stg[wi], (uint32_t)acc);
                        partial = false;
                    } else {
                        stg[wi] = (uint32_t)acc;
                    }
                    ++wi;
                    acc >>= 32;
                    nacc -= 32;
                }
            }
        }
        if (nacc) atomicOr(&stg[wi], (uint32_t)acc);
    }
}

__global__ void mark_stored_kernel(uint64_t n_total, uint32_t chunk, uint32_t nchunks, ChunkInfo* info) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const uint64_t base = (uint64_t)c * chunk;
    info[c].raw_len = (uint32_t)min<uint64_t>(chunk, n_total - base);
    info[c].is_stored = 1;
    info[c].fixed_bits = 0;
}

// ---------------------------------------------------------------------------
// Scan of chunk offsets, checksum combine, framing.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t stored_bits(uint64_t off, uint32_t L) {
    const uint64_t pieces = L == 0 ? 1 : (L + 65534u) / 65535u;
    const uint64_t pad0 = (8 - ((off + 3) & 7)) & 7;
    return 3 + pad0 + 32 * pieces + 8 * (pieces - 1) + 8ull * L;
}

constexpr int FIN_THREADS = 1024;

__global__ __launch_bounds__(FIN_THREADS) void finalize_kernel(ChunkInfo* __restrict__ info, uint32_t nchunks,
                                                               uint64_t n, uint32_t chunk, int format,
                                                               uint8_t* __restrict__ out, uint64_t cap,
                                                               Result* __restrict__ res) {
    __shared__ uint64_t sz[FIN_THREADS];
    __shared__ uint32_t st[FIN_THREADS];
    __shared__ uint32_t ln[FIN_THREADS];
    __shared__ uint32_t red0[FIN_THREADS];
    __shared__ uint32_t red1[FIN_THREADS];
    __shared__ uint64_t carry;
    const int tid = threadIdx.x;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < nchunks; t0 += FIN_THREADS) {
        const uint32_t i = t0 + (uint32_t)tid;
        if (i < nchunks) {
            sz[tid] = info[i].fixed_bits;
            st[tid] = info[i].is_stored;
            ln[tid] = info[i].raw_len;
        }
        __syncthreads();
        if (tid == 0) {
            uint64_t off = carry;
            const uint32_t cnt = min((uint32_t)FIN_THREADS, nchunks - t0);
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint64_t b = st[k] ? stored_bits(off, ln[k]) : sz[k];
                sz[k] = off;
                off += b;
            }
            carry = off;
        }
        __syncthreads();
        if (i < nchunks) info[i].bit_off = sz[tid];
        __syncthreads();
    }
    const uint64_t payload_bits = carry;
    // checksum combine: chunk i covers [i*chunk, i*chunk + len_i)
    uint32_t x0 = 0, x1 = 0;
    for (uint32_t i = (uint32_t)tid; i < nchunks; i += FIN_THREADS) {
        const uint64_t endi = min<uint64_t>(n, (uint64_t)(i + 1) * chunk);
        const uint64_t after = n - endi;
        if (format == 2) {
            x0 ^= gf2_mul(gf2_x8n(after), info[i].crc);
        } else if (format == 1) {
            const uint32_t s1 = info[i].adler_s1;
            x0 = (x0 + s1) % ADLER_MOD;
            x1 = (uint32_t)((x1 + (uint64_t)s1 * (after % ADLER_MOD) + info[i].adler_s2) % ADLER_MOD);
        }
    }
    red0[tid] = x0;
    red1[tid] = x1;
    __syncthreads();
    for (int w = FIN_THREADS / 2; w > 0; w >>= 1) {
        if (tid < w) {
            if (format == 2) {
                red0[tid] ^= red0[tid + w];
            } else {
                red0[tid] = (red0[tid] + red0[tid + w]) % ADLER_MOD;
                red1[tid] = (red1[tid] + red1[tid + w]) % ADLER_MOD;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        uint32_t ck = 0;
        if (format == 2) {
            ck = ~(gf2_mul(gf2_x8n(n), 0xFFFFFFFFu) ^ red0[0]);
        } else if (format == 1) {
            const uint32_t s1 = (1u + red0[0]) % ADLER_MOD;
            const uint32_t s2 = (uint32_t)((n % ADLER_MOD + red1[0]) % ADLER_MOD);
            ck = (s2 << 16) | s1;
        }
        const uint32_t hdr = format == 1 ? 2u : (format == 2 ? 10u : 0u);
        const uint32_t trl = format == 1 ? 4u : (format == 2 ? 8u : 0u);
        const uint64_t payload = (payload_bits + 7) >> 3;
        const uint64_t total = hdr + payload + trl;
        res->payload_bits = payload_bits;
        res->checksum = ck;
        res->total_bytes = total <= cap ? total : 0;
        if (total <= cap) {
            if (format == 1) {
                out[0] = 0x78;
                out[1] = 0x9C;
                uint8_t* t = out + hdr + payload;
                t[0] = (uint8_t)(ck >> 24);
                t[1] = (uint8_t)(ck >> 16);
                t[2] = (uint8_t)(ck >> 8);
                t[3] = (uint8_t)ck;
            } else if (format == 2) {
                const uint8_t g[10] = {0x1F, 0x8B, 8, 0, 0, 0, 0, 0, 0, 255};
                for (int k = 0; k < 10; ++k) out[k] = g[k];
                uint8_t* t = out + hdr + payload;
                for (int k = 0; k < 4; ++k) t[k] = (uint8_t)(ck >> (8 * k));
                for (int k = 0; k < 4; ++k) t[4 + k] = (uint8_t)((uint32_t)n >> (8 * k));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Output assembly.
// ---------------------------------------------------------------------------
struct ChunkView {
    const uint32_t* stg;
    const uint8_t* raw;
    uint64_t off;
    uint32_t len;
    uint32_t stored;
    uint32_t last;
};

__device__ __forceinline__ uint32_t stored_byte(const ChunkView& cv, uint64_t j) {
    const uint32_t L = cv.len;
    const uint32_t L0 = min(L, 65535u);
    const uint32_t pieces = L == 0 ? 1u : (L + 65534u) / 65535u;
    uint32_t k, q, Lk;
    uint64_t dbase;
    if (j < 4ull + L0) {
        k = 0;
        q = (uint32_t)j + 1u;  // piece-relative index with a virtual header byte at 0
        Lk = L0;
        dbase = 0;
    } else {
        const uint64_t r = j - 4 - L0;
        k = 1u + (uint32_t)(r / 65540u);
        q = (uint32_t)(r % 65540u);
        Lk = min(L - k * 65535u, 65535u);
        dbase = (uint64_t)k * 65535u;
    }
    if (q == 0) return (cv.last && k == pieces - 1) ? 1u : 0u;  // header byte of pieces >= 1
    if (q == 1) return Lk & 0xFFu;
    if (q == 2) return (Lk >> 8) & 0xFFu;
    if (q == 3) return (~Lk) & 0xFFu;
    if (q == 4) return ((~Lk) >> 8) & 0xFFu;
    return cv.raw[dbase + q - 5u];
}

// k (<= 8) bits of the chunk's bitstream starting at relative bit r.
__device__ __forceinline__ uint32_t chunk_bits(const ChunkView& cv, uint64_t r, uint32_t k) {
    const uint32_t mask = (1u << k) - 1u;
    if (!cv.stored) {
        const uint64_t w = r >> 5;
        const uint64_t v = (uint64_t)cv.stg[w] | ((uint64_t)cv.stg[w + 1] << 32);
        return (uint32_t)(v >> (r & 31)) & mask;
    }
    const uint64_t pad0 = (8 - ((cv.off + 3) & 7)) & 7;
    if (r < 3 + pad0) {
        const uint32_t pieces = cv.len == 0 ? 1u : (cv.len + 65534u) / 65535u;
        const uint32_t hv = (cv.last && pieces == 1) ? 1u : 0u;
        return (r == 0 ? hv : 0u) & mask;
    }
    return stored_byte(cv, (r - 3 - pad0) >> 3) & mask;
}

__device__ __forceinline__ ChunkView make_view(const uint8_t* src, uint32_t chunk, uint32_t nchunks,
                                               const uint32_t* stage, uint64_t stage_words, const ChunkInfo* info,
                                               uint32_t c) {
    ChunkView v;
    v.stg = stage + (uint64_t)c * stage_words;
    v.raw = src + (uint64_t)c * chunk;
    v.off = info[c].bit_off;
    v.len = info[c].raw_len;
    v.stored = info[c].is_stored;
    v.last = c == nchunks - 1 ? 1u : 0u;
    return v;
}

__global__ __launch_bounds__(256) void assemble_kernel(const uint8_t* __restrict__ src, uint32_t chunk,
                                                       uint32_t nchunks, const uint32_t* __restrict__ stage,
                                                       uint64_t stage_words, const ChunkInfo* __restrict__ info,
                                                       const Result* __restrict__ res, uint8_t* __restrict__ out,
                                                       uint32_t hdr) {
    if (res->total_bytes == 0) return;
    const uint32_t c = blockIdx.x;
    const ChunkView cv = make_view(src, chunk, nchunks, stage, stage_words, info, c);
    const bool has_next = c + 1 < nchunks;
    ChunkView nv = cv;
    if (has_next) nv = make_view(src, chunk, nchunks, stage, stage_words, info, c + 1);
    const uint64_t o = cv.off;
    const uint64_t e = has_next ? nv.off : res->payload_bits;
    const uint64_t b0 = (o + 7) >> 3, b1 = (e + 7) >> 3;
    uint8_t* dst = out + hdr;
    for (uint64_t b = b0 + threadIdx.x; b < b1; b += 256) {
        const uint64_t r = 8 * b - o;
        const uint32_t k = (uint32_t)min<uint64_t>(8, e - 8 * b);
        uint32_t v = chunk_bits(cv, r, k);
        if (k < 8 && has_next) v |= chunk_bits(nv, 0, 8 - k) << k;
        dst[b] = (uint8_t)v;
    }
}

// Whole-buffer checksum with an initial value (deflate_crc32 / deflate_adler32).
__global__ __launch_bounds__(FIN_THREADS) void checksum_combine_kernel(const ChunkInfo* __restrict__ info,
                                                                       uint32_t nchunks, uint64_t n, uint32_t chunk,
                                                                       int kind, uint32_t init, uint32_t* out) {
    __shared__ uint32_t red0[FIN_THREADS];
    __shared__ uint32_t red1[FIN_THREADS];
    const int tid = threadIdx.x;
    uint32_t x0 = 0, x1 = 0;
    for (uint32_t i = (uint32_t)tid; i < nchunks; i += FIN_THREADS) {
        const uint64_t after = n - min<uint64_t>(n, (uint64_t)(i + 1) * chunk);
        if (kind == 0) {
            x0 ^= gf2_mul(gf2_x8n(after), info[i].crc);
        } else {
            const uint32_t s1 = info[i].adler_s1;
            x0 = (x0 + s1) % ADLER_MOD;
            x1 = (uint32_t)((x1 + (uint64_t)s1 * (after % ADLER_MOD) + info[i].adler_s2) % ADLER_MOD);
        }
    }
    red0[tid] = x0;
    red1[tid] = x1;
    __syncthreads();
    for (int w = FIN_THREADS / 2; w > 0; w >>= 1) {
        if (tid < w) {
            if (kind == 0) {
                red0[tid] ^= red0[tid + w];
            } else {
                red0[tid] = (red0[tid] + red0[tid + w]) % ADLER_MOD;
                red1[tid] = (red1[tid] + red1[tid + w]) % ADLER_MOD;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (kind == 0) {
            *out = ~(gf2_mul(gf2_x8n(n), ~init) ^ red0[0]);
        } else {
            const uint32_t a1 = init & 0xFFFFu, a2 = init >> 16;
            const uint32_t s1 = (a1 + red0[0]) % ADLER_MOD;
            const uint32_t s2 = (uint32_t)((a2 + (uint64_t)(n % ADLER_MOD) * a1 + red1[0]) % ADLER_MOD);
            *out = (s2 << 16) | s1;
        }
    }
}

inline void record(hipEvent_t ev, hipStream_t s) {
    if (ev) (void)hipEventRecord(ev, s);
}

}  // namespace

uint32_t max_chunk_for_hash_bits(int hash_bits) { return hash_bits <= 12 ? MAX_CHUNK_L1 : MAX_CHUNK_L1 / 2; }

uint64_t stage_words_for_chunk(uint32_t chunk) { return ((uint64_t)chunk + 64 + 255) / 4; }

hipError_t launch_parallel_compress(const Plan& p) {
    const uint32_t nb = p.nchunks;
    const hipStream_t s = p.stream;
    record(p.ev[0], s);
    if (p.format != 0)
        hipLaunchKernelGGL(checksum_kernel, dim3(nb), dim3(256), 0, s, p.d_in, p.n, p.chunk, p.format == 2 ? 0 : 1,
                           p.d_info);
    record(p.ev[1], s);
    if (p.level == 0) {
        hipLaunchKernelGGL(mark_stored_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, p.n, p.chunk, nb, p.d_info);
    } else if (p.level <= 3) {
        if (p.level == 1) {
            const size_t smem = L1Cfg<12>::TAB_BYTES > L1_PHASE_B_BYTES ? L1Cfg<12>::TAB_BYTES : L1_PHASE_B_BYTES;
            static bool attr12 = false;
            if (!attr12) {
                hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&l1_lz77_kernel<12>),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
                if (e != hipSuccess) return e;
                attr12 = true;
            }
            hipLaunchKernelGGL(l1_lz77_kernel<12>, dim3(nb), dim3(L1_THREADS), smem, s, p.d_in, p.n, p.chunk, nb,
                               p.d_d16, p.d_m1, p.d_stage, p.stage_words, p.d_info);
        } else {
            const size_t smem = L1Cfg<14>::TAB_BYTES > L1_PHASE_B_BYTES ? L1Cfg<14>::TAB_BYTES : L1_PHASE_B_BYTES;
            static bool attr14 = false;
            if (!attr14) {
                hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&l1_lz77_kernel<14>),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
                if (e != hipSuccess) return e;
                attr14 = true;
            }
            hipLaunchKernelGGL(l1_lz77_kernel<14>, dim3(nb), dim3(L1_THREADS), smem, s, p.d_in, p.n, p.chunk, nb,
                               p.d_d16, p.d_m1, p.d_stage, p.stage_words, p.d_info);
        }
    } else {
        return hipErrorNotSupported;
    }
    record(p.ev[2], s);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(FIN_THREADS), 0, s, p.d_info, nb, p.n, p.chunk, p.format,
                       p.d_out, p.cap, p.d_result);
    record(p.ev[3], s);
    const uint32_t hdr = p.format == 1 ? 2u : (p.format == 2 ? 10u : 0u);
    hipLaunchKernelGGL(assemble_kernel, dim3(nb), dim3(256), 0, s, p.d_in, p.chunk, nb, p.d_stage, p.stage_words,
                       p.d_info, p.d_result, p.d_out, hdr);
    record(p.ev[4], s);
    return hipGetLastError();
}

hipError_t launch_checksum(const uint8_t* d_in, uint64_t n, int kind, uint32_t init, uint32_t* d_out_value,
                           ChunkInfo* d_tmp, uint32_t tmp_chunks, hipStream_t stream) {
    const uint32_t chunk = 65536;
    const uint32_t nb = n ? (uint32_t)((n + chunk - 1) / chunk) : 1;
    if (nb > tmp_chunks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(checksum_kernel, dim3(nb), dim3(256), 0, stream, d_in, n, chunk, kind, d_tmp);
    hipLaunchKernelGGL(checksum_combine_kernel, dim3(1), dim3(FIN_THREADS), 0, stream, d_tmp, nb, n, chunk, kind,
                       init, d_out_value);
    return hipGetLastError();
}

}  // namespace orot_gpu
