// device_common.h -- device helpers shared by the gfx950 kernels: CRC-32
// GF(2) arithmetic, unaligned loads, wave scans, RFC 1951 fixed-Huffman
// symbols, match-length helpers, the per-position token word.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orot_gpu {
namespace {

constexpr uint32_t CRC_POLY = 0xEDB88320u;
constexpr uint32_t ADLER_MOD = 65521u;

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


// Per-position token word: literal byte | match length << 8 | distance << 17.
__device__ __forceinline__ uint32_t tok_len(uint32_t t) { return (t >> 8) & 0x1FFu; }
__device__ __forceinline__ uint32_t tok_dist(uint32_t t) { return t >> 17; }
__device__ __forceinline__ uint32_t tok_bits(uint32_t t) {
    const uint32_t l = tok_len(t);
    return l ? match_bits(l, tok_dist(t)) : lit_bits(t & 0xFFu);
}

// Byte j (0..15) of a 16-byte window held as two little-endian u64.
__device__ __forceinline__ uint32_t win_byte(uint64_t w0, uint64_t w1, uint32_t j) {
    return (uint32_t)((j < 8 ? (w0 >> (8 * j)) : (w1 >> (8 * (j - 8)))) & 0xFFu);
}
__device__ __forceinline__ uint64_t win_u64(uint64_t w0, uint64_t w1, uint32_t j) {  // bytes j..j+7, j < 8
    return j == 0 ? w0 : ((w0 >> (8 * j)) | (w1 << (64 - 8 * j)));
}


// Bits of stored block(s) for L raw bytes starting at bit offset off
// (deflate_block.cpp:107-132: 3-bit header, pad to a byte, LEN / NLEN, bytes;
// pieces of at most 65535 bytes).
__device__ __forceinline__ uint64_t stored_bits(uint64_t off, uint32_t L) {
    const uint64_t pieces = L == 0 ? 1 : (L + 65534u) / 65535u;
    const uint64_t pad0 = (8 - ((off + 3) & 7)) & 7;
    return 3 + pad0 + 32 * pieces + 8 * (pieces - 1) + 8ull * L;
}

constexpr uint32_t NO_ITEM = 0xFFFFFFFFu;

// Next item of a persistent workgroup's work queue: one atomic per item,
// broadcast through the LDS word *s_item.  Every thread of the workgroup must
// call it.  The word is live only between the two barriers inside, so it may
// be any word of the workgroup's LDS that holds nothing across the call (the
// persistent kernels pass the first word of their dynamic LDS, which each
// chunk's work writes before it reads).
__device__ __forceinline__ uint32_t next_item(uint32_t* queue, uint32_t count, uint32_t* s_item) {
    if (threadIdx.x == 0) *s_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t k = *s_item;
    __syncthreads();
    return k < count ? k : NO_ITEM;
}

// Grid size for a launch over `count` items with at most `cap` workgroups;
// persistent launches (graph capture) always take `cap`.
inline unsigned grid_for(bool persistent, uint32_t count, uint32_t cap) {
    if (persistent) return cap;
    return count == 0 ? 1u : (count < cap ? count : cap);
}

}  // namespace
}  // namespace orot_gpu
