// deflate_kernels.hip -- gfx950 kernels for orot's parallel DEFLATE compress
// path (src/parallel/parallel_compressor.cpp:36-150).
//
// Pipeline for one input buffer split into chunks of `chunk` bytes (one
// DEFLATE block per chunk, fresh match tables per chunk, as orot does):
//
//   checksum_kernel   one workgroup per chunk: CRC-32 register or Adler-32
//                     partial sums of the chunk (lane slices + GF(2) shifts).
//                     Skipped for zlib on the barrier path, whose staging
//                     pass sums the chunk's bytes as it puts them in LDS.
//   l1_candidate_kernel + l1_barrier_kernel   levels 1-3, chunks up to
//                     128 KiB (orot's defaults): see the section below.
//   l1_lz77_kernel    level 1 with chunks of 128-256 KiB, one workgroup
//                     (8 waves) per chunk:
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
//                     incoming offset mod 8), checksum combine, framing; a
//                     stream piece starts at its stream's bit offset and has
//                     no framing.
//   assemble_kernel   one workgroup per chunk (or per slice of a chunk's
//                     output when a batch has few chunks): shifts the bits to its
//                     offset in the output (stored chunks copy the input).
#include "deflate_gpu.h"
#include "device_common.h"

namespace orot_gpu {
namespace {

constexpr int L1_THREADS = 512;
constexpr int L1_WAVES = L1_THREADS / 64;
constexpr uint32_t MAX_CHUNK_L1 = 256u * 1024u;  // 12-bit tables: 8 sub-ranges of 32 KiB

// ---------------------------------------------------------------------------
// Checksum partials per chunk.
// ---------------------------------------------------------------------------
// CRC-32 register (zero initial value) or Adler sums of one chunk, by 256
// threads.  The chunk is read in 64-byte pieces, piece q by thread q mod 256,
// so every load instruction of a wave covers one contiguous 4 KiB window.
// Partial results combine linearly:
//   CRC:   a thread folds its pieces as acc = acc * x^(8*16384) + crc(piece)
//          (16384 bytes separate its consecutive pieces; the multiply by that
//          constant is four 256-entry LDS tables), then shifts acc by the
//          bytes after its last piece;
//   Adler: s1 = sum b, s2 = sum (len - p) * b over byte positions p, per piece
//          from the byte sum and the position-weighted byte sum (v_dot4).
constexpr uint32_t CK_THREADS = 256;
constexpr uint32_t CK_PIECE = 64;
constexpr uint32_t CK_STRIDE = CK_PIECE * CK_THREADS;

struct CkTables {
    uint32_t crc[4][256];  // slicing-by-4
    uint32_t mx[4][256];   // mx[k][b] = (b << 8k) * x^(8 * CK_STRIDE) mod P
};

__device__ __forceinline__ void load_ck_tables(CkTables& t) {
    const uint32_t xs = gf2_x8n(CK_STRIDE);
    for (int i = threadIdx.x; i < 1024; i += CK_THREADS) {
        t.crc[i >> 8][i & 255] = c_crc.t[i >> 8][i & 255];
        t.mx[i >> 8][i & 255] = gf2_mul((uint32_t)(i & 255) << (8 * (i >> 8)), xs);
    }
    __syncthreads();
}

__device__ __forceinline__ uint32_t crc_word(uint32_t crc, uint32_t w, const CkTables& t) {
    crc ^= w;
    return t.crc[3][crc & 0xFF] ^ t.crc[2][(crc >> 8) & 0xFF] ^ t.crc[1][(crc >> 16) & 0xFF] ^ t.crc[0][crc >> 24];
}

__device__ __forceinline__ uint32_t mul_stride(uint32_t a, const CkTables& t) {
    return t.mx[0][a & 0xFF] ^ t.mx[1][(a >> 8) & 0xFF] ^ t.mx[2][(a >> 16) & 0xFF] ^ t.mx[3][a >> 24];
}

__device__ __forceinline__ void chunk_checksum(const uint8_t* __restrict__ s, uint32_t len, int kind,
                                               ChunkInfo* __restrict__ ci, const CkTables& t, uint32_t* r0,
                                               uint32_t* r1) {
    const uint32_t tid = threadIdx.x;
    if (kind == 0) {
        uint32_t acc = 0, end = 0;
        bool any = false;
        for (uint32_t o = tid * CK_PIECE; o < len; o += CK_STRIDE) {
            const uint32_t L = min(CK_PIECE, len - o);
            uint32_t crc = 0;
            if (L == CK_PIECE) {
                uint32_t w[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) w[k] = ld32u(s + o + 4 * k);
#pragma unroll
                for (int k = 0; k < 16; ++k) crc = crc_word(crc, w[k], t);
                acc = any ? (mul_stride(acc, t) ^ crc) : crc;
            } else {
                for (uint32_t p = o; p < o + L; ++p) crc = (crc >> 8) ^ t.crc[0][(crc ^ s[p]) & 0xFF];
                acc = any ? (gf2_mul(gf2_x8n(CK_STRIDE - CK_PIECE + L), acc) ^ crc) : crc;
            }
            any = true;
            end = o + L;
        }
        r0[tid] = any ? gf2_mul(gf2_x8n(len - end), acc) : 0u;
        __syncthreads();
        for (int w = CK_THREADS / 2; w > 0; w >>= 1) {
            if ((int)tid < w) r0[tid] ^= r0[tid + w];
            __syncthreads();
        }
        if (tid == 0) ci->crc = r0[0];
    } else {
        uint64_t A = 0, B = 0;  // sum b, sum (len - p) * b
        uint32_t it = 0;
        for (uint32_t o = tid * CK_PIECE; o < len; o += CK_STRIDE) {
            const uint32_t L = min(CK_PIECE, len - o);
            uint32_t S = 0, K = 0;  // byte sum, sum of (p - o) * b
            if (L == CK_PIECE) {
                uint32_t w[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) w[k] = ld32u(s + o + 4 * k);
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const uint32_t sb = __builtin_amdgcn_udot4(w[k], 0x01010101u, 0u, false);
                    S += sb;
                    K += 4u * (uint32_t)k * sb + __builtin_amdgcn_udot4(w[k], 0x03020100u, 0u, false);
                }
            } else {
                for (uint32_t p = o; p < o + L; ++p) {
                    S += s[p];
                    K += (p - o) * s[p];
                }
            }
            A += S;
            B += (uint64_t)(len - o) * S - K;
            if (++it == 4096) {  // keep the 64-bit sums far from overflow on huge chunks
                A %= ADLER_MOD;
                B %= ADLER_MOD;
                it = 0;
            }
        }
        r0[tid] = (uint32_t)(A % ADLER_MOD);
        r1[tid] = (uint32_t)(B % ADLER_MOD);
        __syncthreads();
        for (int w = CK_THREADS / 2; w > 0; w >>= 1) {
            if ((int)tid < w) {
                r0[tid] = (r0[tid] + r0[tid + w]) % ADLER_MOD;
                r1[tid] = (r1[tid] + r1[tid + w]) % ADLER_MOD;
            }
            __syncthreads();
        }
        if (tid == 0) {
            ci->adler_s1 = r0[0];
            ci->adler_s2 = r1[0];
        }
    }
    __syncthreads();
}

// Checksum partials of every chunk of the batch (grid-stride over chunks).
__global__ __launch_bounds__(CK_THREADS) void checksum_kernel(const uint8_t* __restrict__ src,
                                                              const ChunkDesc* __restrict__ desc,
                                                              const BatchHdr* __restrict__ hdr, int kind,
                                                              ChunkInfo* __restrict__ info) {
    __shared__ CkTables T;
    __shared__ uint32_t r0[CK_THREADS];
    __shared__ uint32_t r1[CK_THREADS];
    load_ck_tables(T);
    const uint32_t nchunks = hdr->nchunks;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x)
        chunk_checksum(src + desc[c].in_off, desc[c].len, kind, &info[c], T, r0, r1);
}

// Checksum partials of a plain buffer in uniform chunks (whole-buffer entry points).
__global__ __launch_bounds__(CK_THREADS) void checksum_uniform_kernel(const uint8_t* __restrict__ src, uint64_t n,
                                                                      uint32_t chunk, int kind,
                                                                      ChunkInfo* __restrict__ info) {
    __shared__ CkTables T;
    __shared__ uint32_t r0[CK_THREADS];
    __shared__ uint32_t r1[CK_THREADS];
    load_ck_tables(T);
    const uint32_t c = blockIdx.x;
    const uint64_t base = (uint64_t)c * chunk;
    chunk_checksum(src + base, (uint32_t)min<uint64_t>(chunk, n - base), kind, &info[c], T, r0, r1);
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
constexpr uint32_t L1_PHASE_B_BYTES = (L1_BM_WORDS + 2 * L1_THREADS + 16) * 4 + L1_THREADS;

template <int HB>
__device__ __forceinline__ void l1_lz77_body(uint32_t c, uint32_t* smem, const uint8_t* __restrict__ src,
    const ChunkDesc* __restrict__ desc, uint32_t chunk, uint32_t* __restrict__ tokw,
    uint32_t* __restrict__ stage, uint64_t stage_words, ChunkInfo* __restrict__ info) {
    constexpr int W = L1Cfg<HB>::W;
    constexpr uint32_t TAB = L1Cfg<HB>::TAB;
    const uint32_t n = desc[c].len;
    const uint8_t* s = src + desc[c].in_off;
    uint32_t* T = tokw + (uint64_t)c * ((chunk + 7u) & ~7u);  // 32-byte aligned per chunk
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    // ---- Phase A1: latest earlier position with the same hash, within each
    // sub-range (orot's single-probe head table, every position inserted:
    // lz77.cpp:173-176, 400-403).  T[p] = (local index + 1) of that
    // position, 0 if none.  The next step's input word is loaded ahead.
    const uint32_t nh = n >= 4 ? n - 3 : 0;  // positions p with p + 4 <= n are hashed
    const uint32_t SR = (((nh + W - 1) / W) + 63u) & ~63u;
    uint16_t* tabs = reinterpret_cast<uint16_t*>(smem);
    if (wv < W) {
        uint16_t* tab = tabs + (uint32_t)wv * TAB;
        for (uint32_t i = lane; i < TAB / 2; i += 64) reinterpret_cast<uint32_t*>(tab)[i] = 0;
        const uint32_t a = (uint32_t)wv * SR;
        const uint32_t b = min(a + SR, nh);
        uint32_t wnext = (a + (uint32_t)lane < b) ? ld32u(s + a + lane) : 0u;
        for (uint32_t p0 = a; p0 < b; p0 += 64) {
            const uint32_t p = p0 + (uint32_t)lane;
            const bool valid = p < b;
            const uint32_t wcur = wnext;
            wnext = (p + 64 < b) ? ld32u(s + p + 64) : 0u;
            const uint32_t h = valid ? hash4<HB>(wcur) : 0u;
            uint64_t m = ballot(valid);
#pragma unroll
            for (int k = 0; k < HB; ++k) {
                const uint32_t bit = (h >> k) & 1u;
                const uint64_t bk = ballot(bit != 0);
                m &= bk ^ ((uint64_t)bit - 1ull);  // bit ? bk : ~bk
            }
            if (valid) {
                const uint64_t below = m & ((1ull << lane) - 1ull);
                uint32_t res;
                if (below) res = p0 + (uint32_t)(63 - __builtin_clzll(below)) - a + 1u;
                else res = tab[h];
                if ((m >> lane) == 1ull) tab[h] = (uint16_t)(p - a + 1u);
                T[p] = res;
            }
        }
    }
    __syncthreads();

    // ---- Phase A2: per position, the final candidate (first occurrences
    // resolved against earlier sub-ranges; C == 0 is "empty", distance masked
    // to 15 bits, first byte compared at C: lz77.cpp:178-184) and the greedy
    // match length M1 (capped at 258 and the chunk end; < 3 means none).
    // Each thread walks a contiguous run in batches of 8 positions; along a
    // diagonal (same distance) the verified extent is carried, so long runs
    // cost O(1) per position.
    {
        const uint32_t R = (((n + L1_THREADS - 1) / L1_THREADS) + 7u) & ~7u;
        const uint32_t r0 = min(n, (uint32_t)tid * R), r1 = min(n, r0 + R);
        const uint32_t lim = min(n, r1 + 257u);
        uint32_t cd = 0, ce = 0;
        bool cx = false;
        for (uint32_t p0 = r0; p0 < r1; p0 += 8) {
            uint32_t loc[8];
            if (p0 + 8 <= nh) {
                const uint4 v0 = *reinterpret_cast<const uint4*>(T + p0);
                const uint4 v1 = *reinterpret_cast<const uint4*>(T + p0 + 4);
                loc[0] = v0.x; loc[1] = v0.y; loc[2] = v0.z; loc[3] = v0.w;
                loc[4] = v1.x; loc[5] = v1.y; loc[6] = v1.z; loc[7] = v1.w;
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) loc[k] = (p0 + k < nh) ? T[p0 + k] : 0u;
            }
            uint64_t w0, w1;
            if (p0 + 16 <= n) {
                w0 = ld64u(s + p0);
                w1 = ld64u(s + p0 + 8);
            } else {
                w0 = w1 = 0;
                for (uint32_t j = 0; j < 16 && p0 + j < n; ++j) {
                    const uint64_t bv = s[p0 + j];
                    if (j < 8) w0 |= bv << (8 * j);
                    else w1 |= bv << (8 * (j - 8));
                }
            }
            uint32_t dd[8];
            uint64_t cw[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t p = p0 + k;
                dd[k] = 0;
                cw[k] = 0;
                if (p >= r1 || p >= nh) continue;
                const uint32_t w = p / SR;
                uint32_t C = 0;
                bool have = false;
                if (loc[k]) {
                    C = w * SR + loc[k] - 1u;
                    have = true;
                } else {
                    const uint32_t word = (uint32_t)(win_u64(w0, w1, k) & 0xFFFFFFFFull);
                    const uint32_t h = hash4<HB>(word);
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
                    if (d) {
                        const uint32_t first = (p - C == d) ? 0xFFFFFFFFu : s[C];
                        if (first == 0xFFFFFFFFu || first == win_byte(w0, w1, k)) {
                            dd[k] = d;
                            if (p + 8 <= n) cw[k] = ld64u(s + p - d);
                        }
                    }
                }
            }
            uint32_t outw[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t p = p0 + k;
                const uint32_t lit = p < n ? win_byte(w0, w1, k) : 0u;
                uint32_t m1 = 0;
                const uint32_t d = dd[k];
                if (d) {
                    const uint32_t cap = min(258u, n - p);
                    uint32_t L;
                    if (d == cd && p <= ce && (cx || ce >= p + cap)) {
                        L = ce - p;
                    } else {
                        // fresh comparison on diagonal d
                        uint32_t e;
                        bool x;
                        if (p + 8 <= n) {
                            const uint64_t diff = win_u64(w0, w1, k) ^ cw[k];
                            if (diff) {
                                e = p + (uint32_t)(__builtin_ctzll(diff) >> 3);
                                x = true;
                            } else {
                                e = p + 8;
                                x = e >= n;
                            }
                        } else {
                            e = p;
                            while (e < n && s[e] == s[e - d]) ++e;
                            x = true;
                        }
                        // extend 64 bytes per round up to what the rest of the run can use
                        while (!x && e < lim) {
                            if (e + 64 <= n) {
                                uint64_t a8[8], b8[8];
#pragma unroll
                                for (int q = 0; q < 8; ++q) {
                                    a8[q] = ld64u(s + e + 8 * q);
                                    b8[q] = ld64u(s + e - d + 8 * q);
                                }
                                uint32_t adv = 64;
#pragma unroll
                                for (int q = 7; q >= 0; --q) {
                                    const uint64_t df = a8[q] ^ b8[q];
                                    if (df) adv = 8u * q + (uint32_t)(__builtin_ctzll(df) >> 3);
                                }
                                e += adv;
                                if (adv < 64) x = true;
                            } else {
                                while (e < n && s[e] == s[e - d]) ++e;
                                x = true;
                            }
                        }
                        cd = d;
                        ce = e;
                        cx = x;
                        L = e - p;
                    }
                    L = min(L, cap);
                    m1 = L >= 3 ? L : 0u;
                }
                outw[k] = lit | (m1 << 8) | ((m1 ? d : 0u) << 17);
            }
            if (p0 + 8 <= r1) {
                *reinterpret_cast<uint4*>(T + p0) = make_uint4(outw[0], outw[1], outw[2], outw[3]);
                *reinterpret_cast<uint4*>(T + p0 + 4) = make_uint4(outw[4], outw[5], outw[6], outw[7]);
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (p0 + k < r1) T[p0 + k] = outw[k];
            }
        }
    }
    __syncthreads();

    // ---- Phase B: greedy parse (lz77.cpp:378-406) of each lane's segment,
    // speculatively entered at its start; fixed-Huffman bits counted on the way.
    uint32_t* bm = smem;
    uint32_t* exit_spec = smem + L1_BM_WORDS;
    uint32_t* segbits = exit_spec + L1_THREADS;
    uint32_t* scan = segbits + L1_THREADS;
    uint8_t* touched = reinterpret_cast<uint8_t*>(scan + 16);
    const uint32_t S = max(32u, (((n + L1_THREADS - 1) / L1_THREADS) + 31u) & ~31u);
    const uint32_t NS = (n + S - 1) / S;
    const uint32_t nbw = (n + 31) >> 5;
    for (uint32_t i = tid; i < nbw; i += L1_THREADS) bm[i] = 0;
    touched[tid] = 0;
    __syncthreads();
    if ((uint32_t)tid < NS) {
        uint32_t pos = (uint32_t)tid * S;
        const uint32_t end = min(pos + S, n);
        uint32_t bits = 0;
        while (pos < end) {
            bm[pos >> 5] |= 1u << (pos & 31);
            const uint32_t t = T[pos];
            bits += tok_bits(t);
            const uint32_t l = tok_len(t);
            pos += l ? l : 1u;
        }
        exit_spec[tid] = pos;
        segbits[tid] = bits;
    }
    __syncthreads();

    // ---- Serial resolve of segment entries (wave 0, lockstep): a segment
    // entered past its start is re-parsed until it meets a position of its
    // speculative parse (from there on both parses agree).
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
            if (lane == 0) touched[si] = 1;
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
                const uint32_t l = tok_len(T[pos]);
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

    // ---- Phase C: bit counts (re-counted only where the resolve changed the
    // parse), block-wide exclusive scan, stored / fixed choice
    // (deflate_block.cpp:414-421), emission.
    const uint32_t st = (uint32_t)tid * S;
    const uint32_t en = (uint32_t)tid < NS ? min(st + S, n) : st;
    uint32_t mybits = (uint32_t)tid < NS ? segbits[tid] : 0u;
    if ((uint32_t)tid < NS && touched[tid]) {
        mybits = 0;
        uint32_t pk[8];
        int cnt = 0;
        for (uint32_t w = st >> 5; w < ((en + 31) >> 5); ++w) {
            uint32_t x = bm[w];
            while (x) {
                pk[cnt++] = (w << 5) + (uint32_t)__builtin_ctz(x);
                x &= x - 1;
                if (cnt == 8) {
                    uint32_t tv[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) tv[k] = T[pk[k]];
#pragma unroll
                    for (int k = 0; k < 8; ++k) mybits += tok_bits(tv[k]);
                    cnt = 0;
                }
            }
        }
        for (int k = 0; k < cnt; ++k) mybits += tok_bits(T[pk[k]]);
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
    const bool is_last = chunk_final(desc[c].buf);
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
        uint32_t pk[8];
        int cnt = 0;
        auto put = [&](uint32_t t) {
            uint32_t v, len;
            const uint32_t l = tok_len(t);
            if (l) fixed_match(l, tok_dist(t), v, len);
            else fixed_lit(t & 0xFFu, v, len);
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
        };
        for (uint32_t w = st >> 5; w < ((en + 31) >> 5); ++w) {
            uint32_t x = bm[w];
            while (x) {
                pk[cnt++] = (w << 5) + (uint32_t)__builtin_ctz(x);
                x &= x - 1;
                if (cnt == 8) {
                    uint32_t tv[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) tv[k] = T[pk[k]];
                    for (int k = 0; k < 8; ++k) put(tv[k]);
                    cnt = 0;
                }
            }
        }
        for (int k = 0; k < cnt; ++k) put(T[pk[k]]);
        if (nacc) atomicOr(&stg[wi], (uint32_t)acc);
    }
}

template <int HB>
__global__ __launch_bounds__(L1_THREADS) void l1_lz77_kernel(const uint8_t* __restrict__ src,
    const ChunkDesc* __restrict__ desc, const uint32_t* __restrict__ order, BatchHdr* __restrict__ hdr,
    uint32_t chunk, uint32_t* __restrict__ tokw, uint32_t* __restrict__ stage, uint64_t stage_words,
    ChunkInfo* __restrict__ info) {
    extern __shared__ uint32_t smem[];
    for (;;) {
        const uint32_t k = next_item(&hdr->queue[Q_LZ], hdr->nchunks, smem);
        if (k == NO_ITEM) break;
        l1_lz77_body<HB>(order[k], smem, src, desc, chunk, tokw, stage, stage_words, info);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Shared by the L1-class kernels for chunks up to 128 KiB.
// ---------------------------------------------------------------------------
constexpr int LC_THREADS = 1024;
constexpr int LC_WAVES = LC_THREADS / 64;
constexpr uint32_t LC_MAX_CHUNK = 128u * 1024u;
constexpr uint32_t LC_BM_WORDS = LC_MAX_CHUNK / 32;
constexpr uint32_t LC_REGION0 = LC_MAX_CHUNK;  // bytes: the chunk, then length codes

// u8 code of a position: its parse step, and for a literal the length of its
// fixed-Huffman code.  0: literal below 144 (8 bits); 1: literal 144..255 (9
// bits); 2..253: match of length code + 1 (3..254); 254: 258; 255: 255..257,
// with length - 255 in the position's byte of the side array.
__device__ __forceinline__ uint32_t len_code8(uint32_t m1, uint32_t lit) {
    uint32_t c = m1 >= 255u ? 255u : m1 - 1u;
    c = m1 == 258u ? 254u : c;
    return m1 == 0u ? (uint32_t)(lit >= 144u) : c;
}

// ---------------------------------------------------------------------------
// L1-class path, 12-bit (level 1) or 14-bit (levels 2-3) hashes, chunks up
// to 128 KiB (orot's defaults): two launches per batch.
//   l1_candidate_kernel<HB, W>, one W-wave workgroup per chunk (12 bits:
//   W = 2 with four chunks per CU, or W = 6 for small batches; 14 bits:
//   W = 2, one chunk per CU):
//   A1  each wave scans one of W sub-ranges of the chunk's hash positions in
//       order: the latest earlier same-hash position C in its sub-range, one
//       LDS exchange per position (EXCH: same-address exchanges of one wave
//       instruction take effect in lane order, checked at start-up by
//       exchange_order_kernel) or a ballot match-any.  A hash's first
//       position in a later sub-range takes its latest position in the
//       earlier ones.  Output per position, 16 bits: cand_code(p, C).
//   l1_barrier_kernel, one 1024-thread workgroup per chunk:
//   0   the chunk's bytes are staged in LDS.
//   A2  per position: candidate distance from its code (far candidates:
//       C recovered and first byte compared at C, lz77.cpp:178-184) and
//       greedy match length
//       (lz77.cpp:163-191), the verified extent carried along a diagonal,
//       all compares in LDS.  Output per position: a u8 code (len_code8),
//       to HBM in LbLayout order; the candidate codes stay as they are, since
//       a match's distance is its candidate code's low 15 bits.
//   K   the LDS bytes are replaced by the u8 codes.
//   P   parse barriers: q is a barrier when no token starting before q ends
//       after q (max over p < q of p + step(p) == q); every greedy parse
//       visits every barrier.  Lane j parses exactly from the first barrier
//       of its 128-position range to the first barrier of a later range
//       (lz77.cpp:378-406), counting fixed-Huffman bits.
//   C   per-range fixed-Huffman bit counts from the codes in LDS and the
//       distances, block-wide scan, stored / fixed choice
//       (deflate_block.cpp:414-421), emission into the chunk's staging
//       bitstream (literals read from the input, codes from HBM).
// ---------------------------------------------------------------------------
static_assert(LC_MAX_CHUNK / LC_THREADS <= 128, "ranges of at most 136 positions span at most five bitmap words");
constexpr uint32_t LB_BM_OFF = LC_REGION0 / 4;               // words: the token bitmap
constexpr uint32_t LB_FB_OFF = LB_BM_OFF + LC_BM_WORDS;       // first barrier per range
constexpr uint32_t LB_MISC_OFF = LB_FB_OFF + LC_THREADS;
constexpr uint32_t LB_RUNS = 32;                     // long 258-step runs of a chunk (<= 31 fit 128 KiB)
constexpr uint32_t LB_LTAB_OFF = LB_MISC_OFF + 96u + 2u * LB_RUNS;  // fixed-Huffman length symbols by u8 code
constexpr uint32_t LB_LDS_BYTES = (LB_LTAB_OFF + 256u) * 4u;

constexpr uint16_t LA_NONE = 0xFFFFu;

// Ranges and token-word layout of the barrier path.  The chunk is cut into R
// ranges of S positions, one per thread of l1_barrier_kernel.  S is 8 times
// an odd number (136 for 128 KiB chunks): the lanes of a wave then start
// their ranges on LDS words 2 * odd apart, which spreads them over the banks
// in every phase that walks the ranges in order, and each range is whole
// groups of 8 positions starting on a byte of its bitmap word.  Group g of
// range t is stored at (g * R + t) * 8, so those phases read 64 consecutive
// groups per wave instruction.  t = p / S is umulhi(p, ceil(2^32 / S)),
// exact for every p < 2^17 + 256 and S <= 136.
struct LbLayout {
    uint32_t S, R, m;
    __device__ explicit LbLayout(uint32_t n)
        : S(8u * ((((n + LC_THREADS - 1) / LC_THREADS + 7u) >> 3) | 1u)),
          R((n + S - 1) / S),
          m(0xFFFFFFFFu / S + 1u) {}
    __device__ uint32_t range_of(uint32_t p) const { return __umulhi(p, m); }
    // t < 1024, S <= 136, g < 17 and R <= 1024: the products fit the
    // full-rate 24-bit multiplier
    __device__ uint32_t index(uint32_t p) const {
        const uint32_t t = range_of(p), q = p - __umul24(t, S);
        return ((__umul24(q >> 3, R) + t) << 3) | (q & 7u);
    }
    // first word of group g of range t
    __device__ uint32_t group(uint32_t t, uint32_t g) const { return (__umul24(g, R) + t) << 3; }
};
constexpr uint32_t lb_tok_stride(uint32_t chunk) { return ((chunk + 7u) & ~7u) + 256u; }

// Candidate code of position p whose latest earlier same-hash position is
// C (C1 = C + 1, 0 if none).  orot takes head 0 as empty, masks the distance
// to 15 bits and, when the masked distance differs from the true one,
// accepts the candidate only if the bytes at C and p agree
// (lz77.cpp:178-184).  Codes: 0 no candidate; 1..32767 the distance;
// 0x8000 | dd a candidate 32 KiB or more back with masked distance dd != 0.
// For the last, C is the nearest of p - dd - 32768 k (k >= 1) whose hash is
// p's: no same-hash position lies between C and p.
__device__ __forceinline__ uint32_t cand_code(uint32_t p, uint32_t C1) {
    const uint32_t dist = p + 1u - C1, dd = dist & 0x7FFFu;
    const uint32_t far = dd ? (0x8000u | dd) : 0u;
    const uint32_t code = dist == dd ? dd : far;
    return C1 <= 1u ? 0u : code;
}

// LDS of the candidate kernel with HB-bit hashes and W waves: a table per
// sub-range (latest position + 1 of each hash) and, for sub-ranges 1.., the
// offset of each hash's first position.  12 bits: 40 KiB at W = 2, 136 KiB
// at W = 6; 14 bits (levels 2-3): 160 KiB at W = 2.
template <int HB, int W>
constexpr size_t la_lds_bytes() { return (size_t)W * (4u << HB) + (size_t)(W - 1) * (2u << HB); }

template <int HB, int W, bool EXCH>
__global__ __launch_bounds__(64 * W) void l1_candidate_kernel(const uint8_t* __restrict__ src,
    const ChunkDesc* __restrict__ desc, const uint32_t* __restrict__ order, BatchHdr* __restrict__ hdr,
    uint32_t chunk, uint16_t* __restrict__ cand) {
    constexpr uint32_t LA_TAB = 1u << HB;
    constexpr int NT = 64 * W;
    extern __shared__ uint32_t la_smem[];
    uint32_t* tab = la_smem;                                         // W x 4096
    uint16_t* first = reinterpret_cast<uint16_t*>(la_smem + W * LA_TAB);  // (W - 1) x LA_TAB
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // A grid that covers the batch takes one chunk per workgroup.  Otherwise
    // chunks come from the device queue in `order` (full chunks, then the
    // partial ones longest first), so workgroups finish together on uneven
    // batches; with one chunk each, the queue's only effect would be every
    // workgroup's atomic on one counter at the start.
    const uint32_t nchunks = hdr->nchunks;
    const bool queued = gridDim.x < nchunks;
    for (uint32_t it = blockIdx.x;; it += gridDim.x) {
        uint32_t c;
        if (queued) {
            const uint32_t k = next_item(&hdr->queue[Q_SCAN], nchunks, la_smem);
            if (k == NO_ITEM) break;
            c = order[k];
        } else {
            if (it >= nchunks) break;
            c = it;
        }
        const uint32_t n = desc[c].len;
        const uint8_t* s = src + desc[c].in_off;
        uint16_t* Q = cand + (uint64_t)c * lb_tok_stride(chunk);  // LbLayout order
        const LbLayout L(n);
        const uint32_t nh = n >= 4 ? n - 3 : 0;
        // W sub-ranges of 64-position groups; every offset inside a
        // sub-range other than the first is below 65535 (two sub-ranges of
        // a 128 KiB chunk: at most 65532), so it fits 16 bits beside LA_NONE
        const uint32_t SR = ((nh + W - 1) / W + 63u) & ~63u;
        for (uint32_t i = tid; i < W * LA_TAB; i += NT) tab[i] = 0;
        for (uint32_t i = tid; i < (W - 1) * LA_TAB / 2; i += NT)
            reinterpret_cast<uint32_t*>(first)[i] = ((uint32_t)LA_NONE << 16) | LA_NONE;
        __syncthreads();
        uint32_t* my = tab + (uint32_t)wv * LA_TAB;
        const uint32_t a = (uint32_t)wv * SR, b = min(a + SR, nh);
        const uint32_t spare = lb_tok_stride(chunk) - 1u;  // past every index of the layout
        // the 4-byte windows of 512 positions come from global memory; four
        // or five groups of loads (clamped into [a, b)) stay in flight ahead
        // of the exchanges (loads complete in order behind the code stores
        // issued before them, so a shallower queue waits on those stores)
        auto load8 = [&](uint32_t q0, uint32_t* w) {
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = ld32u(s + min(q0 + 64u * j + (uint32_t)lane, b - 1));
        };
        auto step = [&](uint32_t p0, const uint32_t* w) {
            uint32_t h[8], res[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) h[j] = hash4<HB>(w[j]);
            if (EXCH) {
                // A run of lanes with equal hashes makes one exchange: its
                // first lane stores the run's last position, the others take
                // the position just before them.  The hash of the lane below
                // comes from a DPP wave shift (lane 0 gets an impossible
                // hash); with no LDS instruction before them, the eight
                // masked exchanges of a step are issued back to back.
                bool head[8], same[8];
                uint32_t val[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t p = p0 + 64u * j + (uint32_t)lane;
                    const bool valid = p < b;
                    const uint32_t hp =
                        (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)h[j], 0x138, 0xF, 0xF, false);
                    same[j] = valid && hp == h[j];
                    const uint64_t E = ballot(same[j]);
                    const uint32_t extra = (uint32_t)__builtin_ctzll(~((E >> lane) >> 1));
                    head[j] = valid && !same[j];
                    val[j] = p + extra + 1u;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) res[j] = head[j] ? atomicExch(&my[h[j]], val[j]) : 0u;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t p = p0 + 64u * j + (uint32_t)lane;
                    if (same[j]) res[j] = p;
                    if (p >= b) res[j] = 0;
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t q0 = p0 + 64u * j;
                    const uint32_t p = q0 + (uint32_t)lane;
                    const bool valid = p < b;
                    uint64_t m = ballot(valid);
#pragma unroll
                    for (int k = 0; k < HB; ++k) {
                        const uint32_t bit = (h[j] >> k) & 1u;
                        const uint64_t bk = ballot(bit != 0);
                        m &= bk ^ ((uint64_t)bit - 1ull);
                    }
                    res[j] = 0;
                    if (valid) {
                        const uint64_t below = m & ((1ull << lane) - 1ull);
                        if (below) res[j] = q0 + (uint32_t)(63 - __builtin_clzll(below)) + 1u;
                        else res[j] = my[h[j]];
                        if ((m >> lane) == 1ull) my[h[j]] = p + 1u;
                    }
                }
            }
            // Every lane stores (lanes past the half on a spare slot beyond
            // the layout), so the compiler counts the stores exactly and the
            // waits for the prefetched loads do not wait for them.  A first
            // position of the second half stores 0 here and its code after
            // the barrier.
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t p = p0 + 64u * j + (uint32_t)lane;
                const bool valid = p < b;
                if (valid && wv >= 1 && res[j] == 0) first[(wv - 1) * LA_TAB + h[j]] = (uint16_t)(p - a);  // one per hash
                const uint32_t qi = L.index(p);
                Q[valid ? qi : spare] = (uint16_t)cand_code(p, res[j]);
            }
        };
        if constexpr (W <= 2) {
            // one wave per chunk (eight chunks per CU by registers), or two
            // with 14-bit tables (one chunk per CU by LDS): five groups in
            // flight
            uint32_t w0[8], w1[8], w2[8], w3[8], w4[8], w5[8];
            if (a < b) {
                load8(a, w0);
                load8(a + 512u, w1);
                load8(a + 1024u, w2);
                load8(a + 1536u, w3);
                load8(a + 2048u, w4);
            }
            for (uint32_t p0 = a; p0 < b; p0 += 3072u) {
                load8(p0 + 2560u, w5);
                step(p0, w0);
                if (p0 + 512u >= b) break;
                load8(p0 + 3072u, w0);
                step(p0 + 512u, w1);
                if (p0 + 1024u >= b) break;
                load8(p0 + 3584u, w1);
                step(p0 + 1024u, w2);
                if (p0 + 1536u >= b) break;
                load8(p0 + 4096u, w2);
                step(p0 + 1536u, w3);
                if (p0 + 2048u >= b) break;
                load8(p0 + 4608u, w3);
                step(p0 + 2048u, w4);
                if (p0 + 2560u >= b) break;
                load8(p0 + 5120u, w4);
                step(p0 + 2560u, w5);
            }
        } else {
            // several waves per chunk: four groups in flight keep the scan
            // under 168 VGPRs, three waves per SIMD
            uint32_t w0[8], w1[8], w2[8], w3[8], w4[8];
            if (a < b) {
                load8(a, w0);
                load8(a + 512u, w1);
                load8(a + 1024u, w2);
                load8(a + 1536u, w3);
            }
            for (uint32_t p0 = a; p0 < b; p0 += 2560u) {
                load8(p0 + 2048u, w4);
                step(p0, w0);
                if (p0 + 512u >= b) break;
                load8(p0 + 2560u, w0);
                step(p0 + 512u, w1);
                if (p0 + 1024u >= b) break;
                load8(p0 + 3072u, w1);
                step(p0 + 1024u, w2);
                if (p0 + 1536u >= b) break;
                load8(p0 + 3584u, w2);
                step(p0 + 1536u, w3);
                if (p0 + 2048u >= b) break;
                load8(p0 + 4096u, w3);
                step(p0 + 2048u, w4);
            }
        }
        __syncthreads();
        // a hash's first position in sub-range j takes its latest position
        // in the sub-ranges before j
        for (uint32_t i = tid; i < LA_TAB; i += NT) {
            uint32_t last = tab[i];
#pragma unroll
            for (int j = 1; j < W; ++j) {
                const uint32_t f = first[(j - 1) * LA_TAB + i];
                if (f != LA_NONE) Q[L.index(j * SR + f)] = (uint16_t)cand_code(j * SR + f, last);
                const uint32_t t = tab[j * LA_TAB + i];
                if (t) last = t;
            }
        }
        __syncthreads();
    }
}

// Unaligned LDS reads: gfx950 serves them as one ds_read_b32 / ds_read_b64
// at any byte address.
__device__ __forceinline__ uint32_t lds_u32(const uint32_t* L, uint32_t off) {
    uint32_t v;
    __builtin_memcpy(&v, reinterpret_cast<const uint8_t*>(L) + off, 4);
    return v;
}
__device__ __forceinline__ uint64_t lds_u64(const uint32_t* L, uint32_t off) {
    uint64_t v;
    __builtin_memcpy(&v, reinterpret_cast<const uint8_t*>(L) + off, 8);
    return v;
}
__device__ __forceinline__ uint32_t lds_u8(const uint32_t* L, uint32_t off) {
    return (L[off >> 2] >> (8 * (off & 3u))) & 0xFFu;
}
// Fixed-Huffman length symbol of u8 code b >= 2 (len_code8), LSB-first: the
// code bits, then the extra bits, in the low 16 bits; the bit count in bits
// 16..23.  Code 255 is length 255 (symbol 284, extra 28); lengths 256 and 257
// add 1 or 2 to the extra value, which keeps the bit count.
__device__ __forceinline__ uint32_t len_entry(uint32_t b) {
    if (b < 2u) return 0u;
    const uint32_t L = b == 254u ? 258u : (b == 255u ? 255u : b + 1u);
    const LenSym ls = len_sym(L);
    const uint32_t sym = 257u + ls.code;
    const uint32_t cl = sym < 280u ? 7u : 8u;
    const uint32_t code = sym < 280u ? sym - 256u : 0xC0u + (sym - 280u);
    const uint32_t rc = __builtin_bitreverse32(code) >> (32u - cl);
    return (rc | (ls.extra << cl)) | ((cl + ls.ebits) << 16);
}

// Fixed-Huffman distance symbol (5-bit code, then the extra bits) of d in
// 1..32768, without branches.
__device__ __forceinline__ void fixed_dist(uint32_t d, uint32_t& v, uint32_t& len) {
    const uint32_t x = d - 1u;
    const bool small = d <= 4u;
    const uint32_t e = small ? 0u : 30u - (uint32_t)__builtin_clz(x);
    const uint32_t code = small ? x : 2u * e + 2u + ((x >> e) & 1u);
    v = (__builtin_bitreverse32(code) >> 27) | ((x & ((1u << e) - 1u)) << 5);
    len = 5u + e;
}

// Bit j set where byte j of w is zero.
__device__ __forceinline__ uint32_t zero_bytes8(uint64_t w) {
    const uint64_t lo7 = 0x7F7F7F7F7F7F7F7Full;
    const uint64_t y = ~(((w & lo7) + lo7) | w | lo7);
    return (uint32_t)(((y >> 7) * 0x0102040810204080ull) >> 56);
}

// Parse step of the position `pos` with u8 code b (len_code8); XS is the side
// array of lengths 255..257.
__device__ __forceinline__ uint32_t code_step(uint32_t b, const uint8_t* XS, const LbLayout& L, uint32_t pos) {
    if (b == 255u) return 255u + XS[L.index(pos)];
    return b <= 1u ? 1u : (b == 254u ? 258u : b + 1u);
}

template <int HB>
__device__ __forceinline__ void l1_barrier_body(uint32_t c, uint32_t* smem, const uint8_t* __restrict__ src,
    const ChunkDesc* __restrict__ desc, uint32_t chunk, uint16_t* __restrict__ cand,
    uint32_t* __restrict__ tokw, uint32_t* __restrict__ stage, uint64_t stage_words,
    ChunkInfo* __restrict__ info, int adler) {
    const uint32_t n = desc[c].len;
    const uint8_t* s = src + desc[c].in_off;
    // candidate codes from l1_candidate_kernel, the u8 codes written by A2
    // and the side array of lengths 255..257, all in LbLayout order
    const uint16_t* Q = cand + (uint64_t)c * lb_tok_stride(chunk);
    uint8_t* LCB = reinterpret_cast<uint8_t*>(tokw + (uint64_t)c * lb_tok_stride(chunk));
    uint8_t* XS = LCB + lb_tok_stride(chunk);
    const LbLayout L(n);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t* D = smem;                         // chunk bytes, then length codes
    uint8_t* code = reinterpret_cast<uint8_t*>(smem);
    uint32_t* bm = smem + LB_BM_OFF;
    uint32_t* fb = smem + LB_FB_OFF;
    uint32_t* misc = smem + LB_MISC_OFF;

    // ---- 0: stage the chunk in LDS, plus zeroed padding words (which may
    // reach into the bitmap; K clears it)
    // (a thread's 16-byte pieces are all loaded before any is stored, so
    // their loads are in flight together)
    // With `adler`, the chunk's Adler-32 partials come from the staged bytes
    // (zlib_wrapper.cpp's checksum, combined per buffer by finalize_kernel):
    // s1 = sum of bytes, s2 = sum of (n - p) * byte = n * s1 - sum of p * byte.
    uint32_t ad_s = 0;   // this thread's byte sum
    uint64_t ad_p = 0;   // this thread's sum of p * byte
    {
        constexpr int PIECES = LC_MAX_CHUNK / (LC_THREADS * 16u);
        uint64_t va[PIECES], vb[PIECES];
#pragma unroll
        for (int k = 0; k < PIECES; ++k) {
            const uint32_t i = ((uint32_t)tid + (uint32_t)k * LC_THREADS) * 16u;
            va[k] = 0;
            vb[k] = 0;
            if (i + 16 <= n) {
                va[k] = ld64u(s + i);
                vb[k] = ld64u(s + i + 8);
            }
        }
#pragma unroll
        for (int k = 0; k < PIECES; ++k) {
            const uint32_t i = ((uint32_t)tid + (uint32_t)k * LC_THREADS) * 16u;
            if (i + 16 <= n) {
                D[(i >> 2) + 0] = (uint32_t)va[k];
                D[(i >> 2) + 1] = (uint32_t)(va[k] >> 32);
                D[(i >> 2) + 2] = (uint32_t)vb[k];
                D[(i >> 2) + 3] = (uint32_t)(vb[k] >> 32);
            } else if (i < n) {
                for (uint32_t w = i >> 2; w < ((n + 3) >> 2); ++w) {
                    uint32_t v = 0;
                    for (uint32_t b = 0; b < 4; ++b)
                        if (4 * w + b < n) {
                            const uint32_t byte = s[4 * w + b];
                            v |= byte << (8 * b);
                            ad_s += byte;
                            ad_p += (uint64_t)(4 * w + b) * byte;
                        }
                    D[w] = v;
                }
            }
        }
        if (adler) {
            // the thread's own full pieces, read back from LDS: their bytes,
            // and the bytes' offsets inside the piece, by dot products
#pragma unroll
            for (int k = 0; k < PIECES; ++k) {
                const uint32_t i = ((uint32_t)tid + (uint32_t)k * LC_THREADS) * 16u;
                if (i + 16 <= n) {
                    const uint4 q4 = *reinterpret_cast<const uint4*>(D + (i >> 2));
                    const uint32_t q[4] = {q4.x, q4.y, q4.z, q4.w};
                    uint32_t ps = 0, pw = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t sj = __builtin_amdgcn_udot4(q[j], 0x01010101u, 0u, false);
                        ps += sj;
                        pw += 4u * (uint32_t)j * sj + __builtin_amdgcn_udot4(q[j], 0x03020100u, 0u, false);
                    }
                    ad_s += ps;
                    ad_p += (uint64_t)i * ps + pw;
                }
            }
        }
    }
    if (tid < 4) D[((n + 3) >> 2) + tid] = 0;  // past the staged words
    if (adler) {
        // wave sums, then the 16 waves' sums (in the first-barrier words,
        // which P1 writes later) added by thread 0
#pragma unroll
        for (int dlt = 32; dlt; dlt >>= 1) {
            ad_s += (uint32_t)__shfl_xor((int)ad_s, dlt, 64);
            ad_p += (uint64_t)__shfl_xor((long long)ad_p, dlt, 64);
        }
        if (lane == 0) {
            smem[LB_FB_OFF + 3 * wv] = ad_s;
            smem[LB_FB_OFF + 3 * wv + 1] = (uint32_t)ad_p;
            smem[LB_FB_OFF + 3 * wv + 2] = (uint32_t)(ad_p >> 32);
        }
    }
    __syncthreads();
    if (adler && tid == 0) {
        uint64_t a = 0, pp = 0;
        for (int w = 0; w < LC_WAVES; ++w) {
            a += smem[LB_FB_OFF + 3 * w];
            pp += (uint64_t)smem[LB_FB_OFF + 3 * w + 1] | ((uint64_t)smem[LB_FB_OFF + 3 * w + 2] << 32);
        }
        info[c].adler_s1 = (uint32_t)(a % ADLER_MOD);
        info[c].adler_s2 = (uint32_t)(((uint64_t)n * a - pp) % ADLER_MOD);
    }

    // ---- A2: candidate distance from the candidate code (cand_code; for a
    // far candidate C is recovered by hash and its first byte compared,
    // lz77.cpp:178-184) and greedy match length (lz77.cpp:163-191).
    const uint32_t nh = n >= 4 ? n - 3 : 0;
    const uint32_t S = L.S;
    const uint32_t r0 = min(n, (uint32_t)tid * S), r1 = min(n, r0 + S);
    // layout index of the group of this range holding position q (r0 <= q < r1,
    // q - r0 a multiple of 8)
    auto gi = [&](uint32_t q) { return L.group((uint32_t)tid, (q - r0) >> 3); };
    // the range's farthest reach p + step(p), and whether every step is 258,
    // taken from A2's match lengths for P1
    uint32_t runmax = 0;
    bool all258 = r0 < r1;
    {
        const uint32_t lim = min(n, r1 + 257u);
        uint32_t cd = 0, ce = 0;
        bool cx = false;
        uint4 nx;  // the next group's eight 16-bit candidate codes
        // two groups ahead: every load reads a whole slot of the range
        // (a group past the range's end reads its last group again), and
        // the codes of positions at or past nh are cleared where they are used
        const uint32_t glast = r0 < r1 ? r0 + ((r1 - r0 - 1) & ~7u) : r0;
        auto slot = [&](uint32_t q0) -> uint4 {
            return *reinterpret_cast<const uint4*>(Q + L.group((uint32_t)tid, (min(q0, glast) - r0) >> 3));
        };
        nx = make_uint4(0u, 0u, 0u, 0u);
        uint4 nx2 = nx;
        if (r0 < r1) {
            nx = slot(r0);
            nx2 = slot(r0 + 8);
        }
        // the first group's codes are waited for here, so the loop's own
        // waits count only the loads and stores inside it
        asm volatile("" ::"v"(nx.x), "v"(nx.y), "v"(nx.z), "v"(nx.w));
        for (uint32_t p0 = r0; p0 < r1; p0 += 8) {
            uint4 lv = nx;
            if (p0 + 8 > nh) {
                uint32_t v[4] = {lv.x, lv.y, lv.z, lv.w};
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (p0 + k >= nh) v[k >> 1] &= ~(0xFFFFu << (16 * (k & 1)));
                lv = make_uint4(v[0], v[1], v[2], v[3]);
            }
            const uint32_t lw[4] = {lv.x, lv.y, lv.z, lv.w};
            nx = nx2;
            nx2 = slot(p0 + 16);
            const uint64_t w0 = lds_u64(D, p0), w1 = lds_u64(D, p0 + 8);
            uint32_t cw[2] = {0, 0};  // the group's u8 codes
            uint32_t xl = 0;          // 2 bits per position: length - 255 where it is 255..257
            bool big = false;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t p = p0 + k;
                const uint32_t lit = win_byte(w0, w1, k);
                const uint32_t e = (lw[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
                uint32_t d = 0, m1 = 0;
                if (p < r1 && e) {
                    if (!(e & 0x8000u)) {
                        d = e;
                    } else {
                        // the nearest same-hash position among p - dd - 32768 k
                        const uint32_t dd = e & 0x7FFFu;
                        const uint32_t hp = hash4<HB>((uint32_t)(win_u64(w0, w1, k) & 0xFFFFFFFFull));
                        for (uint32_t back = dd + 32768u; back <= p; back += 32768u) {
                            const uint32_t C = p - back;
                            if (hash4<HB>(lds_u32(D, C)) == hp) {
                                if (lds_u8(D, C) == lit) d = dd;
                                break;
                            }
                        }
                    }
                }
                if (d) {
                    const uint32_t cap = min(258u, n - p);
                    uint32_t L;
                    if (d == cd && p <= ce && (cx || ce >= p + cap)) {
                        L = ce - p;
                    } else {
                        // 8-byte compares only: every compare starts below
                        // lim <= n, so it reads at most 11 bytes past n, in
                        // the zero padding; a match that runs into it only
                        // overshoots n, and the cap below cuts it back.
                        const uint64_t diff = win_u64(w0, w1, k) ^ lds_u64(D, p - d);
                        uint32_t e = p + (diff ? (uint32_t)(__builtin_ctzll(diff) >> 3) : 8u);
                        bool x = diff != 0;
                        while (!x && e < lim) {
                            const uint64_t df = lds_u64(D, e) ^ lds_u64(D, e - d);
                            e += df ? (uint32_t)(__builtin_ctzll(df) >> 3) : 8u;
                            x = df != 0;
                        }
                        cd = d;
                        ce = e;
                        cx = x;
                        L = e - p;
                    }
                    L = min(L, cap);
                    m1 = L >= 3 ? L : 0u;
                }
                // a match's distance is its candidate code's low 15 bits
                // (d == e & 0x7FFF whenever m1 != 0), which C reads back
                big |= m1 >= 255u;
                xl |= (m1 >= 255u ? m1 - 255u : 0u) << (2 * k);
                const bool in = p < r1;
                runmax = max(runmax, in ? p + (m1 ? m1 : 1u) : 0u);
                all258 &= !in || m1 == 258u;
                cw[k >> 2] |= len_code8(m1, lit) << (8 * (k & 3));
            }
            if (big) {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (((cw[k >> 2] >> (8 * (k & 3))) & 0xFFu) == 255u) XS[L.index(p0 + k)] = (uint8_t)((xl >> (2 * k)) & 3u);
            }
            // codes of a last group's positions past r1 (literals) are never
            // read: no token starts there
            *reinterpret_cast<uint2*>(LCB + gi(p0)) = make_uint2(cw[0], cw[1]);
        }
    }
    __syncthreads();

    // ---- K: the u8 codes A2 wrote per group, to LDS over the bytes, 8 at a
    // time in range order (codes of a last group's positions past n land
    // below LC_REGION0 and are never read); bitmap cleared.
    // (loaded in batches of six groups, each batch's loads in flight together)
    for (uint32_t p0 = r0; p0 < r1; p0 += 48) {
        uint2 v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const uint32_t q = p0 + 8u * (uint32_t)k;
            v[k] = q < r1 ? *reinterpret_cast<const uint2*>(LCB + gi(q)) : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const uint32_t q = p0 + 8u * (uint32_t)k;
            if (q < r1) *reinterpret_cast<uint2*>(code + q) = v[k];
        }
    }
    for (uint32_t w = tid; w < ((n + 31) >> 5); w += LC_THREADS) bm[w] = 0;
    uint32_t* run258 = misc + 32;  // bit t: every position of range t steps 258
    uint32_t* b0 = misc + 64;      // bit t: range t's first parse barrier is its first position
    uint32_t* runs = misc + 96;    // (first start, starts) of the long 258-step runs; count in misc[24]
    if (tid < (int)(LC_THREADS / 32)) {
        run258[tid] = 0;
        b0[tid] = 0;
    }
    if (tid == 0) misc[24] = 0;
    __syncthreads();

    // ---- P1: first parse barrier of each range.  reach(p) = p + step(p);
    // q is a barrier when max over p < q of reach(p) == q (position 0 is one):
    // every greedy parse visits it.
    {
        if (all258) atomicOr(&run258[tid >> 5], 1u << (tid & 31));
        uint32_t v = runmax;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(v, d, 64);
            if (lane >= d) v = max(v, t);
        }
        if (lane == 63) misc[wv] = v;
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (int i = 0; i < LC_WAVES; ++i) {
                const uint32_t t = misc[i];
                misc[i] = acc;
                acc = max(acc, t);
            }
        }
        __syncthreads();
        const uint32_t prev_lane = __shfl_up(v, 1, 64);
        uint32_t carry = max(misc[wv], lane ? prev_lane : 0u);
        if (tid == 0) carry = 0;
        uint32_t first = 0xFFFFFFFFu;
        // a range of 258-byte steps (shorter than 258 positions): a token at
        // its first position reaches past its end, so only r0 can be a barrier
        const bool skip = all258 && L.S < 258u;
        if (skip) first = carry <= r0 ? r0 : 0xFFFFFFFFu;
        for (uint32_t p = r0; !skip && p < r1 && first == 0xFFFFFFFFu; p += 4) {
            const uint32_t cw = D[p >> 2];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (p + k < r1 && first == 0xFFFFFFFFu) {
                    if (carry <= p + k) first = p + k;
                    else carry = max(carry, p + k + code_step((cw >> (8 * k)) & 0xFFu, XS, L, p + k));
                }
            }
        }
        fb[tid] = r0 < r1 ? first : 0xFFFFFFFFu;
        if (r0 < r1 && first == r0) atomicOr(&b0[tid >> 5], 1u << (tid & 31));
    }
    __syncthreads();

    // ---- P2: exact greedy parse (lz77.cpp:378-406) from this range's
    // barrier to the first barrier of a later range; token starts -> bm.
    // Through ranges whose every step is 258 the token starts are pos + 258 k:
    // the run ends at the first start past n, outside such a range, or on the
    // first barrier of a later range, which in such a range can only be its
    // first position.  Both limits come from the run258 and b0 bitmasks, so
    // the walk only sets the run's bits.
    {
        const uint32_t pstart = fb[tid];
        if (pstart != 0xFFFFFFFFu) {
            uint32_t pos = pstart;
            while (pos < n) {
                if (pos >= r1 && fb[L.range_of(pos)] == pos) break;
                // eight literals inside this range (where no barrier test is
                // needed) with one 8-byte read of their codes
                if ((pos & 7u) == 0 && pos + 8 <= r1) {
                    const uint64_t c8 = *reinterpret_cast<const uint64_t*>(code + pos);
                    if ((c8 & 0xFEFEFEFEFEFEFEFEull) == 0) {
                        atomicOr(&bm[pos >> 5], 0xFFu << (pos & 31));
                        pos += 8;
                        continue;
                    }
                }
                atomicOr(&bm[pos >> 5], 1u << (pos & 31));
                const uint32_t b = code[pos];
                pos += code_step(b, XS, L, pos);
                if (b != 254u || pos >= n) continue;
                const uint32_t t = L.range_of(pos);
                if (!((run258[t >> 5] >> (t & 31)) & 1u)) continue;
                // E: the first range at or after t whose steps are not all 258
                uint32_t E = t;
                for (;;) {
                    const uint32_t zw = ~run258[E >> 5] >> (E & 31);
                    if (zw) {
                        E += (uint32_t)__builtin_ctz(zw);
                        break;
                    }
                    E = ((E >> 5) + 1u) << 5;
                    if (E >= LC_THREADS) break;
                }
                uint32_t lim = min(n, E * S);
                // the first barrier on the run's stride, of a range after this one
                for (uint32_t u = t; u < E;) {
                    const uint32_t bw = b0[u >> 5] >> (u & 31);
                    if (!bw) {
                        u = ((u >> 5) + 1u) << 5;
                        continue;
                    }
                    u += (uint32_t)__builtin_ctz(bw);
                    if (u >= E) break;
                    const uint32_t rs = u * S;
                    if (rs >= pos && rs >= r1 && (rs - pos) % 258u == 0u) {
                        lim = min(lim, rs);
                        break;
                    }
                    ++u;
                }
                const uint32_t cnt = (lim - pos + 257u) / 258u;  // pos < lim, or pos == lim (then 0)
                if (cnt >= 16u) {
                    // a long run (>= 4128 positions; the walks' runs are
                    // disjoint, so at most 31 in a chunk): its bits are set
                    // by the whole workgroup below
                    const uint32_t slot = atomicAdd(&misc[24], 1u);
                    runs[2 * slot] = pos;
                    runs[2 * slot + 1] = cnt;
                } else {
                    for (uint32_t k = 0; k < cnt; ++k) {
                        const uint32_t q = pos + 258u * k;
                        atomicOr(&bm[q >> 5], 1u << (q & 31));
                    }
                }
                pos += 258u * cnt;
            }
        }
    }
    __syncthreads();
    {
        const uint32_t nr = misc[24];
        for (uint32_t r = 0; r < nr; ++r) {
            const uint32_t p0 = runs[2 * r], cnt = runs[2 * r + 1];
            for (uint32_t k = (uint32_t)tid; k < cnt; k += LC_THREADS) {
                const uint32_t q = p0 + 258u * k;
                atomicOr(&bm[q >> 5], 1u << (q & 31));
            }
        }
    }
    __syncthreads();

    // ---- C: per range bit counts, block scan, stored / fixed choice
    // (deflate_block.cpp:414-421).  The block is assembled in LDS over the
    // dead length codes (a fixed block is under n + 5 bytes, so it may reach
    // three words into the bitmap, which every range has copied to registers
    // by then) and copied to the chunk's staging words in one coalesced pass.
    // The range's bitmap words (<= 5; the first and last may hold bits of
    // the neighbouring ranges, which no group of this range selects).
    const uint32_t w0 = r0 >> 5, nwd = r0 < r1 ? ((r1 + 31) >> 5) - w0 : 0u;
    const uint32_t bw0 = nwd > 0 ? bm[w0] : 0u, bw1 = nwd > 1 ? bm[w0 + 1] : 0u, bw2 = nwd > 2 ? bm[w0 + 2] : 0u,
                   bw3 = nwd > 3 ? bm[w0 + 3] : 0u, bw4 = nwd > 4 ? bm[w0 + 4] : 0u;
    // groups of 8 positions are streamed in range order, skipping groups
    // without a token start; the next groups' loads are issued before the
    // current group is used
    auto group_mask = [&](uint32_t p0) -> uint32_t {
        const uint32_t k = (p0 >> 5) - w0;
        const uint32_t wv32 = k == 0 ? bw0 : (k == 1 ? bw1 : (k == 2 ? bw2 : (k == 3 ? bw3 : bw4)));
        return (wv32 >> (p0 & 31)) & 0xFFu;
    };
    auto next_group = [&](uint32_t p0) -> uint32_t {  // first group at or after p0 with a token start
        while (p0 < r1 && group_mask(p0) == 0) p0 += 8;
        return p0;
    };
    // The groups with token starts, in order, with the loads of two more in
    // flight: the group's eight candidate codes (v[0..3], the distances) and,
    // for the emission pass, its eight u8 codes (v[4..5]) and input bytes
    // (v[6..7], the literals).  Every lane loads whole slots (a group past n
    // reads slot entries whose positions have no start bit) and a lane with
    // no such group loads group 0, so the loads carry no branch on the
    // group and the waits for them are exact.  Unrolled by three over fixed
    // buffers.
    auto stream = [&](auto emit, auto&& body) {
        constexpr bool EMIT = decltype(emit)::value;
        auto load = [&](uint32_t g, uint32_t* v) {
            const bool in = g < r1;
            const uint32_t x = in ? gi(g) : 0u;
            const uint4 q = *reinterpret_cast<const uint4*>(Q + x);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
            if (EMIT) {
                const uint2 cb = *reinterpret_cast<const uint2*>(LCB + x);
                v[4] = cb.x; v[5] = cb.y;
                const uint32_t pg = in ? g : 0u;
                uint64_t lt = 0;
                if (pg + 8 <= n) {
                    lt = ld64u(s + pg);
                } else {
                    for (uint32_t k = 0; k < 8; ++k)
                        if (pg + k < n) lt |= (uint64_t)s[pg + k] << (8 * k);
                }
                v[6] = (uint32_t)lt;
                v[7] = (uint32_t)(lt >> 32);
            }
        };
        uint32_t A[8], B[8], C[8];
        uint32_t ga = nwd ? next_group(r0) : r1;
        uint32_t gb = next_group(ga + 8);
        load(ga, A);
        load(gb, B);
        for (;;) {
            if (ga >= r1) break;
            uint32_t gc = next_group(gb + 8);
            load(gc, C);
            body(ga, A);
            ga = gb;
            gb = gc;
            if (ga >= r1) break;
            gc = next_group(gb + 8);
            load(gc, A);
            body(ga, B);
            ga = gb;
            gb = gc;
            if (ga >= r1) break;
            gc = next_group(gb + 8);
            load(gc, B);
            body(ga, C);
            ga = gb;
            gb = gc;
        }
    };
    // distance of position j (0..7) of a group from its candidate code
    auto dist_of = [](const uint32_t* v, uint32_t j) {
        const uint64_t q = j < 4u ? ((uint64_t)v[1] << 32 | v[0]) : ((uint64_t)v[3] << 32 | v[2]);
        return (uint32_t)(q >> (16u * (j & 3u))) & 0x7FFFu;
    };
    const uint32_t* ltab = smem + LB_LTAB_OFF;
    // Literals are counted by popcounts (8 bits, one more for codes 144..255);
    // each lane then visits only its own match starts.
    uint32_t mybits = 0;
    stream(std::false_type{}, [&](uint32_t g, const uint32_t* v) {
        const uint32_t m = group_mask(g);
        const uint2 c2 = *reinterpret_cast<const uint2*>(code + g);  // the group's u8 codes in LDS
        const uint64_t cb = (uint64_t)c2.y << 32 | c2.x;
        const uint32_t lm = zero_bytes8(cb & ~0x0101010101010101ull);  // codes 0 and 1
        const uint32_t hm = zero_bytes8(cb ^ 0x0101010101010101ull);   // code 1
        mybits += 8u * (uint32_t)__popc(m & lm) + (uint32_t)__popc(m & hm);
        for (uint32_t mm = m & ~lm; mm; mm &= mm - 1u) {
            const uint32_t j = (uint32_t)__builtin_ctz(mm);
            const uint32_t b = (uint32_t)(cb >> (8u * j)) & 0xFFu;
            const uint32_t d = dist_of(v, j);
            mybits += (ltab[b] >> 16) + 5u + (d <= 4u ? 0u : 30u - (uint32_t)__builtin_clz(d - 1u));
        }
    });
    const uint32_t incl = wave_incl_scan(mybits, lane);
    if (lane == 63) misc[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < LC_WAVES; ++i) {
            const uint32_t t = misc[i];
            misc[i] = acc;
            acc += t;
        }
        misc[LC_WAVES] = acc;
    }
    __syncthreads();
    const uint32_t excl = misc[wv] + incl - mybits;
    const uint64_t fixed = 3ull + misc[LC_WAVES] + 7ull;
    const bool is_last = chunk_final(desc[c].buf);
    const bool stored = (uint64_t)(n + 5) * 8ull <= fixed;
    if (tid == 0) {
        info[c].fixed_bits = fixed;
        info[c].raw_len = n;
        info[c].is_stored = stored ? 1u : 0u;
    }
    if (stored) return;
    uint32_t* out = smem;
    const uint32_t nw = (uint32_t)((fixed + 31) >> 5) + 1u;
    __syncthreads();  // bitmap words and the scan totals are in registers
    for (uint32_t i = tid; i < nw; i += LC_THREADS) out[i] = 0;
    __syncthreads();
    if (tid == 0) atomicOr(&out[0], (is_last ? 1u : 0u) | 2u);  // BFINAL, BTYPE = 01
    if (mybits) {
        const uint32_t bitpos = 3u + excl;
        uint32_t wi = bitpos >> 5;
        uint32_t nacc = bitpos & 31u;
        uint64_t acc = 0;
        // every output word starts at zero, so each completed word is ORed
        // in (a word shared with a neighbouring range takes both halves);
        // a token that completes no word ORs 0
        auto put = [&](uint32_t v, uint32_t len) {
            acc |= (uint64_t)v << nacc;
            nacc += len;
            const bool fl = nacc >= 32;
            atomicOr(&out[wi], fl ? (uint32_t)acc : 0u);
            wi += fl ? 1u : 0u;
            acc = fl ? (acc >> 32) : acc;
            nacc -= fl ? 32u : 0u;
        };
        stream(std::true_type{}, [&](uint32_t g, const uint32_t* v) {
            const uint32_t m = group_mask(g);
            const uint64_t cb = (uint64_t)v[5] << 32 | v[4];
            const uint64_t lt = (uint64_t)v[7] << 32 | v[6];
            // extra values of lengths 256 and 257 (code 255): 2 bits per position
            uint32_t xadd = 0;
            const uint32_t xm = m & zero_bytes8(~cb);
            for (uint32_t k = xm; k; k &= k - 1u) {
                const uint32_t j = (uint32_t)__builtin_ctz(k);
                xadd |= (uint32_t)XS[L.index(g + j)] << (2u * j);
            }
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if ((m >> j) & 1u) {
                    const uint32_t b = (uint32_t)(cb >> (8u * j)) & 0xFFu;
                    uint32_t val, len;
                    if (b <= 1u) {
                        const uint32_t byte = (uint32_t)(lt >> (8u * j)) & 0xFFu;
                        len = 8u + b;  // code 1: literal 144..255
                        val = __builtin_bitreverse32(byte + (byte < 144u ? 0x30u : 0x100u)) >> (32u - len);
                    } else {
                        const uint32_t ent = ltab[b] + (((xadd >> (2u * j)) & 3u) << 8);
                        uint32_t dv, dl;
                        fixed_dist(dist_of(v, j), dv, dl);
                        const uint32_t ml = ent >> 16;
                        val = (ent & 0xFFFFu) | (dv << ml);
                        len = ml + dl;
                    }
                    put(val, len);
                }
        });
        if (nacc) atomicOr(&out[wi], (uint32_t)acc);
    }
    __syncthreads();
    uint32_t* stg = stage + (uint64_t)c * stage_words;
    for (uint32_t i = tid; i < nw; i += LC_THREADS) stg[i] = out[i];
}

template <int HB>
__global__ __launch_bounds__(LC_THREADS) void l1_barrier_kernel(const uint8_t* __restrict__ src,
    const ChunkDesc* __restrict__ desc, const uint32_t* __restrict__ order, BatchHdr* __restrict__ hdr,
    uint32_t chunk, uint16_t* __restrict__ cand, uint32_t* __restrict__ tokw, uint32_t* __restrict__ stage,
    uint64_t stage_words, ChunkInfo* __restrict__ info, int adler) {
    extern __shared__ uint32_t smem[];
    for (uint32_t i = threadIdx.x; i < 256u; i += LC_THREADS) smem[LB_LTAB_OFF + i] = len_entry(i);
    for (;;) {
        const uint32_t k = next_item(&hdr->queue[Q_LZ], hdr->nchunks, smem);
        if (k == NO_ITEM) break;
        l1_barrier_body<HB>(order[k], smem, src, desc, chunk, cand, tokw, stage, stage_words, info, adler);
        __syncthreads();
    }
}

// Start-up check: same-address LDS exchanges issued by one wave instruction
// take effect in ascending lane order.
__global__ __launch_bounds__(64) void exchange_order_kernel(uint32_t* __restrict__ violations) {
    __shared__ uint32_t tab[64];
    const uint32_t lane = threadIdx.x, pat = blockIdx.x;
    tab[lane] = 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t mask = (pat & 3) == 0 ? 0u : ((pat & 3) == 1 ? 3u : ((pat & 3) == 2 ? 15u : 63u));
    const uint32_t a = ((lane * (pat | 1u) + (pat >> 2)) >> 1) & mask;
    const uint32_t old = atomicExch(&tab[a], lane);
    uint32_t want = 0xFFFFFFFFu;
    for (uint32_t j = 0; j < lane; ++j)
        if ((((j * (pat | 1u) + (pat >> 2)) >> 1) & mask) == a) want = j;
    if (old != want) atomicOr(&violations[0], 1u);
}

// Level 0: every chunk is stored blocks (grid-stride over chunks).
__global__ void mark_stored_kernel(const ChunkDesc* __restrict__ desc, const BatchHdr* __restrict__ hdr,
                                   ChunkInfo* __restrict__ info) {
    const uint32_t nchunks = hdr->nchunks;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        info[c].raw_len = desc[c].len;
        info[c].is_stored = 1;
        info[c].fixed_bits = 0;
    }
}

// ---------------------------------------------------------------------------
// Per buffer: scan of chunk bit offsets, checksum combine, framing.
//
// A chunk's size in bits depends on the incoming offset mod 8 only (stored
// blocks pad to a byte), so each chunk is the table f[r] = bits when it
// starts at an offset = r (mod 8).  Tables compose associatively,
// (A then B)[r] = A[r] + B[(r + A[r]) mod 8], and a Hillis-Steele scan of
// the compositions gives every chunk's offset.
// ---------------------------------------------------------------------------
constexpr int FIN_THREADS = 256;

__device__ __forceinline__ void chunk_table(const ChunkInfo& ci, uint64_t f[8]) {
    if (ci.is_stored == 1) {
#pragma unroll
        for (int r = 0; r < 8; ++r) f[r] = stored_bits((uint64_t)r, ci.raw_len);
    } else if (ci.is_stored == 2) {
#pragma unroll
        for (int r = 0; r < 8; ++r) f[r] = ci.sz8[r];
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) f[r] = ci.fixed_bits;
    }
}

__global__ __launch_bounds__(FIN_THREADS) void finalize_kernel(ChunkInfo* __restrict__ info,
                                                               const ChunkDesc* __restrict__ desc,
                                                               const BufDesc* __restrict__ bufs,
                                                               const BatchHdr* __restrict__ hdr, int format,
                                                               uint8_t* __restrict__ out_arena,
                                                               Result* __restrict__ res) {
    __shared__ uint64_t F[FIN_THREADS][8];
    __shared__ uint32_t red0[FIN_THREADS];
    __shared__ uint32_t red1[FIN_THREADS];
    const int tid = threadIdx.x;
    const uint32_t nbufs = hdr->nbufs;
    for (uint32_t bi = blockIdx.x; bi < nbufs; bi += gridDim.x) {
        const BufDesc B = bufs[bi];
        const bool piece = (B.piece & PIECE) != 0;
        // a stream piece starts at the bit offset its stream has reached
        uint64_t carry = piece ? (B.piece & 7u) : 0u;
        for (uint32_t t0 = 0; t0 < B.nchunks; t0 += FIN_THREADS) {
            const uint32_t i = t0 + (uint32_t)tid;
            const uint32_t cnt = min((uint32_t)FIN_THREADS, B.nchunks - t0);
            uint64_t f[8];
            if (i < B.nchunks) {
                chunk_table(info[B.chunk0 + i], f);
            } else {
#pragma unroll
                for (int r = 0; r < 8; ++r) f[r] = 0;
            }
            for (int d = 1; d < FIN_THREADS; d <<= 1) {
#pragma unroll
                for (int r = 0; r < 8; ++r) F[tid][r] = f[r];
                __syncthreads();
                if (tid >= d) {
                    uint64_t a[8];
#pragma unroll
                    for (int r = 0; r < 8; ++r) a[r] = F[tid - d][r];
                    uint64_t g[8];
#pragma unroll
                    for (int r = 0; r < 8; ++r) g[r] = a[r] + f[(r + a[r]) & 7];
#pragma unroll
                    for (int r = 0; r < 8; ++r) f[r] = g[r];
                }
                __syncthreads();
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) F[tid][r] = f[r];
            __syncthreads();
            const uint32_t r0 = (uint32_t)(carry & 7);
            if (i < B.nchunks) info[B.chunk0 + i].bit_off = carry + (tid ? F[tid - 1][r0] : 0ull);
            carry += F[cnt - 1][r0];
            __syncthreads();
        }
        const uint64_t payload_bits = carry;
        const uint64_t n = B.in_len;
        // checksum combine: a chunk ends `after` bytes before the buffer's end
        uint32_t x0 = 0, x1 = 0;
        for (uint32_t i = (uint32_t)tid; i < B.nchunks; i += FIN_THREADS) {
            const uint32_t c = B.chunk0 + i;
            const uint64_t after = B.in_off + n - (desc[c].in_off + desc[c].len);
            if (format == 2) {
                x0 ^= gf2_mul(gf2_x8n(after), info[c].crc);
            } else if (format == 1) {
                const uint32_t s1 = info[c].adler_s1;
                x0 = (x0 + s1) % ADLER_MOD;
                x1 = (uint32_t)((x1 + (uint64_t)s1 * (after % ADLER_MOD) + info[c].adler_s2) % ADLER_MOD);
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
            const uint32_t hb = piece ? 0u : (format == 1 ? 2u : (format == 2 ? 10u : 0u));
            const uint32_t tb = piece ? 0u : (format == 1 ? 4u : (format == 2 ? 8u : 0u));
            const uint64_t payload = (payload_bits + 7) >> 3;
            const uint64_t total = hb + payload + tb;
            Result& R = res[bi];
            R.payload_bits = payload_bits;
            R.checksum = ck;
            R.total_bytes = total <= B.cap ? total : 0;
            if (total <= B.cap && !piece) {
                uint8_t* out = out_arena + B.out_off;
                if (format == 1) {
                    out[0] = 0x78;
                    out[1] = 0x9C;
                    uint8_t* t = out + hb + payload;
                    t[0] = (uint8_t)(ck >> 24);
                    t[1] = (uint8_t)(ck >> 16);
                    t[2] = (uint8_t)(ck >> 8);
                    t[3] = (uint8_t)ck;
                } else if (format == 2) {
                    const uint8_t g[10] = {0x1F, 0x8B, 8, 0, 0, 0, 0, 0, 0, 255};
                    for (int k = 0; k < 10; ++k) out[k] = g[k];
                    uint8_t* t = out + hb + payload;
                    for (int k = 0; k < 4; ++k) t[k] = (uint8_t)(ck >> (8 * k));
                    for (int k = 0; k < 4; ++k) t[4 + k] = (uint8_t)((uint32_t)n >> (8 * k));
                }
            }
        }
        __syncthreads();
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
    uint32_t last;  // last chunk of its buffer
    uint32_t fin;   // ... and it ends the stream (BFINAL)
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
    if (q == 0) return (cv.fin && k == pieces - 1) ? 1u : 0u;  // header byte of pieces >= 1
    if (q == 1) return Lk & 0xFFu;
    if (q == 2) return (Lk >> 8) & 0xFFu;
    if (q == 3) return (~Lk) & 0xFFu;
    if (q == 4) return ((~Lk) >> 8) & 0xFFu;
    return cv.raw[dbase + q - 5u];
}

// Stored chunk: the four output bytes j..j+3 of the byte-aligned part
// (after the first block header's bits and pad) when they are all data
// bytes of one piece (pieces of up to 65535 bytes, each after a 5-byte
// header, or 4 bytes for the first, whose header byte shares the bits).
__device__ __forceinline__ bool stored_word(const ChunkView& cv, uint32_t j, uint32_t& v) {
    const uint32_t L = cv.len;
    const uint32_t L0 = min(L, 65535u);
    uint32_t q, Lk, dbase;
    if (j < 4u + L0) {
        q = j + 1u;
        Lk = L0;
        dbase = 0;
    } else {
        const uint32_t r = j - 4u - L0;
        const uint32_t k = 1u + r / 65540u;
        q = r - (k - 1u) * 65540u;
        Lk = min(L - k * 65535u, 65535u);
        dbase = k * 65535u;
    }
    if (q < 5u || q + 4u > 5u + Lk) return false;
    v = ld32u(cv.raw + dbase + q - 5u);
    return true;
}

// Stored chunk: the sixteen output bytes j..j+15 of the byte-aligned part
// when they are all data bytes of one piece (as stored_word).
__device__ __forceinline__ bool stored_unit(const ChunkView& cv, uint32_t j, uint32_t* v) {
    const uint32_t L = cv.len;
    const uint32_t L0 = min(L, 65535u);
    uint32_t q, Lk, dbase;
    if (j < 4u + L0) {
        q = j + 1u;
        Lk = L0;
        dbase = 0;
    } else {
        const uint32_t r = j - 4u - L0;
        const uint32_t k = 1u + r / 65540u;
        q = r - (k - 1u) * 65540u;
        Lk = min(L - k * 65535u, 65535u);
        dbase = k * 65535u;
    }
    if (q < 5u || q + 16u > 5u + Lk) return false;
    const uint64_t x = ld64u(cv.raw + dbase + q - 5u), y = ld64u(cv.raw + dbase + q + 3u);
    v[0] = (uint32_t)x;
    v[1] = (uint32_t)(x >> 32);
    v[2] = (uint32_t)y;
    v[3] = (uint32_t)(y >> 32);
    return true;
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
        const uint32_t hv = (cv.fin && pieces == 1) ? 1u : 0u;
        return (r == 0 ? hv : 0u) & mask;
    }
    return stored_byte(cv, (r - 3 - pad0) >> 3) & mask;
}

__device__ __forceinline__ ChunkView make_view(const uint8_t* src, const ChunkDesc* desc, const uint32_t* stage,
                                               uint64_t stage_words, const ChunkInfo* info, uint32_t c) {
    ChunkView v;
    v.stg = stage + (uint64_t)c * stage_words;
    v.raw = src + desc[c].in_off;
    v.off = info[c].bit_off;
    v.len = info[c].raw_len;
    v.stored = info[c].is_stored == 1 ? 1u : 0u;
    v.last = (desc[c].buf & CHUNK_LAST) ? 1u : 0u;
    v.fin = chunk_final(desc[c].buf) ? 1u : 0u;
    return v;
}

// Each chunk's bits, shifted to its offset in its buffer's output (grid-stride
// over chunk x slice: a batch of few chunks cuts each chunk's aligned output
// words into `slices` runs; the byte a chunk shares with the next is written
// by the first).
__global__ __launch_bounds__(256) void assemble_kernel(const uint8_t* __restrict__ src,
                                                       const ChunkDesc* __restrict__ desc,
                                                       const BufDesc* __restrict__ bufs,
                                                       BatchHdr* __restrict__ hdr,
                                                       const uint32_t* __restrict__ stage, uint64_t stage_words,
                                                       const ChunkInfo* __restrict__ info,
                                                       const Result* __restrict__ res, uint8_t* __restrict__ out_arena,
                                                       uint32_t hdr_bytes, uint32_t slices) {
    __shared__ uint32_t s_item;
    const uint32_t items = hdr->nchunks * slices;
    // chunks' outputs differ widely (a stored chunk copies its whole input,
    // a chunk of zeros a few hundred bytes): items come from a device queue
    // unless the grid covers them, one item per workgroup
    const bool queued = gridDim.x < items;
    for (uint32_t nx = blockIdx.x;; nx += gridDim.x) {
        uint32_t it = nx;
        if (queued) {
            it = next_item(&hdr->queue[Q_ASM], items, &s_item);
            if (it == NO_ITEM) break;
        } else if (it >= items) {
            break;
        }
        const uint32_t c = it / slices, sl = it % slices;
        const uint32_t bi = desc[c].buf & CHUNK_BUF;
        if (res[bi].total_bytes == 0) continue;
        const ChunkView cv = make_view(src, desc, stage, stage_words, info, c);
        const bool has_next = !cv.last;
        ChunkView nv = cv;
        if (has_next) nv = make_view(src, desc, stage, stage_words, info, c + 1);
        const uint64_t o = cv.off;
        const uint64_t e = has_next ? nv.off : res[bi].payload_bits;
        const uint64_t b0 = (o + 7) >> 3, b1 = (e + 7) >> 3;
        const uint32_t piece = bufs[bi].piece;
        uint8_t* dst = out_arena + bufs[bi].out_off + ((piece & PIECE) ? 0u : hdr_bytes);
        // a stream piece's first chunk starts inside byte 0, below the held
        // bits of the stream (every chunk is at least 10 bits, so byte 0
        // holds no bits of a second chunk)
        if ((piece & PIECE) && c == bufs[bi].chunk0 && (o & 7) && threadIdx.x == 0 && sl == 0)
            dst[o >> 3] = (uint8_t)(((piece >> 8) & ((1u << (o & 7)) - 1u)) |
                                    (chunk_bits(cv, 0, 8 - (uint32_t)(o & 7)) << (o & 7)));
        auto byte_at = [&](uint64_t b) -> uint32_t {
            const uint64_t r = 8 * b - o;
            const uint32_t k = (uint32_t)min<uint64_t>(8, e - 8 * b);
            uint32_t v = chunk_bits(cv, r, k);
            if (k < 8 && has_next) v |= chunk_bits(nv, 0, 8 - k) << k;
            return v;
        };
        // bytes before the first and after the last 16-byte-aligned output
        // unit one at a time; the aligned units in between with one 16-byte
        // store each
        const uint64_t a0 = (uint64_t)(uintptr_t)(dst + b0);
        const uint64_t ub0 = min(b1, b0 + ((16 - (a0 & 15)) & 15));  // first byte of the aligned run
        const uint64_t nunits = (b1 - ub0) >> 4;
        const uint64_t ub1 = ub0 + 16 * nunits;                         // first byte after it
        if (sl == 0) {
            for (uint64_t b = b0 + threadIdx.x; b < ub0; b += 256) dst[b] = (uint8_t)byte_at(b);
            for (uint64_t b = ub1 + threadIdx.x; b < b1; b += 256) dst[b] = (uint8_t)byte_at(b);
        }
        // a stored chunk's bytes are byte-aligned after its first header
        // bits and pad: output byte b is byte b - jb0 of that part
        const uint64_t jb0 = (o + 3 + ((8 - ((o + 3) & 7)) & 7)) >> 3;
        const uint64_t per = (nunits + slices - 1) / slices;
        const uint64_t ulo = min(nunits, per * sl), uhi = min(nunits, ulo + per);
        for (uint64_t u = ulo + threadIdx.x; u < uhi; u += 256) {
            const uint64_t b = ub0 + 16 * u;
            uint32_t v[4];
            if (!cv.stored && 8 * (b + 16) <= e) {
                // five staging words hold the unit's 128 bits
                const uint64_t r = 8 * b - o;
                const uint32_t* w = cv.stg + (r >> 5);
                const uint32_t sh = (uint32_t)(r & 31);
                const uint32_t s0 = w[0], s1 = w[1], s2 = w[2], s3 = w[3], s4 = w[4];
                v[0] = __builtin_amdgcn_alignbit(s1, s0, sh);
                v[1] = __builtin_amdgcn_alignbit(s2, s1, sh);
                v[2] = __builtin_amdgcn_alignbit(s3, s2, sh);
                v[3] = __builtin_amdgcn_alignbit(s4, s3, sh);
            } else if (!(cv.stored && 8 * (b + 16) <= e && b >= jb0 && stored_unit(cv, (uint32_t)(b - jb0), v))) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint64_t bk = b + 4 * k;
                    if (!(cv.stored && 8 * (bk + 4) <= e && bk >= jb0 && stored_word(cv, (uint32_t)(bk - jb0), v[k])))
                        v[k] = byte_at(bk) | (byte_at(bk + 1) << 8) | (byte_at(bk + 2) << 16) | (byte_at(bk + 3) << 24);
                }
            }
            *reinterpret_cast<uint4*>(dst + b) = make_uint4(v[0], v[1], v[2], v[3]);
        }
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

// Workgroups of `threads` threads and `smem` bytes of LDS resident per CU.
inline uint32_t per_cu(const Plan& p, size_t smem, int threads) {
    const uint32_t by_lds = smem ? (uint32_t)(p.lds_max / smem) : 32u;
    const uint32_t by_waves = (uint32_t)(32 / (threads / 64));
    const uint32_t k = by_lds < by_waves ? by_lds : by_waves;
    return k ? k : 1u;
}

template <class K>
hipError_t allow_lds(K* kernel, size_t smem) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)smem);
}

// Workgroups of `kernel` that stay resident on one CU, registers and LDS both
// counted; `fallback` when HIP cannot say.  A launch with fixed shares of
// equal items sized above this leaves the later workgroups' whole shares to
// run after the first ones finish.
template <class K>
uint32_t resident_per_cu(K* kernel, int threads, size_t smem, uint32_t fallback) {
    int r = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, reinterpret_cast<const void*>(kernel), threads, smem) !=
            hipSuccess ||
        r <= 0)
        return fallback;
    return (uint32_t)r;
}

template <int HB, int W, bool EXCH>
void launch_candidates(const Plan& p) {
    constexpr size_t smem = la_lds_bytes<HB, W>();
    // the candidate scans hold about 180 VGPRs (two waves per SIMD), fewer
    // workgroups than their LDS alone would allow
    static const uint32_t resident =
        resident_per_cu(l1_candidate_kernel<HB, W, EXCH>, 64 * W, smem, per_cu(p, smem, 64 * W));
    const unsigned g = grid_for(p.persistent_grid, p.nchunks, p.cus * resident);
    hipLaunchKernelGGL((l1_candidate_kernel<HB, W, EXCH>), dim3(g), dim3(64 * W), smem, p.stream, p.d_in, p.d_desc,
                       p.d_order, p.d_hdr, p.chunk, p.d_pd);
}

template <int HB, bool EXCH>
hipError_t launch_barrier(const Plan& p) {
    if (HB == 14) {
        // 64 KiB tables: two waves per chunk, one chunk per CU
        launch_candidates<14, 2, EXCH>(p);
    } else {
        // Few chunks: six waves per chunk, one chunk per CU, for the shortest
        // scan per chunk.  Many: one wave per chunk (16 KiB of LDS, no
        // fix-up of the later sub-ranges' first positions), eight chunks per
        // CU, which the scan's registers (two waves per SIMD) allow.  Both
        // give the same codes.
        if (p.nchunks <= 2u * (uint32_t)p.cus) launch_candidates<12, 6, EXCH>(p);
        else launch_candidates<12, 1, EXCH>(p);
    }
    const unsigned g = grid_for(p.persistent_grid, p.nchunks, p.cus * per_cu(p, LB_LDS_BYTES, LC_THREADS));
    hipLaunchKernelGGL((l1_barrier_kernel<HB>), dim3(g), dim3(LC_THREADS), LB_LDS_BYTES, p.stream, p.d_in, p.d_desc,
                       p.d_order, p.d_hdr, p.chunk, p.d_pd, p.d_tok, p.d_stage, p.stage_words, p.d_info,
                       p.format == 1 ? 1 : 0);
    return hipSuccess;
}

template <int HB>
hipError_t launch_wide(const Plan& p) {
    const size_t smem = L1Cfg<HB>::TAB_BYTES > L1_PHASE_B_BYTES ? L1Cfg<HB>::TAB_BYTES : L1_PHASE_B_BYTES;
    const unsigned g = grid_for(p.persistent_grid, p.nchunks, p.cus * per_cu(p, smem, L1_THREADS));
    hipLaunchKernelGGL(l1_lz77_kernel<HB>, dim3(g), dim3(L1_THREADS), smem, p.stream, p.d_in, p.d_desc, p.d_order,
                       p.d_hdr, p.chunk, p.d_tok, p.d_stage, p.stage_words, p.d_info);
    return hipSuccess;
}

// Whether launch_l1_class runs the barrier kernel, which then also makes the
// zlib checksum partials of its chunks.
bool l1_uses_barrier(const Plan& p) {
    if (p.level < 1 || p.level > 3) return false;
    if (p.whole || p.l1_dict || l1_big_chunk(p.chunk, p.hash_bits)) return false;
    return (p.hash_bits == 12 || p.hash_bits == 14) && p.chunk <= LC_MAX_CHUNK;
}

hipError_t launch_l1_class(const Plan& p) {
    if (p.whole || p.l1_dict || l1_big_chunk(p.chunk, p.hash_bits)) return launch_l1_big(p, p.whole);
    if (p.hash_bits == 12) {
        if (p.chunk <= LC_MAX_CHUNK) return p.exch_ok ? launch_barrier<12, true>(p) : launch_barrier<12, false>(p);
        if (p.chunk <= MAX_CHUNK_L1) return launch_wide<12>(p);
        return hipErrorInvalidValue;
    }
    if (p.hash_bits == 14 && p.chunk <= LC_MAX_CHUNK)
        return p.exch_ok ? launch_barrier<14, true>(p) : launch_barrier<14, false>(p);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t check_exchange_order_at(hipStream_t stream, uint32_t* d_word, bool* ok) {
    uint32_t h = 0;
    hipError_t e = hipMemsetAsync(d_word, 0, sizeof(uint32_t), stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(exchange_order_kernel, dim3(1024), dim3(64), 0, stream, d_word);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d_word, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    *ok = e == hipSuccess && h == 0;
    return e;
}

hipError_t check_exchange_order(hipStream_t stream, bool* ok) {
    uint32_t* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(uint32_t));
    if (e != hipSuccess) return e;
    e = check_exchange_order_at(stream, d, ok);
    (void)hipFree(d);
    return e;
}

// LDS limits of the L1-class kernels; set once per process before any launch
// (never while a stream is being captured).
hipError_t set_kernel_attributes_l1(size_t lds_max) {
    hipError_t e = allow_lds(&l1_barrier_kernel<12>, LB_LDS_BYTES);
    if (e == hipSuccess) e = allow_lds(&l1_barrier_kernel<14>, LB_LDS_BYTES);
    if (e == hipSuccess) e = allow_lds(&l1_candidate_kernel<12, 1, true>, la_lds_bytes<12, 1>());
    if (e == hipSuccess) e = allow_lds(&l1_candidate_kernel<12, 1, false>, la_lds_bytes<12, 1>());
    if (e == hipSuccess) e = allow_lds(&l1_candidate_kernel<12, 6, true>, la_lds_bytes<12, 6>());
    if (e == hipSuccess) e = allow_lds(&l1_candidate_kernel<12, 6, false>, la_lds_bytes<12, 6>());
    if (e == hipSuccess && lds_max >= la_lds_bytes<14, 2>())
        e = allow_lds(&l1_candidate_kernel<14, 2, true>, la_lds_bytes<14, 2>());
    if (e == hipSuccess && lds_max >= la_lds_bytes<14, 2>())
        e = allow_lds(&l1_candidate_kernel<14, 2, false>, la_lds_bytes<14, 2>());
    const size_t wide = L1Cfg<12>::TAB_BYTES > L1_PHASE_B_BYTES ? L1Cfg<12>::TAB_BYTES : L1_PHASE_B_BYTES;
    if (e == hipSuccess) e = allow_lds(&l1_lz77_kernel<12>, wide);
    return e;
}

uint32_t max_chunk_for_hash_bits(int hash_bits) { return hash_bits <= 12 ? MAX_CHUNK_L1 : MAX_CHUNK_L1 / 2; }

uint64_t stage_words_for_chunk(uint32_t chunk) { return ((uint64_t)chunk + chunk / 512 + 1024) / 4; }

hipError_t launch_parallel_compress(const Plan& p) {
    const hipStream_t s = p.stream;
    const uint32_t light = p.cus * 8u;
    record(p.ev[0], s);
    // zlib's Adler-32 partials of a barrier-path batch come from the barrier
    // kernel's staging pass instead
    if (p.format != 0 && !(p.format == 1 && l1_uses_barrier(p)))
        hipLaunchKernelGGL(checksum_kernel, dim3(grid_for(p.persistent_grid, p.nchunks, light)), dim3(256), 0, s,
                           p.d_in, p.d_desc, p.d_hdr, p.format == 2 ? 0 : 1, p.d_info);
    record(p.ev[1], s);
    hipError_t e = hipSuccess;
    if (p.level == 0) {
        hipLaunchKernelGGL(mark_stored_kernel, dim3(grid_for(p.persistent_grid, (p.nchunks + 255) / 256, light)),
                           dim3(256), 0, s, p.d_desc, p.d_hdr, p.d_info);
    } else if (p.level <= 3) {
        e = launch_l1_class(p);
    } else {
        e = launch_l6_front(p);
    }
    if (e != hipSuccess) return e;
    record(p.ev[2], s);
    hipLaunchKernelGGL(finalize_kernel, dim3(grid_for(p.persistent_grid, p.nbufs, light)), dim3(FIN_THREADS), 0, s,
                       p.d_info, p.d_desc, p.d_bufs, p.d_hdr, p.format, p.d_out, p.d_result);
    if (p.level >= 4) {
        e = launch_l6_back(p);
        if (e != hipSuccess) return e;
    }
    record(p.ev[3], s);
    const uint32_t hdr = p.format == 1 ? 2u : (p.format == 2 ? 10u : 0u);
    // a batch of few chunks splits each chunk's output over several workgroups
    const uint32_t slices = (p.persistent_grid || p.nchunks == 0) ? 1u : max(1u, min(32u, light / p.nchunks));
    hipLaunchKernelGGL(assemble_kernel, dim3(grid_for(p.persistent_grid, p.nchunks * slices, light)), dim3(256), 0, s,
                       p.d_in, p.d_desc, p.d_bufs, p.d_hdr, p.d_stage, p.stage_words, p.d_info, p.d_result, p.d_out,
                       hdr, slices);
    record(p.ev[4], s);
    return hipGetLastError();
}

hipError_t launch_checksum(const uint8_t* d_in, uint64_t n, int kind, uint32_t init, uint32_t* d_out_value,
                           ChunkInfo* d_tmp, uint32_t tmp_chunks, hipStream_t stream) {
    const uint32_t chunk = 65536;
    const uint32_t nb = n ? (uint32_t)((n + chunk - 1) / chunk) : 1;
    if (nb > tmp_chunks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(checksum_uniform_kernel, dim3(nb), dim3(256), 0, stream, d_in, n, chunk, kind, d_tmp);
    hipLaunchKernelGGL(checksum_combine_kernel, dim3(1), dim3(FIN_THREADS), 0, stream, d_tmp, nb, n, chunk, kind,
                       init, d_out_value);
    return hipGetLastError();
}

}  // namespace orot_gpu
