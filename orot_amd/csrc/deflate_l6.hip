// deflate_l6.hip -- gfx950 kernels of the L6-class path (levels 4..12).
//
// orot's standard path (src/deflate/lz77.cpp:435-496) walks 16-bit hash
// chains whose tables do not fit a workgroup's LDS; this path keeps the
// shape of orot's parse -- hash-chain candidates, longest match, nice length,
// one-position lazy evaluation with the "next > best + 1" rule -- with a
// definition that maps onto the GPU:
//   C(p)   latest q < p (q + 4 <= n) with the same 12-bit hash, every earlier
//          position inserted;
//   M(x)   longest match among at most max_chain chain candidates within
//          32 KiB whose first 4 bytes equal x's (ties keep the nearer);
//   lazy   the probe at pos + 1 is M(pos + 1).
// Chunks keep orot's block sizes (512 KiB for levels 4-6, 1 MiB above).  Tokens
// are cut into blocks of at most 16384 symbols, each the cheapest of stored /
// fixed / dynamic; dynamic codes are optimal length-limited (package-merge).
//
//   l6_pred_kernel    C(p) by 8 sub-range LDS tables (lane-ordered LDS
//                     exchange, or ballot match-any), stored as the in-window
//                     chain step PD[p] = p - C(p) (0: none within 32 KiB).
//   l6_match_kernel   M(x) for every x: the chunk is swept in tiles of 4096
//                     positions with the last 32 KiB of bytes and chain steps
//                     in LDS rings.
//   l6_decide_kernel  lazy decision per position (chunk segments in
//                     separate workgroups).
//   l6_parse_kernel   parse barriers, exact parse, tokens (orot's Token
//                     layout) in order.
//   l6_block_kernel   per block: histograms, package-merge lengths, code-length
//                     run-length header, exact sizes, block type.
//   l6_size_kernel    chunk size for each incoming bit offset mod 8.
//   l6_zero_kernel    clears each chunk's staging words.
//   l6_emit_kernel    per block: header and symbols into the chunk staging.
#include "deflate_gpu.h"
#include "device_common.h"

namespace orot_gpu {
namespace {

constexpr int L6_HB = 12;
constexpr uint32_t L6_TAB = 1u << L6_HB;
constexpr int L6P_THREADS = 512;
constexpr int L6P_W = 8;
constexpr uint32_t L6P_LDS = L6P_W * L6_TAB * 4u;  // 128 KiB
constexpr uint32_t L6P_UNRESOLVED = 0xFFFFu;  // chain step to take from an earlier sub-range (steps are <= 32768)

constexpr int L6M_THREADS = 1024;
constexpr uint32_t L6M_TILE = 4096;
constexpr uint32_t L6M_DRING = 65536;  // bytes, power of two
constexpr uint32_t L6M_PRING = 36864;  // chain steps: 32 KiB + one tile
constexpr uint32_t L6M_MIRROR = 4;     // words of the byte ring repeated past its end
constexpr uint32_t L6M_LDS = L6M_DRING + 4 * L6M_MIRROR + L6M_PRING * 2;
constexpr int L6M_WALKS = 4;  // chain walks in step per thread
constexpr uint32_t L6M_HIST = (32768 + L6M_TILE - 1) / L6M_TILE;  // tiles a walk reaches back
constexpr uint32_t L6M_MAX_SEGS = 16;  // workgroups sharing one chunk's pred / matching

constexpr int L6R_THREADS = 1024;
constexpr int L6R_WAVES = L6R_THREADS / 64;
constexpr uint32_t L6_MAX_CHUNK = 1u << 20;
constexpr uint32_t L6R_BM_WORDS = L6_MAX_CHUNK / 32;
constexpr uint32_t L6R_LDS = (L6R_BM_WORDS + 3 * L6R_THREADS + 64) * 4;
// the parse kernel's LDS for chunks of at most `chunk` bytes: the token
// bitmap, then the per-range arrays (512 KiB chunks: two workgroups per CU)
constexpr uint32_t l6r_bm_words(uint32_t chunk) { return (chunk + 31u) / 32u; }
constexpr uint32_t l6r_lds(uint32_t chunk) { return (l6r_bm_words(chunk) + 3 * L6R_THREADS + 64) * 4; }

constexpr uint32_t BLK_SYMS = 16384;

// Cross-chunk dictionary: a chunk that is not the first of its buffer sees up
// to L6_HALO bytes before it as positions -H..-1 (in its chains and as match
// sources); its tokens cover its own bytes only.  The first chunk of a stream
// piece sees the stream's bytes before the piece (BufDesc::hist of them, laid
// out just before it in the input arena).  Per-position arrays of the
// L6 class are laid out with room for the halo.
constexpr uint32_t L6_HALO = 32768;
__host__ __device__ constexpr uint64_t l6_stride(uint32_t chunk) { return ((uint64_t)chunk + L6_HALO + 7u) & ~7ull; }
__device__ __forceinline__ uint32_t l6_halo(uint32_t c, const ChunkDesc* desc, const BufDesc* bufs) {
    const BufDesc& B = bufs[desc[c].buf & CHUNK_BUF];
    const uint64_t rel = desc[c].in_off - B.in_off + B.hist;
    return (uint32_t)min<uint64_t>(rel, L6_HALO);
}
constexpr int L6B_THREADS = 256;


__device__ __forceinline__ uint32_t ml_len(uint32_t m) { return m & 0xFFFFu; }
__device__ __forceinline__ uint32_t ml_dist(uint32_t m) { return m >> 16; }

// ---------------------------------------------------------------------------
// C(p) and the in-window chain step PD[p].
// ---------------------------------------------------------------------------
// Segment `seg` of `segs` of chunk c writes PD for positions [pa, pb).  Its
// tables cover [lo, pb) with lo = pa - 32 KiB (or 0): a predecessor before lo
// is more than 32 KiB away and gives step 0 all the same.  The positions of
// [lo, pa) only fill the tables, so segments run in separate workgroups.
template <bool EXCH>
__device__ __forceinline__ void l6_pred_body(uint32_t c, uint32_t seg, uint32_t segs, uint32_t* smem,
                                             const uint8_t* __restrict__ src, const ChunkDesc* __restrict__ desc,
                                             const BufDesc* __restrict__ bufs, uint32_t chunk,
                                             uint32_t* __restrict__ tokw, uint16_t* __restrict__ pd) {
    // positions 0 .. n-1 are the halo (H bytes) followed by the chunk
    const uint32_t H = l6_halo(c, desc, bufs);
    const uint32_t n = desc[c].len + H;
    const uint8_t* s = src + desc[c].in_off - H;
    const uint64_t cp = l6_stride(chunk);
    (void)tokw;
    uint16_t* PD = pd + (uint64_t)c * cp;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t nh = n >= 4 ? n - 3 : 0;
    const uint32_t per = (((n + segs - 1) / segs) + 4095u) & ~4095u;
    const uint32_t pa = seg * per, pb = min(n, pa + per);
    if (pa >= pb) return;
    const uint32_t lo = pa > 32768u ? pa - 32768u : 0u;
    const uint32_t he = min(pb, nh);  // hashed positions of the window: [lo, he)
    const uint32_t SR = ((((he > lo ? he - lo : 0u) + L6P_W - 1) / L6P_W) + 63u) & ~63u;
    {
        uint32_t* tab = smem + (uint32_t)wv * L6_TAB;
        for (uint32_t i = lane; i < L6_TAB; i += 64) tab[i] = 0;
        const uint32_t a = lo + (uint32_t)wv * SR, b = min(a + SR, he);
        // Window loads of three more steps stay in flight, and every lane
        // stores (lanes outside the segment's part of the sub-range on
        // position n - 1, whose chain step nothing reads: walks start at
        // n - 4 or before and go back), so the waits are exact.
        auto load8 = [&](uint32_t q0, uint32_t* w) {
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = ld32u(s + min(q0 + 64u * j + (uint32_t)lane, b - 1));
        };
        auto step = [&](uint32_t p0, const uint32_t* wd) {
            uint32_t res[8];
            if (EXCH) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t p = p0 + 64u * j + (uint32_t)lane;
                    res[j] = p < b ? atomicExch(&tab[hash4<L6_HB>(wd[j])], p + 1u) : 0u;
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t q0 = p0 + 64u * j;
                    const uint32_t p = q0 + (uint32_t)lane;
                    const bool valid = p < b;
                    const uint32_t h = hash4<L6_HB>(wd[j]);
                    uint64_t m = ballot(valid);
#pragma unroll
                    for (int k = 0; k < L6_HB; ++k) {
                        const uint32_t bit = (h >> k) & 1u;
                        const uint64_t bk = ballot(bit != 0);
                        m &= bk ^ ((uint64_t)bit - 1ull);
                    }
                    res[j] = 0;
                    if (valid) {
                        const uint64_t below = m & ((1ull << lane) - 1ull);
                        if (below) res[j] = q0 + (uint32_t)(63 - __builtin_clzll(below)) + 1u;
                        else res[j] = tab[h];
                        if ((m >> lane) == 1ull) tab[h] = p + 1u;
                    }
                }
            }
            // the chain step, 0 outside 32 KiB; a position with no earlier
            // occurrence in a sub-range after the first is marked for the
            // resolution below.  History positions (below pa) belong to
            // another segment.
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t p = p0 + 64u * j + (uint32_t)lane;
                const uint32_t dist = p + 1u - res[j];
                const uint32_t v = res[j] ? (dist <= 32768u ? dist : 0u) : (wv > 0 ? L6P_UNRESOLVED : 0u);
                PD[p < b && p >= pa ? p : n - 1] = (uint16_t)v;
            }
        };
        uint32_t w0[8], w1[8], w2[8], w3[8];
        if (a < b) {
            load8(a, w0);
            load8(a + 512u, w1);
            load8(a + 1024u, w2);
        }
        for (uint32_t p0 = a; p0 < b; p0 += 2048u) {
            load8(p0 + 1536u, w3);
            step(p0, w0);
            if (p0 + 512u >= b) break;
            load8(p0 + 2048u, w0);
            step(p0 + 512u, w1);
            if (p0 + 1024u >= b) break;
            load8(p0 + 2560u, w1);
            step(p0 + 1024u, w2);
            if (p0 + 1536u >= b) break;
            load8(p0 + 3072u, w2);
            step(p0 + 1536u, w3);
        }
    }
    __syncthreads();
    // First occurrences of a sub-range take the latest earlier sub-range's
    // entry; positions past the last hashed one get step 0.  Eight
    // positions per thread and pass, read with one 16-byte load (pa and p0
    // are multiples of 8), written back where they lie in the segment.
    for (uint32_t p0 = pa + 8u * (uint32_t)tid; p0 < pb; p0 += 8u * L6P_THREADS) {
        const uint4 v4 = *reinterpret_cast<const uint4*>(PD + p0);
        uint32_t w[4] = {v4.x, v4.y, v4.z, v4.w};
        bool any = false;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t p = p0 + (uint32_t)k;
            const uint32_t cur = (w[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
            any |= p >= nh || cur == L6P_UNRESOLVED;
        }
        if (!any) continue;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t p = p0 + (uint32_t)k;
            uint32_t step = (w[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
            if (p >= nh) {
                step = 0;
            } else if (step == L6P_UNRESOLVED) {
                step = 0;
                const uint32_t h = hash4<L6_HB>(ld32u(s + p));
                for (int w2 = (int)((p - lo) / SR) - 1; w2 >= 0; --w2) {
                    const uint32_t C1 = smem[(uint32_t)w2 * L6_TAB + h];
                    if (C1) {
                        if (p - (C1 - 1u) <= 32768u) step = p - (C1 - 1u);
                        break;
                    }
                }
            }
            w[k >> 1] = (w[k >> 1] & ~(0xFFFFu << (16 * (k & 1)))) | (step << (16 * (k & 1)));
        }
        if (p0 + 8u <= pb) {
            *reinterpret_cast<uint4*>(PD + p0) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
            for (uint32_t k = 0; p0 + k < pb; ++k) PD[p0 + k] = (uint16_t)((w[k >> 1] >> (16 * (k & 1))) & 0xFFFFu);
        }
    }
}

template <bool EXCH>
__global__ __launch_bounds__(L6P_THREADS) void l6_pred_kernel(const uint8_t* __restrict__ src,
                                                              const ChunkDesc* __restrict__ desc,
                                                              const BufDesc* __restrict__ bufs,
                                                              const uint32_t* __restrict__ order,
                                                              BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                              uint32_t* __restrict__ tokw, uint16_t* __restrict__ pd,
                                                              uint32_t segs) {
    extern __shared__ uint32_t smem[];
    for (;;) {
        const uint32_t k = next_item(&hdr->queue[Q_PRED], hdr->nchunks * segs, smem);
        if (k == NO_ITEM) break;
        l6_pred_body<EXCH>(order[k / segs], k % segs, segs, smem, src, desc, bufs, chunk, tokw, pd);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// M(x) for every position: LDS rings of bytes (64 KiB) and chain steps.  The
// byte ring's first L6M_MIRROR words are repeated past its end, so the words
// of an unaligned read are consecutive in LDS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ring_u8(const uint32_t* R, uint32_t a) {
    return (R[(a >> 2) & (L6M_DRING / 4 - 1)] >> (8 * (a & 3u))) & 0xFFu;
}
__device__ __forceinline__ uint32_t ring_u32(const uint32_t* R, uint32_t a) {
    const uint32_t i = (a >> 2) & (L6M_DRING / 4 - 1), k = a & 3u;
    return __builtin_amdgcn_alignbyte(R[i + 1], R[i], k);
}
__device__ __forceinline__ uint64_t ring_u64(const uint32_t* R, uint32_t a) {
    const uint32_t i = (a >> 2) & (L6M_DRING / 4 - 1), k = a & 3u;
    const uint32_t w0 = R[i], w1 = R[i + 1], w2 = R[i + 2];
    return (uint64_t)__builtin_amdgcn_alignbyte(w1, w0, k) | ((uint64_t)__builtin_amdgcn_alignbyte(w2, w1, k) << 32);
}

// Segment `seg` of `segs` of chunk c: a contiguous run of its tiles, whose
// rings are first filled from the L6M_HIST tiles before it (a walk reaches
// back 32 KiB at most), so segments of one chunk run in separate workgroups.
__device__ __forceinline__ void l6_match_body(uint32_t c, uint32_t seg, uint32_t segs, uint32_t* smem,
                                              const uint8_t* __restrict__ src, const ChunkDesc* __restrict__ desc,
                                              const BufDesc* __restrict__ bufs, uint32_t chunk,
                                              const uint16_t* __restrict__ pd, uint32_t* __restrict__ mlw,
                                              const L6Params prm) {
    uint32_t* DR = smem;                                            // byte ring
    uint16_t* PR = reinterpret_cast<uint16_t*>(smem + L6M_DRING / 4 + L6M_MIRROR);  // chain-step ring
    // positions 0 .. n-1 are the halo (H bytes) followed by the chunk; M is
    // computed for the chunk's positions, ML[x - H]
    const uint32_t H = l6_halo(c, desc, bufs);
    const uint32_t n = desc[c].len + H;
    const uint8_t* s = src + desc[c].in_off - H;
    const uint64_t cp = l6_stride(chunk);
    const uint16_t* PD = pd + (uint64_t)c * cp;
    uint32_t* ML = mlw + (uint64_t)c * cp;
    const int tid = threadIdx.x;
    const int max_chain = prm.max_chain, nice = prm.nice_len;
    // tiles [t_lo, t_hi) hold this segment's positions; walks start at tile H / TILE
    const uint32_t tiles = (n + L6M_TILE - 1) / L6M_TILE, t_walk = H / L6M_TILE;
    const uint32_t per = (tiles - t_walk + segs - 1) / segs;
    const uint32_t t_lo = t_walk + seg * per, t_hi = min(tiles, t_lo + per);
    if (t_lo >= t_hi) return;
    const uint32_t x_lo = t_lo * L6M_TILE;
    const uint32_t x_first = (t_lo > L6M_HIST ? t_lo - L6M_HIST : 0u) * L6M_TILE;
    uint32_t dfill = x_first;  // bytes [x_first, dfill) have been loaded into the ring (word granular)
    for (uint32_t x0 = x_first; x0 < t_hi * L6M_TILE; x0 += L6M_TILE) {
        const uint32_t xe = min(n, x0 + L6M_TILE);
        // bytes up to xe + 258 + 8 (rounded to words), chain steps of [x0, xe)
        const uint32_t dneed = min((n + 3u) & ~3u, (xe + 258u + 8u + 3u) & ~3u);
        for (uint32_t a = dfill + 4u * (uint32_t)tid; a < dneed; a += 4u * L6M_THREADS) {
            uint32_t v;
            if (a + 4 <= n) {
                v = ld32u(s + a);
            } else {
                v = 0;
                for (uint32_t k = 0; k < 4; ++k)
                    if (a + k < n) v |= (uint32_t)s[a + k] << (8 * k);
            }
            const uint32_t wi = (a >> 2) & (L6M_DRING / 4 - 1);
            DR[wi] = v;
            if (wi < L6M_MIRROR) DR[L6M_DRING / 4 + wi] = v;
        }
        dfill = max(dfill, dneed);
        for (uint32_t q = x0 + tid; q < xe; q += L6M_THREADS) PR[q % L6M_PRING] = PD[q];
        __syncthreads();
        if (xe <= H || x0 < x_lo) continue;  // halo and history tiles only fill the rings
        // L6M_WALKS positions per thread, x + i * L6M_THREADS, walk their
        // chains in step, so each step has that many independent LDS loads
        // in flight.
        for (uint32_t xp = x0 + tid; xp < xe; xp += L6M_WALKS * L6M_THREADS) {
            struct Walk {
                uint32_t x, cap, cur, ri, best, bd, w, xbw;
                uint64_t xa, xb;
                int k;
                bool live;
            } m[L6M_WALKS];
#pragma unroll
            for (int i = 0; i < L6M_WALKS; ++i) {
                const uint32_t x = xp + (uint32_t)i * L6M_THREADS;
                m[i].x = x;
                m[i].best = 0;
                m[i].bd = 0;
                m[i].k = 0;
                m[i].cur = x;
                m[i].ri = x % L6M_PRING;  // ring slot of cur, kept as cur moves back
                m[i].xbw = 0;             // x's word at best - 3 (best >= 4)
                m[i].live = x < xe && x >= H && x + 4 <= n;
                m[i].cap = m[i].live ? min(258u, n - x) : 0u;
                // bytes x .. x+15 (beyond the chunk end they are stale or zero,
                // which only lengthens a comparison that min(l, cap) cuts back)
                m[i].xa = ring_u64(DR, x);
                m[i].xb = ring_u64(DR, x + 8);
                m[i].w = (uint32_t)m[i].xa;
            }
            // Each round: the chain steps of the walks, then their candidates'
            // first words and words at best - 3, then the comparisons, so the
            // walks' LDS reads overlap.  Dead walks read one fixed address (an
            // LDS broadcast, no bank conflict) and ignore the value.
            for (;;) {
                uint32_t st[L6M_WALKS];
#pragma unroll
                for (int i = 0; i < L6M_WALKS; ++i) st[i] = PR[m[i].live ? m[i].ri : 0u];
                bool any = false;
                uint32_t q[L6M_WALKS], c0[L6M_WALKS], c1[L6M_WALKS];
#pragma unroll
                for (int i = 0; i < L6M_WALKS; ++i) {
                    Walk& a = m[i];
                    q[i] = a.cur - st[i];
                    // bitwise: no exec-mask region per walk
                    const bool ok = a.live & (a.k < max_chain) & (st[i] != 0) & (a.x - q[i] <= 32768u);
                    a.live = ok;
                    any |= ok;
                    // a dead walk's cursor is never read again, so the step
                    // is taken without a branch
                    ++a.k;
                    a.cur = q[i];
                    a.ri = a.ri >= st[i] ? a.ri - st[i] : a.ri + L6M_PRING - st[i];  // st <= 32768 < L6M_PRING
                    // addresses masked to 0 for dead walks, kept opaque so the
                    // select is not turned back into branches around the reads
                    uint32_t a0 = q[i] & (0u - (uint32_t)ok);
                    uint32_t a1 = (q[i] + a.best - 3) & (0u - (uint32_t)(ok && a.best >= 4));
                    asm volatile("" : "+v"(a0), "+v"(a1));
                    c0[i] = ring_u32(DR, a0);
                    c1[i] = ring_u32(DR, a1);
                }
                if (!any) break;
#pragma unroll
                for (int i = 0; i < L6M_WALKS; ++i) {
                    Walk& a = m[i];
                    // only a match that also covers byte `best` can be longer
                    // than best (best < cap here, so these bytes are in the chunk)
                    if (!(a.live & (c0[i] == a.w) & ((a.best < 4) | (c1[i] == a.xbw)))) continue;
                    const uint32_t qq = q[i];
                    // bytes 0..3 are equal (c0): compare 4..7, then 8..15
                    const uint32_t d4 = ring_u32(DR, qq + 4) ^ (uint32_t)(a.xa >> 32);
                    // the length from d4 without a branch (bit 31 keeps ctz defined when
                    // d4 == 0, which the branch below replaces)
                    uint32_t l = 4 + (uint32_t)(__builtin_ctz(d4 | 0x80000000u) >> 3);
                    if (d4 == 0) {
                        const uint64_t df = ring_u64(DR, qq + 8) ^ a.xb;
                        l = df ? 8 + (uint32_t)(__builtin_ctzll(df) >> 3) : 16u;
                        if (df == 0) {
                            // 8-byte steps up to cap: bytes past the chunk's end
                            // in the ring only lengthen a compare that the cap
                            // below cuts back
                            while (l < a.cap) {
                                const uint64_t d2 = ring_u64(DR, a.x + l) ^ ring_u64(DR, qq + l);
                                if (d2) {
                                    l += (uint32_t)(__builtin_ctzll(d2) >> 3);
                                    break;
                                }
                                l += 8;
                            }
                        }
                    }
                    l = min(l, a.cap);
                    // l >= 4: the first four bytes passed
                    const bool better = l > a.best;
                    a.bd = better ? a.x - qq : a.bd;
                    // nothing later can be longer than nice (stop rule) or cap
                    a.live = a.live & !(better & (((int)l >= nice) | (l >= a.cap)));
                    if (better) a.xbw = ring_u32(DR, a.x + l - 3);
                    a.best = better ? l : a.best;
                }
            }
#pragma unroll
            for (int i = 0; i < L6M_WALKS; ++i)
                if (m[i].x < xe && m[i].x >= H) ML[m[i].x - H] = m[i].best >= 3 ? (m[i].best | (m[i].bd << 16)) : 0u;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(L6M_THREADS) void l6_match_kernel(const uint8_t* __restrict__ src,
                                                               const ChunkDesc* __restrict__ desc,
                                                               const BufDesc* __restrict__ bufs,
                                                               const uint32_t* __restrict__ order,
                                                               BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                               const uint16_t* __restrict__ pd,
                                                               uint32_t* __restrict__ mlw, const L6Params prm,
                                                               uint32_t segs) {
    // static LDS: its address is known when the kernel is compiled, so the
    // ring reads take it in their offset field instead of an add per read
    __shared__ uint32_t smem[L6M_LDS / 4];
    for (;;) {
        const uint32_t k = next_item(&hdr->queue[Q_MATCH], hdr->nchunks * segs, smem);
        if (k == NO_ITEM) break;
        l6_match_body(order[k / segs], k % segs, segs, smem, src, desc, bufs, chunk, pd, mlw, prm);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Lazy decisions, parse barriers, exact parse, tokens.
// Decision at a parse position p (lz77.cpp:442-494 with the probe at
// pos + 1 being M(pos + 1)): literal when p + 3 > n or M(p) < 3; otherwise
// up to lazy_depth literals while M(q + 1) > best + 1, then the match.
// code[p] = step (next decision - p) | lazy literals << 12.
// ---------------------------------------------------------------------------
// m0 = ML[p] and m1 = ML[p + 1] (when p + 1 < n), loaded by the caller.
__device__ __forceinline__ uint32_t l6_decide(const uint32_t* ML, uint32_t n, uint32_t p, int lazy, uint32_t m0,
                                              uint32_t m1) {
    if (p + 3 > n) return 1;
    uint32_t b = ml_len(m0);
    if (b < 3) return 1;
    uint32_t q = p;
    int k = lazy;
    while (k > 0 && q + b + 1 < n && q + 4 <= n) {
        const uint32_t nb = ml_len(q == p ? m1 : ML[q + 1]);
        if (nb > b + 1) {
            ++q;
            b = nb;
            --k;
        } else {
            break;
        }
    }
    return (q + b - p) | ((q - p) << 12);
}

// Decisions CODE[p] for every position of segment `seg` of `segs` of each
// chunk (coalesced over positions, eight per thread per pass with their match
// words loaded first, addresses clamped into the chunk).  Items are chunk x
// segment, grid-stride, counts from the batch header.
constexpr int L6D_THREADS = 1024;
// With gbm (the split parse) each item also clears the token-bitmap words of
// its positions in the chunk's staging words.
__global__ __launch_bounds__(L6D_THREADS) void l6_decide_kernel(const ChunkDesc* __restrict__ desc,
                                                                const BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                                const uint32_t* __restrict__ mlw,
                                                                uint16_t* __restrict__ codew, const L6Params prm,
                                                                uint32_t segs, uint32_t* __restrict__ gbm,
                                                                uint64_t gbm_stride) {
    const uint32_t items = hdr->nchunks * segs;
    const int lazy = prm.lazy_depth;
    for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t c = it / segs, seg = it % segs;
        const uint32_t n = desc[c].len;
        const uint32_t per = (((n + segs - 1) / segs) + 4095u) & ~4095u;
        const uint32_t pa = seg * per, pb = min(n, pa + per);
        if (pa >= pb) continue;
        const uint64_t cp = l6_stride(chunk);
        const uint32_t* ML = mlw + (uint64_t)c * cp;
        uint16_t* CODE = codew + (uint64_t)c * cp;
        if (gbm) {  // pa is a multiple of 4096
            uint32_t* G = gbm + (uint64_t)c * gbm_stride;
            for (uint32_t w = (pa >> 5) + threadIdx.x; w < ((pb + 31) >> 5); w += L6D_THREADS) G[w] = 0;
        }
        for (uint32_t p0 = pa + threadIdx.x; p0 < pb; p0 += 8 * L6D_THREADS) {
            uint32_t m0[8], m1[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t p = p0 + (uint32_t)k * L6D_THREADS;
                m0[k] = ML[min(p, n - 1)];
                m1[k] = ML[min(p + 1, n - 1)];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t p = p0 + (uint32_t)k * L6D_THREADS;
                if (p < pb) CODE[p] = (uint16_t)l6_decide(ML, n, p, lazy, m0[k], m1[k]);
            }
        }
    }
}

__device__ __forceinline__ void l6_parse_body(uint32_t c, uint32_t* smem, const uint8_t* __restrict__ src,
                                              const ChunkDesc* __restrict__ desc, uint32_t chunk,
                                              const uint32_t* __restrict__ mlw, uint16_t* __restrict__ codew,
                                              uint32_t* __restrict__ toks, uint32_t* __restrict__ ntok_out,
                                              uint32_t* __restrict__ blkpos, uint32_t maxblk, const L6Params prm) {
    uint32_t* bm = smem;
    uint32_t* fb = smem + l6r_bm_words(chunk);
    uint32_t* rcnt = fb + L6R_THREADS;
    uint32_t* misc = rcnt + L6R_THREADS;
    const uint32_t n = desc[c].len;
    const uint8_t* s = src + desc[c].in_off;
    const uint64_t cp = l6_stride(chunk);
    const uint32_t* ML = mlw + (uint64_t)c * cp;
    uint16_t* CODE = codew + (uint64_t)c * cp;
    uint32_t* TK = toks + (uint64_t)c * cp;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    (void)prm;
    const uint32_t S = max(32u, (((n + L6R_THREADS - 1) / L6R_THREADS) + 31u) & ~31u);
    const uint32_t r0 = min(n, (uint32_t)tid * S), r1 = min(n, r0 + S);
    // bitmap words owned by this range (S is a multiple of 32; an empty range
    // owns none, even when r0 falls inside the last range's final word)
    const uint32_t w0 = r0 >> 5, w1 = r0 < r1 ? (r1 + 31) >> 5 : w0;
    // the decisions CODE[] come from l6_decide_kernel; clear the bitmap and
    // the range's count of lazy literals (added by the exact parse below)
    for (uint32_t w = w0; w < w1; ++w) bm[w] = 0;
    rcnt[tid] = 0;
    __syncthreads();
    // the range's codes 32 at a time (r0 is a multiple of 32, so the 16-byte
    // loads are aligned; codes past n read the chunk's slack and are masked)
    auto codes32 = [&](uint32_t p, uint32_t* w) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint4 v = *reinterpret_cast<const uint4*>(CODE + p + 8 * i);
            w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
        }
    };
    auto code_of = [](const uint32_t* w, int k) -> uint32_t { return (w[k >> 1] >> (16 * (k & 1))) & 0xFFFFu; };
    // first parse barrier per range: max over p < q of (p + step(p)) == q
    {
        uint32_t runmax = 0;
        for (uint32_t p0 = r0; p0 < r1; p0 += 32) {
            uint32_t w[16];
            codes32(p0, w);
#pragma unroll
            for (int k = 0; k < 32; ++k)
                if (p0 + k < r1) runmax = max(runmax, p0 + k + (code_of(w, k) & 0xFFFu));
        }
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
            for (int i = 0; i < L6R_WAVES; ++i) {
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
        for (uint32_t p0 = r0; p0 < r1 && first == 0xFFFFFFFFu; p0 += 32) {
            uint32_t w[16];
            codes32(p0, w);
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const uint32_t p = p0 + k;
                if (p < r1 && first == 0xFFFFFFFFu) {
                    if (carry <= p) first = p;
                    else carry = max(carry, p + (code_of(w, k) & 0xFFFu));
                }
            }
        }
        fb[tid] = r0 < r1 ? first : 0xFFFFFFFFu;
    }
    __syncthreads();
    // exact parse from this range's barrier to the next range's barrier
    {
        const uint32_t pstart = fb[tid];
        if (pstart != 0xFFFFFFFFu) {
            uint32_t pos = pstart, wb = 0xFFFFFFFFu;
            uint32_t cw[16];
            while (pos < n) {
                if (pos >= r1 && fb[pos / S] == pos) break;
                atomicOr(&bm[pos >> 5], 1u << (pos & 31));
                // the decisions of pos's 32-position window, loaded together
                // when the walk enters it: one dependent load per window, not
                // per token
                if ((pos & ~31u) != wb) {
                    wb = pos & ~31u;
                    codes32(wb, cw);
                }
                const uint32_t cd = code_of(cw, (int)(pos & 31u));
                // the token's lazy literals, counted for the range holding pos
                if (cd >> 12) atomicAdd(&rcnt[pos / S], cd >> 12);
                pos += cd & 0xFFFu;
            }
        }
    }
    __syncthreads();
    // tokens per range (decision at p: lazy literals, then a match or one
    // literal): the range's token starts plus their lazy literals, both in
    // LDS, so no decisions are read again
    uint32_t cnt = rcnt[tid];
    for (uint32_t w = w0; w < w1; ++w) cnt += (uint32_t)__popc(bm[w]);
    const uint32_t incl = wave_incl_scan(cnt, lane);
    if (lane == 63) misc[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < L6R_WAVES; ++i) {
            const uint32_t t = misc[i];
            misc[i] = acc;
            acc += t;
        }
        ntok_out[c] = acc;
    }
    __syncthreads();
    uint32_t ti = misc[wv] + incl - cnt;
    // Tokens go out four at a time as one 16-byte store once ti is a multiple
    // of 4; the tokens of a group this range shares with its neighbours are
    // stored one by one.
    uint32_t tb0 = 0, tb1 = 0, tb2 = 0, tb3 = 0;
    bool grouped = (ti & 3u) == 0;
    auto put = [&](uint32_t t) {
        if (!grouped) {
            TK[ti++] = t;
            grouped = (ti & 3u) == 0;
            return;
        }
        const uint32_t k = ti & 3u;
        tb0 = k == 0 ? t : tb0;
        tb1 = k == 1 ? t : tb1;
        tb2 = k == 2 ? t : tb2;
        tb3 = t;
        if (k == 3) *reinterpret_cast<uint4*>(TK + ti - 3) = make_uint4(tb0, tb1, tb2, tb3);
        ++ti;
    };
    for (uint32_t w = w0; w < w1; ++w) {
        uint32_t x = bm[w];
        if (!x) continue;
        // the word's codes, and its 32 input bytes plus 8 more (a token's
        // lazy literals may run past the word), in registers: a token then
        // loads only the match word of a match
        const uint32_t pw = w << 5;
        uint32_t cw[16];
        codes32(pw, cw);
        uint64_t sb[5];
        if (pw + 40 <= n) {
#pragma unroll
            for (int i = 0; i < 5; ++i) sb[i] = ld64u(s + pw + 8 * i);
        } else {
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                uint64_t v = 0;
                for (uint32_t k = 0; k < 8; ++k)
                    if (pw + 8 * i + k < n) v |= (uint64_t)s[pw + 8 * i + k] << (8 * k);
                sb[i] = v;
            }
        }
        // wait for the word's loads here, once: the token loop's stores
        // share the load counter, so a first use inside the loop would
        // wait for every store of the earlier tokens on each iteration
        asm volatile("" ::"v"(cw[0]), "v"(cw[1]), "v"(cw[2]), "v"(cw[3]), "v"(cw[4]), "v"(cw[5]), "v"(cw[6]),
                     "v"(cw[7]), "v"(cw[8]), "v"(cw[9]), "v"(cw[10]), "v"(cw[11]), "v"(cw[12]), "v"(cw[13]),
                     "v"(cw[14]), "v"(cw[15]), "v"(sb[0]), "v"(sb[1]), "v"(sb[2]), "v"(sb[3]), "v"(sb[4]));
        auto byte_at = [&](uint32_t o) -> uint32_t {  // o < 40
            const uint64_t v = o < 8 ? sb[0] : (o < 16 ? sb[1] : (o < 24 ? sb[2] : (o < 32 ? sb[3] : sb[4])));
            return (uint32_t)(v >> (8 * (o & 7))) & 0xFFu;
        };
        while (x) {
            const int kk = __builtin_ctz(x);
            x &= x - 1;
            const uint32_t o = (uint32_t)kk, p = pw + o;
            const uint32_t cd = code_of(cw, kk);
            const uint32_t lits = cd >> 12, step = cd & 0xFFFu;
            for (uint32_t k = 0; k < lits; ++k) {
                if ((ti % BLK_SYMS) == 0 && ti / BLK_SYMS < maxblk) blkpos[(uint64_t)c * maxblk + ti / BLK_SYMS] = p + k;
                put(byte_at(o + k));
            }
            if ((ti % BLK_SYMS) == 0 && ti / BLK_SYMS < maxblk) blkpos[(uint64_t)c * maxblk + ti / BLK_SYMS] = p + lits;
            if (step > lits + 1) {
                const uint32_t m = ML[p + lits];
                put(0x80000000u | (ml_len(m) << 16) | ml_dist(m));
            } else {
                put(byte_at(o + lits));
            }
        }
    }
    // the tokens of the last, partly filled group
    if (grouped) {
        const uint32_t k = ti & 3u;
        if (k > 0) TK[ti - k] = tb0;
        if (k > 1) TK[ti - k + 1] = tb1;
        if (k > 2) TK[ti - k + 2] = tb2;
    }
}

__global__ __launch_bounds__(L6R_THREADS) void l6_parse_kernel(const uint8_t* __restrict__ src,
                                                               const ChunkDesc* __restrict__ desc,
                                                               const uint32_t* __restrict__ order,
                                                               BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                               const uint32_t* __restrict__ mlw,
                                                               uint16_t* __restrict__ codew,
                                                               uint32_t* __restrict__ toks,
                                                               uint32_t* __restrict__ ntok_out,
                                                               uint32_t* __restrict__ blkpos, uint32_t maxblk,
                                                               const L6Params prm) {
    extern __shared__ uint32_t smem[];
    for (;;) {
        const uint32_t k = next_item(&hdr->queue[Q_PARSE], hdr->nchunks, smem);
        if (k == NO_ITEM) break;
        l6_parse_body(order[k], smem, src, desc, chunk, mlw, codew, toks, ntok_out, blkpos, maxblk, prm);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Split parse: for batches of few chunks, each chunk is cut into segments of
// whole bitmap words, one workgroup each.  A decision's step is at most 3 lazy
// literals plus 258, so whether q is a parse barrier depends only on the
// decisions of the L6_REACH positions before q: a segment finds its own
// barriers from the reach of the positions just before it.  Three launches:
//   l6_pseg_parse_kernel  barriers and exact parse of the segment's ranges;
//                         token starts into the chunk's bitmap in global
//                         memory (its staging words, cleared by the decisions
//                         kernel).  The parse of the last range runs on into
//                         the next segment, up to the first barrier at or
//                         after the segment end (looked for over a bounded
//                         stretch; without one it runs to the chunk end,
//                         setting the bits that later segments set as well);
//   l6_pseg_count_kernel  tokens per segment;
//   l6_pseg_write_kernel  the segment's token offset (the counts of the
//                         segments before it) and its tokens, in order.
// ---------------------------------------------------------------------------
constexpr uint32_t L6_REACH = 262;
constexpr uint32_t L6_NEXT_SCAN = 4096;  // positions searched for the next segment's first barrier

__device__ __forceinline__ void pseg_bounds(uint32_t n, uint32_t segs, uint32_t g, uint32_t& pa, uint32_t& pb) {
    const uint32_t per = (((n + segs - 1) / segs) + 31u) & ~31u;
    pa = min(n, g * per);
    pb = min(n, pa + per);
}
__device__ __forceinline__ uint32_t pseg_range_len(uint32_t pa, uint32_t pb) {
    return max(32u, (((pb - pa + L6R_THREADS - 1) / L6R_THREADS) + 31u) & ~31u);
}
// the per-segment token counts follow the chunk's bitmap words
__host__ __device__ constexpr uint64_t pseg_count_word(uint32_t chunk) { return (chunk + 31u) / 32u + 1u; }

// 32 decisions from position p (a multiple of 32: aligned 16-byte loads)
__device__ __forceinline__ void codes32_at(const uint16_t* CODE, uint32_t p, uint32_t* w) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint4 v = *reinterpret_cast<const uint4*>(CODE + p + 8 * i);
        w[4 * i] = v.x;
        w[4 * i + 1] = v.y;
        w[4 * i + 2] = v.z;
        w[4 * i + 3] = v.w;
    }
}
__device__ __forceinline__ uint32_t code_in(const uint32_t* w, int k) { return (w[k >> 1] >> (16 * (k & 1))) & 0xFFFFu; }

// tokens of the decisions whose bits are set in bitmap words [w0, w1)
__device__ __forceinline__ uint32_t pseg_tokens(const uint32_t* G, const uint16_t* CODE, uint32_t w0, uint32_t w1) {
    uint32_t cnt = 0;
    for (uint32_t w = w0; w < w1; ++w) {
        uint32_t x = G[w];
        if (!x) continue;
        uint32_t cw[16];
        codes32_at(CODE, w << 5, cw);
        asm volatile("" ::"v"(cw[0]), "v"(cw[1]), "v"(cw[2]), "v"(cw[3]), "v"(cw[4]), "v"(cw[5]), "v"(cw[6]),
                     "v"(cw[7]), "v"(cw[8]), "v"(cw[9]), "v"(cw[10]), "v"(cw[11]), "v"(cw[12]), "v"(cw[13]),
                     "v"(cw[14]), "v"(cw[15]));
        while (x) {
            const int k = __builtin_ctz(x);
            x &= x - 1;
            cnt += 1u + (code_in(cw, k) >> 12);
        }
    }
    return cnt;
}

__global__ __launch_bounds__(L6R_THREADS) void l6_pseg_parse_kernel(const ChunkDesc* __restrict__ desc,
                                                                    const BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                                    const uint16_t* __restrict__ codew,
                                                                    uint32_t* __restrict__ gbm, uint64_t gbm_stride,
                                                                    uint32_t segs) {
    extern __shared__ uint32_t smem[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t items = hdr->nchunks * segs;
    uint32_t* bm = smem;                          // the segment's bitmap words
    uint32_t* fb = smem + l6r_bm_words(chunk);    // first barrier per range
    uint32_t* misc = fb + L6R_THREADS;
    for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t c = it / segs, g = it % segs;
        const uint32_t n = desc[c].len;
        uint32_t pa, pb;
        pseg_bounds(n, segs, g, pa, pb);
        if (pa >= pb) continue;
        const uint16_t* CODE = codew + (uint64_t)c * l6_stride(chunk);
        uint32_t* G = gbm + (uint64_t)c * gbm_stride;
        const uint32_t nwl = (pb - pa + 31) >> 5;
        const uint32_t S = pseg_range_len(pa, pb);
        const uint32_t r0 = min(pb, pa + (uint32_t)tid * S), r1 = min(pb, r0 + S);
        for (uint32_t w = tid; w < nwl; w += L6R_THREADS) bm[w] = 0;
        if (tid == 0) misc[32] = 0;
        __syncthreads();
        // the farthest reach of the positions before the segment
        if ((uint32_t)tid < L6_REACH && pa > (uint32_t)tid) {
            const uint32_t p = pa - 1 - (uint32_t)tid;
            atomicMax(&misc[32], p + (CODE[p] & 0xFFFu));
        }
        __syncthreads();
        const uint32_t carry_in = misc[32];
        uint32_t runmax = 0;
        for (uint32_t p0 = r0; p0 < r1; p0 += 32) {
            uint32_t w[16];
            codes32_at(CODE, p0, w);
#pragma unroll
            for (int k = 0; k < 32; ++k)
                if (p0 + k < r1) runmax = max(runmax, p0 + k + (code_in(w, k) & 0xFFFu));
        }
        uint32_t v = runmax;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(v, d, 64);
            if (lane >= d) v = max(v, t);
        }
        if (lane == 63) misc[wv] = v;
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = carry_in;
            for (int i = 0; i < L6R_WAVES; ++i) {
                const uint32_t t = misc[i];
                misc[i] = acc;
                acc = max(acc, t);
            }
            misc[33] = acc;  // reach of every position before pb
        }
        __syncthreads();
        const uint32_t prev_lane = __shfl_up(v, 1, 64);
        uint32_t carry = max(misc[wv], lane ? prev_lane : 0u);
        uint32_t first = 0xFFFFFFFFu;
        for (uint32_t p0 = r0; p0 < r1 && first == 0xFFFFFFFFu; p0 += 32) {
            uint32_t w[16];
            codes32_at(CODE, p0, w);
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const uint32_t p = p0 + k;
                if (p < r1 && first == 0xFFFFFFFFu) {
                    if (carry <= p) first = p;
                    else carry = max(carry, p + (code_in(w, k) & 0xFFFu));
                }
            }
        }
        fb[tid] = r0 < r1 ? first : 0xFFFFFFFFu;
        // the first barrier at or after pb
        if (tid == 0) {
            uint32_t nxt = 0xFFFFFFFFu;
            uint32_t cr = misc[33];
            for (uint32_t q = pb; q < min(n, pb + L6_NEXT_SCAN); ++q) {
                if (cr <= q) {
                    nxt = q;
                    break;
                }
                cr = max(cr, q + (CODE[q] & 0xFFFu));
            }
            misc[34] = nxt;
        }
        __syncthreads();
        const uint32_t nxt = misc[34];
        // exact parse from this range's barrier to the next barrier of a later range
        const uint32_t pstart = fb[tid];
        if (pstart != 0xFFFFFFFFu) {
            uint32_t pos = pstart;
            while (pos < n) {
                if (pos >= r1 && (pos < pb ? fb[(pos - pa) / S] == pos : pos == nxt)) break;
                if (pos < pb) atomicOr(&bm[(pos - pa) >> 5], 1u << (pos & 31));
                else atomicOr(&G[pos >> 5], 1u << (pos & 31));
                pos += CODE[pos] & 0xFFFu;
            }
        }
        __syncthreads();
        for (uint32_t w = tid; w < nwl; w += L6R_THREADS)
            if (bm[w]) atomicOr(&G[(pa >> 5) + w], bm[w]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(L6R_THREADS) void l6_pseg_count_kernel(const ChunkDesc* __restrict__ desc,
                                                                    const BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                                    const uint16_t* __restrict__ codew,
                                                                    uint32_t* __restrict__ gbm, uint64_t gbm_stride,
                                                                    uint32_t segs) {
    __shared__ uint32_t red[L6R_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t items = hdr->nchunks * segs;
    for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t c = it / segs, g = it % segs;
        const uint32_t n = desc[c].len;
        uint32_t pa, pb;
        pseg_bounds(n, segs, g, pa, pb);
        uint32_t* G = gbm + (uint64_t)c * gbm_stride;
        uint32_t cnt = 0;
        if (pa < pb) {
            const uint16_t* CODE = codew + (uint64_t)c * l6_stride(chunk);
            const uint32_t S = pseg_range_len(pa, pb);
            const uint32_t r0 = min(pb, pa + (uint32_t)tid * S), r1 = min(pb, r0 + S);
            if (r0 < r1) cnt = pseg_tokens(G, CODE, r0 >> 5, (r1 + 31) >> 5);
        }
        const uint32_t incl = wave_incl_scan(cnt, lane);
        if (lane == 63) red[wv] = incl;
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (int i = 0; i < L6R_WAVES; ++i) acc += red[i];
            G[pseg_count_word(chunk) + g] = acc;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(L6R_THREADS) void l6_pseg_write_kernel(const uint8_t* __restrict__ src,
                                                                    const ChunkDesc* __restrict__ desc,
                                                                    const BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                                    const uint32_t* __restrict__ mlw,
                                                                    const uint16_t* __restrict__ codew,
                                                                    const uint32_t* __restrict__ gbm,
                                                                    uint64_t gbm_stride, uint32_t segs,
                                                                    uint32_t* __restrict__ toks,
                                                                    uint32_t* __restrict__ ntok_out,
                                                                    uint32_t* __restrict__ blkpos, uint32_t maxblk) {
    __shared__ uint32_t misc[L6R_WAVES + 2];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t items = hdr->nchunks * segs;
    for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t c = it / segs, g = it % segs;
        const uint32_t n = desc[c].len;
        uint32_t pa, pb;
        pseg_bounds(n, segs, g, pa, pb);
        if (pa >= pb) {
            if (n == 0 && g == 0 && tid == 0) ntok_out[c] = 0;
            continue;
        }
        const uint8_t* s = src + desc[c].in_off;
        const uint64_t cp = l6_stride(chunk);
        const uint32_t* ML = mlw + (uint64_t)c * cp;
        const uint16_t* CODE = codew + (uint64_t)c * cp;
        uint32_t* TK = toks + (uint64_t)c * cp;
        const uint32_t* G = gbm + (uint64_t)c * gbm_stride;
        const uint32_t S = pseg_range_len(pa, pb);
        const uint32_t r0 = min(pb, pa + (uint32_t)tid * S), r1 = min(pb, r0 + S);
        const uint32_t w0 = r0 >> 5, w1 = r0 < r1 ? (r1 + 31) >> 5 : w0;
        const uint32_t cnt = pseg_tokens(G, CODE, w0, w1);
        const uint32_t incl = wave_incl_scan(cnt, lane);
        if (lane == 63) misc[wv] = incl;
        __syncthreads();
        if (tid == 0) {
            uint32_t base = 0;
            for (uint32_t h = 0; h < g; ++h) base += G[pseg_count_word(chunk) + h];
            uint32_t acc = base;
            for (int i = 0; i < L6R_WAVES; ++i) {
                const uint32_t t = misc[i];
                misc[i] = acc;
                acc += t;
            }
            if (pb == n) ntok_out[c] = acc;
        }
        __syncthreads();
        uint32_t ti = misc[wv] + incl - cnt;
        for (uint32_t w = w0; w < w1; ++w) {
            uint32_t x = G[w];
            if (!x) continue;
            // the word's codes, and its 32 input bytes plus 8 more (a token's
            // lazy literals may run past the word), in registers
            const uint32_t pw = w << 5;
            uint32_t cw[16];
            codes32_at(CODE, pw, cw);
            uint64_t sb[5];
            if (pw + 40 <= n) {
#pragma unroll
                for (int i = 0; i < 5; ++i) sb[i] = ld64u(s + pw + 8 * i);
            } else {
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    uint64_t v = 0;
                    for (uint32_t k = 0; k < 8; ++k)
                        if (pw + 8 * i + k < n) v |= (uint64_t)s[pw + 8 * i + k] << (8 * k);
                    sb[i] = v;
                }
            }
            // wait for the word's loads here, once: the token loop's stores
            // share the load counter, so a first use inside the loop would
            // wait for every store of the earlier tokens on each iteration
            asm volatile("" ::"v"(cw[0]), "v"(cw[1]), "v"(cw[2]), "v"(cw[3]), "v"(cw[4]), "v"(cw[5]), "v"(cw[6]),
                         "v"(cw[7]), "v"(cw[8]), "v"(cw[9]), "v"(cw[10]), "v"(cw[11]), "v"(cw[12]), "v"(cw[13]),
                         "v"(cw[14]), "v"(cw[15]), "v"(sb[0]), "v"(sb[1]), "v"(sb[2]), "v"(sb[3]), "v"(sb[4]));
            auto byte_at = [&](uint32_t o) -> uint32_t {  // o < 40
                const uint64_t v = o < 8 ? sb[0] : (o < 16 ? sb[1] : (o < 24 ? sb[2] : (o < 32 ? sb[3] : sb[4])));
                return (uint32_t)(v >> (8 * (o & 7))) & 0xFFu;
            };
            while (x) {
                const int kk = __builtin_ctz(x);
                x &= x - 1;
                const uint32_t o = (uint32_t)kk, p = pw + o;
                const uint32_t cd = code_in(cw, kk);
                const uint32_t lits = cd >> 12, step = cd & 0xFFFu;
                for (uint32_t k = 0; k < lits; ++k) {
                    if ((ti % BLK_SYMS) == 0 && ti / BLK_SYMS < maxblk) blkpos[(uint64_t)c * maxblk + ti / BLK_SYMS] = p + k;
                    TK[ti++] = byte_at(o + k);
                }
                if ((ti % BLK_SYMS) == 0 && ti / BLK_SYMS < maxblk) blkpos[(uint64_t)c * maxblk + ti / BLK_SYMS] = p + lits;
                if (step > lits + 1) {
                    const uint32_t m = ML[p + lits];
                    TK[ti++] = 0x80000000u | (ml_len(m) << 16) | ml_dist(m);
                } else {
                    TK[ti++] = byte_at(o + lits);
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Blocks: histograms, package-merge lengths, header, sizes, block type.
// ---------------------------------------------------------------------------
struct BlockRec {
    uint32_t t0, t1;        // token range
    uint32_t raw_pos, raw_len;
    uint32_t type;          // 0 stored, 1 fixed, 2 dynamic
    uint32_t hlit, hdist, hclen, nrle;
    uint32_t hdr_bits;      // dynamic: header bits after the 3-bit block header
    uint64_t body_bits;     // fixed / dynamic: bits of header + symbols + EOB
    uint8_t ll[288];
    uint8_t dl[32];
    uint8_t cl[19];
    uint8_t rle_sym[320];
    uint8_t rle_x[320];
};

__constant__ uint8_t c_len_extra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                        2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint8_t c_dist_extra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                                         6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
__constant__ uint8_t c_cl_order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

__device__ __forceinline__ uint32_t fixed_ll_len(uint32_t sym) {
    return sym < 144 ? 8u : (sym < 256 ? 9u : (sym < 280 ? 7u : 8u));
}

// Package-merge (Larmore & Hirschberg), same order as the oracle: symbols
// sorted by (freq, index), a leaf precedes a package of equal weight.  Called
// by every thread of the workgroup; idx / wleaf / wa / wb are LDS scratch and
// leafmask holds one bit per merged item of each level, set for leaves.
constexpr int LM_WORDS = 576 / 32;
__device__ void limited_lengths(const uint32_t* freq, int nsym, int max_bits, uint8_t* lens, uint16_t* idx,
                                uint32_t* wleaf, uint32_t* wa, uint32_t* wb, uint32_t* leafmask /* [16][LM_WORDS] */,
                                uint32_t* s_cnt, uint32_t* s_leaves /* [16] */) {
    const int tid = threadIdx.x;
    if (tid == 0) *s_cnt = 0;
    for (int i = tid; i < nsym; i += L6B_THREADS) lens[i] = 0;
    for (int i = tid; i < 16 * LM_WORDS; i += L6B_THREADS) leafmask[i] = 0;
    __syncthreads();
    // active symbols in (freq, symbol) order: each one's rank among the active
    for (int i = tid; i < nsym; i += L6B_THREADS) {
        const uint32_t f = freq[i];
        if (!f) continue;
        int r = 0;
        for (int j = 0; j < nsym; ++j) {
            const uint32_t g = freq[j];
            r += (g != 0 && (g < f || (g == f && j < i))) ? 1 : 0;
        }
        idx[r] = (uint16_t)i;
        atomicAdd(s_cnt, 1u);
    }
    __syncthreads();
    const int n = (int)*s_cnt;
    if (n == 0) return;
    if (n == 1) {
        if (tid == 0) lens[idx[0]] = 1;
        __syncthreads();
        return;
    }
    for (int i = tid; i < n; i += L6B_THREADS) {
        wleaf[i] = freq[idx[i]];
        wa[i] = wleaf[i];
    }
    for (int w = tid; w < LM_WORDS; w += L6B_THREADS) {
        const int lo = 32 * w;
        leafmask[max_bits * LM_WORDS + w] = n >= lo + 32 ? ~0u : (n > lo ? (1u << (n - lo)) - 1u : 0u);
    }
    __syncthreads();
    // each level: the sorted leaves merged with the sorted packages (pairs of the
    // level below); a leaf goes first on equal weight.  Leaf a lands at
    // a + #{packages lighter than it}, package p at p + #{leaves not heavier}.
    uint32_t* wprev = wa;
    uint32_t* wcur = wb;
    int cntp = n;
    for (int l = max_bits - 1; l >= 1; --l) {
        const int np = cntp / 2;
        for (int a = tid; a < n; a += L6B_THREADS) {
            const uint32_t f = wleaf[a];
            int lo = 0, hi = np;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (wprev[2 * mid] + wprev[2 * mid + 1] < f) lo = mid + 1;
                else hi = mid;
            }
            wcur[a + lo] = f;
            atomicOr(&leafmask[l * LM_WORDS + ((a + lo) >> 5)], 1u << ((a + lo) & 31));
        }
        for (int q = tid; q < np; q += L6B_THREADS) {
            const uint32_t pw = wprev[2 * q] + wprev[2 * q + 1];
            int lo = 0, hi = n;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (wleaf[mid] <= pw) lo = mid + 1;
                else hi = mid;
            }
            wcur[q + lo] = pw;
        }
        __syncthreads();
        uint32_t* t = wprev;
        wprev = wcur;
        wcur = t;
        cntp = n + np;
    }
    // Level l selects its first `take` items; the leaves among them are the
    // lightest leaves_l, and a leaf's length is the number of levels that
    // select it.  One wave walks the levels with popcounts of the masks.
    if (tid < 64) {
        int take = 2 * n - 2;
        for (int l = 1; l <= max_bits; ++l) {
            uint32_t cnt = 0;
            const int lo = 32 * tid;
            if (take > 0 && tid < LM_WORDS && lo < take) {
                uint32_t w = leafmask[l * LM_WORDS + tid];
                if (take - lo < 32) w &= (1u << (take - lo)) - 1u;
                cnt = (uint32_t)__popc(w);
            }
            for (int d = 32; d; d >>= 1) cnt += __shfl_xor(cnt, d, 64);
            if (tid == 0) s_leaves[l] = cnt;
            take = 2 * (take - (int)cnt);
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += L6B_THREADS) {
        int len = 0;
        for (int l = 1; l <= max_bits; ++l) len += i < (int)s_leaves[l] ? 1 : 0;
        lens[idx[i]] = (uint8_t)len;
    }
    __syncthreads();
}

// Workgroup sum of one value per thread; red is LDS scratch of one word per wave.
__device__ __forceinline__ uint32_t l6b_sum(uint32_t v, uint32_t* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int d = 32; d; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < L6B_THREADS / 64; ++w) t += red[w];
    __syncthreads();
    return t;
}

// Code-length RLE items of one maximal run of `m` equal lengths `v`: the
// greedy rule of deflate_block.cpp:347-370 restarted at each item, whose run
// from any position inside the maximal run is min(rest, 138).  emit(sym, x)
// takes x = 0xFF for a plain length.
template <class F>
__device__ __forceinline__ uint32_t rle_run_items(uint32_t v, uint32_t m, F&& emit) {
    uint32_t items = 0;
    while (m) {
        const uint32_t run = m < 138u ? m : 138u;
        if (v == 0 && run >= 3) {
            emit(run <= 10 ? 17u : 18u, run <= 10 ? run - 3 : run - 11);
            ++items;
            m -= run;
        } else if (v != 0 && run >= 4) {
            const uint32_t rep = run - 1 < 6 ? run - 1 : 6u;
            emit(v, 0xFFu);
            emit(16u, rep - 3);
            items += 2;
            m -= 1 + rep;
        } else {
            emit(v, 0xFFu);
            ++items;
            --m;
        }
    }
    return items;
}

__device__ __forceinline__ void l6_block_body(uint32_t c, uint32_t b, const uint32_t* __restrict__ toks,
                                              const uint32_t* __restrict__ ntok_in,
                                              const uint32_t* __restrict__ blkpos, uint32_t chunk, uint32_t maxblk,
                                              BlockRec* __restrict__ recs) {
    __shared__ uint32_t lit[288];
    __shared__ uint32_t dst[32];
    __shared__ uint16_t idx[288];
    __shared__ uint32_t wleaf[288];
    __shared__ uint32_t wa[576];
    __shared__ uint32_t wb[576];
    __shared__ uint32_t leafmask[16 * LM_WORDS];
    __shared__ uint32_t s_leaves[16];
    __shared__ uint32_t clf[19];
    __shared__ uint32_t red[3][L6B_THREADS / 64];
    __shared__ uint32_t s_cnt, s_hl, s_hd;
    __shared__ BlockRec R;
    const uint32_t ntok = ntok_in[c];
    const uint32_t nblk = ntok ? (ntok + BLK_SYMS - 1) / BLK_SYMS : 1;
    if (b >= nblk) return;
    const uint64_t cp = l6_stride(chunk);
    const uint32_t* TK = toks + (uint64_t)c * cp;
    const uint32_t t0 = b * BLK_SYMS, t1 = min(ntok, t0 + BLK_SYMS);
    const int tid = threadIdx.x, lane = tid & 63;
    for (int i = tid; i < 288; i += L6B_THREADS) lit[i] = 0;
    if (tid < 32) dst[tid] = 0;
    if (tid < 19) clf[tid] = 0;
    if (tid == 0) {
        s_hl = 0;
        s_hd = 0;
    }
    __syncthreads();
    // histogram: t0 is a multiple of BLK_SYMS and the token stride a multiple
    // of 8, so 16-byte loads are aligned; the words read past t1 lie inside
    // the chunk's stride and are ignored
    constexpr int HQ = BLK_SYMS / L6B_THREADS / 8;
    const uint4* T4 = reinterpret_cast<const uint4*>(TK + t0);
    uint32_t raw = 0;
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
        uint4 q[HQ];
#pragma unroll
        for (int k = 0; k < HQ; ++k) {
            const uint32_t g = (uint32_t)tid + L6B_THREADS * (uint32_t)(half * HQ + k);
            q[k] = t0 + 4u * g < t1 ? T4[g] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int k = 0; k < HQ; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t g = (uint32_t)tid + L6B_THREADS * (uint32_t)(half * HQ + k);
                if (t0 + 4u * g + j < t1) {
                    const uint32_t v = j == 0 ? q[k].x : (j == 1 ? q[k].y : (j == 2 ? q[k].z : q[k].w));
                    if (v & 0x80000000u) {
                        const uint32_t L = (v >> 16) & 0x1FFu, d = v & 0xFFFFu;
                        atomicAdd(&lit[257 + len_sym(L).code], 1u);
                        atomicAdd(&dst[dist_sym(d).code], 1u);
                        raw += L;
                    } else {
                        atomicAdd(&lit[v & 0xFFu], 1u);
                        raw += 1;
                    }
                }
            }
    }
    const uint32_t rawsum = l6b_sum(raw, red[0]);  // its barriers also publish the histogram
    if (tid == 0) {
        lit[256] += 1;  // EOB
        R.t0 = t0;
        R.t1 = t1;
        R.raw_len = rawsum;
        R.raw_pos = b == 0 ? 0u : blkpos[(uint64_t)c * maxblk + b];
    }
    __syncthreads();
    // lengths
    limited_lengths(lit, 288, 15, R.ll, idx, wleaf, wa, wb, leafmask, &s_cnt, s_leaves);
    limited_lengths(dst, 32, 15, R.dl, idx, wleaf, wa, wb, leafmask, &s_cnt, s_leaves);
    // dynamic header fix-ups: EOB and at least one distance code
    // (deflate_block.cpp:196-290)
    if (tid < 64) {
        const uint64_t any = ballot(tid < 32 && R.dl[tid] != 0);
        if (tid == 0) {
            if (!R.ll[256]) R.ll[256] = 1;
            if (!any) R.dl[0] = 1;
        }
    }
    __syncthreads();
    // fixed and dynamic sizes (deflate_block.cpp:50-92 for the fixed estimate)
    // and the trimmed counts; a block's sizes stay far below 2^32 bits
    uint32_t fx = 0, ex = 0, dd = 0;
    for (int i = tid; i < 288; i += L6B_THREADS) {
        const uint32_t f = lit[i];
        fx += fixed_ll_len((uint32_t)i) * f;
        dd += (uint32_t)R.ll[i] * f;
        if (i >= 257 && i < 286) ex += (uint32_t)c_len_extra[i - 257] * f;
        if (i < 286 && R.ll[i]) atomicMax(&s_hl, (uint32_t)i + 1u);
    }
    if (tid < 32) {
        const uint32_t f = dst[tid];
        fx += 5u * f;
        dd += (uint32_t)R.dl[tid] * f;
        if (tid < 30) ex += (uint32_t)c_dist_extra[tid] * f;
        if (tid < 30 && R.dl[tid]) atomicMax(&s_hd, (uint32_t)tid + 1u);
    }
    const uint64_t extra = l6b_sum(ex, red[0]);
    const uint64_t fixedb = 3ull + l6b_sum(fx, red[1]) + extra;
    const uint64_t dyn_data = l6b_sum(dd, red[2]) + extra;
    const uint32_t hlit = max(257u, s_hl), hdist = max(1u, s_hd);
    // code-length RLE over ll[0..hlit) dl[0..hdist) (deflate_block.cpp:347-370):
    // thread t takes positions 2t and 2t + 1; a maximal run of equal lengths
    // is encoded by the thread holding its first position
    const uint32_t total = hlit + hdist;
    auto seq = [&](uint32_t k) -> uint32_t { return k < hlit ? R.ll[k] : R.dl[k - hlit]; };
    // A run's length is the distance to the next run start: start flags by
    // ballot within a wave, each wave's first start in LDS for the others.
    uint32_t mrun[2] = {0, 0}, vrun[2] = {0, 0}, nitems = 0;
    {
        const uint32_t wv = (uint32_t)tid >> 6;
        uint32_t* fs = red[1];  // first run start of each wave (red[1] is free after the sizes)
        const uint32_t k0 = 2u * (uint32_t)tid, k1 = k0 + 1u;
        const uint32_t v0 = k0 < total ? seq(k0) : 0u, v1 = k1 < total ? seq(k1) : 0u;
        const bool s0 = k0 < total && (k0 == 0 || seq(k0 - 1) != v0);
        const bool s1 = k1 < total && v0 != v1;
        const uint64_t B0 = ballot(s0), B = B0 | ballot(s1);
        auto start_of = [&](uint32_t l) { return 2u * (64u * wv + l) + (((B0 >> l) & 1ull) ? 0u : 1u); };
        if (lane == 0) fs[wv] = B ? start_of((uint32_t)__builtin_ctzll(B)) : 0xFFFFFFFFu;
        __syncthreads();
        // the first run start in a later lane of this wave, or in a later wave
        auto next_start = [&]() -> uint32_t {
            const uint64_t M = lane == 63 ? 0ull : (B >> (lane + 1)) << (lane + 1);
            if (M) return start_of((uint32_t)__builtin_ctzll(M));
            for (uint32_t w = wv + 1; w < L6B_THREADS / 64; ++w)
                if (fs[w] != 0xFFFFFFFFu) return fs[w];
            return total;
        };
        if (s0) {
            mrun[0] = s1 ? 1u : next_start() - k0;
            vrun[0] = v0;
            nitems += rle_run_items(v0, mrun[0], [](uint32_t, uint32_t) {});
        }
        if (s1) {
            mrun[1] = next_start() - k1;
            vrun[1] = v1;
            nitems += rle_run_items(v1, mrun[1], [](uint32_t, uint32_t) {});
        }
    }
    const uint32_t incl = wave_incl_scan(nitems, lane);
    if (lane == 63) red[0][tid >> 6] = incl;
    __syncthreads();
    uint32_t o = incl - nitems;
    for (int w = 0; w < (tid >> 6); ++w) o += red[0][w];
    uint32_t nr = 0;
    for (int w = 0; w < L6B_THREADS / 64; ++w) nr += red[0][w];
    for (int h = 0; h < 2; ++h) {
        if (!mrun[h]) continue;
        rle_run_items(vrun[h], mrun[h], [&](uint32_t sy, uint32_t x) {
            R.rle_sym[o] = (uint8_t)sy;
            R.rle_x[o] = (uint8_t)x;
            atomicAdd(&clf[sy], 1u);
            ++o;
        });
    }
    __syncthreads();
    limited_lengths(clf, 19, 7, R.cl, idx, wleaf, wa, wb, leafmask, &s_cnt, s_leaves);
    // a single code length code gets a partner; HCLEN trimmed in transmission order
    if (tid < 64) {
        const uint64_t act = ballot(tid < 19 && R.cl[tid] != 0);
        if (tid == 0 && __popcll(act) == 1) {
            for (int i = 0; i < 19; ++i)
                if (!R.cl[i]) {
                    R.cl[i] = 1;
                    break;
                }
        }
    }
    __syncthreads();
    uint32_t hclen = 4;
    {
        const uint64_t nz = ballot(lane < 19 && R.cl[c_cl_order[lane < 19 ? lane : 0]] != 0);
        if (nz) hclen = max(4u, 64u - (uint32_t)__clzll(nz));
    }
    uint32_t ib = 0;
    for (uint32_t i = (uint32_t)tid; i < nr; i += L6B_THREADS) {
        const uint32_t sy = R.rle_sym[i];
        ib += R.cl[sy] + (sy == 16 ? 2u : (sy == 17 ? 3u : (sy == 18 ? 7u : 0u)));
    }
    const uint32_t hb = 5 + 5 + 4 + 3u * hclen + l6b_sum(ib, red[0]);
    if (tid == 0) {
        // data bits with the fixed-up lengths (EOB / dist[0] may have changed)
        const uint64_t dynb = 3ull + hb + dyn_data;
        const uint64_t storedb = ((uint64_t)R.raw_len + 5) * 8;
        R.hlit = hlit;
        R.hdist = hdist;
        R.nrle = nr;
        R.hclen = hclen;
        R.hdr_bits = hb;
        if (storedb <= fixedb && storedb <= dynb) {
            R.type = 0;
            R.body_bits = 0;
        } else if (fixedb <= dynb) {
            R.type = 1;
            R.body_bits = fixedb;
        } else {
            R.type = 2;
            R.body_bits = dynb;
        }
    }
    __syncthreads();
    // copy the record out
    const uint32_t* srcw = reinterpret_cast<const uint32_t*>(&R);
    uint32_t* dstw = reinterpret_cast<uint32_t*>(&recs[(uint64_t)c * maxblk + b]);
    for (uint32_t i = tid; i < sizeof(BlockRec) / 4; i += L6B_THREADS) dstw[i] = srcw[i];
}

// Per block (grid-stride over chunk x block slots; slots past a chunk's last
// block are skipped).
__global__ __launch_bounds__(L6B_THREADS, 5) void l6_block_kernel(const uint32_t* __restrict__ toks,
                                                               const uint32_t* __restrict__ ntok_in,
                                                               const uint32_t* __restrict__ blkpos,
                                                               BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                               uint32_t maxblk, BlockRec* __restrict__ recs) {
    // items block-major (block b of every chunk, then b + 1): a chunk has
    // fewer blocks than maxblk, and a chunk-major order with a grid that is a
    // multiple of maxblk would hand every workgroup the same block index
    // Blocks differ in tokens and a chunk's slots past its last block are
    // empty: items come from a device queue unless the grid covers them.
    __shared__ uint32_t s_item;
    const uint32_t nc = hdr->nchunks;
    const uint32_t items = nc * maxblk;
    const bool queued = gridDim.x < items;
    for (uint32_t nx = blockIdx.x;; nx += gridDim.x) {
        uint32_t w = nx;
        if (queued) {
            w = next_item(&hdr->queue[Q_BLOCK], items, &s_item);
            if (w == NO_ITEM) break;
        } else if (w >= items) {
            break;
        }
        l6_block_body(w % nc, w / nc, toks, ntok_in, blkpos, chunk, maxblk, recs);
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t block_bits(const BlockRec& r, uint64_t off) {
    return r.type == 0 ? stored_bits(off, r.raw_len) : r.body_bits;
}

// chunk size for each incoming bit offset mod 8
__global__ void l6_size_kernel(const uint32_t* __restrict__ ntok_in, const BlockRec* __restrict__ recs,
                               const ChunkDesc* __restrict__ desc, const BatchHdr* __restrict__ hdr,
                               uint32_t maxblk, ChunkInfo* __restrict__ info) {
    const uint32_t nchunks = hdr->nchunks;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        const uint32_t ntok = ntok_in[c];
        const uint32_t nblk = ntok ? (ntok + BLK_SYMS - 1) / BLK_SYMS : 1;
        const BlockRec* R = recs + (uint64_t)c * maxblk;
        for (uint32_t r = 0; r < 8; ++r) {
            uint64_t off = r;
            for (uint32_t b = 0; b < nblk; ++b) off += block_bits(R[b], off);
            info[c].sz8[r] = off - r;
        }
        info[c].is_stored = 2;
        info[c].raw_len = desc[c].len;
        info[c].fixed_bits = 0;
    }
}

__global__ __launch_bounds__(256) void l6_zero_kernel(const ChunkInfo* __restrict__ info,
                                                      const BatchHdr* __restrict__ hdr, uint32_t* __restrict__ stage,
                                                      uint64_t stage_words) {
    const uint32_t nchunks = hdr->nchunks;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint64_t bits = info[c].sz8[info[c].bit_off & 7];
        const uint64_t nw = min<uint64_t>(stage_words, (bits + 31) / 32 + 2);
        uint32_t* stg = stage + (uint64_t)c * stage_words;
        for (uint64_t i = threadIdx.x; i < nw; i += 256) stg[i] = 0;
    }
}

// bits [at, at + len) of the chunk staging get v (len <= 32)
__device__ __forceinline__ void put_bits_at(uint32_t* stg, uint64_t at, uint32_t v, uint32_t len) {
    if (!len) return;
    const uint64_t w = at >> 5;
    const uint32_t sh = (uint32_t)(at & 31);
    const uint64_t vv = (uint64_t)(len == 32 ? v : (v & ((1u << len) - 1u))) << sh;
    atomicOr(&stg[w], (uint32_t)vv);
    if (sh + len > 32) atomicOr(&stg[w + 1], (uint32_t)(vv >> 32));
}

__device__ __forceinline__ uint32_t rev_code(uint32_t code, uint32_t len) {
    return len ? __builtin_bitreverse32(code) >> (32 - len) : 0u;
}

constexpr int L6E_THREADS = 256;

__device__ __forceinline__ void l6_emit_body(uint32_t c, uint32_t b, const uint8_t* __restrict__ src,
                                             const ChunkDesc* __restrict__ desc, uint32_t chunk,
                                             const uint32_t* __restrict__ toks, const uint32_t* __restrict__ ntok_in,
                                             const BlockRec* __restrict__ recs, uint32_t maxblk,
                                             const ChunkInfo* __restrict__ info, const Result* __restrict__ res,
                                             uint32_t* __restrict__ stage, uint64_t stage_words) {
    __shared__ uint16_t llc[288];
    __shared__ uint16_t dlc[32];
    __shared__ uint8_t lls[288];
    __shared__ uint8_t dls[32];
    __shared__ uint32_t wsum[L6E_THREADS / 64 + 1];
    __shared__ uint32_t hsum[L6E_THREADS / 64];
    __shared__ uint16_t scnt[6][16];
    __shared__ uint16_t sbase[6][16];
    __shared__ uint8_t scl[19];
    __shared__ uint8_t sclc[19];
    __shared__ uint64_t s_off;
    if (res[desc[c].buf & CHUNK_BUF].total_bytes == 0) return;
    const uint32_t ntok = ntok_in[c];
    const uint32_t nblk = ntok ? (ntok + BLK_SYMS - 1) / BLK_SYMS : 1;
    if (b >= nblk) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const BlockRec* RR = recs + (uint64_t)c * maxblk;
    const BlockRec& R = RR[b];
    const uint64_t cp = l6_stride(chunk);
    const uint32_t* TK = toks + (uint64_t)c * cp;
    const uint8_t* s = src + desc[c].in_off;
    uint32_t* stg = stage + (uint64_t)c * stage_words;
    const bool last = chunk_final(desc[c].buf) && (b == nblk - 1);
    if (tid == 0) {
        const uint64_t r = info[c].bit_off & 7;
        uint64_t off = r;
        for (uint32_t k = 0; k < b; ++k) off += block_bits(RR[k], off);
        s_off = off - r;  // relative to the chunk start
    }
    __syncthreads();
    const uint64_t off = s_off;
    const uint64_t abs_res = info[c].bit_off & 7;
    if (R.type == 0) {
        // stored pieces (deflate_block.cpp:107-132), bytes written 4 at a time
        const uint32_t L = R.raw_len;
        const uint32_t pieces = L == 0 ? 1u : (L + 65534u) / 65535u;
        const uint64_t pad0 = (8 - ((abs_res + off + 3) & 7)) & 7;
        uint64_t at = off + 3 + pad0;  // first byte-aligned bit of piece 0 (after its header)
        for (uint32_t k = 0; k < pieces; ++k) {
            const uint32_t Lk = min(L - k * 65535u, 65535u);
            const bool lastp = last && k == pieces - 1;
            const uint64_t hdr_at = k == 0 ? off : at;
            if (tid == 0) {
                put_bits_at(stg, hdr_at, lastp ? 1u : 0u, 3);
                const uint64_t la = k == 0 ? at : at + 8;
                put_bits_at(stg, la, (Lk & 0xFFFFu) | ((~Lk & 0xFFFFu) << 16), 32);
            }
            const uint64_t da = (k == 0 ? at : at + 8) + 32;
            const uint8_t* rp = s + R.raw_pos + (uint64_t)k * 65535u;
            for (uint32_t i = 4u * (uint32_t)tid; i < Lk; i += 4u * L6E_THREADS) {
                uint32_t v = 0;
                const uint32_t m = min(4u, Lk - i);
                for (uint32_t j = 0; j < m; ++j) v |= (uint32_t)rp[i + j] << (8 * j);
                put_bits_at(stg, da + 8ull * i, v, 8 * m);
            }
            at = da + 8ull * Lk;
        }
        return;
    }
    // fixed or dynamic: code lengths, one symbol per lane.  Thread tid holds
    // literal/length symbol tid (slot a); wave 0 also holds symbols 256..287
    // in lanes 0..31 and distance symbols 0..31 in lanes 32..63 (slot b).
    static_assert(L6E_THREADS == 256, "one literal/length symbol per thread");
    const bool dyn = R.type == 2;
    const uint32_t la = dyn ? (uint32_t)R.ll[tid] : fixed_ll_len((uint32_t)tid);
    uint32_t lb = 0;
    if (wv == 0)
        lb = lane < 32 ? (dyn ? (uint32_t)R.ll[256 + lane] : fixed_ll_len(256u + (uint32_t)lane))
                       : (dyn ? (uint32_t)R.dl[lane - 32] : 5u);
    // canonical codes (RFC 1951 3.2.2): a symbol's code is the first code of
    // its length plus the number of lower symbols of that length.  Segments
    // 0..3 are waves 0..3 of slot a, 4 and 5 the two halves of slot b.
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t ra = 0, rb = 0;
    for (uint32_t L = 1; L <= 15; ++L) {
        const uint64_t ma = ballot(la == L);
        if (la == L) ra = (uint32_t)__popcll(ma & lt);
        if (lane == (int)L) scnt[wv][L] = (uint16_t)__popcll(ma);
        if (wv == 0) {
            const uint64_t mb = ballot(lb == L);
            if (lb == L) rb = (uint32_t)__popcll(mb & lt & (lane < 32 ? 0xFFFFFFFFull : ~0xFFFFFFFFull));
            if (lane == (int)L) {
                scnt[4][L] = (uint16_t)__popcll(mb & 0xFFFFFFFFull);
                scnt[5][L] = (uint16_t)__popcll(mb >> 32);
            }
        }
    }
    if (dyn && wv == 1) {
        // code length code (lengths <= 7), one symbol per lane
        const uint32_t ln = lane < 19 ? (uint32_t)R.cl[lane] : 0u;
        uint32_t code = 0, prev = 0, mine = 0;
        for (uint32_t L = 1; L <= 7; ++L) {
            const uint64_t m = ballot(ln == L);
            code = (code + prev) << 1;
            if (ln == L) mine = code + (uint32_t)__popcll(m & lt);
            prev = (uint32_t)__popcll(m);
        }
        if (lane < 19) {
            scl[lane] = (uint8_t)ln;
            sclc[lane] = (uint8_t)rev_code(mine, ln);
        }
    }
    __syncthreads();
    if (tid < 2) {
        // first code of each length per segment: thread 0 the literal/length
        // table (segments 0..4 in symbol order), thread 1 the distance table
        const int s0 = tid ? 5 : 0, s1 = tid ? 6 : 5;
        uint32_t code = 0, prev = 0;
        for (int L = 1; L <= 15; ++L) {
            code = (code + prev) << 1;
            uint32_t run = code;
            for (int sg = s0; sg < s1; ++sg) {
                sbase[sg][L] = (uint16_t)run;
                run += scnt[sg][L];
            }
            prev = run - code;
        }
    }
    // code-length RLE items (at most 316), two per thread, placed by a scan
    uint32_t hv0 = 0, hn0 = 0, hv1 = 0, hn1 = 0, hincl = 0;
    if (dyn) {
        const uint32_t nr = R.nrle;
        auto item = [&](uint32_t i, uint32_t& v, uint32_t& n) {
            if (i >= nr) return;
            const uint32_t sy = R.rle_sym[i];
            const uint32_t xb = sy == 16 ? 2u : (sy == 17 ? 3u : (sy == 18 ? 7u : 0u));
            n = scl[sy] + xb;
            v = sclc[sy] | (xb ? (uint32_t)R.rle_x[i] << scl[sy] : 0u);
        };
        item(2u * (uint32_t)tid, hv0, hn0);
        item(2u * (uint32_t)tid + 1u, hv1, hn1);
        hincl = wave_incl_scan(hn0 + hn1, lane);
        if (lane == 63) hsum[wv] = hincl;
    }
    __syncthreads();
    llc[tid] = la ? (uint16_t)rev_code(sbase[wv][la] + ra, la) : (uint16_t)0;
    lls[tid] = (uint8_t)la;
    if (wv == 0) {
        const uint16_t cb = lb ? (uint16_t)rev_code(sbase[lane < 32 ? 4 : 5][lb] + rb, lb) : (uint16_t)0;
        if (lane < 32) {
            llc[256 + lane] = cb;
            lls[256 + lane] = (uint8_t)lb;
        } else {
            dlc[lane - 32] = cb;
            dls[lane - 32] = (uint8_t)lb;
        }
    }
    if (dyn) {
        // 3-bit block header, HLIT 5, HDIST 5, HCLEN 4, HCLEN x 3, RLE items
        uint32_t base = 0;
        for (int w = 0; w < wv; ++w) base += hsum[w];
        const uint64_t at = off + 17 + 3ull * R.hclen + base + hincl - (hn0 + hn1);
        put_bits_at(stg, at, hv0, hn0);
        put_bits_at(stg, at + hn0, hv1, hn1);
        if ((uint32_t)tid < R.hclen) put_bits_at(stg, off + 17 + 3ull * tid, scl[c_cl_order[tid]], 3);
    }
    if (tid == 0) {
        const uint32_t h = (last ? 1u : 0u) | (R.type << 1);
        if (dyn)
            put_bits_at(stg, off, h | ((R.hlit - 257) << 3) | ((R.hdist - 1) << 8) | ((R.hclen - 4) << 13), 17);
        else
            put_bits_at(stg, off, h, 3);
    }
    __syncthreads();
    const uint64_t sym0 = off + 3 + (dyn ? R.hdr_bits : 0u);
    // symbols: each thread a contiguous slice of at most 64 tokens (a block
    // holds at most BLK_SYMS), a multiple of 4 long so it starts 16-byte
    // aligned, held in registers for the count and the emission
    constexpr int QMAX = BLK_SYMS / L6E_THREADS / 4;
    const uint32_t nt = R.t1 - R.t0;
    const uint32_t per = ((nt + L6E_THREADS - 1) / L6E_THREADS + 3u) & ~3u;
    const uint32_t a = R.t0 + min(nt, (uint32_t)tid * per), e = R.t0 + min(nt, (uint32_t)tid * per + per);
    uint4 q[QMAX];
    const uint4* T4 = reinterpret_cast<const uint4*>(TK + a);
#pragma unroll
    for (int k = 0; k < QMAX; ++k) q[k] = a + 4u * k < e ? T4[k] : make_uint4(0u, 0u, 0u, 0u);
    auto tok_at = [&](int k, int j) -> uint32_t {
        return j == 0 ? q[k].x : (j == 1 ? q[k].y : (j == 2 ? q[k].z : q[k].w));
    };
    // a match takes up to 15 + 5 + 15 + 13 = 48 bits with dynamic codes
    auto tok_code = [&](uint32_t v, uint64_t& code, uint32_t& len) {
        if (v & 0x80000000u) {
            const uint32_t L = (v >> 16) & 0x1FFu, d = v & 0xFFFFu;
            const LenSym ls = len_sym(L), ds = dist_sym(d);
            const uint32_t sy = 257 + ls.code;
            const uint32_t l1 = lls[sy], l2 = dls[ds.code];
            code = (uint64_t)llc[sy] | ((uint64_t)ls.extra << l1) | ((uint64_t)dlc[ds.code] << (l1 + ls.ebits)) |
                   ((uint64_t)ds.extra << (l1 + ls.ebits + l2));
            len = l1 + ls.ebits + l2 + ds.ebits;
        } else {
            code = llc[v & 0xFFu];
            len = lls[v & 0xFFu];
        }
    };
    auto tok_len = [&](uint32_t v) -> uint32_t {
        if (v & 0x80000000u) {
            const LenSym ls = len_sym((v >> 16) & 0x1FFu), ds = dist_sym(v & 0xFFFFu);
            return lls[257 + ls.code] + ls.ebits + dls[ds.code] + ds.ebits;
        }
        return lls[v & 0xFFu];
    };
    uint32_t mybits = 0;
#pragma unroll
    for (int k = 0; k < QMAX; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (a + 4u * k + j < e) mybits += tok_len(tok_at(k, j));
    const uint32_t incl = wave_incl_scan(mybits, lane);
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < L6E_THREADS / 64; ++i) {
            const uint32_t t = wsum[i];
            wsum[i] = acc;
            acc += t;
        }
        wsum[L6E_THREADS / 64] = acc;
    }
    __syncthreads();
    // Pack the slice in a 64-bit register; words wholly inside it are plain
    // stores, its first and last words (shared with a neighbour, the header
    // or the EOB) are atomicOr.
    // the slice's token loads are complete here (the count used them); this
    // says so to the compiler, which would otherwise wait for them at every
    // token of the emission, behind the stores of the tokens before it
#pragma unroll
    for (int k = 0; k < QMAX; ++k) asm volatile("" ::"v"(q[k].x), "v"(q[k].y), "v"(q[k].z), "v"(q[k].w));
    if (mybits) {
        const uint64_t at = sym0 + wsum[wv] + incl - mybits;
        uint64_t wi = at >> 5;
        uint32_t nacc = (uint32_t)(at & 31);
        bool partial = nacc != 0;
        uint64_t acc = 0;
        auto flush_word = [&]() {
            if (partial) {
                atomicOr(&stg[wi], (uint32_t)acc);
                partial = false;
            } else {
                stg[wi] = (uint32_t)acc;
            }
            ++wi;
            acc >>= 32;
            nacc -= 32;
        };
#pragma unroll
        for (int k = 0; k < QMAX; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (a + 4u * k + j < e) {
                    uint64_t cd;
                    uint32_t ln;
                    tok_code(tok_at(k, j), cd, ln);
                    // at most 48 bits: add in two halves so the register never overflows
                    const uint32_t l0 = min(ln, 32u);
                    acc |= (cd & ((1ull << l0) - 1ull)) << nacc;
                    nacc += l0;
                    if (nacc >= 32) flush_word();
                    if (ln > 32) {
                        acc |= (cd >> 32) << nacc;
                        nacc += ln - 32;
                        if (nacc >= 32) flush_word();
                    }
                }
            }
        if (nacc) atomicOr(&stg[wi], (uint32_t)acc);
    }
    if (tid == 0) put_bits_at(stg, sym0 + wsum[L6E_THREADS / 64], llc[256], lls[256]);  // EOB
}

__global__ __launch_bounds__(L6E_THREADS) void l6_emit_kernel(const uint8_t* __restrict__ src,
                                                              const ChunkDesc* __restrict__ desc,
                                                              BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                              const uint32_t* __restrict__ toks,
                                                              const uint32_t* __restrict__ ntok_in,
                                                              const BlockRec* __restrict__ recs, uint32_t maxblk,
                                                              const ChunkInfo* __restrict__ info,
                                                              const Result* __restrict__ res,
                                                              uint32_t* __restrict__ stage, uint64_t stage_words) {
    // items block-major, from a device queue unless the grid covers them, as
    // in l6_block_kernel
    __shared__ uint32_t s_item;
    const uint32_t nc = hdr->nchunks;
    const uint32_t items = nc * maxblk;
    const bool queued = gridDim.x < items;
    for (uint32_t nx = blockIdx.x;; nx += gridDim.x) {
        uint32_t w = nx;
        if (queued) {
            w = next_item(&hdr->queue[Q_EMIT], items, &s_item);
            if (w == NO_ITEM) break;
        } else if (w >= items) {
            break;
        }
        l6_emit_body(w % nc, w / nc, src, desc, chunk, toks, ntok_in, recs, maxblk, info, res, stage,
                     stage_words);
        __syncthreads();
    }
}

inline uint32_t l6_per_cu(const Plan& p, size_t smem, int threads) {
    const uint32_t by_lds = smem ? (uint32_t)(p.lds_max / smem) : 32u;
    const uint32_t by_waves = (uint32_t)(32 / (threads / 64));
    const uint32_t k = by_lds < by_waves ? by_lds : by_waves;
    return k ? k : 1u;
}

// Workgroups per chunk for the pred and match kernels: a batch of few chunks
// splits each chunk so that the grid fills `cap` workgroups; a captured
// graph (counts only in device memory) keeps one.
inline uint32_t l6_segments(const Plan& p, uint32_t cap) {
    if (p.persistent_grid || p.nchunks == 0) return 1;
    const uint32_t segs = (cap + p.nchunks - 1) / p.nchunks;
    return segs < L6M_MAX_SEGS ? segs : L6M_MAX_SEGS;
}

template <class K>
hipError_t l6_allow_lds(K* kernel, size_t smem) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)smem);
}

template <bool EXCH>
hipError_t launch_pred(const Plan& p) {
    const uint32_t cap = p.cus * l6_per_cu(p, L6P_LDS, L6P_THREADS);
    const uint32_t segs = l6_segments(p, cap);
    const unsigned g = grid_for(p.persistent_grid, p.nchunks * segs, cap);
    hipLaunchKernelGGL((l6_pred_kernel<EXCH>), dim3(g), dim3(L6P_THREADS), L6P_LDS, p.stream, p.d_in, p.d_desc,
                       p.d_bufs, p.d_order, p.d_hdr, p.chunk, p.d_tok, p.d_pd, segs);
    return hipSuccess;
}

}  // namespace

// LDS limits of the L6-class kernels; set once per process before any launch.
hipError_t set_kernel_attributes_l6(size_t lds_max) {
    (void)lds_max;
    hipError_t e = l6_allow_lds(&l6_pred_kernel<true>, L6P_LDS);
    if (e == hipSuccess) e = l6_allow_lds(&l6_pred_kernel<false>, L6P_LDS);
    if (e == hipSuccess) e = l6_allow_lds(&l6_parse_kernel, L6R_LDS);
    if (e == hipSuccess) e = l6_allow_lds(&l6_pseg_parse_kernel, L6R_LDS);
    return e;
}

L6Params l6_level_params(int level) {
    // {max_chain, nice_len, lazy_depth} per level; 0..3 are not L6-class
    static const L6Params t[13] = {
        {0, 0, 0},    {0, 0, 0},     {0, 0, 0},     {0, 0, 0},     {16, 128, 1},   {32, 128, 1},   {32, 258, 1},
        {64, 258, 2}, {128, 258, 2}, {256, 258, 3}, {512, 258, 3}, {1024, 258, 3}, {2048, 258, 3},
    };
    return t[level < 0 ? 0 : (level > 12 ? 12 : level)];
}
uint32_t l6_max_blocks(uint32_t chunk) { return (chunk + BLK_SYMS - 1) / BLK_SYMS; }
uint64_t l6_position_stride(uint32_t chunk) { return l6_stride(chunk); }
uint32_t l6_max_chunk() { return L6_MAX_CHUNK; }
size_t l6_block_record_bytes() { return sizeof(BlockRec); }

// LZ77 and block analysis of the L6-class path (before the offset scan).
hipError_t launch_l6_front(const Plan& p) {
    const uint32_t maxblk = l6_max_blocks(p.chunk);
    hipError_t e = p.exch_ok ? launch_pred<true>(p) : launch_pred<false>(p);
    if (e != hipSuccess) return e;
    const uint32_t mcap = p.cus * l6_per_cu(p, L6M_LDS, L6M_THREADS);
    const uint32_t segs = l6_segments(p, mcap);
    const unsigned gm = grid_for(p.persistent_grid, p.nchunks * segs, mcap);
    hipLaunchKernelGGL(l6_match_kernel, dim3(gm), dim3(L6M_THREADS), 0, p.stream, p.d_in, p.d_desc, p.d_bufs,
                       p.d_order, p.d_hdr, p.chunk, p.d_pd, p.d_tok, p.l6, segs);
    const uint32_t dcap = p.cus * 2u;  // two 1024-thread workgroups per CU
    const uint32_t dsegs = l6_segments(p, dcap);
    const unsigned gd = grid_for(p.persistent_grid, p.nchunks * dsegs, dcap);
    // the parse: one workgroup per chunk, or, for batches of few chunks,
    // segments of each chunk with the token bitmap in the staging words
    const uint32_t rcap = p.cus * l6_per_cu(p, l6r_lds(p.chunk), L6R_THREADS);
    const uint32_t psegs = l6_segments(p, rcap);
    uint32_t* gbm = psegs > 1 ? p.d_stage : nullptr;
    hipLaunchKernelGGL(l6_decide_kernel, dim3(gd), dim3(L6D_THREADS), 0, p.stream, p.d_desc, p.d_hdr, p.chunk, p.d_tok,
                       p.d_pd, p.l6, dsegs, gbm, p.stage_words);
    if (psegs > 1) {
        const unsigned gp = grid_for(false, p.nchunks * psegs, rcap);
        hipLaunchKernelGGL(l6_pseg_parse_kernel, dim3(gp), dim3(L6R_THREADS), l6r_lds(p.chunk), p.stream, p.d_desc,
                           p.d_hdr, p.chunk, p.d_pd, gbm, p.stage_words, psegs);
        hipLaunchKernelGGL(l6_pseg_count_kernel, dim3(gp), dim3(L6R_THREADS), 0, p.stream, p.d_desc, p.d_hdr, p.chunk,
                           p.d_pd, gbm, p.stage_words, psegs);
        hipLaunchKernelGGL(l6_pseg_write_kernel, dim3(gp), dim3(L6R_THREADS), 0, p.stream, p.d_in, p.d_desc, p.d_hdr,
                           p.chunk, p.d_tok, p.d_pd, gbm, p.stage_words, psegs, p.d_aux, p.d_ntok, p.d_blkpos,
                           maxblk);
    } else {
        const unsigned gr = grid_for(p.persistent_grid, p.nchunks, rcap);
        hipLaunchKernelGGL(l6_parse_kernel, dim3(gr), dim3(L6R_THREADS), l6r_lds(p.chunk), p.stream, p.d_in,
                           p.d_desc, p.d_order, p.d_hdr, p.chunk, p.d_tok, p.d_pd, p.d_aux, p.d_ntok, p.d_blkpos,
                           maxblk, p.l6);
    }
    const unsigned gb = grid_for(p.persistent_grid, p.nchunks * maxblk, p.cus * 8u);
    hipLaunchKernelGGL(l6_block_kernel, dim3(gb), dim3(L6B_THREADS), 0, p.stream, p.d_aux, p.d_ntok, p.d_blkpos,
                       p.d_hdr, p.chunk, maxblk, reinterpret_cast<BlockRec*>(p.d_blocks));
    const unsigned gs = grid_for(p.persistent_grid, (p.nchunks + 255) / 256, p.cus);
    hipLaunchKernelGGL(l6_size_kernel, dim3(gs), dim3(256), 0, p.stream, p.d_ntok,
                       reinterpret_cast<const BlockRec*>(p.d_blocks), p.d_desc, p.d_hdr, maxblk, p.d_info);
    return hipGetLastError();
}

// Emission of the L6-class path (after the offset scan).
hipError_t launch_l6_back(const Plan& p) {
    const uint32_t maxblk = l6_max_blocks(p.chunk);
    const unsigned gz = grid_for(p.persistent_grid, p.nchunks, p.cus * 8u);
    hipLaunchKernelGGL(l6_zero_kernel, dim3(gz), dim3(256), 0, p.stream, p.d_info, p.d_hdr, p.d_stage, p.stage_words);
    const unsigned ge = grid_for(p.persistent_grid, p.nchunks * maxblk, p.cus * 8u);
    hipLaunchKernelGGL(l6_emit_kernel, dim3(ge), dim3(L6E_THREADS), 0, p.stream, p.d_in, p.d_desc, p.d_hdr, p.chunk,
                       p.d_aux, p.d_ntok, reinterpret_cast<const BlockRec*>(p.d_blocks), maxblk, p.d_info,
                       p.d_result, p.d_stage, p.stage_words);
    return hipGetLastError();
}

}  // namespace orot_gpu
