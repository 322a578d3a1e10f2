// deflate_big.hip -- levels 1-3 with chunks larger than the single-workgroup
// kernels take (above 128 KiB, or 256 KiB at level 1), up to 1 GiB: orot's
// fast path (src/deflate/lz77.cpp:163-191, 378-406) over one chunk of any
// size, with the same bytes as orot's parallel compressor for that block
// size, and orot's whole-buffer deflate_compress (one chunk per input, an
// all-literal input stored: src/compress/block_compressor.cpp:16-73).
//
// Four launches per batch:
//   big_cand_kernel   one wave per 64 Ki-position segment of a chunk: the
//                     latest earlier position with the same hash inside the
//                     segment (ballot match-any in steps of 64 positions,
//                     one LDS table of positions), and the segment's table.
//   big_carry_kernel  one thread per (chunk, hash): the tables turned into
//                     each segment's carry-in, the latest position of the
//                     hash in the earlier segments (exclusive scan).
//   big_match_kernel  one thread per 64 positions: candidate C (the
//                     segment's own, else the carry-in), orot's rules (C = 0
//                     is empty, the distance masked to 15 bits, the first
//                     byte compared at C) and the greedy match length, the
//                     verified extent carried along a diagonal; the token
//                     word lit | length << 8 | distance << 17 per position.
//   big_parse_kernel  one 1024-thread workgroup per chunk: speculative
//                     greedy parses of 1024 segments, a serial resolve of
//                     their entries on wave 0 (walking from the true entry
//                     until the parses meet, reading only), a parallel
//                     fix-up of the resolved segments, fixed-Huffman bit
//                     counts, a
//                     64-bit block scan, orot's stored / fixed choice
//                     (deflate_block.cpp:414-421) and emission into the
//                     chunk's staging bitstream.  The token-start bitmap
//                     lives in HBM.
#include <algorithm>

#include "deflate_gpu.h"
#include "device_common.h"

namespace orot_gpu {
namespace {

constexpr int BIG_THREADS = 1024;

// Dictionary mode (ext = BIG_DICT): a chunk that is not the first of its
// buffer sees the up to 32 KiB before it (a stream piece's first chunk: the
// stream's bytes before the piece).  Positions then count from the
// dictionary's first byte and per-position arrays are laid out with room for
// it; tokens cover the chunk only.
__device__ __forceinline__ uint32_t big_halo(uint32_t ext, uint32_t c, const ChunkDesc* desc, const BufDesc* bufs) {
    if (!ext) return 0;
    const BufDesc& B = bufs[desc[c].buf & CHUNK_BUF];
    const uint64_t rel = desc[c].in_off - B.in_off + B.hist;
    return (uint32_t)min<uint64_t>(rel, ext);
}
constexpr int BIG_WAVES = BIG_THREADS / 64;
constexpr uint32_t BIG_RUN = 64;  // positions per thread of the match kernel


template <int HB>
__global__ __launch_bounds__(64) void big_cand_kernel(const uint8_t* __restrict__ src, const ChunkDesc* __restrict__ desc,
                                                      const BufDesc* __restrict__ bufs, const BatchHdr* __restrict__ hdr,
                                                      uint32_t chunk, uint32_t ext, uint32_t* __restrict__ cand,
                                                      uint32_t* __restrict__ segtab) {
    constexpr uint32_t TAB = 1u << HB;
    extern __shared__ uint32_t tab[];
    const int lane = threadIdx.x;
    const uint32_t segs = big_segs(chunk + ext);
    const uint64_t items = (uint64_t)hdr->nchunks * segs;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t c = (uint32_t)(it / segs), sg = (uint32_t)(it % segs);
        const uint32_t H = big_halo(ext, c, desc, bufs);
        const uint32_t n = H + desc[c].len;       // positions of the dictionary and the chunk
        const uint32_t nh = n >= 4 ? n - 3 : 0;  // positions p with p + 4 <= n are hashed
        const uint32_t a = sg * BIG_SEG;
        if (a >= nh) continue;  // the carry kernel reads only segments below nh
        const uint32_t b = min(a + BIG_SEG, nh);
        const uint8_t* s = src + desc[c].in_off - H;
        uint32_t* Q = cand + (uint64_t)c * big_position_stride(chunk + ext);
        for (uint32_t i = lane; i < TAB; i += 64) tab[i] = 0;
        __syncthreads();
        uint32_t wnext = a + (uint32_t)lane < b ? ld32u(s + a + lane) : 0u;
        for (uint32_t p0 = a; p0 < b; p0 += 64) {
            const uint32_t p = p0 + (uint32_t)lane;
            const bool valid = p < b;
            const uint32_t wcur = wnext;
            wnext = p + 64 < b ? ld32u(s + p + 64) : 0u;
            const uint32_t h = valid ? hash4<HB>(wcur) : 0u;
            // lanes holding the same hash: one HB-bit match-any by ballots
            uint64_t m = ballot(valid);
#pragma unroll
            for (int k = 0; k < HB; ++k) {
                const uint32_t bit = (h >> k) & 1u;
                const uint64_t bk = ballot(bit != 0);
                m &= bk ^ ((uint64_t)bit - 1ull);
            }
            if (valid) {
                const uint64_t below = m & ((1ull << lane) - 1ull);
                uint32_t res;
                if (below) res = p0 + (uint32_t)(63 - __builtin_clzll(below)) + 1u;
                else res = tab[h];
                if ((m >> lane) == 1ull) tab[h] = p + 1u;  // the group's last lane
                Q[p] = res;
            }
        }
        __syncthreads();
        uint32_t* out = segtab + ((uint64_t)c * segs + sg) * TAB;
        for (uint32_t i = lane; i < TAB; i += 64) out[i] = tab[i];
        __syncthreads();
    }
}

template <int HB>
__global__ __launch_bounds__(256) void big_carry_kernel(const ChunkDesc* __restrict__ desc,
                                                        const BufDesc* __restrict__ bufs,
                                                        const BatchHdr* __restrict__ hdr, uint32_t chunk, uint32_t ext,
                                                        uint32_t* __restrict__ segtab) {
    constexpr uint32_t TAB = 1u << HB;
    const uint32_t segs = big_segs(chunk + ext);
    const uint64_t items = (uint64_t)hdr->nchunks * TAB;
    for (uint64_t it = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; it < items;
         it += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = (uint32_t)(it / TAB), h = (uint32_t)(it % TAB);
        const uint32_t n = big_halo(ext, c, desc, bufs) + desc[c].len;
        const uint32_t nh = n >= 4 ? n - 3 : 0;
        const uint32_t used = (nh + BIG_SEG - 1) / BIG_SEG;
        uint32_t* col = segtab + (uint64_t)c * segs * TAB + h;
        uint32_t run = 0;
        for (uint32_t sg = 0; sg < used; ++sg) {
            const uint32_t v = col[(uint64_t)sg * TAB];
            col[(uint64_t)sg * TAB] = run;
            if (v) run = v;
        }
    }
}

template <int HB>
__global__ __launch_bounds__(256) void big_match_kernel(const uint8_t* __restrict__ src, const ChunkDesc* __restrict__ desc,
                                                        const BufDesc* __restrict__ bufs,
                                                        const BatchHdr* __restrict__ hdr, uint32_t chunk, uint32_t ext,
                                                        const uint32_t* __restrict__ cand,
                                                        const uint32_t* __restrict__ carry,
                                                        uint32_t* __restrict__ tokw) {
    constexpr uint32_t TAB = 1u << HB;
    const uint32_t segs = big_segs(chunk + ext);
    const uint32_t runs = (chunk + BIG_RUN - 1) / BIG_RUN;
    const uint64_t items = (uint64_t)hdr->nchunks * runs;
    for (uint64_t it = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; it < items;
         it += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = (uint32_t)(it / runs), r = (uint32_t)(it % runs);
        const uint32_t H = big_halo(ext, c, desc, bufs);
        const uint32_t n = H + desc[c].len;  // positions count from the dictionary's first byte
        const uint32_t r0 = H + r * BIG_RUN;  // the chunk's positions only
        if (r0 >= n) continue;
        const uint32_t r1 = min(n, r0 + BIG_RUN);
        const uint32_t nh = n >= 4 ? n - 3 : 0;
        const uint32_t lim = min(n, r1 + 257u);  // compares reach past the run for its last tokens
        const uint8_t* s = src + desc[c].in_off - H;
        const uint64_t base = (uint64_t)c * big_position_stride(chunk + ext);
        const uint32_t* Q = cand + base;
        const uint32_t* carry_c = carry + (uint64_t)c * segs * TAB;
        uint32_t* T = tokw + base;
        uint32_t cd = 0, ce = 0;  // last diagonal compared and the end of its verified extent
        bool cx = false;          // ce is a mismatch (or the chunk end), not a limit of the compare
        for (uint32_t p = r0; p < r1; ++p) {
            const uint32_t lit = s[p];
            uint32_t m1 = 0, d = 0;
            if (p < nh) {
                uint32_t C1 = Q[p];
                if (!C1) C1 = carry_c[(uint64_t)(p / BIG_SEG) * TAB + hash4<HB>(ld32u(s + p))];
                // head value 0 (position 0) is "empty"; the distance is masked
                // to 15 bits and the first byte is compared at C itself
                if (C1 > 1u) {
                    const uint32_t C = C1 - 1u;
                    const uint32_t D = (p - C) & 0x7FFFu;
                    if (D && s[C] == lit) {
                        const uint32_t cap = min(258u, n - p);
                        uint32_t L;
                        if (D == cd && p <= ce && (cx || ce >= p + cap)) {
                            L = ce - p;
                        } else {
                            uint32_t e = p;
                            bool x = false;
                            while (e < lim) {
                                if (e + 8 <= n) {
                                    const uint64_t df = ld64u(s + e) ^ ld64u(s + e - D);
                                    if (df) {
                                        e += (uint32_t)(__builtin_ctzll(df) >> 3);
                                        x = true;
                                        break;
                                    }
                                    e += 8;
                                } else {
                                    while (e < n && s[e] == s[e - D]) ++e;
                                    x = true;
                                    break;
                                }
                            }
                            if (e >= n) {
                                e = n;
                                x = true;
                            }
                            cd = D;
                            ce = e;
                            cx = x;
                            L = e - p;
                        }
                        L = min(L, cap);
                        if (L >= 3) {
                            m1 = L;
                            d = D;
                        }
                    }
                }
            }
            T[p] = lit | (m1 << 8) | (d << 17);
        }
    }
}

__device__ __forceinline__ bool bm_test(const uint32_t* bm, uint32_t p) { return (bm[p >> 5] >> (p & 31)) & 1u; }

__global__ __launch_bounds__(BIG_THREADS) void big_parse_kernel(const ChunkDesc* __restrict__ desc,
                                                                const BufDesc* __restrict__ bufs,
                                                                const BatchHdr* __restrict__ hdr, uint32_t chunk,
                                                                uint32_t ext, const uint32_t* __restrict__ tokw,
                                                                uint32_t* __restrict__ bitmap,
                                                                uint32_t* __restrict__ stage, uint64_t stage_words,
                                                                ChunkInfo* __restrict__ info, int whole) {
    __shared__ uint32_t exit_spec[BIG_THREADS];
    __shared__ uint32_t segbits[BIG_THREADS];
    __shared__ uint32_t ntok_seg[BIG_THREADS];
    __shared__ uint8_t touched[BIG_THREADS];
    __shared__ uint32_t entry[BIG_THREADS];    // resolved segments: the true parse's entry
    __shared__ uint32_t meet_at[BIG_THREADS];  // ... and where it meets the speculative one
    __shared__ uint64_t scan[BIG_WAVES + 1];
    __shared__ uint32_t tok_scan[BIG_WAVES + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t nchunks = hdr->nchunks;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint32_t n = desc[c].len;
        // token words of the chunk's positions (after its dictionary's)
        const uint32_t* T = tokw + (uint64_t)c * big_position_stride(chunk + ext) + big_halo(ext, c, desc, bufs);
        uint32_t* bm = bitmap + (uint64_t)c * big_bitmap_words(chunk);
        // segments of S positions (a multiple of 32: each owns whole bitmap words)
        const uint32_t S = max(32u, (((n + BIG_THREADS - 1) / BIG_THREADS) + 31u) & ~31u);
        const uint32_t NS = (n + S - 1) / S;
        const uint32_t nbw = (n + 31) >> 5;
        for (uint32_t i = tid; i < nbw; i += BIG_THREADS) bm[i] = 0;
        touched[tid] = 0;
        __syncthreads();
        // ---- speculative greedy parse of each segment from its start (lz77.cpp:378-406)
        if ((uint32_t)tid < NS) {
            uint32_t pos = (uint32_t)tid * S;
            const uint32_t end = min(pos + S, n);
            uint32_t bits = 0, cnt = 0;
            uint32_t wacc = 0, wi = pos >> 5;
            while (pos < end) {
                if ((pos >> 5) != wi) {
                    bm[wi] = wacc;
                    wi = pos >> 5;
                    wacc = 0;
                }
                wacc |= 1u << (pos & 31);
                const uint32_t t = T[pos];
                bits += tok_bits(t);
                ++cnt;
                const uint32_t l = tok_len(t);
                pos += l ? l : 1u;
            }
            if ((uint32_t)tid * S < end) bm[wi] = wacc;
            exit_spec[tid] = pos;
            segbits[tid] = bits;
            ntok_seg[tid] = cnt;
        }
        __syncthreads();
        // ---- serial resolve on wave 0: a segment entered past its start is
        // walked from its true entry until it meets a position of its
        // speculative parse (from there on both parses agree).  The walk only
        // reads the bitmap; it records the entry and the meeting point, and
        // the fix-up below rewrites the segment's own bitmap words.
        if (wv == 0 && NS > 1) {
            uint32_t e = exit_spec[0];
            uint32_t si = 1;
            while (si < NS) {
                if (e == si * S) {
                    // skip every following segment entered at its start
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
                const uint32_t en = min(si * S + S, n);
                uint32_t pos = e, meet = en, ne = e;
                if (e < en) {
                    for (;;) {
                        if (bm_test(bm, pos)) {
                            meet = pos;
                            ne = exit_spec[si];
                            break;
                        }
                        const uint32_t l = tok_len(T[pos]);
                        const uint32_t np = pos + (l ? l : 1u);
                        if (np >= en) {
                            ne = np;
                            break;
                        }
                        pos = np;
                    }
                }
                if (lane == 0) {
                    touched[si] = 1;
                    entry[si] = e;
                    meet_at[si] = meet;
                }
                e = ne;
                ++si;
            }
        }
        __syncthreads();
        // ---- fix-up of the resolved segments, each by its own thread: the
        // speculative starts before the meeting point are replaced by the
        // true parse's
        if ((uint32_t)tid < NS && touched[tid]) {
            const uint32_t st = (uint32_t)tid * S, meet = meet_at[tid];
            for (uint32_t w = st >> 5; w < ((meet + 31) >> 5); ++w) {
                const uint32_t lo = w == (st >> 5) ? (st & 31u) : 0u;
                const uint32_t hi = (w + 1) * 32 <= meet ? 32u : meet - w * 32;
                const uint32_t mask = (hi == 32u ? 0xFFFFFFFFu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
                bm[w] &= ~mask;
            }
            for (uint32_t pos = entry[tid]; pos < meet;) {
                bm[pos >> 5] |= 1u << (pos & 31);
                const uint32_t l = tok_len(T[pos]);
                pos += l ? l : 1u;
            }
        }
        __syncthreads();
        // ---- bit counts (re-counted where the resolve changed the parse),
        // 64-bit block scan, stored / fixed choice, emission
        const uint32_t st = (uint32_t)tid * S;
        const uint32_t en = (uint32_t)tid < NS ? min(st + S, n) : st;
        uint32_t mybits = (uint32_t)tid < NS ? segbits[tid] : 0u;
        uint32_t mytok = (uint32_t)tid < NS ? ntok_seg[tid] : 0u;
        if ((uint32_t)tid < NS && touched[tid]) {
            mybits = 0;
            mytok = 0;
            for (uint32_t w = st >> 5; w < ((en + 31) >> 5); ++w) {
                uint32_t x = bm[w];
                while (x) {
                    const uint32_t p = (w << 5) + (uint32_t)__builtin_ctz(x);
                    x &= x - 1;
                    mybits += tok_bits(T[p]);
                    ++mytok;
                }
            }
        }
        // inclusive wave scans of the 64-bit bit counts and of the token counts
        uint64_t incl = mybits;
        uint32_t tincl = mytok;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t o = (uint64_t)__shfl_up((long long)incl, d, 64);
            const uint32_t to = (uint32_t)__shfl_up((int)tincl, d, 64);
            if (lane >= d) {
                incl += o;
                tincl += to;
            }
        }
        if (lane == 63) {
            scan[wv] = incl;
            tok_scan[wv] = tincl;
        }
        __syncthreads();
        if (tid == 0) {
            uint64_t acc = 0;
            uint32_t tacc = 0;
            for (int i = 0; i < BIG_WAVES; ++i) {
                const uint64_t t = scan[i];
                scan[i] = acc;
                acc += t;
                tacc += tok_scan[i];
            }
            scan[BIG_WAVES] = acc;
            tok_scan[BIG_WAVES] = tacc;
        }
        __syncthreads();
        const uint64_t excl = scan[wv] + incl - mybits;
        const uint64_t fixed = 3ull + scan[BIG_WAVES] + 7ull;
        const bool is_last = chunk_final(desc[c].buf);
        // whole-buffer calls store an input with no match at all
        // (block_compressor.cpp:60-61); otherwise orot's cost rule
        const bool stored = (whole && tok_scan[BIG_WAVES] == n) || (uint64_t)(n + 5) * 8ull <= fixed;
        if (tid == 0) {
            info[c].fixed_bits = fixed;
            info[c].raw_len = n;
            info[c].is_stored = stored ? 1u : 0u;
        }
        if (!stored) {
            uint32_t* stg = stage + (uint64_t)c * stage_words;
            const uint64_t nw = ((fixed + 31) >> 5) + 1u;
            for (uint64_t i = tid; i < nw; i += BIG_THREADS) stg[i] = 0;
            __syncthreads();
            if (tid == 0) atomicOr(&stg[0], (is_last ? 1u : 0u) | 2u);  // BFINAL, BTYPE = 01
            if (mybits) {
                const uint64_t bitpos = 3ull + excl;
                uint64_t wi = bitpos >> 5;
                uint32_t nacc = (uint32_t)(bitpos & 31u);
                bool partial = nacc != 0;
                uint64_t acc = 0;
                for (uint32_t w = st >> 5; w < ((en + 31) >> 5); ++w) {
                    uint32_t x = bm[w];
                    while (x) {
                        const uint32_t p = (w << 5) + (uint32_t)__builtin_ctz(x);
                        x &= x - 1;
                        const uint32_t t = T[p];
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
                    }
                }
                if (nacc) atomicOr(&stg[wi], (uint32_t)acc);
            }
            __syncthreads();
            // end of block (code 256: seven zero bits) needs no bits set
        }
        __syncthreads();
    }
}

}  // namespace

bool l1_big_chunk(uint32_t chunk, int hash_bits) { return chunk > max_chunk_for_hash_bits(hash_bits); }

hipError_t set_kernel_attributes_big() {
    // the 14-bit segment table takes 64 KiB of LDS
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&big_cand_kernel<14>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)(4u << 14));
}

hipError_t launch_l1_big(const Plan& p, int whole) {
    const uint32_t ext = p.l1_dict ? BIG_DICT : 0u;
    const uint32_t segs = big_segs(p.chunk + ext);
    uint32_t* cand = p.d_aux;
    uint32_t* segtab = reinterpret_cast<uint32_t*>(p.d_blocks);
    uint32_t* bitmap = reinterpret_cast<uint32_t*>(p.d_pd);
    const uint32_t wide = (uint32_t)p.cus * 32u;
    const unsigned gc = grid_for(p.persistent_grid, (uint32_t)std::min<uint64_t>((uint64_t)p.nchunks * segs, wide), wide);
    const unsigned tab = 1u << p.hash_bits;
    const unsigned gk = grid_for(p.persistent_grid, (uint32_t)std::min<uint64_t>(((uint64_t)p.nchunks * tab + 255) / 256, wide), wide);
    const uint64_t runs = ((uint64_t)p.chunk + BIG_RUN - 1) / BIG_RUN;
    const unsigned gm = grid_for(p.persistent_grid, (uint32_t)std::min<uint64_t>((p.nchunks * runs + 255) / 256, wide), wide);
    const unsigned gp = grid_for(p.persistent_grid, p.nchunks, (uint32_t)p.cus);
    if (p.hash_bits == 12) {
        hipLaunchKernelGGL(big_cand_kernel<12>, dim3(gc), dim3(64), 4u << 12, p.stream, p.d_in, p.d_desc, p.d_bufs,
                           p.d_hdr, p.chunk, ext, cand, segtab);
        hipLaunchKernelGGL(big_carry_kernel<12>, dim3(gk), dim3(256), 0, p.stream, p.d_desc, p.d_bufs, p.d_hdr, p.chunk,
                           ext, segtab);
        hipLaunchKernelGGL(big_match_kernel<12>, dim3(gm), dim3(256), 0, p.stream, p.d_in, p.d_desc, p.d_bufs, p.d_hdr,
                           p.chunk, ext, cand, segtab, p.d_tok);
    } else if (p.hash_bits == 14) {
        hipLaunchKernelGGL(big_cand_kernel<14>, dim3(gc), dim3(64), 4u << 14, p.stream, p.d_in, p.d_desc, p.d_bufs,
                           p.d_hdr, p.chunk, ext, cand, segtab);
        hipLaunchKernelGGL(big_carry_kernel<14>, dim3(gk), dim3(256), 0, p.stream, p.d_desc, p.d_bufs, p.d_hdr, p.chunk,
                           ext, segtab);
        hipLaunchKernelGGL(big_match_kernel<14>, dim3(gm), dim3(256), 0, p.stream, p.d_in, p.d_desc, p.d_bufs, p.d_hdr,
                           p.chunk, ext, cand, segtab, p.d_tok);
    } else {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(big_parse_kernel, dim3(gp), dim3(BIG_THREADS), 0, p.stream, p.d_desc, p.d_bufs, p.d_hdr, p.chunk,
                       ext, p.d_tok, bitmap, p.d_stage, p.stage_words, p.d_info, whole);
    return hipGetLastError();
}

}  // namespace orot_gpu
