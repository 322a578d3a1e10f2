// inflate.hip -- gfx950 batched inflate (RFC 1951) with zlib (RFC 1950) and
// gzip (RFC 1952) framing: the decompress side of orot's whole-buffer API
// (deflate_decompress, src/api/deflate_api.cpp:138-215, with the framing
// checks of src/formats/zlib_wrapper.cpp and gzip_wrapper.cpp:76-138).
//
// One wave decodes one stream.  Huffman decoding is sequential, so every lane
// of the wave carries the same decoder state; the lanes share the work that
// is not: building code tables, copying matches (lane i writes byte i of the
// match, from out[pos - dist + i mod dist], which is always before pos),
// copying stored blocks, and the final CRC-32 / Adler-32 over the output (one
// piece per lane, combined with GF(2) shifts).  A batch of many streams
// fills the GPU: streams are taken from a device queue, longest first.
//
// Input is untrusted.  Every read of the compressed stream is bounds-checked
// (bytes past its end read as zero and a stream that consumes them fails),
// every write is checked against the output capacity, every distance against
// the bytes produced, and code lengths are validated as zlib's inflate does,
// so a corrupt stream ends with DEFLATE_DATA_ERROR, never an access outside
// the caller's buffers.
#include "deflate_gpu.h"
#include "device_common.h"

namespace orot_gpu {
namespace {

constexpr int INF_WAVES = 4;  // streams in flight per workgroup
constexpr int INF_THREADS = 64 * INF_WAVES;
constexpr int INF_ROOT = 10;  // bits resolved by one root-table lookup
constexpr uint32_t INF_ROOT_SIZE = 1u << INF_ROOT;

constexpr int32_t INF_OK = 0;
constexpr int32_t INF_DATA_ERROR = -1;   // DEFLATE_DATA_ERROR
constexpr int32_t INF_NEED_OUTPUT = 3;   // DEFLATE_NEED_OUTPUT

// RFC 1951 3.2.5: length and distance bases and extra bits
__constant__ uint16_t c_len_base[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                        31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t c_len_ext[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint16_t c_dist_base[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,    65,    97,    129,
                                         193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
__constant__ uint8_t c_dist_ext[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
// RFC 1951 3.2.7: order of the code length code lengths
__constant__ uint8_t c_cl_perm[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

// One wave's tables: slot 0 literal/length (or the code length code while a
// dynamic header is read), slot 1 distance.
struct InfTables {
    uint16_t root[2][INF_ROOT_SIZE];  // (symbol << 4) | length; 0: longer code or none
    uint16_t sorted[2][320];          // symbols by (length, symbol)
    uint16_t count[2][16];            // codes of each length
    uint8_t lens[320];                // code lengths being read / built
};
constexpr size_t INF_LDS = sizeof(InfTables) * INF_WAVES;

// Memory-ordering point for the wave's own LDS and global traffic: other
// lanes' writes before it are seen by reads after it.
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup"); }

// Builds table `t` from lens[0, nsym).  Returns false for lengths that are not
// a prefix code: over-subscribed, or incomplete where zlib's inflate_table
// rejects it (the code length code, or a longer code than one 1-bit code).
// An empty code is accepted; decoding a symbol from it fails.
__device__ bool inf_build(InfTables& T, int t, const uint8_t* lens, int nsym, bool is_cl, int lane) {
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t total[16];
#pragma unroll
    for (int L = 0; L < 16; ++L) total[L] = 0;
    for (int g = 0; g < nsym; g += 64) {
        const int i = g + lane;
        const uint32_t len = i < nsym ? lens[i] : 0u;
#pragma unroll
        for (int L = 1; L < 16; ++L) total[L] += (uint32_t)__popcll(ballot(len == (uint32_t)L));
    }
    int left = 1, maxlen = 0;
    bool over = false;
#pragma unroll
    for (int L = 1; L < 16; ++L) {
        left = (left << 1) - (int)total[L];
        over |= left < 0;
        if (total[L]) maxlen = L;
    }
    if (over) return false;
    if (maxlen > 0 && left > 0 && (is_cl || maxlen != 1)) return false;
    if (lane == 0) {
#pragma unroll
        for (int L = 0; L < 16; ++L) T.count[t][L] = (uint16_t)(L ? total[L] : 0u);
    }
    for (uint32_t e = (uint32_t)lane; e < INF_ROOT_SIZE; e += 64) T.root[t][e] = 0;
    wave_sync();
    // canonical codes (RFC 1951 3.2.2): first code and first sorted slot of each length
    uint32_t first[16], offs[16], run[16];
    {
        uint32_t code = 0, off = 0;
        first[0] = 0;
        offs[0] = 0;
#pragma unroll
        for (int L = 1; L < 16; ++L) {
            code = (code + total[L - 1]) << 1;
            first[L] = code;
            offs[L] = off;
            off += total[L];
            run[L] = 0;
        }
    }
    for (int g = 0; g < nsym; g += 64) {
        const int i = g + lane;
        const uint32_t len = i < nsym ? lens[i] : 0u;
        uint32_t code = 0, slot = 0;
#pragma unroll
        for (int L = 1; L < 16; ++L) {
            const uint64_t m = ballot(len == (uint32_t)L);
            if (len == (uint32_t)L) {
                const uint32_t r = run[L] + (uint32_t)__popcll(m & lt);
                code = first[L] + r;
                slot = offs[L] + r;
            }
            run[L] += (uint32_t)__popcll(m);
        }
        if (len) {
            T.sorted[t][slot] = (uint16_t)i;
            if (len <= (uint32_t)INF_ROOT) {
                const uint32_t rev = __builtin_bitreverse32(code) >> (32 - len);
                for (uint32_t e = rev; e < INF_ROOT_SIZE; e += 1u << len) T.root[t][e] = (uint16_t)((i << 4) | len);
            }
        }
    }
    wave_sync();
    return true;
}

// Bit reader over the stream bytes [0, n): a 64-bit buffer refilled 8 bytes
// at a time (F. Giesen's branchless refill), so at least 56 bits are held
// after refill().  Bytes past n read as zero; consumed_bits() tells whether
// the decoder used them.
struct BitReader {
    const uint8_t* p;
    uint64_t n;
    uint64_t ip;  // next byte to load
    uint64_t bb;
    uint32_t bc;

    __device__ void refill() {
        uint64_t v;
        if (ip + 8 <= n) {
            v = ld64u(p + ip);
        } else {
            v = 0;
            for (uint32_t k = 0; k < 8; ++k)
                if (ip + k < n) v |= (uint64_t)p[ip + k] << (8 * k);
        }
        bb |= v << bc;
        ip += (63u - bc) >> 3;
        bc |= 56u;
    }
    __device__ uint32_t bits(uint32_t k) {  // k <= 32, after a refill that covers them
        const uint32_t v = (uint32_t)(bb & ((1ull << k) - 1ull));
        bb >>= k;
        bc -= k;
        return v;
    }
    __device__ uint64_t consumed_bits() const { return ip * 8 - bc; }
    __device__ bool runaway() const { return ip > n + 16; }
};

// One symbol of table t (at least 15 bits held).  The root table resolves
// codes of up to INF_ROOT bits; longer ones take the canonical decode of
// M. Adler's puff (one bit per step against the per-length counts).
__device__ __forceinline__ int inf_decode(const InfTables& T, int t, BitReader& br) {
    const uint32_t e = T.root[t][br.bb & (INF_ROOT_SIZE - 1)];
    if (e) {
        br.bb >>= (e & 15u);
        br.bc -= (e & 15u);
        return (int)(e >> 4);
    }
    int code = 0, first = 0, index = 0;
    for (int len = 1; len < 16; ++len) {
        code |= (int)(br.bb & 1ull);
        br.bb >>= 1;
        br.bc -= 1;
        const int count = T.count[t][len];
        if (code - count < first) return T.sorted[t][index + (code - first)];
        index += count;
        first += count;
        first <<= 1;
        code <<= 1;
    }
    return -1;
}

// Inflates the raw stream src[0, n) into out[0, cap).  Returns the status;
// *produced is the output size on success.
__device__ int32_t inflate_raw(InfTables& T, const uint8_t* src, uint64_t n, uint8_t* out, uint64_t cap,
                               uint64_t* produced, int lane) {
    BitReader br{src, n, 0, 0, 0};
    uint64_t pos = 0, fenced = 0;  // out[0, fenced) is visible to the wave's loads
    int fixed_built = 0;            // tables hold the fixed code
    // output overflow of a stream that has already run past its end is a
    // truncated stream, not a short output buffer
    auto need_output = [&]() { return br.consumed_bits() > n * 8 ? INF_DATA_ERROR : INF_NEED_OUTPUT; };
    for (;;) {
        br.refill();
        if (br.runaway()) return INF_DATA_ERROR;
        const uint32_t hdr = br.bits(3);
        const bool final_block = hdr & 1u;
        const uint32_t type = hdr >> 1;
        if (type == 0) {
            // stored (3.2.4): byte boundary, LEN, NLEN, LEN bytes
            br.bits(br.bc & 7u);
            const uint32_t len = br.bits(16), nlen = br.bits(16);
            if ((len ^ 0xFFFFu) != nlen) return INF_DATA_ERROR;
            // the data starts at the exact stream position (the buffer may
            // hold bytes past the end of the stream, read as zero)
            const uint64_t start = br.consumed_bits() >> 3;
            if (start + len > n) return INF_DATA_ERROR;
            if (pos + len > cap) return need_output();
            for (uint32_t i = (uint32_t)lane; i < len; i += 64) out[pos + i] = src[start + i];
            br.ip = start + len;
            br.bb = 0;
            br.bc = 0;
            pos += len;
        } else if (type == 3) {
            return INF_DATA_ERROR;
        } else {
            if (type == 1) {
                if (!fixed_built) {
                    // 3.2.6: 0-143 8 bits, 144-255 9, 256-279 7, 280-287 8; 32 distance codes of 5 bits
                    for (int i = lane; i < 288; i += 64)
                        T.lens[i] = (uint8_t)(i < 144 ? 8 : (i < 256 ? 9 : (i < 280 ? 7 : 8)));
                    wave_sync();
                    inf_build(T, 0, T.lens, 288, false, lane);
                    if (lane < 32) T.lens[lane] = 5;
                    wave_sync();
                    inf_build(T, 1, T.lens, 32, false, lane);
                    fixed_built = 1;
                }
            } else {
                // dynamic (3.2.7)
                fixed_built = 0;
                const uint32_t hlit = br.bits(5) + 257, hdist = br.bits(5) + 1, hclen = br.bits(4) + 4;
                if (hlit > 286 || hdist > 30) return INF_DATA_ERROR;
                for (int i = lane; i < 19; i += 64) T.lens[i] = 0;
                wave_sync();
                for (uint32_t k = 0; k < hclen; ++k) {
                    if ((k & 15u) == 0) br.refill();
                    const uint32_t v = br.bits(3);
                    if (lane == 0) T.lens[c_cl_perm[k]] = (uint8_t)v;
                }
                wave_sync();
                if (!inf_build(T, 0, T.lens, 19, true, lane)) return INF_DATA_ERROR;
                const uint32_t total = hlit + hdist;
                uint32_t k = 0;
                while (k < total) {
                    br.refill();
                    if (br.runaway()) return INF_DATA_ERROR;
                    const int sym = inf_decode(T, 0, br);
                    if (sym < 0) return INF_DATA_ERROR;
                    if (sym < 16) {
                        if (lane == 0) T.lens[k] = (uint8_t)sym;
                        ++k;
                        continue;
                    }
                    uint32_t rep, val = 0;
                    if (sym == 16) {
                        if (k == 0) return INF_DATA_ERROR;
                        wave_sync();
                        val = T.lens[k - 1];
                        rep = 3 + br.bits(2);
                    } else if (sym == 17) {
                        rep = 3 + br.bits(3);
                    } else {
                        rep = 11 + br.bits(7);
                    }
                    if (k + rep > total) return INF_DATA_ERROR;
                    // rep reaches 138 (symbol 18): more entries than lanes
                    for (uint32_t i = (uint32_t)lane; i < rep; i += 64) T.lens[k + i] = (uint8_t)val;
                    k += rep;
                }
                wave_sync();
                if (T.lens[256] == 0) return INF_DATA_ERROR;  // no end-of-block code
                if (!inf_build(T, 0, T.lens, (int)hlit, false, lane)) return INF_DATA_ERROR;
                if (!inf_build(T, 1, T.lens + hlit, (int)hdist, false, lane)) return INF_DATA_ERROR;
            }
            // symbols until end of block; a refill holds 56 bits, enough for a
            // length code (15) + extra (5) + distance code (15) + extra (13)
            for (;;) {
                br.refill();
                if (br.runaway()) return INF_DATA_ERROR;
                int sym = inf_decode(T, 0, br);
                if (sym < 0) return INF_DATA_ERROR;
                if (sym < 256) {
                    if (pos >= cap) return need_output();
                    if (lane == 0) out[pos] = (uint8_t)sym;
                    ++pos;
                    continue;
                }
                if (sym == 256) break;
                sym -= 257;
                if (sym >= 29) return INF_DATA_ERROR;
                const uint32_t len = c_len_base[sym] + br.bits(c_len_ext[sym]);
                const int ds = inf_decode(T, 1, br);
                if (ds < 0 || ds >= 30) return INF_DATA_ERROR;
                const uint32_t dist = c_dist_base[ds] + br.bits(c_dist_ext[ds]);
                if (dist > pos) return INF_DATA_ERROR;
                if (pos + len > cap) return need_output();
                // the sources out[pos - dist, pos - dist + min(len, dist)) must be visible
                if (pos - dist + (len < dist ? len : dist) > fenced) {
                    wave_sync();
                    fenced = pos;
                }
                for (uint32_t i = (uint32_t)lane; i < len; i += 64) {
                    const uint32_t j = i < dist ? i : i % dist;
                    out[pos + i] = out[pos - dist + j];
                }
                pos += len;
            }
        }
        if (final_block) break;
    }
    if (br.consumed_bits() > n * 8) return INF_DATA_ERROR;  // ran past the end of the stream
    *produced = pos;
    return INF_OK;
}

// CRC-32 (gzip) of p[0, n): one piece per lane, combined with GF(2) shifts.
__device__ uint32_t wave_crc32(const uint8_t* p, uint64_t n, int lane) {
    const uint64_t per = (((n + 63) / 64) + 3) & ~3ull;
    const uint64_t a = min(n, (uint64_t)lane * per), b = min(n, a + per);
    uint32_t r = 0;
    uint64_t i = a;
    for (; i + 4 <= b; i += 4) {
        r ^= ld32u(p + i);
        r = c_crc.t[3][r & 0xFF] ^ c_crc.t[2][(r >> 8) & 0xFF] ^ c_crc.t[1][(r >> 16) & 0xFF] ^ c_crc.t[0][r >> 24];
    }
    for (; i < b; ++i) r = (r >> 8) ^ c_crc.t[0][(r ^ p[i]) & 0xFF];
    uint32_t x = (b > a) ? gf2_mul(gf2_x8n(n - b), r) : 0u;
    for (int d = 32; d; d >>= 1) x ^= (uint32_t)__shfl_xor((int)x, d, 64);
    return ~(gf2_mul(gf2_x8n(n), 0xFFFFFFFFu) ^ x);
}

// Adler-32 (zlib) of p[0, n): per-lane sums combined by piece offsets.
__device__ uint32_t wave_adler32(const uint8_t* p, uint64_t n, int lane) {
    const uint64_t per = (n + 63) / 64;
    const uint64_t a = min(n, (uint64_t)lane * per), b = min(n, a + per);
    uint64_t s1 = 0, s2 = 0;
    uint64_t i = a;
    while (i < b) {
        const uint64_t e = min(b, i + 4096);  // s2 stays far below 2^64 between reductions
        for (; i < e; ++i) {
            s1 += p[i];
            s2 += s1;
        }
        s1 %= ADLER_MOD;
        s2 %= ADLER_MOD;
    }
    // bytes after the piece add s1 once each to s2
    uint64_t x1 = s1, x2 = (s2 + ((n - b) % ADLER_MOD) * s1) % ADLER_MOD;
    for (int d = 32; d; d >>= 1) {
        x1 += (uint64_t)__shfl_xor((long long)x1, d, 64);
        x2 += (uint64_t)__shfl_xor((long long)x2, d, 64);
    }
    const uint32_t a1 = (uint32_t)((1 + x1) % ADLER_MOD);
    const uint32_t a2 = (uint32_t)((n % ADLER_MOD + x2) % ADLER_MOD);
    return (a2 << 16) | a1;
}

__device__ __forceinline__ uint32_t rd_le32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// One stream with its framing (format 0 raw, 1 zlib, 2 gzip).
__device__ int32_t inflate_item(InfTables& T, const uint8_t* src, uint64_t n, uint8_t* out, uint64_t cap,
                                int format, uint64_t* produced, int lane) {
    uint64_t off = 0, rn = n;
    if (format == 1) {
        // zlib_wrapper.cpp: CM 8, FCHECK, no preset dictionary; Adler-32 trailer (big-endian)
        if (n < 6) return INF_DATA_ERROR;
        if ((src[0] & 0x0F) != 8) return INF_DATA_ERROR;
        if ((((uint32_t)src[0] << 8) | src[1]) % 31 != 0) return INF_DATA_ERROR;
        if (src[1] & 0x20) return INF_DATA_ERROR;
        off = 2;
        rn = n - 6;
    } else if (format == 2) {
        // gzip_wrapper.cpp:76-138: magic, method 8, optional fields; CRC-32 and ISIZE trailer
        if (n < 18) return INF_DATA_ERROR;
        if (src[0] != 0x1F || src[1] != 0x8B || src[2] != 8) return INF_DATA_ERROR;
        const uint32_t flg = src[3];
        off = 10;
        if (flg & 0x04) {
            if (off + 2 > n) return INF_DATA_ERROR;
            off += 2 + ((uint32_t)src[off] | ((uint32_t)src[off + 1] << 8));
        }
        if (flg & 0x08) {
            while (off < n && src[off] != 0) ++off;
            ++off;
        }
        if (flg & 0x10) {
            while (off < n && src[off] != 0) ++off;
            ++off;
        }
        if (flg & 0x02) off += 2;
        if (off + 8 > n) return INF_DATA_ERROR;
        rn = n - off - 8;
    }
    uint64_t m = 0;
    const int32_t st = inflate_raw(T, src + off, rn, out, cap, &m, lane);
    if (st != INF_OK) return st;
    wave_sync();  // the output is complete before the checksum reads it
    if (format == 1) {
        const uint8_t* tr = src + n - 4;
        const uint32_t want = ((uint32_t)tr[0] << 24) | ((uint32_t)tr[1] << 16) | ((uint32_t)tr[2] << 8) | tr[3];
        if (wave_adler32(out, m, lane) != want) return INF_DATA_ERROR;
    } else if (format == 2) {
        const uint8_t* tr = src + n - 8;
        if (wave_crc32(out, m, lane) != rd_le32(tr)) return INF_DATA_ERROR;
        if (rd_le32(tr + 4) != (uint32_t)(m & 0xFFFFFFFFu)) return INF_DATA_ERROR;
    }
    *produced = m;
    return INF_OK;
}

// Persistent: each wave takes the next stream of the queue (order[] is the
// streams longest first) until the queue is empty.
__global__ __launch_bounds__(INF_THREADS) void inflate_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                              const InflateItem* __restrict__ items,
                                                              const uint32_t* __restrict__ order, uint32_t count,
                                                              int format, InflateResult* __restrict__ res,
                                                              uint32_t* __restrict__ queue) {
    extern __shared__ uint32_t smem[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    InfTables& T = reinterpret_cast<InfTables*>(smem)[wv];
    for (;;) {
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(queue, 1u);
        k = (uint32_t)__shfl((int)k, 0, 64);
        if (k >= count) break;
        const uint32_t it = order[k];
        const InflateItem I = items[it];
        uint64_t m = 0;
        const int32_t st = inflate_item(T, in + I.in_off, I.in_size, out + I.out_off, I.out_cap, format, &m, lane);
        if (lane == 0) {
            res[it].out_size = st == INF_OK ? m : 0;
            res[it].status = st;
        }
    }
}

}  // namespace

hipError_t launch_inflate(const uint8_t* d_in, uint8_t* d_out, const InflateItem* d_items, const uint32_t* d_order,
                          uint32_t count, int format, InflateResult* d_res, uint32_t* d_queue, int cus,
                          hipStream_t stream) {
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&inflate_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)INF_LDS);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipError_t e = hipMemsetAsync(d_queue, 0, sizeof(uint32_t), stream);
    if (e != hipSuccess) return e;
    if (count == 0) return hipSuccess;
    const uint32_t waves = (count + INF_WAVES - 1) / INF_WAVES;
    const uint32_t cap = (uint32_t)(cus > 0 ? cus : 1) * 8u;
    const unsigned g = waves < cap ? waves : cap;
    hipLaunchKernelGGL(inflate_kernel, dim3(g), dim3(INF_THREADS), INF_LDS, stream, d_in, d_out, d_items, d_order,
                       count, format, d_res, d_queue);
    return hipGetLastError();
}

}  // namespace orot_gpu
