// inflate_host.cpp -- the host inflater of inflate_host.h.
//
// Decoding follows RFC 1951 directly: canonical codes (3.2.2) decoded with a
// 9-bit lookup table and, for longer codes, a walk over the per-length code
// counts (M. Adler's puff); code-length validation as zlib's inflate_table
// does (over-subscribed codes and incomplete ones are rejected, except a
// single one-bit code and an empty distance code); RFC 1950 / 1952 headers
// and trailers (Adler-32 big-endian; CRC-32 and ISIZE little-endian).
//
// The decoder reads bits from an input view it may rewind: a symbol, a
// length / distance pair or a header field is decoded whole or, when the
// view ends inside it, the read position and bit accumulator go back to its
// start and the call returns for more input.  A stored block's payload is
// taken from the bit accumulator's whole bytes first and then from the
// input, so no input byte is lost at the switch from a Huffman block.
#include "inflate_host.h"

#include <cstring>

#include "host_common.h"

namespace orot_gpu {
namespace {

constexpr uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                   31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
constexpr uint8_t LEN_EXT[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
constexpr uint16_t DIST_BASE[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                    33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                    1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
constexpr uint8_t DIST_EXT[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
constexpr uint8_t CL_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

// CRC-32 (IEEE, reflected 0xEDB88320), slicing by 8.
struct CrcTables {
    uint32_t t[8][256];
    CrcTables() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
            t[0][i] = c;
        }
        for (uint32_t i = 0; i < 256; ++i)
            for (int s = 1; s < 8; ++s) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xFF];
    }
};
const CrcTables& crc_tables() {
    static const CrcTables T;
    return T;
}

uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    const CrcTables& T = crc_tables();
    uint32_t c = ~crc;
    while (n >= 8) {
        uint32_t a, b;
        std::memcpy(&a, p, 4);
        std::memcpy(&b, p + 4, 4);
        a ^= c;
        c = T.t[7][a & 0xFF] ^ T.t[6][(a >> 8) & 0xFF] ^ T.t[5][(a >> 16) & 0xFF] ^ T.t[4][a >> 24] ^
            T.t[3][b & 0xFF] ^ T.t[2][(b >> 8) & 0xFF] ^ T.t[1][(b >> 16) & 0xFF] ^ T.t[0][b >> 24];
        p += 8;
        n -= 8;
    }
    while (n--) c = (c >> 8) ^ T.t[0][(c ^ *p++) & 0xFF];
    return ~c;
}

// Adler-32 (RFC 1950 8.2): sums reduced every 5552 bytes, before they can
// overflow 32 bits.
uint32_t adler32_update(uint32_t adler, const uint8_t* p, size_t n) {
    uint32_t a = adler & 0xFFFF, b = adler >> 16;
    while (n) {
        size_t k = n < 5552 ? n : 5552;
        n -= k;
        while (k--) {
            a += *p++;
            b += a;
        }
        a %= 65521u;
        b %= 65521u;
    }
    return (b << 16) | a;
}

// Builds c from lens[0, n).  0: a code (possibly empty); -1: not a prefix
// code that inflate accepts.  `is_cl`: the code-length code, which must be
// complete.
int build(HostInflater::Code& c, const uint8_t* lens, int n, bool is_cl) {
    std::memset(&c, 0, sizeof(c));
    for (int i = 0; i < n; ++i) c.count[lens[i]]++;
    c.count[0] = 0;
    int left = 1, maxlen = 0;
    for (int L = 1; L < 16; ++L) {
        left = (left << 1) - c.count[L];
        if (left < 0) return -1;  // over-subscribed
        if (c.count[L]) maxlen = L;
    }
    if (maxlen > 0 && left > 0 && (is_cl || maxlen != 1)) return -1;  // incomplete
    uint16_t offs[16];
    offs[1] = 0;
    for (int L = 1; L < 15; ++L) offs[L + 1] = (uint16_t)(offs[L] + c.count[L]);
    for (int i = 0; i < n; ++i)
        if (lens[i]) c.sym[offs[lens[i]]++] = (uint16_t)i;
    // canonical codes of up to 9 bits into the lookup table, bit-reversed
    uint32_t code = 0;
    int k = 0;
    for (int L = 1; L <= 9; ++L) {
        for (int j = 0; j < c.count[L]; ++j, ++k, ++code) {
            uint32_t rev = 0;
            for (int b = 0; b < L; ++b) rev |= ((code >> b) & 1u) << (L - 1 - b);
            for (uint32_t e = rev; e < 512; e += 1u << L) c.fast[e] = (uint16_t)((c.sym[k] << 4) | L);
        }
        code <<= 1;
    }
    return 0;
}

}  // namespace

HostInflater::HostInflater(int format, const deflate_allocator& alloc)
    : format_(format), alloc_(alloc), held_(HookAlloc<uint8_t>(alloc)) {
    st_ = format == DEFLATE_FORMAT_ZLIB ? ZHDR : (format == DEFLATE_FORMAT_GZIP ? GHDR : BLOCK);
    check_ = format == DEFLATE_FORMAT_ZLIB ? 1u : 0u;
}

HostInflater* HostInflater::create(int format) {
    const deflate_allocator a = host_allocator();
    void* p = a.alloc(a.opaque, sizeof(HostInflater), alignof(HostInflater) < 64 ? 64 : alignof(HostInflater));
    if (!p) return nullptr;
    return new (p) HostInflater(format, a);
}

void HostInflater::destroy(HostInflater* p) {
    if (!p) return;
    const deflate_allocator a = p->alloc_;
    p->~HostInflater();
    a.free(a.opaque, p);
}

bool HostInflater::need(int k) {
    while (cnt_ < k) {
        if (pos_ == len_) return false;
        acc_ |= (uint64_t)base_[pos_++] << cnt_;
        cnt_ += 8;
    }
    return true;
}

uint32_t HostInflater::take(int k) {
    const uint32_t v = (uint32_t)(acc_ & ((1ull << k) - 1ull));
    acc_ >>= k;
    cnt_ -= k;
    return v;
}

// One symbol: >= 0, -1 for a bit pattern no code starts, -2 when the input
// view ends before the code does (nothing consumed).
int HostInflater::decode(const Code& c) {
    (void)need(15);
    const uint16_t e = c.fast[acc_ & 511];
    if (e) {
        const int L = e & 15;
        if (L > cnt_) return -2;
        acc_ >>= L;
        cnt_ -= L;
        return e >> 4;
    }
    int code = 0, first = 0, index = 0;
    for (int L = 1; L < 16; ++L) {
        if (L > cnt_) return -2;
        code |= (int)((acc_ >> (L - 1)) & 1ull);
        const int count = c.count[L];
        if (code - count < first) {
            acc_ >>= L;
            cnt_ -= L;
            return c.sym[index + (code - first)];
        }
        index += count;
        first += count;
        first <<= 1;
        code <<= 1;
    }
    return -1;
}

void HostInflater::put(uint8_t b) {
    *o_++ = b;
    --on_;
    win_[wpos_++ & 32767u] = b;
    ++total_out_;
}

void HostInflater::put_bytes(const uint8_t* s, size_t m) {
    std::memcpy(o_, s, m);
    o_ += m;
    on_ -= m;
    total_out_ += m;
    // the last 32 KiB of the copy into the history ring
    const size_t h = m < 32768 ? m : 32768;
    const uint8_t* t = s + (m - h);
    wpos_ += (uint32_t)(m - h);
    for (size_t i = 0; i < h;) {
        const uint32_t at = wpos_ & 32767u;
        const size_t run = (32768 - at) < (h - i) ? (32768 - at) : (h - i);
        std::memcpy(win_ + at, t + i, run);
        wpos_ += (uint32_t)run;
        i += run;
    }
}

void HostInflater::checksum_out() {
    if (o_ > chk_from_) {
        if (format_ == DEFLATE_FORMAT_ZLIB) check_ = adler32_update(check_, chk_from_, (size_t)(o_ - chk_from_));
        if (format_ == DEFLATE_FORMAT_GZIP) check_ = crc32_update(check_, chk_from_, (size_t)(o_ - chk_from_));
    }
    chk_from_ = o_;
}

deflate_result HostInflater::decode_loop() {
    for (;;) {
        // a rewind point: the state of the bit reader before the current item
        const size_t pos0 = pos_;
        const uint64_t acc0 = acc_;
        const int cnt0 = cnt_;
        auto more = [&]() {
            pos_ = pos0;
            acc_ = acc0;
            cnt_ = cnt0;
            return DEFLATE_OK;
        };
        switch (st_) {
        case ZHDR: {  // RFC 1950: CM 8, CINFO <= 7, FCHECK, no preset dictionary
            if (!need(16)) return more();
            const uint32_t cmf = take(8), flg = take(8);
            if ((cmf & 0x0F) != 8 || (cmf >> 4) > 7 || ((cmf << 8) | flg) % 31 != 0 || (flg & 0x20)) {
                st_ = BAD;
                continue;
            }
            st_ = BLOCK;
            continue;
        }
        case GHDR: {  // RFC 1952: ID1 ID2 CM FLG MTIME(4) XFL OS
            if (!need(32)) return more();
            const uint32_t w = take(32);
            if ((w & 0xFFFF) != 0x8B1F || ((w >> 16) & 0xFF) != 8 || (w >> 29)) {
                st_ = BAD;
                continue;
            }
            gflags_ = (uint8_t)(w >> 24);
            if (!need(32)) return more();
            (void)take(32);  // MTIME
            if (!need(16)) return more();
            (void)take(16);  // XFL, OS
            st_ = (gflags_ & 0x04) ? GEXLEN : GNAME;
            continue;
        }
        case GEXLEN:
            if (!need(16)) return more();
            skip_ = take(16);
            st_ = GEXTRA;
            continue;
        case GEXTRA:
            while (skip_) {
                if (!need(8)) return DEFLATE_OK;
                (void)take(8);
                --skip_;
            }
            st_ = GNAME;
            continue;
        case GNAME:
        case GCOMMENT: {
            const uint8_t bit = st_ == GNAME ? 0x08 : 0x10;
            if (gflags_ & bit) {
                for (;;) {
                    if (!need(8)) return DEFLATE_OK;
                    if (take(8) == 0) break;
                }
                gflags_ &= (uint8_t)~bit;
            }
            st_ = st_ == GNAME ? GCOMMENT : GHCRC;
            continue;
        }
        case GHCRC:
            if (gflags_ & 0x02) {
                if (!need(16)) return more();
                (void)take(16);
            }
            st_ = BLOCK;
            continue;
        case BLOCK: {
            if (!need(3)) return more();
            final_ = take(1) != 0;
            const uint32_t type = take(2);
            if (type == 0) {
                (void)take(cnt_ & 7);  // to the byte boundary
                st_ = STORED_LEN;
            } else if (type == 1) {
                uint8_t l[320];
                for (int i = 0; i < 288; ++i) l[i] = (uint8_t)(i < 144 ? 8 : (i < 256 ? 9 : (i < 280 ? 7 : 8)));
                // 32 five-bit distance codes make a complete code; 30 and 31 never occur
                for (int i = 0; i < 32; ++i) l[288 + i] = 5;
                (void)build(lit_, l, 288, false);
                (void)build(dist_, l + 288, 32, false);
                st_ = CODES;
            } else if (type == 2) {
                st_ = DYN_HDR;
            } else {
                st_ = BAD;
            }
            continue;
        }
        case STORED_LEN: {
            if (!need(32)) return more();
            const uint32_t len = take(16), nlen = take(16);
            if ((len ^ 0xFFFFu) != nlen) {
                st_ = BAD;
                continue;
            }
            stored_left_ = len;
            st_ = STORED_COPY;
            continue;
        }
        case STORED_COPY:
            // whole bytes already in the accumulator come first
            while (stored_left_ && cnt_ >= 8) {
                if (!on_) return DEFLATE_NEED_OUTPUT;
                put((uint8_t)take(8));
                --stored_left_;
            }
            while (stored_left_) {
                if (!on_) return DEFLATE_NEED_OUTPUT;
                if (pos_ == len_) return DEFLATE_OK;
                size_t m = stored_left_;
                if (m > on_) m = on_;
                if (m > len_ - pos_) m = len_ - pos_;
                put_bytes(base_ + pos_, m);
                pos_ += m;
                stored_left_ -= (uint32_t)m;
            }
            st_ = final_ ? (format_ == DEFLATE_FORMAT_RAW ? DONE : TRAILER) : BLOCK;
            continue;
        case DYN_HDR:
            if (!need(14)) return more();
            hlit_ = (int)take(5) + 257;
            hdist_ = (int)take(5) + 1;
            hclen_ = (int)take(4) + 4;
            if (hlit_ > 286 || hdist_ > 30) {
                st_ = BAD;
                continue;
            }
            std::memset(lens_, 0, sizeof(lens_));
            idx_ = 0;
            st_ = DYN_CL;
            continue;
        case DYN_CL:
            while (idx_ < hclen_) {
                if (!need(3)) return DEFLATE_OK;
                lens_[CL_ORDER[idx_++]] = (uint8_t)take(3);
            }
            if (build(cl_, lens_, 19, true) != 0) {
                st_ = BAD;
                continue;
            }
            std::memset(lens_, 0, sizeof(lens_));
            idx_ = 0;
            st_ = DYN_LENS;
            continue;
        case DYN_LENS: {
            const int total = hlit_ + hdist_;
            while (idx_ < total) {
                const size_t p0 = pos_;
                const uint64_t a0 = acc_;
                const int c0 = cnt_;
                const int sym = decode(cl_);
                if (sym == -2) return DEFLATE_OK;
                if (sym < 0) {
                    st_ = BAD;
                    break;
                }
                if (sym < 16) {
                    lens_[idx_++] = (uint8_t)sym;
                    continue;
                }
                int rep, val = 0;
                const int eb = sym == 16 ? 2 : (sym == 17 ? 3 : 7);
                if (!need(eb)) {
                    pos_ = p0;
                    acc_ = a0;
                    cnt_ = c0;
                    return DEFLATE_OK;
                }
                if (sym == 16) {
                    if (idx_ == 0) {
                        st_ = BAD;
                        break;
                    }
                    val = lens_[idx_ - 1];
                    rep = 3 + (int)take(2);
                } else if (sym == 17) {
                    rep = 3 + (int)take(3);
                } else {
                    rep = 11 + (int)take(7);
                }
                if (idx_ + rep > total) {
                    st_ = BAD;
                    break;
                }
                while (rep--) lens_[idx_++] = (uint8_t)val;
            }
            if (st_ == BAD) continue;
            if (lens_[256] == 0 || build(lit_, lens_, hlit_, false) != 0 ||
                build(dist_, lens_ + hlit_, hdist_, false) != 0) {
                st_ = BAD;
                continue;
            }
            st_ = CODES;
            continue;
        }
        case CODES:
            for (;;) {
                const size_t p0 = pos_;
                const uint64_t a0 = acc_;
                const int c0 = cnt_;
                auto rewind = [&](deflate_result r) {
                    pos_ = p0;
                    acc_ = a0;
                    cnt_ = c0;
                    return r;
                };
                int sym = decode(lit_);
                if (sym == -2) return DEFLATE_OK;
                if (sym < 0) {
                    st_ = BAD;
                    break;
                }
                if (sym < 256) {
                    if (!on_) return rewind(DEFLATE_NEED_OUTPUT);
                    put((uint8_t)sym);
                    continue;
                }
                if (sym == 256) {
                    st_ = final_ ? (format_ == DEFLATE_FORMAT_RAW ? DONE : TRAILER) : BLOCK;
                    break;
                }
                sym -= 257;
                if (sym >= 29) {
                    st_ = BAD;
                    break;
                }
                if (!need(LEN_EXT[sym])) return rewind(DEFLATE_OK);
                const uint32_t len = LEN_BASE[sym] + take(LEN_EXT[sym]);
                const int ds = decode(dist_);
                if (ds == -2) return rewind(DEFLATE_OK);
                if (ds < 0 || ds >= 30) {
                    st_ = BAD;
                    break;
                }
                if (!need(DIST_EXT[ds])) return rewind(DEFLATE_OK);
                const uint32_t dist = DIST_BASE[ds] + take(DIST_EXT[ds]);
                if (dist > total_out_) {
                    st_ = BAD;
                    break;
                }
                copy_len_ = len;
                copy_dist_ = dist;
                st_ = COPY;
                break;
            }
            continue;
        case COPY:
            while (copy_len_) {
                if (!on_) return DEFLATE_NEED_OUTPUT;
                put(win_[(wpos_ - copy_dist_) & 32767u]);
                --copy_len_;
            }
            st_ = CODES;
            continue;
        case TRAILER: {
            (void)take(cnt_ & 7);  // to the byte boundary
            const int nb = format_ == DEFLATE_FORMAT_ZLIB ? 4 : 8;
            if (!need(8 * nb)) return more();
            checksum_out();
            if (format_ == DEFLATE_FORMAT_ZLIB) {
                const uint32_t v = take(32);
                const uint32_t want = ((v & 0xFF) << 24) | ((v & 0xFF00) << 8) | ((v >> 8) & 0xFF00) | (v >> 24);
                if (want != check_) {
                    st_ = BAD;
                    continue;
                }
            } else {
                const uint32_t crc = take(32), isize = take(32);
                if (crc != check_ || isize != (uint32_t)(total_out_ & 0xFFFFFFFFu)) {
                    st_ = BAD;
                    continue;
                }
            }
            st_ = DONE;
            continue;
        }
        case DONE:
            return DEFLATE_STREAM_END;
        case BAD:
        default:
            return DEFLATE_DATA_ERROR;
        }
    }
}

deflate_result HostInflater::stream(const uint8_t** next_in, size_t* avail_in, uint8_t** next_out,
                                    size_t* avail_out) {
    if (st_ == BAD) return DEFLATE_DATA_ERROR;
    if (st_ == DONE) return DEFLATE_STREAM_END;
    // the held bytes not read yet, then this call's input
    const size_t fresh = *avail_in;
    held_.erase(held_.begin(), held_.begin() + (ptrdiff_t)pos_);
    if (fresh) held_.insert(held_.end(), *next_in, *next_in + fresh);
    base_ = held_.data();
    len_ = held_.size();
    pos_ = 0;
    o_ = *next_out;
    on_ = *avail_out;
    chk_from_ = o_;
    deflate_result r = decode_loop();
    checksum_out();
    size_t unused = 0;
    if (r == DEFLATE_STREAM_END) {
        // bytes after the stream go back to the caller (those of this call)
        const size_t whole = (size_t)cnt_ / 8;
        size_t after = (len_ - pos_) + whole;
        if (after > fresh) after = fresh;
        unused = after;
        held_.clear();
        pos_ = 0;
        acc_ = 0;
        cnt_ = 0;
    }
    *next_in += fresh - unused;
    *avail_in = unused;
    *next_out = o_;
    *avail_out = on_;
    return r;
}

deflate_result HostInflater::whole(const uint8_t* in, size_t n, uint8_t* out, size_t cap, size_t* produced) {
    base_ = in;
    len_ = n;
    pos_ = 0;
    o_ = out;
    on_ = cap;
    chk_from_ = o_;
    const deflate_result r = decode_loop();
    checksum_out();
    if (produced) *produced = (size_t)(o_ - out);
    return r;
}

}  // namespace orot_gpu
