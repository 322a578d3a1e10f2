// inflate_host.h -- resumable host inflater (RFC 1951, with RFC 1950 / 1952
// framing) behind orot's decompress entry points: inflate_stream_new /
// inflate_stream_decompress (orot src/api/stream_api.cpp:253-380) and
// deflate_decompress (orot src/api/deflate_api.cpp:138-215).  Decompression
// is not the GPU compress path; orot's host-buffer decompress API is served
// on the host.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <new>
#include <vector>

#include "../../include/orot/deflate.h"

namespace orot_gpu {

// std::allocator over a deflate_allocator (orot include/orot/deflate.h:49-54).
template <class T>
struct HookAlloc {
    using value_type = T;
    deflate_allocator a;
    explicit HookAlloc(const deflate_allocator& x) : a(x) {}
    template <class U>
    HookAlloc(const HookAlloc<U>& o) : a(o.a) {}
    T* allocate(size_t n) {
        void* p = a.alloc(a.opaque, n * sizeof(T), alignof(T) < 16 ? 16 : alignof(T));
        if (!p) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { a.free(a.opaque, p); }
    template <class U>
    bool operator==(const HookAlloc<U>& o) const { return a.alloc == o.a.alloc && a.free == o.a.free && a.opaque == o.a.opaque; }
    template <class U>
    bool operator!=(const HookAlloc<U>& o) const { return !(*this == o); }
};

class HostInflater {
public:
    HostInflater(int format, const deflate_allocator& alloc);

    // Allocated through the deflate_set_allocator hook; nullptr on failure.
    static HostInflater* create(int format);
    static void destroy(HostInflater* p);

    // orot's inflate_stream_decompress contract.  Every input byte is taken
    // (held inside until it is decoded), except after the end of the stream.
    // DEFLATE_OK: all input taken, more is needed; DEFLATE_NEED_OUTPUT: the
    // output space ran out first; DEFLATE_STREAM_END: the stream and its
    // trailer are complete, and the bytes of this call after it are left in
    // *next_in; DEFLATE_DATA_ERROR: invalid data (the inflater stays in the
    // error state).
    deflate_result stream(const uint8_t** next_in, size_t* avail_in, uint8_t** next_out, size_t* avail_out);

    // One whole stream in one call, decoded from the caller's buffer without
    // a copy: DEFLATE_STREAM_END with *produced set, DEFLATE_NEED_OUTPUT,
    // DEFLATE_OK (the input ended before the stream did) or
    // DEFLATE_DATA_ERROR.
    deflate_result whole(const uint8_t* in, size_t n, uint8_t* out, size_t cap, size_t* produced);

    struct Code {
        uint16_t count[16];   // codes of each length
        uint16_t sym[320];    // symbols by (length, symbol)
        uint16_t fast[512];   // 9-bit lookup: symbol << 4 | length, 0 for a longer code
    };

private:
    enum State : int {
        ZHDR, GHDR, GEXLEN, GEXTRA, GNAME, GCOMMENT, GHCRC, BLOCK, STORED_LEN, STORED_COPY, DYN_HDR, DYN_CL,
        DYN_LENS, CODES, COPY, TRAILER, DONE, BAD
    };

    deflate_result decode_loop();
    bool need(int k);
    uint32_t take(int k);
    int decode(const Code& c);
    void put(uint8_t b);
    void put_bytes(const uint8_t* s, size_t m);
    void checksum_out();

    int format_;
    State st_;
    // input view: base_[pos_, len_) not yet read into the bit accumulator
    const uint8_t* base_ = nullptr;
    size_t len_ = 0, pos_ = 0;
    uint64_t acc_ = 0;
    int cnt_ = 0;
    deflate_allocator alloc_;
    std::vector<uint8_t, HookAlloc<uint8_t>> held_;  // streamed input not decoded yet
    // output of the current call
    uint8_t* o_ = nullptr;
    size_t on_ = 0;
    uint8_t* chk_from_ = nullptr;
    // block state
    bool final_ = false;
    uint32_t stored_left_ = 0;
    uint32_t skip_ = 0;
    int hlit_ = 0, hdist_ = 0, hclen_ = 0, idx_ = 0;
    uint8_t lens_[320] = {};
    Code lit_{}, dist_{}, cl_{};
    uint32_t copy_len_ = 0, copy_dist_ = 0;
    uint8_t gflags_ = 0;
    // history: the last 32 KiB of output
    uint8_t win_[32768] = {};
    uint32_t wpos_ = 0;
    uint64_t total_out_ = 0;
    uint32_t check_ = 0;
};

}  // namespace orot_gpu
