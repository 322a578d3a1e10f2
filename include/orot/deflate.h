/*
 * orot/deflate.h -- drop-in C API of orot's DEFLATE compressor, served by the
 * MI355X (gfx950) implementation in liborot_amd.so.
 *
 * The types, values and entry points below keep the meaning of orot's
 * include/orot/deflate.h; each declaration cites the orot line it replaces.
 * Compression runs on the GPU.  The decompress side of orot's host-buffer
 * API (deflate_decompress, the streaming decompressor) is served by a host
 * inflater, since it is not the GPU compress path.
 *
 * Deviations from orot, all deliberate:
 *   - gzip trailers carry the IEEE CRC-32 on every host (orot's x86 build
 *     writes CRC-32C, src/simd/x86/match_sse42.cpp:76-99);
 *   - deflate_parallel_compress returns 0 instead of writing past
 *     out_capacity when the compressed stream does not fit;
 *   - num_threads is accepted and ignored (the GPU decides its parallelism);
 *   - block sizes are accepted up to 1 GiB; levels 4-12 (whose bytes are
 *     not orot's at any block size) run a block above 1 MiB as 1 MiB pieces,
 *     each with the 32 KiB before it as its dictionary;
 *   - deflate_decompress reports an output buffer that is too small as
 *     DEFLATE_NEED_OUTPUT, as orot does with its libdeflate backend (its
 *     built-in inflater reports DEFLATE_DATA_ERROR);
 *   - the streaming compressor cuts its input at whole chunks of the
 *     level's block size and ends a sync or full flush with an empty stored
 *     block, so its streams stay valid past orot's single 128 KiB block.
 */
#ifndef OROT_DEFLATE_H
#define OROT_DEFLATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result codes (orot include/orot/deflate.h:15-25). */
typedef enum deflate_result {
    DEFLATE_OK = 0,
    DEFLATE_STREAM_END = 1,
    DEFLATE_NEED_INPUT = 2,
    DEFLATE_NEED_OUTPUT = 3,
    DEFLATE_DATA_ERROR = -1,
    DEFLATE_MEM_ERROR = -2,
    DEFLATE_BUF_ERROR = -3,
    DEFLATE_VERSION_ERROR = -4,
    DEFLATE_PARAM_ERROR = -5
} deflate_result;

/* Framing (orot include/orot/deflate.h:28-32). */
typedef enum deflate_format {
    DEFLATE_FORMAT_RAW = 0,  /* RFC 1951 */
    DEFLATE_FORMAT_ZLIB = 1, /* RFC 1950, Adler-32 trailer */
    DEFLATE_FORMAT_GZIP = 2  /* RFC 1952, CRC-32 + ISIZE trailer */
} deflate_format;

/* Flush modes (orot include/orot/deflate.h:42-47). */
typedef enum deflate_flush {
    DEFLATE_NO_FLUSH = 0,   /* collect input */
    DEFLATE_SYNC_FLUSH = 1, /* all input so far decodable, output byte-aligned */
    DEFLATE_FULL_FLUSH = 2, /* as SYNC_FLUSH, and later matches do not reach back past it */
    DEFLATE_FINISH = 3      /* final block and trailer */
} deflate_flush;

/* Custom allocator (orot include/orot/deflate.h:49-54). */
typedef struct deflate_allocator {
    void* (*alloc)(void* opaque, size_t size, size_t alignment);
    void (*free)(void* opaque, void* ptr);
    void* opaque;
} deflate_allocator;

/* Levels (orot include/orot/deflate.h:35-39); values are clamped to 0..12. */
#define DEFLATE_LEVEL_STORE 0
#define DEFLATE_LEVEL_FAST 1
#define DEFLATE_LEVEL_DEFAULT 6
#define DEFLATE_LEVEL_BETTER 9
#define DEFLATE_LEVEL_MAX 12

/* orot include/orot/deflate.h:61, src/api/deflate_api.cpp:75-77: a non-NULL
 * allocator (with both functions set) replaces the current one; NULL is
 * ignored.  Set it before first use.  It serves the host memory of the
 * decompress side (inflate_stream_new, deflate_decompress); device memory and
 * pinned staging come from the HIP runtime. */
void deflate_set_allocator(const deflate_allocator* alloc);

/* Upper bound of the compressed size (orot include/orot/deflate.h:66,
 * src/api/deflate_api.cpp:81-98): n + n/4096 + n/16384 + n/2^25 + 13, plus 6
 * (zlib) or 18 (gzip). */
size_t deflate_compress_bound(size_t in_size, deflate_format format);

/* Whole-buffer compression (orot include/orot/deflate.h:69-77,
 * src/api/deflate_api.cpp:100-136, src/compress/block_compressor.cpp:16-73).
 * Returns 0 when out_capacity is below deflate_compress_bound(in_size,
 * format), as orot does; the gzip header's XFL byte is 2 for levels >= 9, 4
 * for levels <= 2, else 0 (src/formats/gzip_wrapper.cpp:31).  Levels 0-3 give
 * orot's bytes: the whole input is one block (stored blocks at level 0, and
 * an input with no match at all is stored), for inputs of up to 1 GiB.
 * Levels 4-12 give the L6-class stream of deflate_parallel_compress with the
 * level's default block size. */
size_t deflate_compress(const void* in, size_t in_size, void* out, size_t out_capacity, int level,
                        deflate_format format);

/* Whole-buffer inflate (orot include/orot/deflate.h:79-89,
 * src/api/deflate_api.cpp:138-215 with the framing checks of
 * src/formats/zlib_wrapper.cpp and gzip_wrapper.cpp:76-138).  DEFLATE_OK with
 * *actual_out_size set; DEFLATE_DATA_ERROR for a corrupt or truncated stream,
 * a bad header or a checksum / ISIZE mismatch; DEFLATE_NEED_OUTPUT when the
 * output does not fit out_capacity; DEFLATE_PARAM_ERROR for bad arguments;
 * DEFLATE_MEM_ERROR when host memory runs out.  Served by the host inflater;
 * bytes after the end of the stream are ignored. */
deflate_result deflate_decompress(const void* in, size_t in_size, void* out, size_t out_capacity,
                                  size_t* actual_out_size, deflate_format format);

/* Parallel compressor (orot include/orot/deflate.h:120-145,
 * src/api/stream_api.cpp:384-406, src/parallel/parallel_compressor.cpp:22-28).
 * block_size 0 selects 128 KiB for levels <= 3, 512 KiB for levels <= 6 and
 * 1 MiB above; any block size up to 1 GiB is accepted (levels 0-3 give orot's
 * bytes at every one).  Returns NULL when no GPU is usable, on allocation
 * failure or for a block size above 1 GiB. */
typedef struct deflate_parallel_ctx deflate_parallel_ctx;

deflate_parallel_ctx* deflate_parallel_new(int level, deflate_format format, int num_threads, size_t block_size);
void deflate_parallel_free(deflate_parallel_ctx* ctx);

/* Compress in[0..in_size) (host memory) into out[0..out_capacity).  Blocks
 * until done.  Returns the compressed size or 0 on failure. */
size_t deflate_parallel_compress(deflate_parallel_ctx* ctx, const void* in, size_t in_size, void* out,
                                 size_t out_capacity);

/* Streaming compressor (orot include/orot/deflate.h:92-108,
 * src/api/stream_api.cpp:134-247, src/compress/compressor.cpp:21-97).
 * Input is collected in pinned host memory and compressed on the GPU in
 * pieces of whole chunks (32 MiB), each piece's bits continuing the last;
 * the next piece fills while the GPU compresses the previous one, and host
 * copies of 1 MiB or more are split over three helper threads that the
 * library starts on first use.
 * deflate_stream_free keeps up to two freed streams (about 200 MiB of pinned
 * and device memory each) for reuse by deflate_stream_new with the same
 * level and format, until the process ends.
 * Fed with DEFLATE_NO_FLUSH up to DEFLATE_FINISH, the stream is byte for byte
 * what deflate_parallel_compress makes of the concatenated input with the
 * level's default block size; for 1000 bytes to 128 KiB of input that is
 * also orot's own stream.  SYNC and FULL flushes compress what is pending and
 * append an empty stored block (00 00 FF FF); a FULL flush also drops the
 * dictionary of levels >= 4.  Results: DEFLATE_OK (input consumed, output
 * drained), DEFLATE_NEED_OUTPUT (call again with more output space),
 * DEFLATE_STREAM_END (after DEFLATE_FINISH, once everything is out),
 * DEFLATE_PARAM_ERROR (NULL pointers, unknown flush), DEFLATE_MEM_ERROR (a
 * GPU failure; the stream is unusable afterwards).  deflate_stream_new
 * returns NULL when no GPU is usable. */
typedef struct deflate_stream deflate_stream;

deflate_stream* deflate_stream_new(int level, deflate_format format);
void deflate_stream_free(deflate_stream* s);
deflate_result deflate_stream_compress(deflate_stream* s, const uint8_t** next_in, size_t* avail_in,
                                       uint8_t** next_out, size_t* avail_out, deflate_flush flush);

/* Streaming decompressor (orot include/orot/deflate.h:111-118,
 * src/api/stream_api.cpp:253-380), on the host.  inflate_stream_new returns
 * NULL on allocation failure or an unknown format.  inflate_stream_decompress
 * takes all of the input it is given (held until decoded) and returns
 * DEFLATE_OK (more input needed), DEFLATE_NEED_OUTPUT (call again with more
 * output space), DEFLATE_STREAM_END (the stream and its trailer are
 * complete; bytes of this call after the end stay in *next_in and
 * *avail_in), DEFLATE_DATA_ERROR (corrupt data, a bad header or a checksum /
 * ISIZE mismatch; the stream stays failed) or DEFLATE_PARAM_ERROR (NULL
 * pointers, or a compressor stream).  Unlike orot's inflater, a stored
 * block that follows a Huffman block decodes correctly. */
deflate_stream* inflate_stream_new(deflate_format format);
void inflate_stream_free(deflate_stream* s);
deflate_result inflate_stream_decompress(deflate_stream* s, const uint8_t** next_in, size_t* avail_in,
                                         uint8_t** next_out, size_t* avail_out);

/* Checksums (orot include/orot/deflate.h:149-150). */
uint32_t deflate_adler32(uint32_t initial, const void* data, size_t len);
uint32_t deflate_crc32(uint32_t initial, const void* data, size_t len);

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
#include <span>
#include <stdexcept>
#include <vector>

namespace orot {
namespace deflate {

/* C++ wrappers with the names and behaviour of orot's header-only wrappers
 * (orot include/orot/deflate.h:188-389). */
enum class Format : int { Raw = DEFLATE_FORMAT_RAW, Zlib = DEFLATE_FORMAT_ZLIB, Gzip = DEFLATE_FORMAT_GZIP };

struct Level {
    static constexpr int Store = DEFLATE_LEVEL_STORE;
    static constexpr int Fast = DEFLATE_LEVEL_FAST;
    static constexpr int Default = DEFLATE_LEVEL_DEFAULT;
    static constexpr int Better = DEFLATE_LEVEL_BETTER;
    static constexpr int Max = DEFLATE_LEVEL_MAX;
};

class Error : public std::runtime_error {
public:
    Error(deflate_result code, const char* what) : std::runtime_error(what), code_(code) {}
    deflate_result code() const noexcept { return code_; }

private:
    deflate_result code_;
};

class ParallelCompressor {
public:
    explicit ParallelCompressor(int level = Level::Default, Format fmt = Format::Gzip, int threads = 0,
                                size_t block_size = 0)
        : ctx_(deflate_parallel_new(level, static_cast<deflate_format>(fmt), threads, block_size)) {
        if (!ctx_) throw Error(DEFLATE_MEM_ERROR, "ParallelCompressor: allocation failed");
    }
    ~ParallelCompressor() { deflate_parallel_free(ctx_); }
    ParallelCompressor(const ParallelCompressor&) = delete;
    ParallelCompressor& operator=(const ParallelCompressor&) = delete;

    /* Output sized with the gzip bound whatever the format, as orot does. */
    std::vector<uint8_t> compress(std::span<const uint8_t> input) {
        std::vector<uint8_t> out(deflate_compress_bound(input.size(), DEFLATE_FORMAT_GZIP));
        const size_t n = deflate_parallel_compress(ctx_, input.data(), input.size(), out.data(), out.size());
        if (n == 0) throw Error(DEFLATE_BUF_ERROR, "ParallelCompressor::compress: buffer overflow");
        out.resize(n);
        return out;
    }

private:
    deflate_parallel_ctx* ctx_;
};

/* orot include/orot/deflate.h:258-304: feed() with DEFLATE_NO_FLUSH, finish()
 * with DEFLATE_FINISH; each returns the bytes written to out. */
class Compressor {
public:
    explicit Compressor(int level = Level::Default, Format fmt = Format::Zlib)
        : s_(deflate_stream_new(level, static_cast<deflate_format>(fmt))) {
        if (!s_) throw Error(DEFLATE_MEM_ERROR, "Compressor: allocation failed");
    }
    ~Compressor() { deflate_stream_free(s_); }
    Compressor(const Compressor&) = delete;
    Compressor& operator=(const Compressor&) = delete;

    size_t feed(std::span<const uint8_t> in, std::span<uint8_t> out) {
        const uint8_t* next_in = in.data();
        size_t avail_in = in.size();
        uint8_t* next_out = out.data();
        size_t avail_out = out.size();
        const deflate_result r =
            deflate_stream_compress(s_, &next_in, &avail_in, &next_out, &avail_out, DEFLATE_NO_FLUSH);
        if (r < 0) throw Error(r, "Compressor::feed failed");
        return out.size() - avail_out;
    }

    size_t finish(std::span<uint8_t> out) {
        const uint8_t* next_in = nullptr;
        size_t avail_in = 0;
        uint8_t* next_out = out.data();
        size_t avail_out = out.size();
        const deflate_result r =
            deflate_stream_compress(s_, &next_in, &avail_in, &next_out, &avail_out, DEFLATE_FINISH);
        if (r != DEFLATE_OK && r != DEFLATE_STREAM_END) throw Error(r, "Compressor::finish failed");
        return out.size() - avail_out;
    }

private:
    deflate_stream* s_;
};

inline uint32_t adler32(std::span<const uint8_t> data, uint32_t init = 1) {
    return deflate_adler32(init, data.data(), data.size());
}
inline uint32_t crc32(std::span<const uint8_t> data, uint32_t init = 0) {
    return deflate_crc32(init, data.data(), data.size());
}

/* orot include/orot/deflate.h:219-236: the whole-buffer bound, then one call. */
inline std::vector<uint8_t> compress(std::span<const uint8_t> input, int level = Level::Default,
                                     Format format = Format::Zlib) {
    std::vector<uint8_t> out(deflate_compress_bound(input.size(), static_cast<deflate_format>(format)));
    const size_t n = deflate_compress(input.data(), input.size(), out.data(), out.size(), level,
                                      static_cast<deflate_format>(format));
    if (n == 0) throw Error(DEFLATE_BUF_ERROR, "deflate::compress: output buffer overflow");
    out.resize(n);
    return out;
}

/* orot include/orot/deflate.h:308-342: feed() returns the bytes written and
 * sets done at the end of the stream. */
class Decompressor {
public:
    explicit Decompressor(Format fmt = Format::Zlib) : s_(inflate_stream_new(static_cast<deflate_format>(fmt))) {
        if (!s_) throw Error(DEFLATE_MEM_ERROR, "Decompressor: allocation failed");
    }
    ~Decompressor() { inflate_stream_free(s_); }
    Decompressor(const Decompressor&) = delete;
    Decompressor& operator=(const Decompressor&) = delete;

    size_t feed(std::span<const uint8_t> in, std::span<uint8_t> out, bool& done) {
        const uint8_t* next_in = in.data();
        size_t avail_in = in.size();
        uint8_t* next_out = out.data();
        size_t avail_out = out.size();
        const deflate_result r = inflate_stream_decompress(s_, &next_in, &avail_in, &next_out, &avail_out);
        if (r == DEFLATE_STREAM_END) {
            done = true;
        } else if (r < 0) {
            throw Error(r, "Decompressor::feed failed");
        } else {
            done = false;
        }
        return out.size() - avail_out;
    }

private:
    deflate_stream* s_;
};

/* orot include/orot/deflate.h:238-256: one call into a max_output buffer. */
inline std::vector<uint8_t> decompress(std::span<const uint8_t> input, Format format = Format::Zlib,
                                       size_t max_output = 256ULL * 1024 * 1024) {
    std::vector<uint8_t> out(max_output);
    size_t actual = 0;
    const deflate_result r = deflate_decompress(input.data(), input.size(), out.data(), out.size(), &actual,
                                                static_cast<deflate_format>(format));
    if (r != DEFLATE_OK) throw Error(r, "deflate::decompress failed");
    out.resize(actual);
    return out;
}

}  // namespace deflate
}  // namespace orot
#endif /* __cplusplus */

#endif /* OROT_DEFLATE_H */
