// stream_api.cpp -- orot's streaming compressor (include/orot/deflate.h:92-108,
// src/api/stream_api.cpp:134-247 over src/compress/compressor.cpp:21-97) on
// the GPU pipeline.
//
// Input collects in pinned host memory.  When a full piece of PIECE_TARGET bytes
// (a whole number of chunks) is held and more input arrives, or on a flush,
// the pending bytes go to the device and compress as one buffer in stream
// piece mode: raw bitstream, no framing, bits continuing at the stream's
// current bit offset, and the last block final only at DEFLATE_FINISH.
// Two input buffers alternate: the next piece fills while the GPU compresses
// the last one, which is collected (its bit offset and checksum) before the
// next launch.
// Pieces are cut at chunk boundaries of the whole stream, so a stream fed
// without flushes is, byte for byte, what deflate_parallel_compress makes of
// the concatenated input (the L6 class gets the 32 KiB before each piece as
// its dictionary, as a chunk inside one buffer does).
//
// orot's Compressor accumulates 128 KiB blocks with a BitWriter flushed to a
// byte after each one and no marker, so its streams are valid only while the
// whole input fits one block and no flush is requested; a sync or full flush
// here appends the empty stored block of RFC 1951 3.2.4 (00 00 FF FF), as
// zlib's Z_SYNC_FLUSH does, and the stream stays valid across any number of
// pieces and flushes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/orot/deflate.h"
#include "deflate_gpu.h"
#include "engine.h"
#include "host_common.h"
#include "inflate_host.h"

using orot_gpu::BufferSpec;
using orot_gpu::Engine;
using orot_gpu::EngineConfig;
using orot_gpu::set_error;
using orot_gpu::set_error_msg;

namespace {

constexpr size_t HIST = 32768;              // dictionary bytes kept before a piece
constexpr size_t PIECE_TARGET = 32u << 20;  // bytes per GPU piece (rounded to whole chunks)
constexpr uint32_t MOD_ADLER = 65521u;

// CRC-32 of A||B from crc(A), crc(B) and |B|: crc(A) * x^(8|B|) + crc(B) over
// GF(2) modulo the reflected polynomial 0xEDB88320 (bit 31 = x^0).
uint32_t gf2_mul_host(uint32_t a, uint32_t b) {
    uint32_t p = 0;
    for (int i = 0; i < 32; ++i) {
        if (a & 0x80000000u) p ^= b;
        a <<= 1;
        b = (b >> 1) ^ (0xEDB88320u & (0u - (b & 1u)));
    }
    return p;
}
uint32_t crc32_concat(uint32_t ca, uint32_t cb, uint64_t nb) {
    uint32_t r = 0x80000000u, x = 0x00800000u;  // x^0, x^8
    for (; nb; nb >>= 1, x = gf2_mul_host(x, x))
        if (nb & 1) r = gf2_mul_host(x, r);
    return gf2_mul_host(r, ca) ^ cb;
}
// Adler-32 of A||B: s1 = s1a + s1b - 1, s2 = s2a + s2b + |B| (s1a - 1), mod 65521.
uint32_t adler32_concat(uint32_t aa, uint32_t ab, uint64_t nb) {
    const uint64_t m = MOD_ADLER, n = nb % m;
    const uint64_t s1a = aa & 0xFFFFu, s2a = aa >> 16, s1b = ab & 0xFFFFu, s2b = ab >> 16;
    const uint64_t s1 = (s1a + s1b + m - 1) % m;
    const uint64_t s2 = (s2a + s2b + n * ((s1a + m - 1) % m)) % m;
    return (uint32_t)((s2 << 16) | s1);
}

// Host copies of at least COPY_SPLIT bytes (input into a pinned piece, output
// out of the pinned staging) are split over the caller and COPY_THREADS
// helper threads: one thread's memcpy does not reach the host's memory
// bandwidth.  The helpers start on first use and live until the process
// ends; one split copy runs at a time.
constexpr size_t COPY_SPLIT = 1u << 20;
constexpr int COPY_THREADS = 3;

class CopyPool {
public:
    static CopyPool& get() {
        static auto* p = new CopyPool();
        return *p;
    }

    void copy(uint8_t* d, const uint8_t* s, size_t n) {
        if (n < COPY_SPLIT) {
            std::memcpy(d, s, n);
            return;
        }
        std::lock_guard<std::mutex> one(call_mu_);
        const size_t per = ((n + COPY_THREADS) / (COPY_THREADS + 1) + 63) & ~(size_t)63;
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = d;
            src_ = s;
            n_ = n;
            per_ = per;
            pending_ = COPY_THREADS;
            ++gen_;
        }
        cv_.notify_all();
        std::memcpy(d, s, std::min(n, per));  // part 0 on the caller
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool() {
        for (int k = 1; k <= COPY_THREADS; ++k) std::thread([this, k] { run(k); }).detach();
    }

    void run(int k) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            uint8_t* d = dst_;
            const uint8_t* s = src_;
            const size_t a = std::min(n_, (size_t)k * per_), b = std::min(n_, a + per_);
            lk.unlock();
            if (b > a) std::memcpy(d + a, s + a, b - a);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }

    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    uint64_t gen_ = 0;
    int pending_ = 0;
    uint8_t* dst_ = nullptr;
    const uint8_t* src_ = nullptr;
    size_t n_ = 0, per_ = 0;
};

}  // namespace

struct deflate_stream {
    // a streaming decompressor (inflate_stream_new) holds only this; the
    // engine and buffers below stay unused
    orot_gpu::HostInflater* inflater = nullptr;
    Engine eng;
    int level_arg = 0, format_arg = 0;  // as passed to deflate_stream_new (the reuse key)
    int format = 0;
    size_t piece_cap = 0;       // bytes per piece, a multiple of the chunk
    uint8_t* h_buf[2] = {};     // pinned: HIST dictionary bytes, then a piece
    uint8_t* h_in = nullptr;    // the buffer being filled
    size_t pend = 0;            // pending input bytes at h_in + HIST
    size_t hist = 0;            // stream bytes kept just before them (<= HIST)
    bool inflight = false;      // a piece is on the GPU (from the other buffer)
    size_t fly_n = 0;           // its input bytes
    uint8_t* d_in = nullptr;    // device copy of the same layout
    uint8_t* d_out = nullptr;
    size_t out_cap = 0;         // device output capacity of one piece
    uint8_t* h_out = nullptr;   // pinned staging of finished output bytes
    size_t out_len = 0, out_pos = 0;
    uint32_t held = 0, nheld = 0;  // bits past the last whole output byte
    uint64_t total_in = 0;
    uint64_t dict_in = 0;       // bytes since the stream start or the last full flush
    uint32_t check = 0;         // running Adler-32 (zlib) or CRC-32 (gzip)
    bool header = false;        // header queued
    bool unsynced = false;      // blocks written since the last flush marker
    bool ended = false;         // final block and trailer queued
    bool failed = false;

    ~deflate_stream() {
        orot_gpu::HostInflater::destroy(inflater);
        if (inflight) {
            uint64_t got = 0;
            (void)eng.wait(&got);
        }
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        for (uint8_t* b : h_buf)
            if (b) (void)hipHostFree(b);
        if (h_out) (void)hipHostFree(h_out);
    }

    // State of a new stream; the engine, buffers and arenas stay.
    void reset() {
        h_in = h_buf[0];
        pend = hist = 0;
        inflight = false;
        fly_n = 0;
        out_len = out_pos = 0;
        held = nheld = 0;
        total_in = dict_in = 0;
        check = format == DEFLATE_FORMAT_ZLIB ? 1u : 0u;
        header = unsynced = ended = failed = false;
    }

    void put_byte(uint8_t b) { h_out[out_len++] = b; }
    void put_bits(uint32_t v, uint32_t n) {  // LSB first, n <= 24
        held |= v << nheld;
        nheld += n;
        while (nheld >= 8) {
            put_byte((uint8_t)held);
            held >>= 8;
            nheld -= 8;
        }
    }
    void align() {
        if (nheld) put_bits(0, 8 - nheld);
    }

    // Starts the pending bytes as one piece (final unless `open`) on the GPU
    // and switches filling to the other buffer, whose dictionary is the last
    // HIST bytes since the start or the last full flush.  The previous piece
    // must have been collected: this one continues at its bit offset.
    bool launch(bool open) {
        const size_t n = pend;
        const hipStream_t st = eng.stream();
        hipError_t e = hipMemcpyAsync(d_in + HIST - hist, h_in + HIST - hist, hist + n, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            set_error("deflate_stream_compress: hipMemcpy(in)", e);
            return false;
        }
        BufferSpec b{HIST, n, 0, out_cap};
        b.piece = orot_gpu::PIECE | (open ? orot_gpu::PIECE_OPEN : 0u) | (held << 8) | nheld;
        b.hist = (uint32_t)hist;
        if (!eng.submit(d_in, d_out, &b, 1, false)) return false;
        inflight = true;
        fly_n = n;
        dict_in += n;
        const size_t keep = std::min<uint64_t>(HIST, dict_in);
        uint8_t* next = h_in == h_buf[0] ? h_buf[1] : h_buf[0];
        std::memcpy(next + HIST - keep, h_in + HIST + n - keep, keep);
        h_in = next;
        hist = keep;
        pend = 0;
        return true;
    }

    // Waits for the piece on the GPU and queues its whole bytes; the staging
    // buffer must be empty.
    bool collect() {
        uint64_t got = 0;
        inflight = false;
        if (!eng.wait(&got)) return false;
        if (got == 0) {
            set_error_msg("deflate_stream_compress: piece output exceeds its bound");
            return false;
        }
        const orot_gpu::Result R = eng.results()[0];
        const uint64_t whole = R.payload_bits >> 3;
        const hipStream_t st = eng.stream();
        hipError_t e = hipMemcpyAsync(h_out + out_len, d_out, got, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_error("deflate_stream_compress: hipMemcpy(out)", e);
            return false;
        }
        nheld = (uint32_t)(R.payload_bits & 7);
        held = nheld ? h_out[out_len + whole] & ((1u << nheld) - 1u) : 0u;
        out_len += whole;
        if (format == DEFLATE_FORMAT_ZLIB) check = adler32_concat(check, R.checksum, fly_n);
        if (format == DEFLATE_FORMAT_GZIP) check = crc32_concat(check, R.checksum, fly_n);
        total_in += fly_n;
        unsynced = true;
        return true;
    }

    void sync_marker() {  // empty stored block, BFINAL = 0
        put_bits(0, 3);
        align();
        put_byte(0x00);
        put_byte(0x00);
        put_byte(0xFF);
        put_byte(0xFF);
        unsynced = false;
    }

    bool finish() {  // no piece in flight
        if (!launch(false) || !collect()) return false;  // also with no pending bytes: the final (empty) block
        align();
        if (format == DEFLATE_FORMAT_ZLIB) {
            for (int k = 3; k >= 0; --k) put_byte((uint8_t)(check >> (8 * k)));
        } else if (format == DEFLATE_FORMAT_GZIP) {
            for (int k = 0; k < 4; ++k) put_byte((uint8_t)(check >> (8 * k)));
            for (int k = 0; k < 4; ++k) put_byte((uint8_t)(total_in >> (8 * k)));
        }
        ended = true;
        return true;
    }

    void drain(uint8_t** next_out, size_t* avail_out) {
        const size_t c = std::min(*avail_out, out_len - out_pos);
        if (c) CopyPool::get().copy(*next_out, h_out + out_pos, c);
        *next_out += c;
        *avail_out -= c;
        out_pos += c;
        if (out_pos == out_len) out_pos = out_len = 0;
    }
};

namespace {

// Freed streams kept for reuse by the next deflate_stream_new with the same
// level and format: a stream's pinned buffers are first touched on its first
// pieces, which costs as much as compressing a few of them.  At most
// STREAM_POOL are kept (about 100 MiB of pinned and 100 MiB of device memory
// each); the pool lives until the process ends.
constexpr size_t STREAM_POOL = 2;
std::mutex g_pool_mu;
std::vector<deflate_stream*>& pool() {
    static auto* p = new std::vector<deflate_stream*>();
    return *p;
}

}  // namespace

extern "C" {

deflate_stream* deflate_stream_new(int level, deflate_format format) {
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto& P = pool();
        for (size_t i = 0; i < P.size(); ++i) {
            if (P[i]->level_arg == level && P[i]->format_arg == (int)format) {
                deflate_stream* s = P[i];
                P.erase(P.begin() + (long)i);
                s->reset();
                return s;
            }
        }
    }
    auto* s = new (std::nothrow) deflate_stream;
    if (!s) return nullptr;
    s->level_arg = level;
    s->format_arg = (int)format;
    EngineConfig cfg;
    cfg.level = level;
    cfg.format = (int)format;
    if (!s->eng.init(cfg)) {
        delete s;
        return nullptr;
    }
    s->format = s->eng.config().format;
    const size_t chunk = s->eng.config().chunk;
    s->piece_cap = std::max<size_t>(1, PIECE_TARGET / chunk) * chunk;
    const size_t n = s->piece_cap, nch = n / chunk;
    s->out_cap = std::max(deflate_compress_bound(n, DEFLATE_FORMAT_RAW) + 64, n + 6 * (nch + n / 65535 + 2) + 64);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->d_in), HIST + n + 64);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_out), s->out_cap);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_buf[0]), HIST + n, 0);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_buf[1]), HIST + n, 0);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_out), s->out_cap + 64, 0);
    if (e != hipSuccess) {
        set_error("deflate_stream_new: allocation", e);
        delete s;
        return nullptr;
    }
    s->h_in = s->h_buf[0];
    s->check = s->format == DEFLATE_FORMAT_ZLIB ? 1u : 0u;
    return s;
}

void deflate_stream_free(deflate_stream* s) {
    if (!s) return;
    if (s->inflater) {
        delete s;
        return;
    }
    if (s->inflight) {
        uint64_t got = 0;
        s->inflight = false;
        if (!s->eng.wait(&got)) s->failed = true;
    }
    if (!s->failed) {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (pool().size() < STREAM_POOL) {
            pool().push_back(s);
            return;
        }
    }
    delete s;
}

deflate_result deflate_stream_compress(deflate_stream* s, const uint8_t** next_in, size_t* avail_in,
                                       uint8_t** next_out, size_t* avail_out, deflate_flush flush) {
    if (!s || s->inflater || !next_in || !avail_in || !next_out || !avail_out || (*avail_in && !*next_in) ||
        (*avail_out && !*next_out) || flush < DEFLATE_NO_FLUSH || flush > DEFLATE_FINISH)
        return DEFLATE_PARAM_ERROR;
    if (s->failed) return DEFLATE_MEM_ERROR;
    if (!s->header) {
        if (s->format == DEFLATE_FORMAT_ZLIB) {
            s->put_byte(0x78);
            s->put_byte(0x9C);
        } else if (s->format == DEFLATE_FORMAT_GZIP) {
            const uint8_t g[10] = {0x1F, 0x8B, 8, 0, 0, 0, 0, 0, 0, 255};
            for (uint8_t b : g) s->put_byte(b);
        }
        s->header = true;
    }
    for (;;) {
        s->drain(next_out, avail_out);
        if (s->out_len) return DEFLATE_NEED_OUTPUT;
        if (s->ended) return DEFLATE_STREAM_END;
        bool ok = true;
        if (*avail_in) {
            if (s->pend == s->piece_cap) {
                // full, and the stream goes on: the piece before it first
                ok = s->inflight ? s->collect() : s->launch(true);
            } else {
                const size_t c = std::min(*avail_in, s->piece_cap - s->pend);
                CopyPool::get().copy(s->h_in + HIST + s->pend, *next_in, c);
                s->pend += c;
                *next_in += c;
                *avail_in -= c;
                continue;
            }
        } else if (flush != DEFLATE_NO_FLUSH && s->inflight) {
            ok = s->collect();  // a flush or the end writes everything before it
        } else if (flush == DEFLATE_FINISH) {
            ok = s->finish();
        } else if ((flush == DEFLATE_SYNC_FLUSH || flush == DEFLATE_FULL_FLUSH) && (s->pend || s->unsynced)) {
            ok = !s->pend || (s->launch(true) && s->collect());
            if (ok) s->sync_marker();
            if (ok && flush == DEFLATE_FULL_FLUSH) s->hist = s->dict_in = 0;  // no matches reach back past it
        } else {
            // a full flush with nothing to write (right after a sync flush)
            // still ends the dictionary: the marker is already out
            if (flush == DEFLATE_FULL_FLUSH) s->hist = s->dict_in = 0;
            return DEFLATE_OK;
        }
        if (!ok) {
            s->failed = true;
            return DEFLATE_MEM_ERROR;
        }
    }
}

// ---- streaming decompressor (orot src/api/stream_api.cpp:253-380) ----------

deflate_stream* inflate_stream_new(deflate_format format) {
    if (format != DEFLATE_FORMAT_RAW && format != DEFLATE_FORMAT_ZLIB && format != DEFLATE_FORMAT_GZIP) {
        set_error_msg("inflate_stream_new: unknown format");
        return nullptr;
    }
    auto* s = new (std::nothrow) deflate_stream;
    if (!s) return nullptr;
    s->inflater = orot_gpu::HostInflater::create((int)format);
    if (!s->inflater) {
        delete s;
        set_error_msg("inflate_stream_new: allocation failed");
        return nullptr;
    }
    return s;
}

void inflate_stream_free(deflate_stream* s) { deflate_stream_free(s); }

deflate_result inflate_stream_decompress(deflate_stream* s, const uint8_t** next_in, size_t* avail_in,
                                         uint8_t** next_out, size_t* avail_out) {
    if (!s || !s->inflater || !next_in || !avail_in || !next_out || !avail_out || (*avail_in && !*next_in) ||
        (*avail_out && !*next_out))
        return DEFLATE_PARAM_ERROR;
    try {
        return s->inflater->stream(next_in, avail_in, next_out, avail_out);
    } catch (const std::bad_alloc&) {
        return DEFLATE_MEM_ERROR;
    }
}

}  // extern "C"
