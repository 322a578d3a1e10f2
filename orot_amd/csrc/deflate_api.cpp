// deflate_api.cpp -- C-ABI host layer (C++20) of liborot_amd.so.
//
// Implements orot's parallel compress entry points (src/api/stream_api.cpp:
// 380-406 -> src/parallel/parallel_compressor.cpp:22-150), the bound
// (src/api/deflate_api.cpp:81-98) and the checksum entry points
// (src/api/deflate_api.cpp:223-231) on top of the gfx950 kernels.  The
// context owns its HIP stream, its device arenas (grown on demand, kept across
// calls) and optional timing events.  There is no CPU compression path: a
// missing or failing GPU makes every entry point fail (NULL / 0) with a
// message in orot_gpu_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/orot/deflate.h"
#include "../../include/orot/deflate_gpu.h"
#include "deflate_gpu.h"

using orot_gpu::ChunkInfo;
using orot_gpu::Result;

namespace {

thread_local std::string g_last_error;

void set_error(const char* where, hipError_t e) {
    g_last_error = std::string(where) + ": " + hipGetErrorString(e);
}
void set_error_msg(const std::string& m) { g_last_error = m; }

size_t default_block_size(int level) {  // src/parallel/parallel_compressor.cpp:16-20
    if (level <= 3) return 131072;
    if (level <= 6) return 524288;
    return 1048576;
}

int clamp_level(int level) { return level < 0 ? 0 : (level > 12 ? 12 : level); }

// Result of the start-up lane-order check of LDS exchanges: -1 not run yet,
// 0 failed (the L1 kernels use their ballot variant), 1 passed.
std::mutex g_exch_mu;
int g_exch_state = -1;

bool exchange_order_ok(hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_exch_mu);
    if (g_exch_state < 0) {
        bool ok = false;
        hipError_t e = orot_gpu::check_exchange_order(stream, &ok);
        if (e != hipSuccess) return false;
        g_exch_state = ok ? 1 : 0;
    }
    return g_exch_state == 1;
}

template <class T>
bool grow(T*& p, size_t& cap, size_t need, const char* what) {
    if (need <= cap && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t bytes = std::max<size_t>(need, 1) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e != hipSuccess) {
        p = nullptr;
        set_error(what, e);
        return false;
    }
    cap = std::max<size_t>(need, 1);
    return true;
}

}  // namespace

struct deflate_parallel_ctx {
    int level = 1;
    int format = 0;
    size_t block = 131072;
    hipStream_t stream = nullptr;
    uint8_t* d_in = nullptr;
    size_t in_cap = 0;
    uint8_t* d_out = nullptr;
    size_t out_cap = 0;
    uint32_t* d_tok = nullptr;
    size_t tok_cap = 0;
    uint16_t* d_pd = nullptr;
    size_t pd_cap = 0;
    uint32_t* d_aux = nullptr;
    size_t aux_cap = 0;
    uint32_t* d_ntok = nullptr;
    size_t ntok_cap = 0;
    uint32_t* d_blkpos = nullptr;
    size_t blkpos_cap = 0;
    uint8_t* d_blocks = nullptr;
    size_t blocks_cap = 0;
    uint32_t* d_stage = nullptr;
    size_t stage_cap = 0;
    ChunkInfo* d_info = nullptr;
    size_t info_cap = 0;
    Result* d_res = nullptr;
    size_t res_cap = 0;
    bool timing = false;
    hipEvent_t ev[8] = {};
    float stage_ms[4] = {};
    int nstage = 0;
    bool exch_ok = false;
    size_t lds_max = 65536;
    std::mutex mu;

    ~deflate_parallel_ctx() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {(void*)d_in, (void*)d_out, (void*)d_tok, (void*)d_pd, (void*)d_aux, (void*)d_ntok, (void*)d_blkpos, (void*)d_blocks, (void*)d_stage, (void*)d_info,
                        (void*)d_res})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Runs the pipeline on device buffers; returns the framed size or 0.
    size_t run(const uint8_t* in, size_t n, uint8_t* out, size_t cap) {
        const int lv = clamp_level(level);
        if (lv >= 4 && block > orot_gpu::l6_max_chunk()) {
            set_error_msg("deflate_parallel_compress: block_size " + std::to_string(block) +
                          " exceeds the GPU limit " + std::to_string(orot_gpu::l6_max_chunk()) + " for this level");
            return 0;
        }
        if (lv >= 1 && lv <= 3) {
            const uint32_t maxc = orot_gpu::max_chunk_for_hash_bits(lv == 1 ? 12 : 14);
            if (block > maxc) {
                set_error_msg("deflate_parallel_compress: block_size " + std::to_string(block) +
                              " exceeds the GPU limit " + std::to_string(maxc) + " for this level");
                return 0;
            }
        }
        if (block > 0xFFFFFFFFull) {
            set_error_msg("deflate_parallel_compress: block_size too large");
            return 0;
        }
        const uint32_t chunk = (uint32_t)block;
        const uint64_t nb64 = n ? (n + chunk - 1) / chunk : 1;
        if (nb64 > 0x7FFFFFFF) {
            set_error_msg("deflate_parallel_compress: too many chunks");
            return 0;
        }
        const uint32_t nb = (uint32_t)nb64;
        const uint64_t sw = orot_gpu::stage_words_for_chunk(chunk);
        const size_t padded = (size_t)nb * ((chunk + 7u) & ~7u) + 8;
        if (lv >= 4) {
            const size_t mb = orot_gpu::l6_max_blocks(chunk);
            if (!grow(d_pd, pd_cap, padded, "hipMalloc(pd)") || !grow(d_aux, aux_cap, padded, "hipMalloc(tokens)") ||
                !grow(d_ntok, ntok_cap, nb, "hipMalloc(ntok)") ||
                !grow(d_blkpos, blkpos_cap, (size_t)nb * mb, "hipMalloc(blkpos)") ||
                !grow(d_blocks, blocks_cap, (size_t)nb * mb * orot_gpu::l6_block_record_bytes(), "hipMalloc(blocks)"))
                return 0;
        }
        if (!grow(d_tok, tok_cap, padded, "hipMalloc(tok)") ||
            !grow(d_stage, stage_cap, (size_t)(sw * nb), "hipMalloc(stage)") ||
            !grow(d_info, info_cap, nb, "hipMalloc(info)") || !grow(d_res, res_cap, 1, "hipMalloc(result)"))
            return 0;
        if (timing && !ev[0]) {
            for (auto& e : ev) {
                hipError_t r = hipEventCreate(&e);
                if (r != hipSuccess) {
                    set_error("hipEventCreate", r);
                    return 0;
                }
            }
        }
        orot_gpu::Plan p{};
        p.d_in = in;
        p.n = n;
        p.chunk = chunk;
        p.nchunks = nb;
        p.level = lv;
        p.format = format;
        p.d_out = out;
        p.cap = cap;
        p.d_tok = d_tok;
        p.d_pd = d_pd;
        p.d_aux = d_aux;
        p.d_ntok = d_ntok;
        p.d_blkpos = d_blkpos;
        p.d_blocks = d_blocks;
        p.d_stage = d_stage;
        p.stage_words = sw;
        p.d_info = d_info;
        p.d_result = d_res;
        p.stream = stream;
        p.exch_ok = exch_ok;
        p.lds_max = lds_max;
        for (int i = 0; i < 8; ++i) p.ev[i] = timing ? ev[i] : nullptr;
        hipError_t e = orot_gpu::launch_parallel_compress(p);
        if (e != hipSuccess) {
            set_error("launch_parallel_compress", e);
            return 0;
        }
        Result r{};
        e = hipMemcpyAsync(&r, d_res, sizeof(Result), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            set_error("deflate_parallel_compress", e);
            return 0;
        }
        nstage = 0;
        if (timing) {
            for (int i = 0; i < 4; ++i) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
                stage_ms[i] = ms;
            }
            nstage = 4;
        }
        if (r.total_bytes == 0) set_error_msg("deflate_parallel_compress: output does not fit out_capacity");
        return (size_t)r.total_bytes;
    }
};

extern "C" {

size_t deflate_compress_bound(size_t n, deflate_format format) {
    size_t b = n + (n >> 12) + (n >> 14) + (n >> 25) + 13;
    if (format == DEFLATE_FORMAT_ZLIB) b += 6;
    if (format == DEFLATE_FORMAT_GZIP) b += 18;
    return b;
}

deflate_parallel_ctx* deflate_parallel_new(int level, deflate_format format, int num_threads, size_t block_size) {
    (void)num_threads;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        set_error_msg("deflate_parallel_new: no usable GPU");
        return nullptr;
    }
    auto* ctx = new (std::nothrow) deflate_parallel_ctx;
    if (!ctx) return nullptr;
    ctx->level = level;
    ctx->format = (format == DEFLATE_FORMAT_ZLIB || format == DEFLATE_FORMAT_GZIP) ? (int)format : 0;
    ctx->block = block_size ? block_size : default_block_size(level);
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error("hipStreamCreate", e);
        ctx->stream = nullptr;
        delete ctx;
        return nullptr;
    }
    ctx->exch_ok = exchange_order_ok(ctx->stream);
    {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && v > 0)
            ctx->lds_max = (size_t)v;
    }
    return ctx;
}

void deflate_parallel_free(deflate_parallel_ctx* ctx) { delete ctx; }

size_t deflate_parallel_compress(deflate_parallel_ctx* ctx, const void* in, size_t n, void* out, size_t cap) {
    if (!ctx) return 0;
    std::lock_guard<std::mutex> lk(ctx->mu);
    // orot checks only the empty-input bound and the framing bytes
    // (parallel_compressor.cpp:40-45, 91-108); this path also refuses to
    // overflow the payload.
    // Arena for the device output: the orot bound, or the stored-block worst
    // case when small chunks make per-block headers exceed it.
    const size_t nchunks = n ? (n + ctx->block - 1) / ctx->block : 1;
    const size_t bound = std::max(deflate_compress_bound(n, (deflate_format)ctx->format) + 64,
                                  n + 6 * (nchunks + n / 65535 + 2) + 64);
    if (!grow(ctx->d_in, ctx->in_cap, n, "hipMalloc(in)") || !grow(ctx->d_out, ctx->out_cap, bound, "hipMalloc(out)"))
        return 0;
    if (n) {
        hipError_t e = hipMemcpyAsync(ctx->d_in, in, n, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            set_error("hipMemcpy(in)", e);
            return 0;
        }
    }
    const size_t got = ctx->run(ctx->d_in, n, ctx->d_out, std::min(cap, bound));
    if (!got) return 0;
    hipError_t e = hipMemcpyAsync(out, ctx->d_out, got, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_error("hipMemcpy(out)", e);
        return 0;
    }
    return got;
}

size_t deflate_parallel_compress_device(deflate_parallel_ctx* ctx, const void* d_in, size_t n, void* d_out,
                                        size_t cap) {
    if (!ctx) return 0;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return ctx->run(static_cast<const uint8_t*>(d_in), n, static_cast<uint8_t*>(d_out), cap);
}

void deflate_parallel_set_timing(deflate_parallel_ctx* ctx, int on) {
    if (!ctx) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->timing = on != 0;
}

int deflate_parallel_stage_times(const deflate_parallel_ctx* ctx, float* ms, int max_values) {
    if (!ctx || !ms) return 0;
    const int k = std::min(ctx->nstage, max_values);
    for (int i = 0; i < k; ++i) ms[i] = ctx->stage_ms[i];
    return k;
}

}  // extern "C"

namespace {

// Process-wide state for the whole-buffer checksum entry points.
struct ChecksumState {
    std::mutex mu;
    hipStream_t stream = nullptr;
    uint8_t* d_in = nullptr;
    size_t in_cap = 0;
    ChunkInfo* d_info = nullptr;
    size_t info_cap = 0;
    uint32_t* d_val = nullptr;
    size_t val_cap = 0;
    bool init() {
        if (stream) return true;
        hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error("hipStreamCreate", e);
            stream = nullptr;
            return false;
        }
        return true;
    }
};
ChecksumState& cks() {
    static ChecksumState s;
    return s;
}

uint32_t gpu_checksum(int kind, uint32_t init, const void* data, size_t len) {
    if (len == 0) return init;  // both checksums are the identity on empty input
    auto& s = cks();
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.init()) return 0;
    const uint32_t nb = (uint32_t)((len + 65535) / 65536);
    if (!grow(s.d_in, s.in_cap, len, "hipMalloc(checksum)") || !grow(s.d_info, s.info_cap, nb, "hipMalloc") ||
        !grow(s.d_val, s.val_cap, 1, "hipMalloc"))
        return 0;
    hipError_t e = hipMemcpyAsync(s.d_in, data, len, hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = orot_gpu::launch_checksum(s.d_in, len, kind, init, s.d_val, s.d_info, nb, s.stream);
    uint32_t v = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&v, s.d_val, 4, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) {
        set_error("checksum", e);
        return 0;
    }
    return v;
}

}  // namespace

extern "C" {

uint32_t deflate_crc32(uint32_t initial, const void* data, size_t len) { return gpu_checksum(0, initial, data, len); }

uint32_t deflate_adler32(uint32_t initial, const void* data, size_t len) {
    return gpu_checksum(1, initial, data, len);
}

void* orot_gpu_malloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) {
        set_error("hipMalloc", e);
        return nullptr;
    }
    return p;
}

int orot_gpu_free(void* p) {
    hipError_t e = hipFree(p);
    if (e != hipSuccess) set_error("hipFree", e);
    return (int)e;
}

int orot_gpu_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    const hipMemcpyKind k =
        kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    hipError_t e = hipMemcpy(dst, src, bytes, k);
    if (e != hipSuccess) set_error("hipMemcpy", e);
    return (int)e;
}

int orot_gpu_memset(void* dst, int value, size_t bytes) {
    hipError_t e = hipMemset(dst, value, bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) set_error("hipMemset", e);
    return (int)e;
}

int orot_gpu_synchronize(void) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) set_error("hipDeviceSynchronize", e);
    return (int)e;
}

int orot_gpu_device_info(char* name, size_t name_len, int* compute_units) {
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, 0);
    if (e != hipSuccess) {
        set_error("hipGetDeviceProperties", e);
        return (int)e;
    }
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return 0;
}

const char* orot_gpu_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
