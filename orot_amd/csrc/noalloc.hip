// noalloc.hip -- deflate_parallel_compress_device_noalloc: the single-buffer
// device pipeline of deflate_parallel_compress_device, run out of a
// file-scope device array instead of engine arenas, for callers that cannot
// allocate device memory (code under graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The array `scratch` holds the batch tables of one buffer, every per-chunk
// working array of the kernels and the result record.  Each GPU the library
// runs on holds its own copy until the process ends.  Its device address is
// taken once per device with hipGetSymbolAddress (the array's name in host
// code is only a host-side stand-in) and only that address is handed to
// kernels and copies.  Calls on one device share the array: a mutex per
// device serializes them inside this process, and callers must not overlap
// calls on one GPU from elsewhere.  The tables are written by a kernel, so no
// host staging (pinned or pageable) is needed; the only host transfer is the
// 24-byte result record.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/orot/deflate.h"
#include "../../include/orot/deflate_gpu.h"
#include "deflate_gpu.h"
#include "engine.h"
#include "host_common.h"

namespace orot_gpu {

hipError_t set_kernel_attributes_l1(size_t lds_max);
hipError_t set_kernel_attributes_l6(size_t lds_max);

namespace {

// 64 MiB: the working set of a 4 MiB input at every level with the level's
// default block size (about 7.0 B per input byte at levels 1-3 and 11.7 B at
// levels 4-12, so 8 MiB fits at levels 0-3).
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr size_t ALIGN = 256;
inline size_t align_up(size_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

// Writes the batch tables of one buffer of n bytes cut into `nchunks` chunks:
// header (queue heads zero), BufDesc[0], ChunkDesc[i] and the queue order.
// Only the last chunk can be short, so the order "full chunks first, then
// the rest longest first" is the index order.
__global__ __launch_bounds__(256) void single_buffer_tables_kernel(BatchHdr* __restrict__ hdr, BufDesc* __restrict__ bd,
                                                                   ChunkDesc* __restrict__ cd,
                                                                   uint32_t* __restrict__ order, uint64_t n,
                                                                   uint64_t cap, uint32_t chunk, uint32_t nchunks) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        BatchHdr h{};
        h.nchunks = nchunks;
        h.nbufs = 1;
        *hdr = h;
        BufDesc b{};
        b.in_off = 0;
        b.in_len = n;
        b.out_off = 0;
        b.cap = cap;
        b.chunk0 = 0;
        b.nchunks = nchunks;
        b.piece = 0;
        b.hist = 0;
        *bd = b;
    }
    for (uint32_t i = tid; i < nchunks; i += gridDim.x * blockDim.x) {
        const uint64_t off = (uint64_t)i * chunk;
        ChunkDesc c;
        c.in_off = off;
        c.len = (uint32_t)(n - off < chunk ? n - off : chunk);
        c.buf = i + 1 == nchunks ? CHUNK_LAST : 0u;
        cd[i] = c;
        order[i] = i;
    }
}

// Byte offsets of the arrays of one call inside the scratch array.
struct Layout {
    size_t hdr, bufs, desc, order, tok, pd, aux, ntok, blkpos, blocks, stage, info, res, total;
};

Layout layout_for(const EngineConfig& cfg, uint32_t nchunks) {
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    const bool l1 = cfg.level >= 1, l6 = cfg.level >= 4;
    const bool big = l1 && !l6 && l1_big_chunk(cfg.chunk, cfg.hash_bits);
    const size_t per_pos =
        (size_t)nchunks *
            (l6 ? l6_position_stride(cfg.chunk) : (big ? big_position_stride(cfg.chunk) : l1_position_stride(cfg.chunk))) +
        8;
    const size_t pd = big ? (size_t)nchunks * big_bitmap_words(cfg.chunk) * 4 + 16 : per_pos * sizeof(uint16_t);
    const size_t mb = l6 ? l6_max_blocks(cfg.chunk) : 0;
    L.hdr = take(sizeof(BatchHdr));
    L.bufs = take(sizeof(BufDesc));
    L.desc = take((size_t)nchunks * sizeof(ChunkDesc));
    L.order = take((size_t)nchunks * sizeof(uint32_t));
    L.tok = l1 ? take(per_pos * sizeof(uint32_t)) : 0;
    L.pd = l1 ? take(pd) : 0;
    L.stage = l1 ? take((size_t)nchunks * stage_words_for_chunk(cfg.chunk) * sizeof(uint32_t)) : 0;
    L.aux = (l6 || big) ? take(per_pos * sizeof(uint32_t)) : 0;
    L.ntok = l6 ? take((size_t)nchunks * sizeof(uint32_t)) : 0;
    L.blkpos = l6 ? take((size_t)nchunks * mb * sizeof(uint32_t)) : 0;
    L.blocks = l6    ? take((size_t)nchunks * mb * l6_block_record_bytes())
               : big ? take((size_t)nchunks * big_segtab_words(cfg.chunk, cfg.hash_bits) * 4)
                     : 0;
    L.info = take((size_t)nchunks * sizeof(ChunkInfo));
    L.res = take(sizeof(Result));
    L.total = o;
    return L;
}

// Per-device state: the device address of `scratch` and the device facts the
// launches need, set up once with that device current.
struct DeviceScratch {
    std::mutex mu;  // one call at a time on this device's array
    bool ready = false;
    uint8_t* ptr = nullptr;
    bool exch_ok = false;
    int cus = 256;
    size_t lds_max = 65536;
};

constexpr int MAX_DEVICES = 64;
DeviceScratch g_dev[MAX_DEVICES];
std::mutex g_setup_mu;

// Returns the current device's state, set up on first use; nullptr on
// failure (message in the last error).
DeviceScratch* device_scratch() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("deflate_parallel_compress_device_noalloc: hipGetDevice", e);
        return nullptr;
    }
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error_msg("deflate_parallel_compress_device_noalloc: device index beyond 63");
        return nullptr;
    }
    DeviceScratch& d = g_dev[dev];
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (d.ready) return &d;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("deflate_parallel_compress_device_noalloc: hipGetSymbolAddress(scratch)",
                  e != hipSuccess ? e : hipErrorInvalidSymbol);
        return nullptr;
    }
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) d.cus = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && v > 0)
        d.lds_max = (size_t)v;
    e = set_kernel_attributes_l1(d.lds_max);
    if (e == hipSuccess) e = set_kernel_attributes_l6(d.lds_max);
    if (e == hipSuccess) e = set_kernel_attributes_big();
    if (e != hipSuccess) {
        set_error("deflate_parallel_compress_device_noalloc: hipFuncSetAttribute", e);
        return nullptr;
    }
    // the lane-order check runs on the first word of this device's array
    bool ok = false;
    e = check_exchange_order_at(nullptr, reinterpret_cast<uint32_t*>(p), &ok);
    if (e != hipSuccess) {
        set_error("deflate_parallel_compress_device_noalloc: exchange check", e);
        return nullptr;
    }
    d.ptr = static_cast<uint8_t*>(p);
    d.exch_ok = ok;
    d.ready = true;
    return &d;
}

}  // namespace
}  // namespace orot_gpu

using namespace orot_gpu;

extern "C" {

size_t deflate_parallel_compress_device_noalloc(int level, deflate_format format, size_t block_size, const void* d_in,
                                                size_t in_size, void* d_out, size_t out_capacity) {
    if ((!d_in && in_size) || !d_out) {
        set_error_msg("deflate_parallel_compress_device_noalloc: null buffer");
        return 0;
    }
    if (block_size > 0xFFFFFFFFull) {
        set_error_msg("deflate_parallel_compress_device_noalloc: block_size too large");
        return 0;
    }
    EngineConfig cfg;
    cfg.level = level;
    cfg.format = (int)format;
    cfg.chunk = (uint32_t)block_size;
    if (!resolve_config(cfg)) return 0;
    const uint64_t nch = in_size ? (in_size + cfg.chunk - 1) / cfg.chunk : 1;
    if (nch > 0x7FFFFFFFull) {
        set_error_msg("deflate_parallel_compress_device_noalloc: too many chunks");
        return 0;
    }
    const uint32_t nchunks = (uint32_t)nch;
    const Layout L = layout_for(cfg, nchunks);
    if (L.total > NOALLOC_SCRATCH_BYTES) {
        set_error_msg("deflate_parallel_compress_device_noalloc: the input needs " + std::to_string(L.total) +
                      " bytes of working memory, above the fixed scratch of " +
                      std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes (4 MiB of input fits at every level)");
        return 0;
    }
    DeviceScratch* ds = device_scratch();
    if (!ds) return 0;
    std::lock_guard<std::mutex> lk(ds->mu);
    uint8_t* base = ds->ptr;
    const hipStream_t s = nullptr;

    Plan p{};
    p.d_in = static_cast<const uint8_t*>(d_in);
    p.d_out = static_cast<uint8_t*>(d_out);
    p.d_hdr = reinterpret_cast<BatchHdr*>(base + L.hdr);
    p.d_bufs = reinterpret_cast<const BufDesc*>(base + L.bufs);
    p.d_desc = reinterpret_cast<const ChunkDesc*>(base + L.desc);
    p.d_order = reinterpret_cast<const uint32_t*>(base + L.order);
    p.nchunks = nchunks;
    p.nbufs = 1;
    p.persistent_grid = false;
    p.chunk = cfg.chunk;
    p.level = cfg.level;
    p.hash_bits = cfg.hash_bits;
    p.l6 = cfg.l6;
    p.format = cfg.format;
    p.d_tok = cfg.level >= 1 ? reinterpret_cast<uint32_t*>(base + L.tok) : nullptr;
    p.d_pd = cfg.level >= 1 ? reinterpret_cast<uint16_t*>(base + L.pd) : nullptr;
    p.d_stage = cfg.level >= 1 ? reinterpret_cast<uint32_t*>(base + L.stage) : nullptr;
    const bool big = cfg.level >= 1 && cfg.level <= 3 && l1_big_chunk(cfg.chunk, cfg.hash_bits);
    p.d_aux = (cfg.level >= 4 || big) ? reinterpret_cast<uint32_t*>(base + L.aux) : nullptr;
    p.d_ntok = cfg.level >= 4 ? reinterpret_cast<uint32_t*>(base + L.ntok) : nullptr;
    p.d_blkpos = cfg.level >= 4 ? reinterpret_cast<uint32_t*>(base + L.blkpos) : nullptr;
    p.d_blocks = (cfg.level >= 4 || big) ? static_cast<void*>(base + L.blocks) : nullptr;
    p.whole = 0;
    p.l1_dict = 0;
    p.stage_words = stage_words_for_chunk(cfg.chunk);
    p.d_info = reinterpret_cast<ChunkInfo*>(base + L.info);
    p.d_result = reinterpret_cast<Result*>(base + L.res);
    p.stream = s;
    p.exch_ok = ds->exch_ok;
    p.lds_max = ds->lds_max;
    p.cus = ds->cus;

    const unsigned tg = (unsigned)((nchunks + 255) / 256);
    hipLaunchKernelGGL(single_buffer_tables_kernel, dim3(tg), dim3(256), 0, s, p.d_hdr,
                       reinterpret_cast<BufDesc*>(base + L.bufs), reinterpret_cast<ChunkDesc*>(base + L.desc),
                       reinterpret_cast<uint32_t*>(base + L.order), (uint64_t)in_size, (uint64_t)out_capacity,
                       cfg.chunk, nchunks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_parallel_compress(p);
    Result r{};
    if (e == hipSuccess) e = hipMemcpyAsync(&r, p.d_result, sizeof(Result), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error("deflate_parallel_compress_device_noalloc", e);
        return 0;
    }
    if (r.total_bytes == 0) set_error_msg("deflate_parallel_compress_device_noalloc: output does not fit out_capacity");
    return (size_t)r.total_bytes;
}

}  // extern "C"
