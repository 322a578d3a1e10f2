// engine.cpp -- batch tables, arenas, direct and graph submission.
#include "engine.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "host_common.h"

namespace orot_gpu {

hipError_t set_kernel_attributes_l1(size_t lds_max);
hipError_t set_kernel_attributes_l6(size_t lds_max);

namespace {

constexpr size_t HDR_BYTES = 64;
static_assert(sizeof(BatchHdr) <= HDR_BYTES, "batch header");

size_t default_block_size(int level) {  // src/parallel/parallel_compressor.cpp:16-20
    if (level <= 3) return 131072;
    if (level <= 6) return 524288;
    return 1048576;
}

// Lane-order check of LDS exchanges, once per process.
std::mutex g_exch_mu;
int g_exch_state = -1;

bool exchange_order_ok(hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_exch_mu);
    if (g_exch_state < 0) {
        bool ok = false;
        if (check_exchange_order(stream, &ok) != hipSuccess) return false;
        g_exch_state = ok ? 1 : 0;
    }
    return g_exch_state == 1;
}

template <class T>
bool dev_alloc(T*& p, size_t n, const char* what) {
    p = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        set_error(what, e);
        return false;
    }
    return true;
}

template <class T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

size_t tables_bytes(uint32_t buf_cap, uint32_t chunk_cap) {
    return HDR_BYTES + (size_t)buf_cap * sizeof(BufDesc) + (size_t)chunk_cap * (sizeof(ChunkDesc) + 4);
}
size_t desc_offset(uint32_t buf_cap) { return HDR_BYTES + (size_t)buf_cap * sizeof(BufDesc); }
size_t order_offset(uint32_t buf_cap, uint32_t chunk_cap) {
    return desc_offset(buf_cap) + (size_t)chunk_cap * sizeof(ChunkDesc);
}

}  // namespace

bool resolve_config(EngineConfig& cfg) {
    cfg.level = cfg.level < 0 ? 0 : (cfg.level > 12 ? 12 : cfg.level);
    cfg.format = (cfg.format == 1 || cfg.format == 2) ? cfg.format : 0;
    if (cfg.chunk == 0) cfg.chunk = (uint32_t)default_block_size(cfg.level);
    const int lv = cfg.level;
    if (lv >= 1 && lv <= 3) {
        if (cfg.hash_bits == 0) cfg.hash_bits = lv == 1 ? 12 : 14;
        if (cfg.hash_bits != 12 && cfg.hash_bits != 14) {
            set_error_msg("hash_bits must be 12 or 14 for levels 1-3");
            return false;
        }
        // above the single-workgroup kernels' limit the big-chunk path runs
        if (cfg.chunk > MAX_CHUNK_BIG) {
            set_error_msg("block_size " + std::to_string(cfg.chunk) + " exceeds the GPU limit of 1 GiB");
            return false;
        }
    } else if (lv >= 4) {
        if (cfg.l6.max_chain == 0) cfg.l6 = l6_level_params(lv);
        if (cfg.l6.max_chain < 1 || cfg.l6.nice_len < 3 || cfg.l6.nice_len > 258 || cfg.l6.lazy_depth < 0 ||
            cfg.l6.lazy_depth > 3) {
            set_error_msg("invalid matcher parameters");
            return false;
        }
        // Larger L6-class blocks run as pieces of the kernels' largest chunk,
        // each with the 32 KiB before it as its dictionary (the class's own
        // definition; its bytes are not orot's at any block size).
        if (cfg.chunk > l6_max_chunk()) cfg.chunk = l6_max_chunk();
    } else if (cfg.chunk > MAX_CHUNK_BIG) {
        set_error_msg("block_size exceeds 1 GiB");
        return false;
    }
    return true;
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (done_) (void)hipEventDestroy(done_);
    dev_free(d_tok_);
    dev_free(d_pd_);
    dev_free(d_aux_);
    dev_free(d_ntok_);
    dev_free(d_blkpos_);
    dev_free(d_blocks_);
    dev_free(d_stage_);
    dev_free(d_info_);
    dev_free(d_res_);
    dev_free(d_tables_);
    if (h_tables_) (void)hipHostFree(h_tables_);
    if (h_res_) (void)hipHostFree(h_res_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

bool Engine::init(const EngineConfig& cfg) {
    cfg_ = cfg;
    if (!resolve_config(cfg_)) return false;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        set_error_msg("no usable GPU");
        return false;
    }
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus_ = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && v > 0)
            lds_max_ = (size_t)v;
    }
    e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
    if (e != hipSuccess) {
        stream_ = nullptr;
        set_error("hipStreamCreate", e);
        return false;
    }
    e = hipEventCreateWithFlags(&done_, hipEventDisableTiming);
    if (e != hipSuccess) {
        done_ = nullptr;
        set_error("hipEventCreate", e);
        return false;
    }
    // OROT_AMD_BALLOT_SCAN=1 takes the ballot match-any variant of the
    // candidate scans even where same-address exchanges apply in lane order
    // (both variants give the same codes; this keeps the fallback tested)
    const char* force_ballot = std::getenv("OROT_AMD_BALLOT_SCAN");
    exch_ok_ = !(force_ballot && force_ballot[0] == '1') && exchange_order_ok(stream_);
    e = set_kernel_attributes_l1(lds_max_);
    if (e == hipSuccess) e = set_kernel_attributes_l6(lds_max_);
    if (e == hipSuccess) e = set_kernel_attributes_big();
    if (e != hipSuccess) {
        set_error("hipFuncSetAttribute", e);
        return false;
    }
    return true;
}

bool Engine::set_config(const EngineConfig& cfg) {
    EngineConfig c = cfg;
    if (!resolve_config(c)) return false;
    if (stream_ && !done()) (void)hipEventSynchronize(done_);
    cfg_ = c;
    drop_graph();
    return true;
}

void Engine::drop_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    gexec_ = nullptr;
    g_gen_ = ~0ull;
}

bool Engine::ensure_arenas(uint32_t nchunks, uint32_t nbufs) {
    const bool l6 = cfg_.level >= 4;
    const bool l1 = cfg_.level >= 1;
    const bool dict = l1 && !l6 && cfg_.l1_dict;
    const bool big = l1 && !l6 && (cfg_.whole || dict || l1_big_chunk(cfg_.chunk, cfg_.hash_bits));
    if (nchunks <= chunk_cap_ && nbufs <= buf_cap_ && arena_chunk_ == cfg_.chunk && (!l6 || arena_l6_) &&
        (!l1 || d_tok_) && (!big || arena_big_) && (!dict || arena_dict_))
        return true;
    // grow by at least half again, so a stream of growing batches settles quickly
    const uint32_t ncap = std::max(nchunks, (arena_chunk_ == cfg_.chunk ? chunk_cap_ + chunk_cap_ / 2 : 0u));
    const uint32_t bcap = std::max(nbufs, buf_cap_ + buf_cap_ / 2);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    dev_free(d_tok_);
    dev_free(d_pd_);
    dev_free(d_aux_);
    dev_free(d_ntok_);
    dev_free(d_blkpos_);
    dev_free(d_blocks_);
    dev_free(d_stage_);
    dev_free(d_info_);
    dev_free(d_res_);
    dev_free(d_tables_);
    if (h_tables_) (void)hipHostFree(h_tables_);
    if (h_res_) (void)hipHostFree(h_res_);
    h_tables_ = nullptr;
    h_res_ = nullptr;
    chunk_cap_ = buf_cap_ = 0;
    const uint32_t chunk = cfg_.chunk;
    const uint32_t ext = dict ? BIG_DICT : 0u;  // dictionary room per chunk
    const size_t per_pos =
        (size_t)ncap *
            (l6 ? l6_position_stride(chunk) : (big ? big_position_stride(chunk + ext) : l1_position_stride(chunk))) +
        8;
    bool ok = true;
    if (l1) {
        // big chunks keep their token-start bitmap where the others keep 16-bit codes
        const size_t pd = big ? std::max(per_pos, (size_t)ncap * big_bitmap_words(chunk) * 2 + 8) : per_pos;
        ok = ok && dev_alloc(d_tok_, per_pos, "hipMalloc(token words)") &&
             dev_alloc(d_pd_, pd, "hipMalloc(candidates / chain steps)") &&
             dev_alloc(d_stage_, (size_t)ncap * stage_words_for_chunk(chunk), "hipMalloc(staging)");
    }
    if (big) {
        ok = ok && dev_alloc(d_aux_, per_pos, "hipMalloc(candidates)") &&
             dev_alloc(d_blocks_, (size_t)ncap * big_segtab_words(chunk + ext, cfg_.hash_bits) * 4,
                       "hipMalloc(segment tables)");
    }
    if (l6) {
        const size_t mb = l6_max_blocks(chunk);
        ok = ok && dev_alloc(d_aux_, per_pos, "hipMalloc(tokens)") && dev_alloc(d_ntok_, ncap, "hipMalloc(token counts)") &&
             dev_alloc(d_blkpos_, (size_t)ncap * mb, "hipMalloc(block starts)") &&
             dev_alloc(d_blocks_, (size_t)ncap * mb * l6_block_record_bytes(), "hipMalloc(block records)");
    }
    ok = ok && dev_alloc(d_info_, ncap, "hipMalloc(chunk info)") && dev_alloc(d_res_, bcap, "hipMalloc(results)") &&
         dev_alloc(d_tables_, tables_bytes(bcap, ncap), "hipMalloc(batch tables)");
    if (ok) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&h_tables_), tables_bytes(bcap, ncap), 0);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&h_res_), (size_t)bcap * sizeof(Result), 0);
        if (e != hipSuccess) {
            set_error("hipHostMalloc", e);
            ok = false;
        }
    }
    if (!ok) return false;
    chunk_cap_ = ncap;
    buf_cap_ = bcap;
    arena_chunk_ = chunk;
    arena_l6_ = l6;
    arena_big_ = big;
    arena_dict_ = dict;
    ++gen_;
    return true;
}

bool Engine::build_tables(const BufferSpec* bufs, size_t count, uint32_t& nchunks) {
    const uint32_t chunk = cfg_.chunk;
    auto* hdr = reinterpret_cast<BatchHdr*>(h_tables_);
    auto* bd = reinterpret_cast<BufDesc*>(h_tables_ + HDR_BYTES);
    auto* cd = reinterpret_cast<ChunkDesc*>(h_tables_ + desc_offset(buf_cap_));
    auto* order = reinterpret_cast<uint32_t*>(h_tables_ + order_offset(buf_cap_, chunk_cap_));
    uint32_t c = 0;
    for (size_t i = 0; i < count; ++i) {
        const BufferSpec& b = bufs[i];
        const uint64_t nch = b.in_len ? (b.in_len + chunk - 1) / chunk : 1;
        bd[i].in_off = b.in_off;
        bd[i].in_len = b.in_len;
        bd[i].out_off = b.out_off;
        bd[i].cap = b.cap;
        bd[i].chunk0 = c;
        bd[i].nchunks = (uint32_t)nch;
        bd[i].piece = b.piece;
        bd[i].hist = b.hist;
        const uint32_t last = CHUNK_LAST | ((b.piece & PIECE_OPEN) ? CHUNK_OPEN : 0u);
        for (uint64_t j = 0; j < nch; ++j, ++c) {
            cd[c].in_off = b.in_off + j * chunk;
            cd[c].len = (uint32_t)std::min<uint64_t>(chunk, b.in_len - j * chunk);
            cd[c].buf = (uint32_t)i | (j + 1 == nch ? last : 0u);
        }
    }
    nchunks = c;
    // queue order: full chunks in index order, then the shorter ones longest first
    uint32_t k = 0;
    for (uint32_t i = 0; i < c; ++i)
        if (cd[i].len == chunk) order[k++] = i;
    const uint32_t full = k;
    for (uint32_t i = 0; i < c; ++i)
        if (cd[i].len != chunk) order[k++] = i;
    std::stable_sort(order + full, order + c, [&](uint32_t a, uint32_t b) { return cd[a].len > cd[b].len; });
    std::memset(hdr, 0, HDR_BYTES);
    hdr->nchunks = c;
    hdr->nbufs = (uint32_t)count;
    return true;
}

Plan Engine::make_plan(const uint8_t* d_in, uint8_t* d_out, uint32_t nchunks, uint32_t nbufs, bool persistent) {
    Plan p{};
    p.d_in = d_in;
    p.d_out = d_out;
    p.d_hdr = reinterpret_cast<BatchHdr*>(d_tables_);
    p.d_bufs = reinterpret_cast<const BufDesc*>(d_tables_ + HDR_BYTES);
    p.d_desc = reinterpret_cast<const ChunkDesc*>(d_tables_ + desc_offset(buf_cap_));
    p.d_order = reinterpret_cast<const uint32_t*>(d_tables_ + order_offset(buf_cap_, chunk_cap_));
    p.nchunks = nchunks;
    p.nbufs = nbufs;
    p.persistent_grid = persistent;
    p.chunk = cfg_.chunk;
    p.level = cfg_.level;
    p.hash_bits = cfg_.hash_bits;
    p.l6 = cfg_.l6;
    p.format = cfg_.format;
    p.whole = cfg_.whole ? 1 : 0;
    p.l1_dict = (cfg_.l1_dict && cfg_.level >= 1 && cfg_.level <= 3) ? 1 : 0;
    p.d_tok = d_tok_;
    p.d_pd = d_pd_;
    p.d_aux = d_aux_;
    p.d_ntok = d_ntok_;
    p.d_blkpos = d_blkpos_;
    p.d_blocks = d_blocks_;
    p.d_stage = d_stage_;
    p.stage_words = stage_words_for_chunk(cfg_.chunk);
    p.d_info = d_info_;
    p.d_result = d_res_;
    p.stream = stream_;
    p.exch_ok = exch_ok_;
    p.lds_max = lds_max_;
    p.cus = cus_;
    return p;
}

bool Engine::capture(const uint8_t* d_in, uint8_t* d_out) {
    drop_graph();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("hipStreamBeginCapture", e);
        return false;
    }
    hipError_t le = hipMemcpyAsync(d_tables_, h_tables_, tables_bytes(buf_cap_, chunk_cap_), hipMemcpyHostToDevice,
                                   stream_);
    if (le == hipSuccess) le = launch_parallel_compress(make_plan(d_in, d_out, chunk_cap_, buf_cap_, true));
    if (le == hipSuccess)
        le = hipMemcpyAsync(h_res_, d_res_, (size_t)buf_cap_ * sizeof(Result), hipMemcpyDeviceToHost, stream_);
    e = hipStreamEndCapture(stream_, &g);
    if (le != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error("graph capture", le != hipSuccess ? le : e);
        return false;
    }
    e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        gexec_ = nullptr;
        set_error("hipGraphInstantiate", e);
        return false;
    }
    g_in_ = d_in;
    g_out_ = d_out;
    g_gen_ = gen_;
    return true;
}

bool Engine::submit(const uint8_t* d_in, uint8_t* d_out, const BufferSpec* bufs, size_t count, bool use_graph) {
    // the pinned tables of the previous submit may still be in flight
    (void)hipEventSynchronize(done_);
    last_nbufs_ = 0;
    timed_ = false;
    if (count == 0) return true;
    if (count > 0x7FFFFFFFull) {
        set_error_msg("too many buffers in one batch");
        return false;
    }
    uint64_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        total += bufs[i].in_len ? (bufs[i].in_len + cfg_.chunk - 1) / cfg_.chunk : 1;
        if (total > 0x7FFFFFFFull) {
            set_error_msg("too many chunks in one batch");
            return false;
        }
    }
    if (!ensure_arenas((uint32_t)total, (uint32_t)count)) return false;
    uint32_t nchunks = 0;
    build_tables(bufs, count, nchunks);
    if (timing_ && !ev_[0]) {
        for (auto& ev : ev_) {
            hipError_t e = hipEventCreate(&ev);
            if (e != hipSuccess) {
                ev = nullptr;
                set_error("hipEventCreate", e);
                return false;
            }
        }
    }
    hipError_t e = hipSuccess;
    if (use_graph) {
        if ((!gexec_ || g_in_ != d_in || g_out_ != d_out || g_gen_ != gen_) && !capture(d_in, d_out)) return false;
        if (timing_) (void)hipEventRecord(ev_[0], stream_);
        e = hipGraphLaunch(gexec_, stream_);
        if (timing_) (void)hipEventRecord(ev_[1], stream_);
    } else {
        const size_t head = HDR_BYTES + count * sizeof(BufDesc);
        e = hipMemcpyAsync(d_tables_, h_tables_, head, hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tables_ + desc_offset(buf_cap_), h_tables_ + desc_offset(buf_cap_),
                               (size_t)nchunks * sizeof(ChunkDesc), hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tables_ + order_offset(buf_cap_, chunk_cap_),
                               h_tables_ + order_offset(buf_cap_, chunk_cap_), (size_t)nchunks * 4,
                               hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) {
            Plan p = make_plan(d_in, d_out, nchunks, (uint32_t)count, false);
            for (int i = 0; i < 8; ++i) p.ev[i] = timing_ ? ev_[i] : nullptr;
            e = launch_parallel_compress(p);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_res_, d_res_, count * sizeof(Result), hipMemcpyDeviceToHost, stream_);
    }
    if (e == hipSuccess) e = hipEventRecord(done_, stream_);
    if (e != hipSuccess) {
        set_error("submit", e);
        return false;
    }
    last_nbufs_ = (uint32_t)count;
    last_graph_ = use_graph;
    timed_ = timing_;
    return true;
}

uint64_t Engine::arena_bytes_per_chunk(const EngineConfig& c) {
    const bool l6 = c.level >= 4;
    const bool l1 = c.level >= 1;
    const bool dict = l1 && !l6 && c.l1_dict;
    const bool big = l1 && !l6 && (c.whole || dict || l1_big_chunk(c.chunk, c.hash_bits));
    const uint32_t ext = dict ? BIG_DICT : 0u;
    const uint64_t stride =
        l6 ? l6_position_stride(c.chunk) : (big ? big_position_stride(c.chunk + ext) : l1_position_stride(c.chunk));
    uint64_t b = sizeof(ChunkInfo) + sizeof(ChunkDesc) + 4;
    if (l1) {
        b += 4 * stride + 2 * (big ? std::max<uint64_t>(stride, 2 * big_bitmap_words(c.chunk)) : stride) +
             4 * stage_words_for_chunk(c.chunk);
    }
    if (big) b += 4 * stride + 4 * big_segtab_words(c.chunk + ext, c.hash_bits);
    if (l6) {
        const uint64_t mb = l6_max_blocks(c.chunk);
        b += 4 * stride + 4 + 4 * mb + mb * l6_block_record_bytes();
    }
    return b;
}

bool Engine::prepare_device_tables(uint32_t chunk_cap, uint32_t buf_cap, DeviceTables* out) {
    if (!ensure_arenas(chunk_cap, buf_cap)) return false;
    out->hdr = reinterpret_cast<BatchHdr*>(d_tables_);
    out->bufs = reinterpret_cast<BufDesc*>(d_tables_ + HDR_BYTES);
    out->desc = reinterpret_cast<ChunkDesc*>(d_tables_ + desc_offset(buf_cap_));
    out->order = reinterpret_cast<uint32_t*>(d_tables_ + order_offset(buf_cap_, chunk_cap_));
    out->results = d_res_;
    return true;
}

bool Engine::launch_device_tables(const uint8_t* d_in, uint8_t* d_out, uint32_t nchunks, uint32_t nbufs) {
    last_nbufs_ = 0;
    timed_ = false;
    if (nchunks > chunk_cap_ || nbufs > buf_cap_) {
        set_error_msg("device-built batch exceeds the arenas");
        return false;
    }
    if (nbufs == 0) return true;
    hipError_t e = launch_parallel_compress(make_plan(d_in, d_out, nchunks, nbufs, false));
    if (e == hipSuccess) e = hipEventRecord(done_, stream_);
    if (e != hipSuccess) {
        set_error("launch (device tables)", e);
        return false;
    }
    last_graph_ = false;
    return true;
}

bool Engine::wait(uint64_t* sizes) {
    hipError_t e = hipEventSynchronize(done_);
    if (e != hipSuccess) {
        set_error("hipEventSynchronize", e);
        return false;
    }
    for (uint32_t i = 0; i < last_nbufs_; ++i) sizes[i] = h_res_[i].total_bytes;
    return true;
}

bool Engine::done() const { return hipEventQuery(done_) == hipSuccess; }

int Engine::stage_times(float* ms, int max_values) const {
    if (!timed_ || !ms) return 0;
    const int n = last_graph_ ? 1 : 4;
    const int k = std::min(n, max_values);
    for (int i = 0; i < k; ++i) {
        float v = 0;
        (void)hipEventElapsedTime(&v, ev_[i], ev_[i + 1]);
        ms[i] = v;
    }
    return k;
}

}  // namespace orot_gpu
