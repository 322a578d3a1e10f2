// engine.h -- one GPU compression pipeline: a HIP stream, device arenas, the
// batch tables and their pinned staging.  Every C entry point (single buffer,
// batch, server) drives an Engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "deflate_gpu.h"

namespace orot_gpu {

// One buffer of a batch: bytes [in_off, in_off + in_len) of the input arena
// compress into [out_off, out_off + cap) of the output arena.
struct BufferSpec {
    uint64_t in_off;
    uint64_t in_len;
    uint64_t out_off;
    uint64_t cap;
    uint32_t piece = 0;  // BufDesc::piece: 0 for a whole stream
    uint32_t hist = 0;   // stream bytes before in_off in the arena (L6 class dictionary)
};

struct EngineConfig {
    int level = 1;        // clamped 0..12
    int format = 0;       // 0 raw, 1 zlib, 2 gzip
    uint32_t chunk = 0;   // DEFLATE block size (0: orot's default for the level)
    int hash_bits = 0;    // L1 class: 12 or 14 (0: the level's own)
    L6Params l6{};        // L6 class (max_chain 0: the level's own)
    bool whole = false;   // levels 0-3: orot's whole-buffer rules (chunk must cover each buffer)
    bool l1_dict = false; // levels 1-3: each chunk sees the 32 KiB before it
};

// Fills the zero fields of cfg with the level's defaults and checks the
// limits of the kernels; false (with the reason in the last error) if cfg
// cannot run.
bool resolve_config(EngineConfig& cfg);

class Engine {
public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Creates the stream and reads the device limits.
    bool init(const EngineConfig& cfg);
    bool set_config(const EngineConfig& cfg);
    const EngineConfig& config() const { return cfg_; }

    // Enqueues the batch on the engine's stream (returns before the GPU is
    // done).  With use_graph the whole submission -- table upload, kernels,
    // result download -- replays one captured hipGraph while the arenas and
    // the d_in / d_out bases stay the same.
    bool submit(const uint8_t* d_in, uint8_t* d_out, const BufferSpec* bufs, size_t count, bool use_graph);
    // Waits for the last submit; sizes[i] = framed size of buffer i, 0 if it
    // did not fit its capacity.
    bool wait(uint64_t* sizes);
    // True when the last submit has finished (non-blocking).
    bool done() const;
    // Results of the last submit, valid after wait().
    const Result* results() const { return h_res_; }

    void set_timing(bool on) { timing_ = on; }
    // Device time of the last submit: 4 stages without a graph, the whole
    // submission (1 value) with one.
    int stage_times(float* ms, int max_values) const;

    hipStream_t stream() const { return stream_; }
    int compute_units() const { return cus_; }

    // Batches whose tables a kernel writes in device memory (the server's
    // admission).  prepare_device_tables sizes the arenas for exactly
    // chunk_cap chunks and buf_cap buffers and says where the tables go;
    // launch_device_tables enqueues the pipeline over them, nchunks / nbufs
    // being the counts the tables hold (they size the grids).
    struct DeviceTables {
        BatchHdr* hdr;
        BufDesc* bufs;
        ChunkDesc* desc;
        uint32_t* order;
        const Result* results;
    };
    bool prepare_device_tables(uint32_t chunk_cap, uint32_t buf_cap, DeviceTables* out);
    bool launch_device_tables(const uint8_t* d_in, uint8_t* d_out, uint32_t nchunks, uint32_t nbufs);
    // Device bytes the arenas take per chunk of a resolved configuration.
    static uint64_t arena_bytes_per_chunk(const EngineConfig& cfg);

private:
    bool ensure_arenas(uint32_t nchunks, uint32_t nbufs);
    bool build_tables(const BufferSpec* bufs, size_t count, uint32_t& nchunks);
    Plan make_plan(const uint8_t* d_in, uint8_t* d_out, uint32_t nchunks, uint32_t nbufs, bool persistent);
    bool capture(const uint8_t* d_in, uint8_t* d_out);
    void drop_graph();

    EngineConfig cfg_{};
    hipStream_t stream_ = nullptr;
    int cus_ = 256;
    size_t lds_max_ = 65536;
    bool exch_ok_ = false;

    // arenas (capacities in elements)
    uint32_t* d_tok_ = nullptr;
    uint16_t* d_pd_ = nullptr;
    uint32_t* d_aux_ = nullptr;
    uint32_t* d_ntok_ = nullptr;
    uint32_t* d_blkpos_ = nullptr;
    uint8_t* d_blocks_ = nullptr;
    uint32_t* d_stage_ = nullptr;
    ChunkInfo* d_info_ = nullptr;
    Result* d_res_ = nullptr;
    uint8_t* d_tables_ = nullptr;  // BatchHdr | BufDesc[buf_cap] | ChunkDesc[chunk_cap] | order[chunk_cap]
    uint8_t* h_tables_ = nullptr;  // pinned mirror
    Result* h_res_ = nullptr;      // pinned
    uint32_t chunk_cap_ = 0;       // chunks the arenas hold
    uint32_t buf_cap_ = 0;
    uint32_t arena_chunk_ = 0;     // chunk size the arenas were sized for
    bool arena_l6_ = false;
    bool arena_big_ = false;
    bool arena_dict_ = false;
    uint64_t gen_ = 0;             // bumped when any arena moves

    // last submit
    uint32_t last_nbufs_ = 0;
    bool last_graph_ = false;
    hipEvent_t ev_[8] = {};
    hipEvent_t done_ = nullptr;
    bool timing_ = false;
    bool timed_ = false;

    // captured graph
    hipGraphExec_t gexec_ = nullptr;
    const uint8_t* g_in_ = nullptr;
    uint8_t* g_out_ = nullptr;
    uint64_t g_gen_ = ~0ull;
};

}  // namespace orot_gpu
