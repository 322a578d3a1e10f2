// deflate_gpu.h -- internal interface between the C-ABI host layer
// (deflate_api.cpp) and the gfx950 kernels (deflate_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace orot_gpu {

// Per-chunk record in HBM.  Written by the checksum and LZ77 kernels, read by
// the scan (finalize) and assembly kernels.
struct ChunkInfo {
    uint64_t fixed_bits;  // size of the chunk's fixed-Huffman block, header..EOB
    uint64_t bit_off;     // offset of the chunk inside the raw DEFLATE payload
    uint32_t raw_len;     // input bytes of the chunk
    uint32_t is_stored;   // 1: the chunk is emitted as stored block(s)
    uint32_t crc;         // CRC-32 register over the chunk, zero initial value
    uint32_t adler_s1;    // sum of bytes mod 65521
    uint32_t adler_s2;    // sum of (raw_len - j) * byte_j mod 65521
    uint32_t pad_;
};

// Result words written by the finalize kernel.
struct Result {
    uint64_t total_bytes;    // framed output size, 0 when it exceeds the capacity
    uint64_t payload_bits;   // raw DEFLATE bit length
    uint32_t checksum;       // CRC-32 (gzip) or Adler-32 (zlib) of the input
    uint32_t pad_;
};

struct Plan {
    const uint8_t* d_in;
    uint64_t n;             // input bytes
    uint32_t chunk;         // chunk (DEFLATE block) size
    uint32_t nchunks;
    int level;              // clamped 0..12
    int format;             // 0 raw, 1 zlib, 2 gzip
    uint8_t* d_out;
    uint64_t cap;
    uint32_t* d_tok;        // n entries: per-position token word
    uint32_t* d_stage;      // nchunks * stage_words
    uint64_t stage_words;   // per chunk
    ChunkInfo* d_info;      // nchunks
    Result* d_result;
    hipStream_t stream;
    hipEvent_t ev[8];       // optional timing events (nullptr = off)
    bool exch_ok;           // LDS exchanges of one wave apply in lane order
    size_t lds_max;         // largest dynamic LDS allocation per workgroup
};

// One-time start-up check behind Plan::exch_ok.
hipError_t check_exchange_order(hipStream_t stream, bool* ok);

// Largest chunk the L1-class kernels accept for the given hash width.
uint32_t max_chunk_for_hash_bits(int hash_bits);

// Stage capacity (32-bit words) needed per chunk of `chunk` bytes.
uint64_t stage_words_for_chunk(uint32_t chunk);

// Enqueue the whole pipeline on plan.stream.  Returns a hipError_t.
hipError_t launch_parallel_compress(const Plan& plan);

// Checksums of a device buffer (whole-buffer entry points).
hipError_t launch_checksum(const uint8_t* d_in, uint64_t n, int kind /*0 crc,1 adler*/, uint32_t init,
                           uint32_t* d_out_value, ChunkInfo* d_tmp, uint32_t tmp_chunks, hipStream_t stream);

}  // namespace orot_gpu
