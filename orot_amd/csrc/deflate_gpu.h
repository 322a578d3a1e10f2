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
    uint32_t is_stored;   // 0: one block of fixed_bits, 1: stored block(s),
                          // 2: several blocks, size by incoming bit offset mod 8 in sz8
    uint32_t crc;         // CRC-32 register over the chunk, zero initial value
    uint32_t adler_s1;    // sum of bytes mod 65521
    uint32_t adler_s2;    // sum of (raw_len - j) * byte_j mod 65521
    uint32_t pad_;
    uint64_t sz8[8];      // is_stored == 2: chunk bits when it starts at bit offset = r (mod 8)
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
    uint32_t* d_tok;        // per position (chunk-padded): token word / match word
    uint16_t* d_pd;         // L6 class: per position chain step, then decision code
    uint32_t* d_aux;        // L6 class: tokens in order (chunk-padded)
    uint32_t* d_ntok;       // L6 class: tokens per chunk
    uint32_t* d_blkpos;     // L6 class: raw position of each block's first token
    void* d_blocks;         // L6 class: block records
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

// L6-class path (levels 4..12): LZ77 + block analysis, then emission.
uint32_t l6_max_blocks(uint32_t chunk);
uint32_t l6_max_chunk();
size_t l6_block_record_bytes();
hipError_t launch_l6_front(const Plan& plan);
hipError_t launch_l6_back(const Plan& plan);

// Checksums of a device buffer (whole-buffer entry points).
hipError_t launch_checksum(const uint8_t* d_in, uint64_t n, int kind /*0 crc,1 adler*/, uint32_t init,
                           uint32_t* d_out_value, ChunkInfo* d_tmp, uint32_t tmp_chunks, hipStream_t stream);

}  // namespace orot_gpu
