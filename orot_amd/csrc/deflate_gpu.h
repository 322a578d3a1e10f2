// deflate_gpu.h -- internal interface between the host layer (deflate_api.cpp,
// deflate_engine.cpp) and the gfx950 kernels (deflate_kernels.hip,
// deflate_l6.hip).
//
// Every launch compresses a batch: a table of buffers, each cut into chunks
// (DEFLATE blocks) of at most `chunk` bytes.  The single-buffer entry points
// are batches of one.  Work counts live in device memory (BatchHdr), so a
// captured graph replays for any batch that fits the arenas.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace orot_gpu {

constexpr uint32_t CHUNK_LAST = 0x80000000u;  // ChunkDesc::buf flag: last chunk of its buffer
constexpr uint32_t CHUNK_OPEN = 0x40000000u;  // ... of a stream piece that does not end the stream
constexpr uint32_t CHUNK_BUF = 0x3FFFFFFFu;   // the buffer index bits

// The chunk's last block carries BFINAL: it ends its buffer and the buffer is
// not an open stream piece.
__host__ __device__ constexpr bool chunk_final(uint32_t buf) {
    return (buf & (CHUNK_LAST | CHUNK_OPEN)) == CHUNK_LAST;
}

// One chunk of one buffer.
struct ChunkDesc {
    uint64_t in_off;  // first byte of the chunk in the input arena
    uint32_t len;     // bytes
    uint32_t buf;     // buffer index | CHUNK_LAST | CHUNK_OPEN
};

// BufDesc::piece: the buffer is one piece of a stream.  Its output is the raw
// bitstream continuing a previous piece: no header or trailer, bits starting
// at offset (piece & 7) of byte 0, whose lower bits are the held bits
// (piece >> 8) & 0xFF; the result's checksum is the piece's own.
constexpr uint32_t PIECE = 0x80000000u;
constexpr uint32_t PIECE_OPEN = 0x40000000u;  // the piece does not end the stream

// One buffer of the batch.
struct BufDesc {
    uint64_t in_off;   // first byte of the buffer in the input arena
    uint64_t in_len;
    uint64_t out_off;  // first byte of its framed output in the output arena
    uint64_t cap;      // output capacity
    uint32_t chunk0;   // its chunks are [chunk0, chunk0 + nchunks)
    uint32_t nchunks;
    uint32_t piece;    // 0, or PIECE | PIECE_OPEN? | held bits << 8 | bit offset
    uint32_t hist;     // bytes before in_off that precede the buffer in its stream (<= 32 KiB)
};

// Work-queue heads of the persistent kernels.
enum Queue { Q_LZ = 0, Q_PRED = 1, Q_MATCH = 2, Q_PARSE = 3, Q_SCAN = 4, Q_ASM = 5, Q_BLOCK = 6, Q_EMIT = 7, Q_COUNT = 8 };

// Batch header in device memory, rewritten (queue heads zero) at each submit.
struct BatchHdr {
    uint32_t nchunks;
    uint32_t nbufs;
    uint32_t pad_[6];
    uint32_t queue[Q_COUNT];
};

// Per-chunk record in HBM.  Written by the checksum and LZ77 kernels, read by
// the scan (finalize) and assembly kernels.
struct ChunkInfo {
    uint64_t fixed_bits;  // size of the chunk's fixed-Huffman block, header..EOB
    uint64_t bit_off;     // offset of the chunk inside its buffer's raw DEFLATE payload
    uint32_t raw_len;     // input bytes of the chunk
    uint32_t is_stored;   // 0: one block of fixed_bits, 1: stored block(s),
                          // 2: several blocks, size by incoming bit offset mod 8 in sz8
    uint32_t crc;         // CRC-32 register over the chunk, zero initial value
    uint32_t adler_s1;    // sum of bytes mod 65521
    uint32_t adler_s2;    // sum of (raw_len - j) * byte_j mod 65521
    uint32_t pad_;
    uint64_t sz8[8];      // is_stored == 2: chunk bits when it starts at bit offset = r (mod 8)
};

// Per-buffer result written by the finalize kernel.
struct Result {
    uint64_t total_bytes;    // framed output size, 0 when it exceeds the capacity
    uint64_t payload_bits;   // raw DEFLATE bit length
    uint32_t checksum;       // CRC-32 (gzip) or Adler-32 (zlib) of the input
    uint32_t pad_;
};

// Matcher parameters of the L6 class (levels 4..12).
struct L6Params {
    int max_chain;   // chain candidates examined per position
    int nice_len;    // stop the walk at this length
    int lazy_depth;  // literals the lazy rule may defer
};

struct Plan {
    const uint8_t* d_in;    // input arena
    uint8_t* d_out;         // output arena
    const ChunkDesc* d_desc;
    const uint32_t* d_order;  // queue order of the chunks (longest first)
    const BufDesc* d_bufs;
    BatchHdr* d_hdr;
    uint32_t nchunks;       // host copy of the counts: grid sizes of direct launches
    uint32_t nbufs;
    bool persistent_grid;   // launch at resident capacity (graph capture): counts come from d_hdr only
    uint32_t chunk;         // chunk capacity (per-chunk scratch stride)
    int level;              // 0 stored, 1..3 L1 class, >= 4 L6 class
    int hash_bits;          // L1 class: 12 or 14
    L6Params l6;
    int format;             // 0 raw, 1 zlib, 2 gzip
    uint32_t* d_tok;        // per position (chunk-padded): token word / match word
    uint16_t* d_pd;         // per position: L1 class candidate distance code, L6 class chain step then decision
    uint32_t* d_aux;        // L6 class: tokens in order (chunk-padded)
    uint32_t* d_ntok;       // L6 class: tokens per chunk
    uint32_t* d_blkpos;     // L6 class: raw position of each block's first token
    void* d_blocks;         // L6 class: block records
    uint32_t* d_stage;      // nchunks * stage_words
    uint64_t stage_words;   // per chunk
    ChunkInfo* d_info;      // nchunks
    Result* d_result;       // nbufs
    hipStream_t stream;
    hipEvent_t ev[8];       // optional timing events (nullptr = off)
    int whole;              // levels 1-3: one chunk per buffer with orot's whole-buffer rules (all-literal -> stored)
    int l1_dict;            // levels 1-3: each chunk sees the 32 KiB before it (big-chunk path, arrays with BIG_DICT room)
    bool exch_ok;           // LDS exchanges of one wave apply in lane order
    size_t lds_max;         // largest dynamic LDS allocation per workgroup
    int cus;                // compute units
};

// One-time start-up check behind Plan::exch_ok (allocates one device word).
hipError_t check_exchange_order(hipStream_t stream, bool* ok);
// The same check on a caller-provided device word (no allocation).
hipError_t check_exchange_order_at(hipStream_t stream, uint32_t* d_word, bool* ok);

// Largest chunk the L1-class kernels accept for the given hash width.
uint32_t max_chunk_for_hash_bits(int hash_bits);

// Levels 1-3 with chunks above the single-workgroup kernels' limit (or a
// whole-buffer call) take the big-chunk path of deflate_big.hip: token words
// and candidates (d_tok, d_aux: big_position_stride per chunk), the
// token-start bitmap (d_pd: big_bitmap_words 32-bit words per chunk) and the
// segment tables (d_blocks: big_segtab_words 32-bit words per chunk).
constexpr uint32_t MAX_CHUNK_BIG = 1u << 30;
constexpr uint32_t BIG_SEG = 65536;  // positions per candidate segment
constexpr uint32_t BIG_DICT = 32768;  // dictionary room per chunk in dictionary mode
bool l1_big_chunk(uint32_t chunk, int hash_bits);
__host__ __device__ constexpr uint64_t big_position_stride(uint32_t chunk) { return ((uint64_t)chunk + 7u) & ~7ull; }
__host__ __device__ constexpr uint64_t big_bitmap_words(uint32_t chunk) { return ((uint64_t)chunk + 31u) / 32u + 1u; }
__host__ __device__ constexpr uint32_t big_segs(uint32_t chunk) { return (chunk + BIG_SEG - 1) / BIG_SEG; }
__host__ __device__ constexpr uint64_t big_segtab_words(uint32_t chunk, int hash_bits) {
    return (uint64_t)big_segs(chunk) << hash_bits;
}
hipError_t set_kernel_attributes_big();
hipError_t launch_l1_big(const Plan& p, int whole);

// Per-chunk stride (32-bit words) of the L1-class token words: the chunk
// plus one partial range of the interleaved layout.
inline uint64_t l1_position_stride(uint32_t chunk) { return (((uint64_t)chunk + 7) & ~7ull) + 256; }

// Stage capacity (32-bit words) needed per chunk of `chunk` bytes.
uint64_t stage_words_for_chunk(uint32_t chunk);

// Enqueue the whole pipeline on plan.stream.  Returns a hipError_t.
hipError_t launch_parallel_compress(const Plan& plan);

// L6-class path (levels 4..12): LZ77 + block analysis, then emission.
L6Params l6_level_params(int level);
uint32_t l6_max_blocks(uint32_t chunk);
uint64_t l6_position_stride(uint32_t chunk);  // per-chunk stride of d_tok / d_pd / d_aux (chunk + halo)
uint32_t l6_max_chunk();
size_t l6_block_record_bytes();
hipError_t launch_l6_front(const Plan& plan);
hipError_t launch_l6_back(const Plan& plan);

// Batched inflate (inflate.hip): one stream per item, order[] the items
// longest first; *d_queue is the device work-queue head (zeroed here).
struct InflateItem {
    uint64_t in_off, in_size, out_off, out_cap;
};
struct InflateResult {
    uint64_t out_size;
    int32_t status;  // deflate_result value
    uint32_t pad_;
};
hipError_t launch_inflate(const uint8_t* d_in, uint8_t* d_out, const InflateItem* d_items, const uint32_t* d_order,
                          uint32_t count, int format, InflateResult* d_res, uint32_t* d_queue, int cus,
                          hipStream_t stream);

// Checksums of a device buffer (whole-buffer entry points).
hipError_t launch_checksum(const uint8_t* d_in, uint64_t n, int kind /*0 crc,1 adler*/, uint32_t init,
                           uint32_t* d_out_value, ChunkInfo* d_tmp, uint32_t tmp_chunks, hipStream_t stream);

}  // namespace orot_gpu
