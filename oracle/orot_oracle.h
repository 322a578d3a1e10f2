/*
 * orot_oracle.h -- CPU restatement of orot's parallel DEFLATE compress path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * (tests/, __graft_entry__.smoke(), and bench.py's cpu_baseline leg).  No
 * product code links, loads or calls it.
 *
 * Every function cites the orot source (paths relative to the orot tree) whose
 * behaviour it restates.  Pinning of each function (golden vectors, zlib,
 * closed forms, brute force) is listed in the header comment of
 * orot_oracle.c.
 */
#ifndef OROT_ORACLE_H
#define OROT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Token encoding, same bit layout as orot's Token (src/deflate/lz77.hpp:26-43):
 * literal = byte value, match = bit31 | len << 16 | dist. */
#define ORACLE_TOK_MATCH 0x80000000u

/* ---- checksums (src/simd/simd_detect.cpp:33-70) ---- */
uint32_t oracle_crc32(uint32_t crc, const uint8_t* p, size_t n);
uint32_t oracle_adler32(uint32_t adler, const uint8_t* p, size_t n);
/* crc32(A||B) from crc32(A), crc32(B), |B| (GF(2) polynomial arithmetic). */
uint32_t oracle_crc32_combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b);
uint32_t oracle_adler32_combine(uint32_t adler_a, uint32_t adler_b, uint64_t len_b);

/* ---- bounds (src/formats/raw_deflate.cpp:9-12, src/api/deflate_api.cpp:81-98) ---- */
size_t oracle_compress_bound(size_t n, int format);

/* ---- LZ77 over one chunk with fresh tables (src/deflate/lz77.cpp:352-499).
 * tokens must hold n entries (at least 1).  Returns the token count.
 * Levels 0..9 are restated; levels 10..12 (BT4) return (size_t)-1. */
size_t oracle_lz77(const uint8_t* src, size_t n, int level, uint32_t* tokens);

/* Per-position L1-class candidate used by the data-parallel formulation:
 * d16[p] = (p - C(p)) & 0x7FFF when C(p) != 0, the distance is non-zero and
 * src[C(p)] == src[p], else 0; C(p) = latest q < p, q + 4 <= n, with the same
 * hash (every earlier position is in the table; lz77.cpp:163-191,397-403). */
void oracle_l1_candidates(const uint8_t* src, size_t n, int hash_bits, uint16_t* d16);

/* ---- orot parallel compressor (src/parallel/parallel_compressor.cpp:36-150).
 * threads <= 0 means one thread per online CPU.  Levels 0..3 reproduce orot's
 * bytes exactly (stored / fixed-Huffman blocks); the gzip trailer uses IEEE
 * CRC-32.  Levels 4..9 run orot's exact LZ77 parse and encode each chunk with
 * the block encoder of the HIP L6-class path (oracle_encode_blocks_dyn).
 * Returns bytes written, 0 on failure (including out_cap too small). */
size_t oracle_parallel_compress(int level, int format, int threads, size_t block_size,
                                const uint8_t* src, size_t n, uint8_t* dst, size_t cap);

/* The HIP path's own L6-class matcher (levels 4..12; see orot_oracle.c):
 * match length / distance per position, and the lazy token stream. */
void oracle_l6_match(const uint8_t* src, size_t n, int level, uint32_t* mlen, uint32_t* mdist);
size_t oracle_l6_lz77(const uint8_t* src, size_t n, int level, uint32_t* tokens);
/* The same with a cross-chunk dictionary: src points at the halo bytes that
 * precede the chunk (halo <= 32768), the chunk follows them; results are for
 * the chunk's n positions. */
void oracle_l6_match_h(const uint8_t* src, size_t halo, size_t n, int level, uint32_t* mlen, uint32_t* mdist);
size_t oracle_l6_lz77_h(const uint8_t* src, size_t halo, size_t n, int level, uint32_t* tokens);

/* What the HIP path computes: levels 0..3 as oracle_parallel_compress
 * (orot's bytes); levels 4..12 with the L6-class matcher above, orot's chunk
 * sizes (src/parallel/parallel_compressor.cpp:16-20), each chunk after the
 * first with the up to 32 KiB before it as dictionary, and
 * oracle_encode_blocks_dyn. */
size_t oracle_gpu_parallel_compress(int level, int format, int threads, size_t block_size, const uint8_t* src,
                                    size_t n, uint8_t* dst, size_t cap);

/* Levels 1-3 in the HIP path's dictionary mode: each chunk after the first
 * of its buffer sees the up to 32 KiB before it (its positions are in the
 * hash table before the chunk's first probe; tokens cover the chunk only).
 * Other levels: as oracle_gpu_parallel_compress. */
size_t oracle_l1_dict_compress(int level, int format, int threads, size_t block_size, const uint8_t* src, size_t n,
                               uint8_t* dst, size_t cap);

/* The dictionary mode's token stream of one chunk (levels 1-3): src points
 * at the halo bytes before the chunk (halo <= 32768), the chunk follows. */
size_t oracle_l1_lz77_h(const uint8_t* src, size_t halo, size_t n, int level, uint32_t* tokens);

/* orot's whole-buffer deflate_compress (src/api/deflate_api.cpp:100-136,
 * src/compress/block_compressor.cpp:16-73, the zlib and gzip wrappers):
 * 0 when cap is below the bound; levels 0-3: the whole input as one chunk,
 * stored when it holds no match, orot's bytes (IEEE CRC-32); levels 4-12: the
 * HIP path's L6-class stream with the level's default chunks.  gzip XFL by
 * level. */
size_t oracle_whole_compress(int level, int format, const uint8_t* src, size_t n, uint8_t* dst, size_t cap);

/* Single-chunk block emission with orot's Fixed/Stored hint
 * (src/deflate/deflate_block.cpp:389-474, hint Fixed).  Appends to a bit
 * buffer starting at bit offset *bitpos of dst (dst must be zero-filled).
 * Returns 1 if the chunk was emitted as stored, 0 if fixed. */
int oracle_encode_block_fixed(const uint32_t* tokens, size_t ntok, const uint8_t* raw, size_t raw_len,
                              int is_last, uint8_t* dst, uint64_t* bitpos);

/* HIP L6-class block encoder (documented deviation from orot's single
 * dynamic block and its clamped heap Huffman): tokens are split into blocks of
 * at most max_syms tokens, each block takes the cheapest of stored / fixed /
 * dynamic by exact size, dynamic codes are optimal length-limited
 * (package-merge, 15 / 7 bits).  Same bit-buffer convention as above. */
void oracle_encode_blocks_dyn(const uint32_t* tokens, size_t ntok, const uint8_t* raw, size_t raw_len,
                              int is_last, int max_syms, uint8_t* dst, uint64_t* bitpos);

/* One block of the given type (0 stored, 1 fixed, 2 dynamic with the
 * package-merge lengths oracle_encode_blocks_dyn uses), whatever the choice
 * rule would pick: the emitted sizes the choice is checked against.  Same
 * bit-buffer convention as above. */
void oracle_encode_block_forced(const uint32_t* tokens, size_t ntok, const uint8_t* raw, int is_last, int btype,
                                uint8_t* dst, uint64_t* bitpos);

/* Optimal length-limited Huffman code lengths (package-merge) with the
 * deterministic tie order (freq, symbol).  One active symbol gets length 1.
 * Returns the number of active symbols. */
int oracle_limited_lengths(const uint32_t* freq, int nsym, int max_bits, uint8_t* lens);

/* Exact bit size of a dynamic block (header + data + EOB) for given lengths. */
uint64_t oracle_dynamic_block_bits(const uint32_t* lit_freq, const uint32_t* dist_freq,
                                   const uint8_t* ll, const uint8_t* dl);

#ifdef __cplusplus
}
#endif
#endif
