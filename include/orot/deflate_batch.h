/*
 * orot/deflate_batch.h -- batched compression on one MI355X.
 *
 * No counterpart in orot: orot compresses one buffer per call
 * (deflate_parallel_compress).  A batch compresses many independent buffers
 * in one submission -- every chunk of every buffer is work for the same
 * persistent kernels -- with the output bytes of deflate_parallel_compress
 * on each buffer alone (same level, format and block size).
 */
#ifndef OROT_DEFLATE_BATCH_H
#define OROT_DEFLATE_BATCH_H

#include "deflate.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One buffer of a device-resident batch (offsets relative to the d_in / d_out
 * arguments of deflate_batch_compress_device). */
typedef struct deflate_batch_item {
    uint64_t in_offset;
    uint64_t in_size;
    uint64_t out_offset;
    uint64_t out_capacity;
} deflate_batch_item;

/* Flags of deflate_batch_new. */
#define DEFLATE_BATCH_GRAPH 1u    /* replay one captured hipGraph per submission */
#define DEFLATE_BATCH_AUTOTUNE 2u /* with block_size 0: tune on the first batch */

typedef struct deflate_batch_ctx deflate_batch_ctx;

/* level / format / block_size as deflate_parallel_new (block_size 0: orot's
 * default for the level, or the autotuner's choice with
 * DEFLATE_BATCH_AUTOTUNE).  NULL when no GPU is usable or the block size
 * exceeds the level's limit. */
deflate_batch_ctx* deflate_batch_new(int level, deflate_format format, size_t block_size, unsigned flags);
void deflate_batch_free(deflate_batch_ctx* ctx);

/* Compress count buffers that live in device memory.  out_sizes[i] receives
 * the framed size of item i, or 0 if it did not fit out_capacity.  Blocks
 * until done.  Returns 0 on success, -1 on failure (orot_gpu_last_error). */
int deflate_batch_compress_device(deflate_batch_ctx* ctx, const void* d_in, void* d_out,
                                  const deflate_batch_item* items, size_t count, uint64_t* out_sizes);

/* Same for host buffers: in[i] (in_sizes[i] bytes) into out[i]
 * (out_capacities[i] bytes). */
int deflate_batch_compress(deflate_batch_ctx* ctx, const void* const* in, const size_t* in_sizes, void* const* out,
                           const size_t* out_capacities, size_t* out_sizes, size_t count);

/* Chunk size, hash width (levels 1-3) and matcher depth (levels 4-12). */
typedef struct deflate_batch_config {
    size_t block_size;
    int hash_bits;
    int max_chain;
    int nice_len;
    int lazy_depth;
    int dictionary; /* levels 1-3: 1 = each chunk sees the 32 KiB before it (not orot's bytes); 0 = off */
} deflate_batch_config;

int deflate_batch_get_config(deflate_batch_ctx* ctx, deflate_batch_config* cfg);
int deflate_batch_set_config(deflate_batch_ctx* ctx, const deflate_batch_config* cfg);

/* Times each candidate configuration on (a prefix of up to 256 MiB of) the
 * given batch and keeps the fastest whose compressed size stays within 1.5%
 * (levels 1-3, whose default is orot's own output) or 0.5% (levels 4-12) of
 * the level's default configuration.  Outputs go to d_out as in
 * deflate_batch_compress_device.  Returns 0 on success. */
int deflate_batch_autotune(deflate_batch_ctx* ctx, const void* d_in, void* d_out, const deflate_batch_item* items,
                           size_t count);

/* Device time of the last batch (see deflate_parallel_stage_times; one
 * value, the whole submission, with DEFLATE_BATCH_GRAPH). */
void deflate_batch_set_timing(deflate_batch_ctx* ctx, int on);
int deflate_batch_stage_times(deflate_batch_ctx* ctx, float* ms, int max_values);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* OROT_DEFLATE_BATCH_H */
