/*
 * orot/deflate_gpu.h -- device-resident extensions of liborot_amd.so.
 *
 * None of these has a counterpart in orot; they exist for callers whose data
 * already lives in GPU memory (benchmarks, pipelines that produce the input on
 * the GPU) and for test harnesses that hold no GPU runtime of their own.
 */
#ifndef OROT_DEFLATE_GPU_H
#define OROT_DEFLATE_GPU_H

#include "deflate.h"

#ifdef __cplusplus
extern "C" {
#endif

/* deflate_parallel_compress with d_in / d_out in device memory.  Same output
 * bytes as the host entry point; blocks until the result size is known. */
size_t deflate_parallel_compress_device(deflate_parallel_ctx* ctx, const void* d_in, size_t in_size, void* d_out,
                                        size_t out_capacity);

/* No counterpart in orot: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same output bytes as deflate_parallel_compress_device on a
 * context of this level, format and block size (0: the level's default),
 * computed without any context and without allocating: the working memory
 * is a fixed 64 MiB device array inside the library, one per GPU, held
 * until the process ends.  Inputs whose working set exceeds it (above
 * 4 MiB at some levels; 8 MiB fits at levels 0-3) return 0 with the limit in
 * orot_gpu_last_error().  Runs on the current device's null stream and
 * blocks until the size is known.  Calls on one GPU share the array: this
 * library serializes them within a process; calls that overlap on one GPU
 * from several processes race.  Returns the compressed size, or 0. */
size_t deflate_parallel_compress_device_noalloc(int level, deflate_format format, size_t block_size, const void* d_in,
                                                size_t in_size, void* d_out, size_t out_capacity);

/* No counterpart in orot's parallel API (orot's lz77_insert_dict,
 * src/deflate/lz77.cpp:43-55, is never called by its parallel path).  At
 * levels 1-3, on = 1 lets each chunk after the first of its buffer match into
 * the 32 KiB before it: the chunk's LZ77 runs with those bytes in its hash
 * table, its tokens cover its own bytes, and back-references may reach into
 * the previous chunk.  Smaller output, not orot's bytes; off by default.
 * Levels 4-12 always use the dictionary.  Returns 0, or -1 for a NULL ctx. */
int deflate_parallel_set_dictionary(deflate_parallel_ctx* ctx, int on);

/* Per-stage device time of the last call on ctx, in milliseconds, measured
 * with HIP events on the context's stream: [0] checksum, [1] LZ77 + block
 * emission, [2] offset scan + framing, [3] assembly.  Timing is recorded
 * only after deflate_parallel_set_timing(ctx, 1).  Returns the number of
 * values written. */
void deflate_parallel_set_timing(deflate_parallel_ctx* ctx, int on);
int deflate_parallel_stage_times(const deflate_parallel_ctx* ctx, float* ms, int max_values);

/* deflate_decompress with d_in / d_out in device memory; blocks until done. */
deflate_result deflate_decompress_device(const void* d_in, size_t in_size, void* d_out, size_t out_capacity,
                                         size_t* actual_out_size, deflate_format format);

/* Batched inflate of `count` independent streams of one format, all in
 * device memory: stream i is d_in[in_off, in_off + in_size) and inflates into
 * d_out[out_off, out_off + out_capacity).  results[i] receives the
 * deflate_decompress result of stream i and out_sizes[i] its size (0 unless
 * DEFLATE_OK); both are host arrays.  Returns 0, or -1 when the batch could
 * not run (message in orot_gpu_last_error()).  In this build the GPU inflate
 * is switched off: both entry points fail after their argument checks. */
typedef struct deflate_inflate_item {
    uint64_t in_off;
    uint64_t in_size;
    uint64_t out_off;
    uint64_t out_capacity;
} deflate_inflate_item;
int deflate_batch_decompress_device(const void* d_in, void* d_out, const deflate_inflate_item* items, size_t count,
                                    deflate_format format, uint64_t* out_sizes, int32_t* results);

/* Device memory helpers (hipMalloc / hipFree / hipMemcpy / sync).
 * kind: 0 host->device, 1 device->host, 2 device->device.  Return 0 on
 * success. */
void* orot_gpu_malloc(size_t bytes);
int orot_gpu_free(void* ptr);
int orot_gpu_memcpy(void* dst, const void* src, size_t bytes, int kind);
int orot_gpu_memset(void* dst, int value, size_t bytes);
int orot_gpu_synchronize(void);

/* Device name and compute-unit count; returns 0 on success. */
int orot_gpu_device_info(char* name, size_t name_len, int* compute_units);

/* Message of the last failure on the calling thread ("" if none). */
const char* orot_gpu_last_error(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* OROT_DEFLATE_GPU_H */
