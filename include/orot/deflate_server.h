/*
 * orot/deflate_server.h -- a compression server on one MI355X.
 *
 * No counterpart in orot.  Requests arrive one at a time (submit returns at
 * once).  The queues (one ring per level and format), the pool of
 * `hbm_budget` bytes with its allocator, and the admission counters live in
 * device memory: whenever a lane (one HIP stream and its pipeline, per level
 * and format) is idle, an admission kernel takes the longest run of that
 * queue's oldest requests that fits the pool and the batch limits, places
 * their buffers in the pool and writes the lane's batch tables; the pipeline
 * then runs on the lane's stream.  Up to `lanes_per_level` batches per level
 * and format are on the device at once; requests that arrive while others
 * run join a later batch of their level: continuous batching.  Each output is
 * the bytes deflate_parallel_compress gives for that input alone.
 */
#ifndef OROT_DEFLATE_SERVER_H
#define OROT_DEFLATE_SERVER_H

#include "deflate.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct deflate_server deflate_server;

typedef struct deflate_server_config {
    size_t hbm_budget;         /* device bytes for request buffers (host inputs + outputs); 0: 32 GiB */
    size_t max_batch_bytes;    /* input bytes per batch; 0: 256 MiB */
    int max_batch_requests;    /* 0: 4096 */
    int lanes_per_level;       /* batches of one level in flight at once; 0: 2 */
    size_t lane_arena_bytes;   /* device working memory of one lane; it bounds the
                                  blocks of one batch (and of one request); 0: 2 GiB */
} deflate_server_config;

/* Completion record. */
typedef struct deflate_server_completion {
    uint64_t id;
    uint64_t in_size;
    uint64_t out_size;         /* framed size; 0 on failure */
    int32_t status;            /* 0 ok, -1 did not fit / failed */
    int32_t level;
    uint64_t submit_ns;        /* server clock (steady, ns since deflate_server_new) */
    uint64_t launch_ns;        /* its batch was enqueued */
    uint64_t done_ns;          /* the server saw its batch finish */
} deflate_server_completion;

deflate_server* deflate_server_new(const deflate_server_config* cfg);
/* Waits for every request in flight and runs the queued ones that fit, then
 * frees the server and its pool.  Queued requests that cannot be admitted
 * (the pool is held by outputs not yet released) are dropped. */
void deflate_server_free(deflate_server* srv);

/* Queue one request.  d_in: input already in device memory (it must stay
 * valid until the request completes).  Returns its id (> 0), or 0 when the
 * request can never fit (the budget, or a lane's blocks) or the arguments are
 * invalid. */
uint64_t deflate_server_submit_device(deflate_server* srv, const void* d_in, size_t in_size, int level,
                                      deflate_format format);
/* Same with host input: copied into the server's pool when admitted, where
 * the copy stays until release (the host buffer must stay valid until the
 * request completes). */
uint64_t deflate_server_submit(deflate_server* srv, const void* in, size_t in_size, int level,
                               deflate_format format);

/* Up to max completions, waiting up to timeout_ms for the first one.
 * Returns the number written. */
int deflate_server_poll(deflate_server* srv, deflate_server_completion* out, int max, int timeout_ms);

/* Device address of a completed request's output (valid until release). */
const void* deflate_server_output(deflate_server* srv, uint64_t id);
/* Copy a completed request's output to host memory; returns bytes or 0. */
size_t deflate_server_read(deflate_server* srv, uint64_t id, void* out, size_t out_capacity);
/* Return a completed request's memory to the pool. */
int deflate_server_release(deflate_server* srv, uint64_t id);

/* Counters: [0] submitted, [1] completed, [2] queued (not yet admitted),
 * [3] in flight, [4] pool bytes in use (read from the device allocator after
 * the releases so far are applied), [5] batches launched, [6] largest queue
 * seen.  Returns the number written. */
int deflate_server_stats(deflate_server* srv, uint64_t* values, int max_values);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* OROT_DEFLATE_SERVER_H */
