// server_dev.h -- the compression server's device-resident state
// (server_kernels.hip) and the interface the host worker (server.cpp) drives.
//
// Device memory holds, per server:
//   * one request ring per (level, format) queue: SrvSlot records in arrival
//     order, the head (next request to admit) advanced by the admission
//     kernel; the host appends records and passes the tail as an argument;
//   * the pool allocator: virtual byte offsets [vtail, vhead) are allocated
//     (physical offset = virtual % budget), one SrvRec per admitted request
//     in admission order; released records are reclaimed from the tail;
//   * the release ring: record numbers the client released, applied by the
//     next admission;
//   * counters (requests admitted and completed).
// Each lane (one Engine) gets its batch tables written by the admission
// kernel and its completion records by the completion kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "deflate_gpu.h"

namespace orot_gpu {

constexpr uint32_t SRV_RECS = 1u << 16;       // allocation records (admitted, not yet reclaimed)
constexpr uint32_t SRV_RELS = SRV_RECS;       // release ring
constexpr uint32_t SRV_ADMIT_THREADS = 1024;  // one workgroup
constexpr uint32_t SRV_ADMIT_PER_THREAD = 4;
constexpr uint32_t SRV_MAX_BATCH = SRV_ADMIT_THREADS * SRV_ADMIT_PER_THREAD;  // requests per admission
constexpr uint64_t SRV_ALIGN = 256;

constexpr uint32_t SLOT_HOST_INPUT = 1;  // SrvSlot::flags: the input comes from host memory

// One queued request.
struct SrvSlot {
    uint64_t in_addr;  // device address of the input (host input: unused)
    uint64_t in_size;
    uint64_t cap;      // output capacity
    uint64_t id;       // request id
    uint32_t flags;
    uint32_t pad_;
};

// One pool allocation, in admission order.
struct SrvRec {
    uint64_t vend;  // virtual end of the allocation (and of any skip before it)
    uint32_t live;  // 0 once released
    uint32_t pad_;
};

// Allocator state and counters.
struct SrvPool {
    uint64_t budget;
    uint64_t vhead;
    uint64_t vtail;
    uint32_t rec_head;  // records [rec_tail, rec_head) are not yet reclaimed (mod SRV_RECS)
    uint32_t rec_tail;
    uint32_t rel_head;  // releases applied
    uint32_t pad_;
    uint64_t admitted;
    uint64_t completed;
};

// Result of one admission, and per admitted request its placement and, once
// the batch has run, its size.  Written by the kernels into pinned host
// memory, read by the host after the kernel's stream reaches an event.
struct LaneHead {
    uint32_t count;     // requests admitted
    uint32_t nchunks;   // chunks of the batch
    uint64_t first_seq; // ring sequence number of the first one (the rest follow in order)
    uint64_t pending;   // requests still queued after this admission
    uint64_t in_use;    // pool bytes between vtail and vhead after it
};
struct LaneItem {
    uint64_t id;
    uint64_t out_off;   // pool offset of the output
    uint64_t in_off;    // pool offset of a host input's copy
    uint64_t out_size;  // framed size, 0 when it did not fit
    uint32_t rec;       // allocation record number (for release)
    int32_t status;     // 0 ok, -1 failed
};

struct AdmitArgs {
    SrvPool* pool;
    SrvRec* recs;
    const uint32_t* rels;  // release ring (SRV_RELS)
    uint32_t rel_tail;     // releases pushed so far
    const SrvSlot* ring;   // this queue's ring (nullptr: apply releases only)
    uint64_t* head;
    uint64_t tail;         // requests pushed so far
    uint32_t ring_cap;     // power of two
    uint32_t chunk;        // the lane's block size
    BatchHdr* hdr;         // the lane's batch tables
    BufDesc* bufs;
    ChunkDesc* desc;
    uint32_t* order;
    uint32_t buf_cap;      // at most SRV_MAX_BATCH
    uint32_t chunk_cap;
    uint64_t max_bytes;    // input bytes per batch (the first request is always taken)
    uint64_t pool_addr;    // device address of the pool: base of every in_off
    LaneHead* lane;
    LaneItem* items;
};

hipError_t launch_srv_admit(const AdmitArgs& a, hipStream_t s);
hipError_t launch_srv_complete(const BatchHdr* hdr, const Result* res, LaneItem* items, SrvPool* pool,
                               hipStream_t s);

}  // namespace orot_gpu
