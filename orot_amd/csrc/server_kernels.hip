// server_kernels.hip -- admission and completion kernels of the compression
// server (server_dev.h holds the state they work on).
//
// srv_admit_kernel (one workgroup, on the server's control stream):
//   1. marks the records the client released (release ring) dead;
//   2. reclaims dead records from the tail of the allocation ring, which
//      advances vtail; an empty pool restarts at a physical offset of 0;
//   3. reads the next requests of one queue (at most 4096) and takes the
//      longest prefix that fits: the pool (an allocation never straddles the
//      end of the pool: the first request that would is moved to offset 0,
//      and the bytes it skips are reclaimed with it), the batch byte and
//      request limits, and the lane's chunk capacity;
//   4. writes the lane's batch tables (buffer and chunk descriptors, queue
//      order, header) and the placement of each admitted request.
// All of it is prefix sums and minimum reductions over the candidates; the
// sequential first-fit of a host allocator has no place here.
//
// srv_complete_kernel (on the lane's stream, after the pipeline) copies each
// buffer's framed size into the lane's completion records.
#include <hip/hip_runtime.h>

#include "server_dev.h"

namespace orot_gpu {

namespace {

constexpr uint32_t T = SRV_ADMIT_THREADS;
constexpr uint32_t IPT = SRV_ADMIT_PER_THREAD;

__device__ __forceinline__ uint64_t align_srv(uint64_t v) { return (v + SRV_ALIGN - 1) & ~(SRV_ALIGN - 1); }

// Exclusive prefix sum of one value per thread over the workgroup
// (Hillis-Steele over two LDS buffers); *total receives the sum.
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* buf, uint64_t* total) {
    const uint32_t t = threadIdx.x;
    uint64_t* a = buf;
    uint64_t* b = buf + T;
    a[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < T; d <<= 1) {
        b[t] = a[t] + (t >= d ? a[t - d] : 0ull);
        __syncthreads();
        uint64_t* x = a;
        a = b;
        b = x;
    }
    const uint64_t incl = a[t];
    *total = a[T - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(T) void srv_admit_kernel(AdmitArgs a) {
    __shared__ uint64_t scan[2 * T];
    __shared__ uint32_t s_min;
    __shared__ uint32_t s_n;
    __shared__ uint64_t s_vhead, s_vtail, s_head, s_skip, s_end, s_full, s_chunks;
    __shared__ uint32_t s_rec_tail;
    const uint32_t t = threadIdx.x;
    SrvPool* P = a.pool;
    const uint64_t B = P->budget;
    const uint32_t rel_head = P->rel_head;
    const uint32_t rt = P->rec_tail;
    const uint32_t rh = P->rec_head;

    // 1. releases
    const uint32_t nrel = a.rel_tail - rel_head;
    for (uint32_t k = t; k < nrel; k += T) a.recs[a.rels[(rel_head + k) % SRV_RELS] % SRV_RECS].live = 0;
    if (t == 0) s_min = ~0u;
    __syncthreads();

    // 2. the first live record from the tail
    const uint32_t span = rh - rt;
    for (uint32_t base = 0; base < span; base += T) {
        const uint32_t k = base + t;
        if (k < span && a.recs[(rt + k) % SRV_RECS].live) atomicMin(&s_min, k);
        __syncthreads();
        const uint32_t f = s_min;
        __syncthreads();
        if (f != ~0u) break;
    }
    if (t == 0) {
        uint64_t vh = P->vhead, vt = P->vtail;
        const uint32_t first = s_min == ~0u ? span : s_min;
        if (first == span) {
            vh = (vh + B - 1) / B * B;
            vt = vh;
        } else if (first > 0) {
            vt = a.recs[(rt + first - 1) % SRV_RECS].vend;
        }
        s_vhead = vh;
        s_vtail = vt;
        s_rec_tail = rt + first;
        s_head = a.ring ? *a.head : 0ull;
        s_min = ~0u;
        s_n = ~0u;
        s_skip = 0;
    }
    __syncthreads();
    const uint64_t vh = s_vhead, vt = s_vtail, head = s_head;
    const uint64_t pending = a.ring ? a.tail - head : 0ull;
    const uint32_t W = (uint32_t)(pending < a.buf_cap ? pending : a.buf_cap);

    // 3. candidates: thread t holds requests t*IPT .. t*IPT+IPT-1
    SrvSlot sl[IPT];
    uint64_t need[IPT];
    uint32_t nch[IPT], full[IPT];
    uint64_t sum_need = 0, sum_in = 0, sum_ch = 0, sum_full = 0;
    for (uint32_t q = 0; q < IPT; ++q) {
        const uint32_t j = t * IPT + q;
        need[q] = 0;
        nch[q] = full[q] = 0;
        sl[q] = SrvSlot{};
        if (j < W) {
            sl[q] = a.ring[(head + j) & (a.ring_cap - 1)];
            const uint64_t n = sl[q].in_size;
            need[q] = align_srv(sl[q].cap ? sl[q].cap : 1) +
                      ((sl[q].flags & SLOT_HOST_INPUT) ? align_srv(n ? n : 1) : 0ull);
            nch[q] = n ? (uint32_t)((n + a.chunk - 1) / a.chunk) : 1u;
            full[q] = (uint32_t)(n / a.chunk);
        }
        sum_need += need[q];
        sum_in += sl[q].in_size;
        sum_ch += nch[q];
        sum_full += full[q];
    }
    uint64_t tot;
    uint64_t p_need = block_exclusive_scan(sum_need, scan, &tot);
    uint64_t p_in = block_exclusive_scan(sum_in, scan, &tot);
    uint64_t p_ch = block_exclusive_scan(sum_ch, scan, &tot);
    uint64_t p_full = block_exclusive_scan(sum_full, scan, &tot);

    // 4. placement: the first request that would straddle the end of the pool
    // moves to the next physical 0, and every later one with it
    uint64_t start[IPT];
    {
        uint64_t s = vh + p_need;
        for (uint32_t q = 0; q < IPT; ++q) {
            start[q] = s;
            const uint32_t j = t * IPT + q;
            if (j < W && (s % B) + need[q] > B) atomicMin(&s_min, j);
            s += need[q];
        }
    }
    __syncthreads();
    const uint32_t cross = s_min;
    for (uint32_t q = 0; q < IPT; ++q)
        if (t * IPT + q == cross) s_skip = B - (start[q] % B);
    __syncthreads();
    const uint64_t skip = s_skip;
    const uint32_t recs_used = rh - s_rec_tail;
    {
        uint64_t pin = p_in, pch = p_ch;
        for (uint32_t q = 0; q < IPT; ++q) {
            const uint32_t j = t * IPT + q;
            if (j >= cross) start[q] += skip;
            const uint64_t end = start[q] + need[q];
            const bool ok = end - vt <= B && (j == 0 || pin + sl[q].in_size <= a.max_bytes) &&
                            pch + nch[q] <= a.chunk_cap && (uint64_t)recs_used + j + 1 <= SRV_RECS;
            if (j < W && !ok) atomicMin(&s_n, j);
            pin += sl[q].in_size;
            pch += nch[q];
        }
    }
    __syncthreads();
    const uint32_t n = s_n < W ? s_n : W;
    {
        uint64_t pch = p_ch, pf = p_full;
        for (uint32_t q = 0; q < IPT; ++q) {
            const uint32_t j = t * IPT + q;
            if (j + 1 == n) {
                s_end = start[q] + need[q];
                s_full = pf + full[q];
                s_chunks = pch + nch[q];
            }
            pch += nch[q];
            pf += full[q];
        }
    }
    __syncthreads();

    // 5. tables and placements of the admitted requests
    const uint64_t F = n ? s_full : 0ull;
    {
        uint64_t pch = p_ch, pf = p_full;
        for (uint32_t q = 0; q < IPT; ++q) {
            const uint32_t j = t * IPT + q;
            if (j < n) {
                const SrvSlot& r = sl[q];
                const bool host = (r.flags & SLOT_HOST_INPUT) != 0;
                const uint64_t out_off = start[q] % B;
                const uint64_t pool_in = out_off + align_srv(r.cap ? r.cap : 1);
                // offsets from the pool base, modulo 2^64 for inputs below it
                const uint64_t in_off = host ? pool_in : r.in_addr - a.pool_addr;
                BufDesc bd;
                bd.in_off = in_off;
                bd.in_len = r.in_size;
                bd.out_off = out_off;
                bd.cap = r.cap;
                bd.chunk0 = (uint32_t)pch;
                bd.nchunks = nch[q];
                bd.piece = 0;
                bd.hist = 0;
                a.bufs[j] = bd;
                for (uint32_t c = 0; c < nch[q]; ++c) {
                    ChunkDesc cd;
                    const uint64_t off = (uint64_t)c * a.chunk;
                    const uint64_t left = r.in_size - off;
                    cd.in_off = in_off + off;
                    cd.len = (uint32_t)(r.in_size == 0 ? 0ull : (left < a.chunk ? left : a.chunk));
                    cd.buf = j | (c + 1 == nch[q] ? CHUNK_LAST : 0u);
                    a.desc[pch + c] = cd;
                    // queue order: full chunks first, then the short last ones
                    if (c < full[q])
                        a.order[pf + c] = (uint32_t)(pch + c);
                    else
                        a.order[F + (pch - pf)] = (uint32_t)(pch + c);
                }
                a.recs[(rh + j) % SRV_RECS] = SrvRec{start[q] + need[q], 1u, 0u};
                LaneItem it;
                it.id = r.id;
                it.out_off = out_off;
                it.in_off = host ? pool_in : ~0ull;
                it.out_size = 0;
                it.rec = rh + j;
                it.status = 0;
                a.items[j] = it;
            }
            pch += nch[q];
            pf += full[q];
        }
    }

    // 6. commit
    if (t == 0) {
        const uint64_t chunks = n ? s_chunks : 0ull;
        const uint64_t new_head = n ? s_end : vh;
        P->vhead = new_head;
        P->vtail = vt;
        P->rec_tail = s_rec_tail;
        P->rec_head = rh + n;
        P->rel_head = a.rel_tail;
        P->admitted += n;
        if (a.ring) *a.head = head + n;
        if (a.hdr) {
            BatchHdr h{};
            h.nchunks = (uint32_t)chunks;
            h.nbufs = n;
            *a.hdr = h;
        }
        LaneHead lh;
        lh.count = n;
        lh.nchunks = (uint32_t)chunks;
        lh.first_seq = head;
        lh.pending = pending - n;
        lh.in_use = new_head - vt;
        *a.lane = lh;
    }
}

__global__ __launch_bounds__(256) void srv_complete_kernel(const BatchHdr* __restrict__ hdr,
                                                           const Result* __restrict__ res, LaneItem* items,
                                                           SrvPool* pool) {
    const uint32_t n = hdr->nbufs;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const uint64_t sz = res[j].total_bytes;
        items[j].out_size = sz;
        items[j].status = sz ? 0 : -1;
    }
    if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&pool->completed), (unsigned long long)n);
}

}  // namespace

hipError_t launch_srv_admit(const AdmitArgs& a, hipStream_t s) {
    if (a.buf_cap > SRV_MAX_BATCH || (a.ring && (a.ring_cap & (a.ring_cap - 1)) != 0) || (a.ring && !a.hdr))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(srv_admit_kernel, dim3(1), dim3(T), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_srv_complete(const BatchHdr* hdr, const Result* res, LaneItem* items, SrvPool* pool,
                               hipStream_t s) {
    hipLaunchKernelGGL(srv_complete_kernel, dim3(1), dim3(256), 0, s, hdr, res, items, pool);
    return hipGetLastError();
}

}  // namespace orot_gpu
