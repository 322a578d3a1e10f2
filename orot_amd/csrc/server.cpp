// server.cpp -- continuous-batching compression server (deflate_server.h).
//
// The queues, the pool allocator and the admission live in device memory
// (server_dev.h, server_kernels.hip).  API threads only append to the
// worker's inboxes (submit, release) and read completed records; the worker
// thread, which holds the server's mutex only to exchange those lists,
// repeats:
//   push   -- new requests go to their queue's ring in device memory and new
//             releases to the release ring (copies from pinned staging on
//             the control stream);
//   admit  -- for each queue with requests and an idle lane the admission
//             kernel takes the longest run of its oldest requests that fits
//             the pool and the batch limits and writes the lane's tables;
//   launch -- host inputs are copied to the pool offsets the admission chose,
//             then the lane's pipeline and completion kernel run on the
//             lane's own stream;
//   retire -- lanes whose batch finished hand their requests to the
//             completion queue.
// Lanes of different levels and formats, and several lanes of one, run on
// the GPU at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/orot/deflate_server.h"
#include "engine.h"
#include "host_common.h"
#include "server_dev.h"

using namespace orot_gpu;

namespace {

constexpr uint32_t RING = 8192;          // requests per queue ring
constexpr uint32_t STAGE_SLOTS = RING;   // new requests pushed per round
constexpr int NKEYS = 13 * 4;            // level * 4 + format

uint64_t align_up(uint64_t v) { return (v + SRV_ALIGN - 1) & ~(SRV_ALIGN - 1); }

struct Request {
    uint64_t id = 0;
    int level = 1;
    int format = 0;
    const uint8_t* d_in = nullptr;  // device input (caller's)
    const void* h_in = nullptr;     // host input, copied at admission
    uint64_t in_size = 0;
    uint64_t cap = 0;
    uint64_t need = 0;              // pool bytes
    uint64_t out_off = 0;
    uint64_t in_off = 0;
    uint32_t rec = 0;               // device allocation record
    uint64_t out_size = 0;
    int status = 0;
    bool done = false;
    uint64_t t_submit = 0, t_launch = 0, t_done = 0;
};

// Device address of pinned host memory the kernels write.
template <class T>
T* device_view(T* h) {
    void* d = nullptr;
    return hipHostGetDevicePointer(&d, h, 0) == hipSuccess ? static_cast<T*>(d) : h;
}

template <class T>
bool pinned_alloc(T*& p, size_t n) {
    p = nullptr;
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T), 0);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipHostMalloc(server)", e);
        return false;
    }
    return true;
}

struct Lane {
    Engine eng;
    Engine::DeviceTables tabs{};
    LaneHead* h_head = nullptr;   // pinned: written by the admission kernel
    LaneItem* h_items = nullptr;  // pinned: admission, then completion
    hipEvent_t ev_done = nullptr;
    bool busy = false;
    std::vector<Request*> reqs;
    ~Lane() {
        if (eng.stream()) (void)hipStreamSynchronize(eng.stream());
        if (ev_done) (void)hipEventDestroy(ev_done);
        if (h_head) (void)hipHostFree(h_head);
        if (h_items) (void)hipHostFree(h_items);
    }
};

struct Key {
    EngineConfig cfg{};          // resolved
    uint32_t chunk_cap = 0;      // chunks one batch (and one request) may have
    uint32_t buf_cap = 0;        // requests one batch may have
    SrvSlot* d_ring = nullptr;
    uint64_t pushed = 0;         // requests copied to the device ring
    uint64_t seen = 0;           // requests the host has seen admitted
    bool stalled = false;        // the last admission took nothing: wait for a release
    std::deque<Request*> backlog;  // submitted, not yet in the ring
    std::vector<Request*> by_seq;  // the request at each ring position
    std::vector<std::unique_ptr<Lane>> lanes;
};

uint64_t out_capacity(uint64_t n, uint32_t chunk, int format) {
    const uint64_t nchunks = n ? (n + chunk - 1) / chunk : 1;
    const uint64_t bound = deflate_compress_bound(n, (deflate_format)format) + 64;
    return std::max(bound, n + 6 * (nchunks + n / 65535 + 2) + 64);
}

}  // namespace

struct deflate_server {
    deflate_server_config cfg{};
    int device = 0;
    uint8_t* d_pool = nullptr;
    uint8_t* d_state = nullptr;   // SrvPool | SrvRec[SRV_RECS] | rels[SRV_RELS] | heads[NKEYS]
    SrvPool* d_pool_state = nullptr;
    SrvRec* d_recs = nullptr;
    uint32_t* d_rels = nullptr;
    uint64_t* d_heads = nullptr;
    hipStream_t control = nullptr;
    // pinned staging of pushes, double-buffered
    SrvSlot* h_slots[2] = {};
    uint32_t* h_rels[2] = {};
    hipEvent_t ev_stage[2] = {};
    int stage_cur = 0;
    LaneHead* h_stats = nullptr;
    uint32_t rel_tail = 0;
    Key keys[NKEYS];

    // shared with the API threads (mu)
    std::mutex mu;
    std::condition_variable cv_work;
    std::condition_variable cv_done;
    std::vector<Request*> inbox;
    std::vector<uint32_t> rel_inbox;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> live;
    std::deque<Request*> completed;
    uint64_t next_id = 1;
    uint64_t n_submitted = 0, n_completed = 0, n_inflight = 0, n_admitted = 0, n_batches = 0, max_queue = 0;
    uint64_t pool_in_use = 0;
    uint64_t stats_req = 0, stats_done = 0;
    bool stop = false;
    bool worker_exited = false;
    uint64_t wake = 0;  // bumped by submit / release / stats / stop
    std::thread worker;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    uint64_t now() const {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
            .count();
    }

    ~deflate_server() {
        for (auto& k : keys) {
            k.lanes.clear();
            if (k.d_ring) (void)hipFree(k.d_ring);
        }
        if (control) (void)hipStreamSynchronize(control);
        for (int i = 0; i < 2; ++i) {
            if (ev_stage[i]) (void)hipEventDestroy(ev_stage[i]);
            if (h_slots[i]) (void)hipHostFree(h_slots[i]);
            if (h_rels[i]) (void)hipHostFree(h_rels[i]);
        }
        if (h_stats) (void)hipHostFree(h_stats);
        if (control) (void)hipStreamDestroy(control);
        if (d_state) (void)hipFree(d_state);
        if (d_pool) (void)hipFree(d_pool);
    }

    // Device state, staging and the per-queue limits (caller's thread).
    bool init() {
        for (int key = 0; key < NKEYS; ++key) {
            Key& k = keys[key];
            k.cfg.level = key / 4;
            k.cfg.format = key % 4;
            if (k.cfg.format == 3 || !resolve_config(k.cfg)) continue;
            k.buf_cap = (uint32_t)std::min<uint64_t>({(uint64_t)cfg.max_batch_requests, SRV_MAX_BATCH, RING});
            const uint64_t per = Engine::arena_bytes_per_chunk(k.cfg);
            const uint64_t by_mem = std::max<uint64_t>(cfg.lane_arena_bytes / per, 1);
            const uint64_t by_batch = (cfg.max_batch_bytes + k.cfg.chunk - 1) / k.cfg.chunk + k.buf_cap;
            k.chunk_cap = (uint32_t)std::min<uint64_t>({by_mem, by_batch, 0x3FFFFFFFull});
        }
        const size_t state_bytes = sizeof(SrvPool) + (size_t)SRV_RECS * sizeof(SrvRec) + (size_t)SRV_RELS * 4 +
                                   (size_t)NKEYS * 8;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_pool), cfg.hbm_budget);
        if (e != hipSuccess) {
            d_pool = nullptr;
            set_error("hipMalloc(server pool)", e);
            return false;
        }
        e = hipMalloc(reinterpret_cast<void**>(&d_state), state_bytes);
        if (e != hipSuccess) {
            d_state = nullptr;
            set_error("hipMalloc(server state)", e);
            return false;
        }
        d_pool_state = reinterpret_cast<SrvPool*>(d_state);
        d_recs = reinterpret_cast<SrvRec*>(d_state + sizeof(SrvPool));
        d_rels = reinterpret_cast<uint32_t*>(d_state + sizeof(SrvPool) + (size_t)SRV_RECS * sizeof(SrvRec));
        d_heads = reinterpret_cast<uint64_t*>(d_state + sizeof(SrvPool) + (size_t)SRV_RECS * sizeof(SrvRec) +
                                              (size_t)SRV_RELS * 4);
        SrvPool init{};
        init.budget = cfg.hbm_budget;
        e = hipMemset(d_state, 0, state_bytes);
        if (e == hipSuccess) e = hipMemcpy(d_pool_state, &init, sizeof(init), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&control, hipStreamNonBlocking);
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&ev_stage[i], hipEventDisableTiming);
        if (e != hipSuccess) {
            set_error("server init", e);
            return false;
        }
        for (int i = 0; i < 2; ++i)
            if (!pinned_alloc(h_slots[i], STAGE_SLOTS) || !pinned_alloc(h_rels[i], SRV_RELS)) return false;
        return pinned_alloc(h_stats, 1);
    }

    // ---- worker side -----------------------------------------------------

    Lane* idle_lane(Key& k) {
        for (auto& lp : k.lanes)
            if (!lp->busy) return lp.get();
        if ((int)k.lanes.size() >= cfg.lanes_per_level) return nullptr;
        auto lp = std::make_unique<Lane>();
        if (!lp->eng.init(k.cfg) || !lp->eng.prepare_device_tables(k.chunk_cap, k.buf_cap, &lp->tabs) ||
            !pinned_alloc(lp->h_head, 1) || !pinned_alloc(lp->h_items, k.buf_cap) ||
            hipEventCreateWithFlags(&lp->ev_done, hipEventDisableTiming) != hipSuccess)
            return nullptr;
        k.lanes.push_back(std::move(lp));
        return k.lanes.back().get();
    }

    // Copies new requests into their rings and new releases into the release
    // ring (control stream).  Returns whether anything was pushed.
    bool push(std::vector<uint32_t>& rels) {
        bool work = !rels.empty();
        for (const Key& k : keys) work = work || (!k.backlog.empty() && k.pushed - k.seen < RING);
        if (!work) return false;
        const int cur = stage_cur;
        (void)hipEventSynchronize(ev_stage[cur]);  // the staging's previous copies are done
        SrvSlot* st = h_slots[cur];
        uint32_t used = 0;
        hipError_t e = hipSuccess;
        for (Key& k : keys) {
            if (k.backlog.empty()) continue;
            if (!k.d_ring) {
                e = hipMalloc(reinterpret_cast<void**>(&k.d_ring), (size_t)RING * sizeof(SrvSlot));
                if (e != hipSuccess) {
                    k.d_ring = nullptr;
                    break;
                }
                k.by_seq.assign(RING, nullptr);
            }
            const uint64_t first = k.pushed;
            const uint32_t base = used;
            while (!k.backlog.empty() && used < STAGE_SLOTS && k.pushed - k.seen < RING) {
                Request* r = k.backlog.front();
                k.backlog.pop_front();
                SrvSlot s{};
                s.in_addr = (uint64_t)(uintptr_t)r->d_in;
                s.in_size = r->in_size;
                s.cap = r->cap;
                s.id = r->id;
                s.flags = r->h_in ? SLOT_HOST_INPUT : 0u;
                st[used++] = s;
                k.by_seq[k.pushed % RING] = r;
                ++k.pushed;
            }
            // the new range [first, pushed) of the ring, in at most two pieces
            uint64_t seq = first;
            uint32_t src = base;
            while (seq < k.pushed && e == hipSuccess) {
                const uint32_t pos = (uint32_t)(seq % RING);
                const uint32_t m = (uint32_t)std::min<uint64_t>(k.pushed - seq, RING - pos);
                e = hipMemcpyAsync(k.d_ring + pos, st + src, (size_t)m * sizeof(SrvSlot), hipMemcpyHostToDevice,
                                   control);
                seq += m;
                src += m;
            }
            if (e != hipSuccess) break;
        }
        uint32_t nrel = 0;
        if (e == hipSuccess && !rels.empty()) {
            nrel = (uint32_t)std::min<size_t>(rels.size(), SRV_RELS);
            std::copy(rels.begin(), rels.begin() + nrel, h_rels[cur]);
            uint32_t done = 0;
            while (done < nrel && e == hipSuccess) {
                const uint32_t pos = (rel_tail + done) % SRV_RELS;
                const uint32_t m = std::min(nrel - done, SRV_RELS - pos);
                e = hipMemcpyAsync(d_rels + pos, h_rels[cur] + done, (size_t)m * 4, hipMemcpyHostToDevice, control);
                done += m;
            }
            rel_tail += nrel;
            rels.erase(rels.begin(), rels.begin() + nrel);
        }
        if (e == hipSuccess) e = hipEventRecord(ev_stage[cur], control);
        stage_cur ^= 1;
        if (e != hipSuccess) set_error("server push", e);
        return used > 0 || nrel > 0;
    }

    AdmitArgs admit_args(Key& k, int key, Lane* L) {
        AdmitArgs a{};
        a.pool = d_pool_state;
        a.recs = d_recs;
        a.rels = d_rels;
        a.rel_tail = rel_tail;
        a.ring = k.d_ring;
        a.head = d_heads + key;
        a.tail = k.pushed;
        a.ring_cap = RING;
        a.chunk = k.cfg.chunk;
        a.hdr = L->tabs.hdr;
        a.bufs = L->tabs.bufs;
        a.desc = L->tabs.desc;
        a.order = L->tabs.order;
        a.buf_cap = k.buf_cap;
        a.chunk_cap = k.chunk_cap;
        a.max_bytes = cfg.max_batch_bytes;
        a.pool_addr = (uint64_t)(uintptr_t)d_pool;
        a.lane = device_view(L->h_head);
        a.items = device_view(L->h_items);
        return a;
    }

    // Completes a lane's requests (status from the completion records, or
    // failed) and hands them to the completion queue.
    void finish(Lane& L, bool failed) {
        const uint64_t t = now();
        for (size_t i = 0; i < L.reqs.size(); ++i) {
            Request* r = L.reqs[i];
            r->out_size = failed ? 0 : L.h_items[i].out_size;
            r->status = (!failed && L.h_items[i].status == 0 && r->out_size) ? 0 : -1;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Request* r : L.reqs) {
                r->done = true;
                r->t_done = t;
                completed.push_back(r);
            }
            n_completed += L.reqs.size();
            n_inflight -= L.reqs.size();
        }
        cv_done.notify_all();
        L.reqs.clear();
        L.busy = false;
    }

    bool retire() {
        bool any = false;
        for (Key& k : keys)
            for (auto& lp : k.lanes) {
                Lane& L = *lp;
                if (!L.busy) continue;
                const hipError_t q = hipEventQuery(L.ev_done);
                if (q == hipErrorNotReady) continue;
                finish(L, q != hipSuccess);
                any = true;
            }
        return any;
    }

    // One round of admissions: every queue with requests and an idle lane.
    // Returns the number of requests admitted.
    uint64_t admit_round() {
        std::vector<std::pair<Key*, Lane*>> started;
        for (int key = 0; key < NKEYS; ++key) {
            Key& k = keys[key];
            if (k.pushed == k.seen || k.stalled) continue;
            Lane* L = idle_lane(k);
            if (!L) continue;
            if (launch_srv_admit(admit_args(k, key, L), control) != hipSuccess) continue;
            L->busy = true;  // reserved until the admission is read
            started.push_back({&k, L});
        }
        if (started.empty()) return 0;
        const uint64_t t = now();
        const hipError_t se = hipStreamSynchronize(control);
        uint64_t admitted = 0, batches = 0;
        for (auto& [kp, L] : started) {
            Key& k = *kp;
            L->busy = false;
            if (se != hipSuccess) continue;
            const LaneHead h = *L->h_head;
            {
                std::lock_guard<std::mutex> lk(mu);
                pool_in_use = h.in_use;
            }
            if (h.count == 0) {
                k.stalled = h.pending > 0;
                continue;
            }
            L->reqs.resize(h.count);
            bool ok = h.first_seq == k.seen;
            for (uint32_t i = 0; i < h.count && ok; ++i) {
                Request* r = k.by_seq[(h.first_seq + i) % RING];
                ok = r && r->id == L->h_items[i].id;
                L->reqs[i] = r;
                if (!ok) break;
                r->out_off = L->h_items[i].out_off;
                r->in_off = L->h_items[i].in_off;
                r->rec = L->h_items[i].rec;
                r->t_launch = t;
            }
            k.seen += h.count;
            admitted += h.count;
            ++batches;
            {
                std::lock_guard<std::mutex> lk(mu);
                n_admitted += h.count;
                n_inflight += h.count;
                ++n_batches;
            }
            L->busy = true;
            hipError_t e = ok ? hipSuccess : hipErrorInvalidValue;
            if (!ok) set_error_msg("server: admission records do not match the queue");
            const hipStream_t s = L->eng.stream();
            for (Request* r : L->reqs)
                if (e == hipSuccess && r->h_in && r->in_size)
                    e = hipMemcpyAsync(d_pool + r->in_off, r->h_in, r->in_size, hipMemcpyHostToDevice, s);
            if (e == hipSuccess && !L->eng.launch_device_tables(d_pool, d_pool, h.nchunks, h.count))
                e = hipErrorLaunchFailure;
            if (e == hipSuccess)
                e = launch_srv_complete(L->tabs.hdr, L->tabs.results, device_view(L->h_items), d_pool_state, s);
            if (e == hipSuccess) e = hipEventRecord(L->ev_done, s);
            if (e != hipSuccess) {
                // the batch is not (wholly) enqueued: finish it as failed once the stream is idle
                (void)hipStreamSynchronize(s);
                if (ok) set_error("server launch", e);
                finish(*L, true);
            }
        }
        (void)batches;
        return admitted;
    }

    void refresh_stats() {
        AdmitArgs a{};
        a.pool = d_pool_state;
        a.recs = d_recs;
        a.rels = d_rels;
        a.rel_tail = rel_tail;
        a.lane = device_view(h_stats);
        if (launch_srv_admit(a, control) == hipSuccess && hipStreamSynchronize(control) == hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            pool_in_use = h_stats->in_use;
        }
    }

    bool any_busy() const {
        for (const Key& k : keys)
            for (auto& lp : k.lanes)
                if (lp->busy) return true;
        return false;
    }

    bool all_admitted_or_stalled() const {
        for (const Key& k : keys)
            if (!k.stalled && (!k.backlog.empty() || k.pushed != k.seen)) return false;
        return true;
    }

    void run() {
        (void)hipSetDevice(device);
        std::vector<Request*> new_reqs;
        std::vector<uint32_t> new_rels;
        uint64_t stats_served = 0;
        for (;;) {
            uint64_t seen_wake, stats_ticket;
            bool stopping;
            {
                std::lock_guard<std::mutex> lk(mu);
                new_reqs.insert(new_reqs.end(), inbox.begin(), inbox.end());
                inbox.clear();
                new_rels.insert(new_rels.end(), rel_inbox.begin(), rel_inbox.end());
                rel_inbox.clear();
                stopping = stop;
                stats_ticket = stats_req;
                seen_wake = wake;
            }
            for (Request* r : new_reqs) keys[r->level * 4 + r->format].backlog.push_back(r);
            new_reqs.clear();
            if (!new_rels.empty())
                for (Key& k : keys) k.stalled = false;
            bool progress = push(new_rels);
            progress = retire() || progress;
            while (admit_round() > 0) progress = true;
            if (stats_ticket > stats_served) {
                refresh_stats();
                stats_served = stats_ticket;
                std::lock_guard<std::mutex> lk(mu);
                stats_done = std::max(stats_done, stats_ticket);
                cv_done.notify_all();
            }
            const bool busy = any_busy();
            if (stopping && !busy && new_rels.empty() && all_admitted_or_stalled()) break;
            std::unique_lock<std::mutex> lk(mu);
            if (busy) {
                // a batch is on the GPU: look again shortly, or at once on news
                cv_work.wait_for(lk, std::chrono::microseconds(50), [&] { return wake != seen_wake; });
            } else if (!progress && new_rels.empty()) {
                cv_work.wait(lk, [&] { return stop || wake != seen_wake; });
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        worker_exited = true;
        cv_done.notify_all();
    }

    // ---- API side --------------------------------------------------------

    uint64_t submit(const void* d_in, const void* h_in, uint64_t n, int level, int format) {
        if (level < 0) level = 0;
        if (level > 12) level = 12;
        if (format != 1 && format != 2) format = 0;
        const Key& k = keys[level * 4 + format];
        auto r = std::make_unique<Request>();
        r->level = level;
        r->format = format;
        r->d_in = static_cast<const uint8_t*>(d_in);
        r->h_in = h_in;
        r->in_size = n;
        r->cap = out_capacity(n, k.cfg.chunk, format);
        r->need = align_up(r->cap) + (h_in ? align_up(std::max<uint64_t>(n, 1)) : 0);
        if (r->need > cfg.hbm_budget) {
            set_error_msg("deflate_server_submit: request larger than the HBM budget");
            return 0;
        }
        const uint64_t nch = n ? (n + k.cfg.chunk - 1) / k.cfg.chunk : 1;
        if (nch > k.chunk_cap) {
            set_error_msg("deflate_server_submit: request of " + std::to_string(nch) + " blocks; a lane holds " +
                          std::to_string(k.chunk_cap) + " (lane_arena_bytes)");
            return 0;
        }
        std::lock_guard<std::mutex> lk(mu);
        r->id = next_id++;
        r->t_submit = now();
        Request* p = r.get();
        live[p->id] = std::move(r);
        inbox.push_back(p);
        ++n_submitted;
        max_queue = std::max(max_queue, n_submitted - n_admitted);
        ++wake;
        cv_work.notify_one();
        return p->id;
    }
};

extern "C" {

deflate_server* deflate_server_new(const deflate_server_config* cfg) {
    auto* s = new (std::nothrow) deflate_server;
    if (!s) return nullptr;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.hbm_budget) s->cfg.hbm_budget = 32ull << 30;
    if (!s->cfg.max_batch_bytes) s->cfg.max_batch_bytes = 256ull << 20;
    if (s->cfg.max_batch_requests <= 0) s->cfg.max_batch_requests = 4096;
    if (s->cfg.lanes_per_level <= 0) s->cfg.lanes_per_level = 2;
    if (!s->cfg.lane_arena_bytes) s->cfg.lane_arena_bytes = 2ull << 30;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || hipGetDevice(&s->device) != hipSuccess) {
        set_error_msg("deflate_server_new: no usable GPU");
        delete s;
        return nullptr;
    }
    if (!s->init()) {
        delete s;
        return nullptr;
    }
    s->worker = std::thread([s] { s->run(); });
    return s;
}

void deflate_server_free(deflate_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        ++s->wake;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    delete s;
}

uint64_t deflate_server_submit_device(deflate_server* s, const void* d_in, size_t n, int level,
                                      deflate_format format) {
    if (!s || (!d_in && n)) return 0;
    return s->submit(d_in, nullptr, n, level, (int)format);
}

uint64_t deflate_server_submit(deflate_server* s, const void* in, size_t n, int level, deflate_format format) {
    if (!s || (!in && n)) return 0;
    static const uint8_t empty = 0;
    return s->submit(nullptr, in ? in : &empty, n, level, (int)format);
}

int deflate_server_poll(deflate_server* s, deflate_server_completion* out, int max, int timeout_ms) {
    if (!s || !out || max <= 0) return 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->completed.empty() && timeout_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->completed.empty(); });
    int k = 0;
    while (k < max && !s->completed.empty()) {
        Request* r = s->completed.front();
        s->completed.pop_front();
        out[k].id = r->id;
        out[k].in_size = r->in_size;
        out[k].out_size = r->out_size;
        out[k].status = r->status;
        out[k].level = r->level;
        out[k].submit_ns = r->t_submit;
        out[k].launch_ns = r->t_launch;
        out[k].done_ns = r->t_done;
        ++k;
    }
    return k;
}

const void* deflate_server_output(deflate_server* s, uint64_t id) {
    if (!s) return nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end() || !it->second->done || it->second->status != 0) return nullptr;
    return s->d_pool + it->second->out_off;
}

size_t deflate_server_read(deflate_server* s, uint64_t id, void* out, size_t cap) {
    if (!s || !out) return 0;
    const uint8_t* src = nullptr;
    size_t n = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end() || !it->second->done || it->second->status != 0) return 0;
        src = s->d_pool + it->second->out_off;
        n = it->second->out_size;
    }
    if (n > cap) return 0;
    hipError_t e = hipMemcpy(out, src, n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("hipMemcpy(server output)", e);
        return 0;
    }
    return n;
}

int deflate_server_release(deflate_server* s, uint64_t id) {
    if (!s) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end() || !it->second->done) return -1;
    s->rel_inbox.push_back(it->second->rec);
    s->live.erase(it);
    ++s->wake;
    s->cv_work.notify_one();  // freed memory may admit a queue head
    return 0;
}

int deflate_server_stats(deflate_server* s, uint64_t* v, int max) {
    if (!s || !v) return 0;
    std::unique_lock<std::mutex> lk(s->mu);
    // the pool figure comes from the device, after the releases so far
    const uint64_t ticket = ++s->stats_req;
    ++s->wake;
    s->cv_work.notify_one();
    s->cv_done.wait(lk, [&] { return s->stats_done >= ticket || s->worker_exited; });
    const uint64_t vals[7] = {s->n_submitted, s->n_completed, s->n_submitted - s->n_admitted, s->n_inflight,
                              s->pool_in_use, s->n_batches, s->max_queue};
    const int k = std::min(max, 7);
    for (int i = 0; i < k; ++i) v[i] = vals[i];
    return k;
}

}  // extern "C"
