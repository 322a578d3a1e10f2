// server.cpp -- continuous-batching compression server (deflate_server.h).
//
// One device pool of hbm_budget bytes holds every admitted request's output
// (and its input, for host submissions).  A worker thread repeats:
//   retire  -- lanes whose batch finished hand their requests to the
//              completion queue (sizes come from the device-side results);
//   launch  -- every idle lane takes the requests of its (level, format)
//              queue that fit the pool, in arrival order, up to the batch
//              limits, and enqueues them as one batch on its own stream.
// Lanes of different levels, and several lanes of one level, run on the GPU
// at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/orot/deflate_server.h"
#include "engine.h"
#include "host_common.h"

using orot_gpu::BufferSpec;
using orot_gpu::Engine;
using orot_gpu::EngineConfig;

namespace {

constexpr uint64_t ALIGN = 256;

uint64_t align_up(uint64_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

// First-fit allocator over [0, size).
class Pool {
public:
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(uint64_t n, uint64_t& off) {
        n = align_up(std::max<uint64_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const uint64_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = align_up(std::max<uint64_t>(n, 1));
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const { return used_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t used_ = 0;
};

struct Request {
    uint64_t id = 0;
    int level = 1;
    int format = 0;
    const uint8_t* d_in = nullptr;  // device input (caller's, or inside the pool)
    const void* h_in = nullptr;     // host input, copied at admission
    uint64_t in_size = 0;
    uint64_t in_off = 0;            // pool offset of a copied host input
    bool in_pool = false;
    uint64_t out_off = 0;
    uint64_t cap = 0;
    uint64_t out_size = 0;
    int status = 0;
    bool done = false;
    uint64_t t_submit = 0, t_launch = 0, t_done = 0;
};

struct Lane {
    Engine eng;
    int level = 1;
    int format = 0;
    bool busy = false;
    std::vector<Request*> reqs;
    std::vector<uint64_t> sizes;
};

size_t default_chunk(int level) { return level <= 3 ? 131072 : (level <= 6 ? 524288 : 1048576); }

uint64_t out_capacity(uint64_t n, int level, int format) {
    const uint64_t chunk = default_chunk(level);
    const uint64_t nchunks = n ? (n + chunk - 1) / chunk : 1;
    const uint64_t bound = deflate_compress_bound(n, (deflate_format)format) + 64;
    return std::max(bound, n + 6 * (nchunks + n / 65535 + 2) + 64);
}

}  // namespace

struct deflate_server {
    deflate_server_config cfg{};
    uint8_t* d_pool = nullptr;
    Pool pool;
    std::mutex mu;
    std::condition_variable cv_work;
    std::condition_variable cv_done;
    std::map<int, std::deque<Request*>> queues;  // key = level * 4 + format
    std::map<int, std::vector<std::unique_ptr<Lane>>> lanes;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> live;
    std::deque<Request*> completed;
    uint64_t next_id = 1;
    uint64_t n_submitted = 0, n_completed = 0, n_queued = 0, n_inflight = 0, n_batches = 0, max_queue = 0;
    bool stop = false;
    uint64_t wake = 0;  // bumped by submit / release / stop: the worker's reason to look again
    std::thread worker;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    uint64_t now() const {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
            .count();
    }

    // retire finished batches (mu held)
    bool retire() {
        bool any = false;
        for (auto& kv : lanes)
            for (auto& lp : kv.second) {
                Lane& L = *lp;
                if (!L.busy || !L.eng.done()) continue;
                L.sizes.resize(L.reqs.size());
                const bool ok = L.eng.wait(L.sizes.data());
                const uint64_t t = now();
                for (size_t i = 0; i < L.reqs.size(); ++i) {
                    Request* r = L.reqs[i];
                    r->out_size = ok ? L.sizes[i] : 0;
                    r->status = r->out_size ? 0 : -1;
                    r->done = true;
                    r->t_done = t;
                    if (r->in_pool) {
                        pool.release(r->in_off, r->in_size);
                        r->in_pool = false;
                    }
                    completed.push_back(r);
                }
                n_completed += L.reqs.size();
                n_inflight -= L.reqs.size();
                L.reqs.clear();
                L.busy = false;
                any = true;
            }
        return any;
    }

    Lane* idle_lane(int key) {
        auto& v = lanes[key];
        for (auto& lp : v)
            if (!lp->busy) return lp.get();
        if ((int)v.size() >= cfg.lanes_per_level) return nullptr;
        auto lp = std::make_unique<Lane>();
        EngineConfig ec;
        ec.level = key / 4;
        ec.format = key % 4;
        if (!lp->eng.init(ec)) return nullptr;
        lp->level = ec.level;
        lp->format = ec.format;
        v.push_back(std::move(lp));
        return v.back().get();
    }

    // form and enqueue batches on idle lanes (mu held)
    bool launch() {
        bool any = false;
        for (auto& kv : queues) {
            auto& q = kv.second;
            while (!q.empty()) {
                Lane* L = idle_lane(kv.first);
                if (!L) break;
                std::vector<Request*> take;
                uint64_t bytes = 0;
                while (!q.empty() && (int)take.size() < cfg.max_batch_requests &&
                       (take.empty() || bytes + q.front()->in_size <= cfg.max_batch_bytes)) {
                    Request* r = q.front();
                    uint64_t out_off = 0, in_off = 0;
                    if (!pool.alloc(r->cap, out_off)) break;
                    if (r->h_in) {
                        if (!pool.alloc(r->in_size, in_off)) {
                            pool.release(out_off, r->cap);
                            break;
                        }
                        r->in_off = in_off;
                        r->in_pool = true;
                        r->d_in = d_pool + in_off;
                    }
                    r->out_off = out_off;
                    bytes += r->in_size;
                    take.push_back(r);
                    q.pop_front();
                    --n_queued;
                }
                if (take.empty()) break;  // the queue head waits for memory
                const uint8_t* base = take[0]->d_in;
                for (Request* r : take) base = std::min(base, r->d_in);
                std::vector<BufferSpec> specs(take.size());
                bool ok = true;
                for (size_t i = 0; i < take.size(); ++i) {
                    Request* r = take[i];
                    if (r->h_in && r->in_size) {
                        ok = ok && hipMemcpyAsync(const_cast<uint8_t*>(r->d_in), r->h_in, r->in_size,
                                                  hipMemcpyHostToDevice, L->eng.stream()) == hipSuccess;
                    }
                    specs[i] = {(uint64_t)(r->d_in - base), r->in_size, r->out_off, r->cap};
                }
                ok = ok && L->eng.submit(base, d_pool, specs.data(), specs.size(), false);
                const uint64_t t = now();
                for (Request* r : take) r->t_launch = t;
                n_inflight += take.size();
                ++n_batches;
                L->reqs = std::move(take);
                L->busy = true;
                if (!ok) {
                    // the lane's done event was not recorded: finish the batch as failed now
                    (void)hipStreamSynchronize(L->eng.stream());
                    for (Request* r : L->reqs) {
                        r->status = -1;
                        r->out_size = 0;
                        r->done = true;
                        r->t_done = now();
                        if (r->in_pool) {
                            pool.release(r->in_off, r->in_size);
                            r->in_pool = false;
                        }
                        completed.push_back(r);
                    }
                    n_completed += L->reqs.size();
                    n_inflight -= L->reqs.size();
                    L->reqs.clear();
                    L->busy = false;
                }
                any = true;
            }
        }
        return any;
    }

    bool any_busy() const {
        for (auto& kv : lanes)
            for (auto& lp : kv.second)
                if (lp->busy) return true;
        return false;
    }

    void run() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            const uint64_t seen = wake;
            const bool retired = retire();
            if (retired) cv_done.notify_all();
            const bool launched = launch();
            const bool busy = any_busy();
            // at stop, requests that can never be admitted (the pool is full of
            // outputs nobody will release) are dropped with the server
            if (stop && !busy && (n_queued == 0 || !launched)) break;
            if (busy) {
                // a batch is on the GPU: check again shortly
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            } else if (!launched && !retired) {
                // nothing runs and the queue heads wait for memory (or the
                // queues are empty): sleep until a submit, release or stop
                cv_work.wait(lk, [&] { return stop || wake != seen; });
            }
        }
    }

    uint64_t submit(const void* d_in, const void* h_in, uint64_t n, int level, int format) {
        if (level < 0) level = 0;
        if (level > 12) level = 12;
        if (format != 1 && format != 2) format = 0;
        auto r = std::make_unique<Request>();
        r->level = level;
        r->format = format;
        r->d_in = static_cast<const uint8_t*>(d_in);
        r->h_in = h_in;
        r->in_size = n;
        r->cap = out_capacity(n, level, format);
        const uint64_t need = align_up(r->cap) + (h_in ? align_up(std::max<uint64_t>(n, 1)) : 0);
        if (need > cfg.hbm_budget) {
            orot_gpu::set_error_msg("deflate_server_submit: request larger than the HBM budget");
            return 0;
        }
        std::lock_guard<std::mutex> lk(mu);
        r->id = next_id++;
        r->t_submit = now();
        Request* p = r.get();
        live[p->id] = std::move(r);
        queues[level * 4 + format].push_back(p);
        ++n_submitted;
        ++n_queued;
        ++wake;
        max_queue = std::max(max_queue, n_queued);
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
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        orot_gpu::set_error_msg("deflate_server_new: no usable GPU");
        delete s;
        return nullptr;
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->d_pool), s->cfg.hbm_budget);
    if (e != hipSuccess) {
        s->d_pool = nullptr;
        orot_gpu::set_error("hipMalloc(server pool)", e);
        delete s;
        return nullptr;
    }
    s->pool.reset(s->cfg.hbm_budget);
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
    s->lanes.clear();
    if (s->d_pool) (void)hipFree(s->d_pool);
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
        orot_gpu::set_error("hipMemcpy(server output)", e);
        return 0;
    }
    return n;
}

int deflate_server_release(deflate_server* s, uint64_t id) {
    if (!s) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end() || !it->second->done) return -1;
    s->pool.release(it->second->out_off, it->second->cap);
    s->live.erase(it);
    ++s->wake;
    s->cv_work.notify_one();  // freed memory may admit the queue head
    return 0;
}

int deflate_server_stats(deflate_server* s, uint64_t* v, int max) {
    if (!s || !v) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t vals[7] = {s->n_submitted, s->n_completed, s->n_queued, s->n_inflight,
                              s->pool.used(), s->n_batches, s->max_queue};
    const int k = std::min(max, 7);
    for (int i = 0; i < k; ++i) v[i] = vals[i];
    return k;
}

}  // extern "C"
