// inflate_api.cpp -- C-ABI host layer of the decompress side: orot's
// deflate_decompress (src/api/deflate_api.cpp:138-215) on the host inflater
// of inflate_host.cpp, and the device-resident and batched GPU inflate entry
// points of deflate_gpu.h.
//
// The GPU inflate kernel is switched off (kInflateEnabled): on its first GPU
// run it did not finish on a stored-block stream that the same kernel source
// decodes correctly under the lane-emulating CPU harness of
// tests/test_inflate_emu.py.  Until the cause is found the two GPU entry
// points return their failure value after the argument checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/orot/deflate.h"
#include "../../include/orot/deflate_gpu.h"
#include "deflate_gpu.h"
#include "host_common.h"
#include "inflate_host.h"

using orot_gpu::InflateItem;
using orot_gpu::InflateResult;
using orot_gpu::set_error;
using orot_gpu::set_error_msg;

namespace {

template <class T>
bool grow(T*& p, size_t& cap, size_t need, const char* what) {
    if (need <= cap && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t n = std::max<size_t>(need, 1);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        set_error(what, e);
        return false;
    }
    cap = n;
    return true;
}

// Process-wide inflate state: one stream, work arrays grown on demand.
struct InflateState {
    std::mutex mu;
    hipStream_t stream = nullptr;
    int cus = 0;
    uint8_t* d_in = nullptr;
    size_t in_cap = 0;
    uint8_t* d_out = nullptr;
    size_t out_cap = 0;
    InflateItem* d_items = nullptr;
    size_t items_cap = 0;
    uint32_t* d_order = nullptr;
    size_t order_cap = 0;
    InflateResult* d_res = nullptr;
    size_t res_cap = 0;
    uint32_t* d_queue = nullptr;
    size_t queue_cap = 0;

    bool init() {
        if (stream) return true;
        int dev = 0;
        hipDeviceProp_t prop;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error("inflate: no usable GPU", e);
            stream = nullptr;
            return false;
        }
        cus = prop.multiProcessorCount;
        return true;
    }

    // Runs `items` (device offsets into d_in / d_out); results on the host.
    bool run(const uint8_t* din, uint8_t* dout, const std::vector<InflateItem>& items, int format,
             std::vector<InflateResult>& res) {
        const size_t n = items.size();
        res.assign(n, InflateResult{0, 0, 0});
        if (n == 0) return true;
        std::vector<uint32_t> order(n);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return items[a].in_size > items[b].in_size; });
        if (!grow(d_items, items_cap, n, "hipMalloc(inflate items)") ||
            !grow(d_order, order_cap, n, "hipMalloc(inflate order)") ||
            !grow(d_res, res_cap, n, "hipMalloc(inflate results)") ||
            !grow(d_queue, queue_cap, 1, "hipMalloc(inflate queue)"))
            return false;
        hipError_t e = hipMemcpyAsync(d_items, items.data(), n * sizeof(InflateItem), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_order, order.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess)
            e = orot_gpu::launch_inflate(din, dout, d_items, d_order, (uint32_t)n, format, d_res, d_queue, cus, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(res.data(), d_res, n * sizeof(InflateResult), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            set_error("inflate", e);
            return false;
        }
        return true;
    }
};

InflateState& inf() {
    static InflateState s;
    return s;
}

constexpr bool kInflateEnabled = false;

bool inflate_disabled() {
    if (kInflateEnabled) return false;
    set_error_msg("deflate_decompress: the GPU inflate is switched off in this build");
    return true;
}

bool valid_format(deflate_format f) {
    return f == DEFLATE_FORMAT_RAW || f == DEFLATE_FORMAT_ZLIB || f == DEFLATE_FORMAT_GZIP;
}

deflate_result finish(const InflateResult& r, size_t* actual) {
    if (r.status == DEFLATE_OK && actual) *actual = (size_t)r.out_size;
    if (r.status == DEFLATE_DATA_ERROR) set_error_msg("deflate_decompress: corrupt or truncated stream");
    if (r.status == DEFLATE_NEED_OUTPUT) set_error_msg("deflate_decompress: output does not fit out_capacity");
    return (deflate_result)r.status;
}

}  // namespace

extern "C" {

// orot's whole-buffer inflate of host buffers, on the host inflater
// (inflate_host.cpp): the decompress side is not the GPU compress path.
deflate_result deflate_decompress(const void* in, size_t in_size, void* out, size_t out_capacity,
                                  size_t* actual_out_size, deflate_format format) {
    if ((!in && in_size) || (!out && out_capacity) || !valid_format(format)) return DEFLATE_PARAM_ERROR;
    orot_gpu::HostInflater* h = orot_gpu::HostInflater::create((int)format);
    if (!h) {
        set_error_msg("deflate_decompress: allocation failed");
        return DEFLATE_MEM_ERROR;
    }
    size_t got = 0;
    const deflate_result r = h->whole(static_cast<const uint8_t*>(in), in_size, static_cast<uint8_t*>(out),
                                      out_capacity, &got);
    orot_gpu::HostInflater::destroy(h);
    if (r == DEFLATE_STREAM_END) {
        if (actual_out_size) *actual_out_size = got;
        return DEFLATE_OK;
    }
    if (r == DEFLATE_NEED_OUTPUT) {
        set_error_msg("deflate_decompress: output does not fit out_capacity");
        return DEFLATE_NEED_OUTPUT;
    }
    set_error_msg(r == DEFLATE_OK ? "deflate_decompress: truncated stream" : "deflate_decompress: corrupt stream");
    return DEFLATE_DATA_ERROR;
}

deflate_result deflate_decompress_device(const void* d_in, size_t in_size, void* d_out, size_t out_capacity,
                                         size_t* actual_out_size, deflate_format format) {
    if ((!d_in && in_size) || (!d_out && out_capacity) || !valid_format(format)) return DEFLATE_PARAM_ERROR;
    if (inflate_disabled()) return DEFLATE_MEM_ERROR;
    auto& s = inf();
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.init()) return DEFLATE_MEM_ERROR;
    std::vector<InflateResult> res;
    if (!s.run(static_cast<const uint8_t*>(d_in), static_cast<uint8_t*>(d_out),
               {InflateItem{0, in_size, 0, out_capacity}}, (int)format, res))
        return DEFLATE_MEM_ERROR;
    return finish(res[0], actual_out_size);
}

int deflate_batch_decompress_device(const void* d_in, void* d_out, const deflate_inflate_item* items, size_t count,
                                    deflate_format format, uint64_t* out_sizes, int32_t* results) {
    if ((count && (!items || !out_sizes || !results)) || !valid_format(format) || count > 0xFFFFFFFFull) {
        set_error_msg("deflate_batch_decompress_device: invalid arguments");
        return -1;
    }
    if (inflate_disabled()) return -1;
    auto& s = inf();
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.init()) return -1;
    std::vector<InflateItem> v(count);
    for (size_t i = 0; i < count; ++i) v[i] = {items[i].in_off, items[i].in_size, items[i].out_off, items[i].out_capacity};
    std::vector<InflateResult> res;
    if (!s.run(static_cast<const uint8_t*>(d_in), static_cast<uint8_t*>(d_out), v, (int)format, res)) return -1;
    for (size_t i = 0; i < count; ++i) {
        results[i] = res[i].status;
        out_sizes[i] = res[i].status == DEFLATE_OK ? res[i].out_size : 0;
    }
    return 0;
}

}  // extern "C"
