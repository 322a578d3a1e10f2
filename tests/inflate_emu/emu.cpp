// emu.cpp -- runs the inflate kernel's source (orot_amd/csrc/inflate.hip,
// unchanged) on the CPU in the shape launch_inflate gives it: workgroups of
// INF_WAVES waves, each wave 64 host threads that play its lanes, the waves of
// a workgroup running at once and taking streams from the one shared queue,
// over the grid launch_inflate computes (one CU's worth of workgroups, run one
// workgroup after another; they share nothing but the queue and the result
// and output arrays, whose slots are disjoint).
//
// Lanes of a wave execute in lockstep on the GPU.  Between two exchange
// points (ballot, shuffle, the wave fence) the lane threads here run freely,
// and at each exchange point the wave's 64 threads meet at a barrier of their
// own; so every exchange sees all 64 lanes' values, as in lockstep, and a
// kernel that reached an exchange point with only some of its lanes would
// stop here instead of passing.  There is no workgroup barrier in the shim:
// the kernel must not need one.
//
// The dynamic LDS is filled with bytes 0x07 before each workgroup (a valid
// code length, so a length read before it is written changes the code), and
// the output buffers with 0x5A before the launch: the GPU gives neither
// zeroed, and a kernel that read them before writing would pass on zeros.
//
//   emu <format 0|1|2> <count> (<stream file> <output capacity> <output file>)...
// prints "item I: status S size N" per stream and writes each output.
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>
#include "hip/hip_runtime.h"
thread_local EmuIdx threadIdx;
EmuIdx blockIdx, blockDim, gridDim;
thread_local int emu_lane;

// one wave's exchange: a barrier of its 64 lane threads and a slot per lane
struct EmuWave {
    std::barrier<> bar{64};
    long long x[64];
};
static thread_local EmuWave* emu_w;
uint64_t emu_ballot(bool b) {
    emu_w->x[emu_lane] = b;
    emu_w->bar.arrive_and_wait();
    uint64_t m = 0;
    for (int i = 0; i < 64; ++i)
        if (emu_w->x[i]) m |= 1ull << i;
    emu_w->bar.arrive_and_wait();
    return m;
}
long long emu_shfl_ll(long long v, int src) {
    emu_w->x[emu_lane] = v;
    emu_w->bar.arrive_and_wait();
    const long long r = emu_w->x[src & 63];
    emu_w->bar.arrive_and_wait();
    return r;
}
void emu_wave_fence() {
    std::atomic_thread_fence(std::memory_order_acq_rel);
    emu_w->bar.arrive_and_wait();
}
// the kernel's dynamic LDS, declared where its block-scope extern finds it
namespace orot_gpu { namespace { uint32_t smem[8192]; } }
#include "../../orot_amd/csrc/inflate.hip"
namespace orot_gpu {
static_assert(INF_LDS <= sizeof(smem), "the emulated LDS holds the kernel's dynamic LDS");
constexpr uint32_t EMU_CUS = 1;  // launch_inflate's grid cap is 8 workgroups per CU

int emu_main(int argc, char** argv) {
    if (argc < 3) return 2;
    const int fmt = atoi(argv[1]);
    const uint32_t count = (uint32_t)atoi(argv[2]);
    if (argc != 3 + 3 * (int)count) return 2;
    std::vector<uint8_t> in, out;
    std::vector<InflateItem> items(count);
    for (uint32_t i = 0; i < count; ++i) {
        FILE* f = fopen(argv[3 + 3 * i], "rb");
        if (!f) return 2;
        std::vector<uint8_t> s(1 << 24);
        const size_t n = fread(s.data(), 1, s.size(), f);
        fclose(f);
        const uint64_t cap = strtoull(argv[4 + 3 * i], nullptr, 10);
        items[i] = InflateItem{in.size(), n, out.size(), cap};
        in.insert(in.end(), s.begin(), s.begin() + n);
        out.resize(out.size() + cap + 64);  // + 64: a gap that must stay untouched
    }
    memset(out.data(), 0x5A, out.size());
    // launch_inflate's queue order: longest stream first
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return items[a].in_size > items[b].in_size; });
    std::vector<InflateResult> res(count, InflateResult{0, 12345, 0});
    uint32_t queue = 0;
    const uint32_t waves = (count + INF_WAVES - 1) / INF_WAVES;
    const uint32_t grid = waves < EMU_CUS * 8u ? waves : EMU_CUS * 8u;
    gridDim = {grid, 1, 1};
    blockDim = {(unsigned)INF_THREADS, 1, 1};
    for (uint32_t g = 0; g < grid; ++g) {
        blockIdx = {g, 0, 0};
        memset(smem, 0x07, sizeof(smem));
        std::vector<std::unique_ptr<EmuWave>> wv;
        for (int w = 0; w < INF_WAVES; ++w) wv.emplace_back(new EmuWave);
        std::vector<std::thread> th;
        for (int t = 0; t < INF_THREADS; ++t)
            th.emplace_back([&, t] {
                emu_lane = t & 63;
                emu_w = wv[t >> 6].get();
                threadIdx = {(unsigned)t, 0, 0};
                inflate_kernel(in.data(), out.data(), items.data(), order.data(), count, fmt, res.data(), &queue);
            });
        for (auto& t : th) t.join();
    }
    for (uint32_t i = 0; i < count; ++i) {
        const InflateItem& I = items[i];
        printf("item %u: status %d size %llu\n", i, res[i].status, (unsigned long long)res[i].out_size);
        for (uint32_t k = 0; k < 64; ++k)
            if (out[I.out_off + I.out_cap + k] != 0x5A) {
                printf("item %u: write past its output capacity\n", i);
                return 3;
            }
        FILE* o = fopen(argv[5 + 3 * i], "wb");
        fwrite(out.data() + I.out_off, 1, res[i].status == 0 ? res[i].out_size : 0, o);
        fclose(o);
    }
    printf("queue %u\n", queue);
    return 0;
}
}  // namespace orot_gpu
int main(int argc, char** argv) { return orot_gpu::emu_main(argc, argv); }
