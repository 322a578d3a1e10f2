// emu.cpp -- runs the inflate kernel's source (orot_amd/csrc/inflate.hip,
// unchanged) on the CPU: 64 host threads are the lanes of one wave, and the
// wave intrinsics are barrier exchanges between them (shim/hip/hip_runtime.h).
//   emu <stream file> <format 0|1|2> <output capacity> <output file>
// prints "kernel: status S size N" and writes the output bytes.
#include <cstdio>
#include <thread>
#include <vector>
#include "hip/hip_runtime.h"
thread_local EmuIdx threadIdx;
EmuIdx blockIdx, blockDim, gridDim;
thread_local int emu_lane;
static std::barrier<> g_bar(64);
static long long g_x[64];
uint64_t emu_ballot(bool b) {
    g_x[emu_lane] = b;
    g_bar.arrive_and_wait();
    uint64_t m = 0;
    for (int i = 0; i < 64; ++i) if (g_x[i]) m |= 1ull << i;
    g_bar.arrive_and_wait();
    return m;
}
long long emu_shfl_ll(long long v, int src) {
    g_x[emu_lane] = v;
    g_bar.arrive_and_wait();
    const long long r = g_x[src & 63];
    g_bar.arrive_and_wait();
    return r;
}
void emu_sync() { g_bar.arrive_and_wait(); }
// the kernel's dynamic LDS, declared where its block-scope extern finds it
namespace orot_gpu { namespace { uint32_t smem[8192]; } }
#include "../../orot_amd/csrc/inflate.hip"
namespace orot_gpu {
int emu_main(int argc, char** argv) {
    FILE* f = fopen(argv[1], "rb");
    std::vector<uint8_t> s(1 << 24);
    const size_t n = fread(s.data(), 1, s.size(), f);
    fclose(f);
    const int fmt = atoi(argv[2]);
    const uint64_t cap = strtoull(argv[3], nullptr, 10);
    std::vector<uint8_t> out(cap + 1);
    InflateItem item{0, n, 0, cap};
    uint32_t order = 0, queue = 0;
    InflateResult res{0, 12345, 0};
    blockIdx = {0, 0, 0};
    gridDim = {1, 1, 1};
    blockDim = {64, 1, 1};
    std::vector<std::thread> th;
    for (int l = 0; l < 64; ++l)
        th.emplace_back([&, l] {
            emu_lane = l;
            threadIdx = {(unsigned)l, 0, 0};
            inflate_kernel(s.data(), out.data(), &item, &order, 1u, fmt, &res, &queue);
        });
    for (auto& t : th) t.join();
    printf("kernel: status %d size %llu queue %u\n", res.status, (unsigned long long)res.out_size, queue);
    FILE* o = fopen(argv[4], "wb");
    fwrite(out.data(), 1, res.out_size, o);
    fclose(o);
    return 0;
}
}  // namespace orot_gpu
int main(int argc, char** argv) { return orot_gpu::emu_main(argc, argv); }
