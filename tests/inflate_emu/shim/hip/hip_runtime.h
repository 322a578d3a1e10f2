// Host stand-in for the HIP runtime header, used only by tests/inflate_emu:
// device qualifiers vanish, and the wave intrinsics the inflate kernel uses
// become exchanges between the 64 host threads that play one wave's lanes.
// A workgroup barrier ends the program: the inflate kernel's waves run
// independently and must never need one.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstdint>
#define __device__
#define __global__
#define __host__
#define __forceinline__ inline
#define __constant__ static
#define __shared__
#define __launch_bounds__(x)
typedef int hipError_t;
enum { hipSuccess = 0, hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
struct EmuIdx { unsigned x, y, z; };
extern thread_local EmuIdx threadIdx;
extern EmuIdx blockIdx, blockDim, gridDim;
extern thread_local int emu_lane;
uint64_t emu_ballot(bool b);
long long emu_shfl_ll(long long v, int src);
void emu_wave_fence();
#define __builtin_amdgcn_ballot_w64(b) emu_ballot(b)
#define __builtin_amdgcn_fence(a, b) emu_wave_fence()
#define hipFuncSetAttribute(...) 0
#define hipMemsetAsync(...) 0
#define hipLaunchKernelGGL(...) ((void)0)
#define hipGetLastError() 0
inline int __shfl(int v, int src, int = 64) { return (int)emu_shfl_ll(v, src); }
inline int __shfl_xor(int v, int m, int = 64) { return (int)emu_shfl_ll(v, emu_lane ^ m); }
inline long long __shfl_xor(long long v, int m, int = 64) { return emu_shfl_ll(v, emu_lane ^ m); }
inline unsigned __shfl_up(unsigned v, int d, int = 64) {
    const long long r = emu_shfl_ll(v, emu_lane >= d ? emu_lane - d : emu_lane);
    return (unsigned)r;
}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline void __syncthreads() {
    fprintf(stderr, "emu: workgroup barrier reached\n");
    abort();
}
template <class T> T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
using std::max;
using std::min;
