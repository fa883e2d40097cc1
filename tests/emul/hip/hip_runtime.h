// Host stand-in for the HIP device dialect used by the library's generated
// expectation kernels, so that their text runs on the CPU in tests: one
// std::thread per lane, __syncthreads as a barrier over the block, and
// __shared__ arrays as function statics (blocks run one at a time).
#pragma once
#include <barrier>
#include <cmath>
#include <cstdint>

struct double2 { double x, y; };
struct float2 { float x, y; };
struct emu_dim3 { unsigned x = 0, y = 0, z = 0; };

inline thread_local emu_dim3 threadIdx, blockIdx;
inline emu_dim3 gridDim, blockDim;
inline std::barrier<>* emu_barrier = nullptr;

#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))

inline void __syncthreads() { emu_barrier->arrive_and_wait(); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
