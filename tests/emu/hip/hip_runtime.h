// CPU emulation of the small HIP subset the backend's kernels and host code
// use, for running the real kernel sources on the host in tests (bounds and
// index bugs show up as wrong numbers or sanitizer reports, never as GPU
// faults).  Each block runs its threads as std::threads sharing one
// std::barrier (__syncthreads); blocks run one after another, so __shared__
// becomes a function-local static.  Test infrastructure only.
#pragma once
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__
#define __align__(n) alignas(n)
#define HIP_SYMBOL(X) X

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
inline thread_local dim3 tIdx, bIdx, bDim, gDim;
inline thread_local std::barrier<>* bar = nullptr;

template <typename K, typename... A>
void launch(K kernel, dim3 grid, dim3 block, A... args) {
  const unsigned nt = block.x * block.y * block.z;
  for (unsigned bz = 0; bz < grid.z; ++bz)
    for (unsigned by = 0; by < grid.y; ++by)
      for (unsigned bx = 0; bx < grid.x; ++bx) {
        std::barrier<> b(nt);
        std::vector<std::thread> th;
        th.reserve(nt);
        for (unsigned t = 0; t < nt; ++t)
          th.emplace_back([&, t] {
            tIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
            bIdx = dim3(bx, by, bz);
            bDim = block;
            gDim = grid;
            bar = &b;
            kernel(args...);
            b.arrive_and_drop();
          });
        for (auto& x : th) x.join();
      }
}
}  // namespace emu

#define threadIdx (::emu::tIdx)
#define blockIdx (::emu::bIdx)
#define blockDim (::emu::bDim)
#define gridDim (::emu::gDim)
inline void __syncthreads() { ::emu::bar->arrive_and_wait(); }
#define __builtin_amdgcn_s_barrier() __syncthreads()
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))

#define hipLaunchKernelGGL(K, G, B, SHM, ST, ...) ::emu::launch(K, G, B, __VA_ARGS__)

typedef void* hipStream_t;
typedef void* hipEvent_t;
enum hipError_t { hipSuccess = 0, hipErrorUnknown = 999 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
inline const char* hipGetErrorString(hipError_t) { return "emulated error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
template <typename P>
inline hipError_t hipMalloc(P** p, size_t n) {
  *p = static_cast<P*>(std::aligned_alloc(256, (n + 255) / 256 * 256));
  return *p ? hipSuccess : hipErrorUnknown;
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = nullptr;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
template <typename T>
inline hipError_t hipGetSymbolAddress(void** p, const T& symbol) {
  *p = const_cast<void*>(static_cast<const void*>(&symbol));
  return hipSuccess;
}
