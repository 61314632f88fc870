// CPU emulation of the small HIP subset the backend's kernels and host code
// use, for running the real kernel sources on the host in tests (bounds and
// index bugs show up as wrong numbers or sanitizer reports, never as GPU
// faults).  The threads of a block run as fibers (ucontext) on the calling
// OS thread: each runs until its next __syncthreads or its end, and a round
// over all live fibers completes the barrier (threads that return early drop
// out).  Blocks run one after another, so __shared__ becomes a
// function-local static.  Test infrastructure only.
#pragma once
#include <ucontext.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
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
inline dim3 tIdx, bIdx, bDim, gDim;

struct Fiber {
  ucontext_t ctx;
  std::unique_ptr<char[]> stack;
  bool done = false;
};
inline ucontext_t sched_ctx;
inline std::vector<Fiber> fibers;
inline int cur = 0;
inline std::function<void()>* body = nullptr;
constexpr size_t kStack = size_t(256) << 10;

inline void trampoline() {
  (*body)();
  fibers[cur].done = true;  // returns to sched_ctx through uc_link
}
inline void sync() { swapcontext(&fibers[cur].ctx, &sched_ctx); }

template <typename K, typename... A>
void launch(K kernel, dim3 grid, dim3 block, A... args) {
  const unsigned nt = block.x * block.y * block.z;
  const unsigned nblocks = grid.x * grid.y * grid.z;
  if (fibers.size() < nt) fibers.resize(nt);
  std::function<void()> f = [&] { kernel(args...); };
  body = &f;
  bDim = block;
  gDim = grid;
  for (unsigned bi = 0; bi < nblocks; ++bi) {
    bIdx = dim3(bi % grid.x, (bi / grid.x) % grid.y, bi / (grid.x * grid.y));
    for (unsigned t = 0; t < nt; ++t) {
      Fiber& fb = fibers[t];
      if (!fb.stack) fb.stack.reset(new char[kStack]);
      getcontext(&fb.ctx);
      fb.ctx.uc_stack.ss_sp = fb.stack.get();
      fb.ctx.uc_stack.ss_size = kStack;
      fb.ctx.uc_link = &sched_ctx;
      makecontext(&fb.ctx, trampoline, 0);
      fb.done = false;
    }
    // rounds: every live fiber runs to its next barrier (or its end)
    for (bool live = true; live;) {
      live = false;
      for (unsigned t = 0; t < nt; ++t) {
        if (fibers[t].done) continue;
        cur = int(t);
        tIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
        swapcontext(&sched_ctx, &fibers[t].ctx);
        live = live || !fibers[t].done;
      }
    }
  }
}
}  // namespace emu

#define threadIdx (::emu::tIdx)
#define blockIdx (::emu::bIdx)
#define blockDim (::emu::bDim)
#define gridDim (::emu::gDim)
inline void __syncthreads() { ::emu::sync(); }
// a lane exchange within a wavefront of `width` lanes: every thread of the
// block must call it (as at a barrier); each deposits its value, then reads
// its partner's
namespace emu {
inline std::vector<double> lane_slots;
}
inline double __shfl_xor(double v, int mask, int width = 64) {
  const int t = ::emu::cur;
  if (::emu::lane_slots.size() < ::emu::fibers.size()) ::emu::lane_slots.resize(::emu::fibers.size());
  ::emu::lane_slots[t] = v;
  ::emu::sync();
  const double r = ::emu::lane_slots[(t & ~(width - 1)) | ((t & (width - 1)) ^ mask)];
  ::emu::sync();
  return r;
}
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
