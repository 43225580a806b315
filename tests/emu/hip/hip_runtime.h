// Host stand-in for HIP used by the emulation tests: every block runs as
// blockDim OS threads with a shared barrier; __ballot is wave-wide for 64-thread blocks.
#pragma once
#include <stdint.h>
#include <string.h>
#include <pthread.h>
#include <thread>
#include <vector>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
struct dim3 { unsigned x = 1, y = 1, z = 1; dim3() {} dim3(unsigned a) : x(a) {} };
inline thread_local dim3 blockIdx, threadIdx, blockDim, gridDim;
typedef void* hipStream_t;
inline pthread_barrier_t g_bar;
inline unsigned long long g_acc;
inline void __syncthreads() { pthread_barrier_wait(&g_bar); }
inline unsigned long long __ballot(int p) {
  pthread_barrier_wait(&g_bar);
  if (threadIdx.x == 0) g_acc = 0;
  pthread_barrier_wait(&g_bar);
  if (p) __atomic_fetch_or(&g_acc, 1ull << (threadIdx.x & 63), __ATOMIC_SEQ_CST);
  pthread_barrier_wait(&g_bar);
  const unsigned long long m = __atomic_load_n(&g_acc, __ATOMIC_SEQ_CST);
  pthread_barrier_wait(&g_bar);
  return m;
}
#define __builtin_amdgcn_wave_barrier() pthread_barrier_wait(&g_bar)
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
template <class K, class... A>
void emu_launch(K k, dim3 g, dim3 b, A... args) {
  for (unsigned blk = 0; blk < g.x; ++blk) {
    pthread_barrier_init(&g_bar, nullptr, b.x);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < b.x; ++t)
      th.emplace_back([=]() { blockIdx.x = blk; threadIdx.x = t; blockDim = b; gridDim = g; k(args...); });
    for (auto& x : th) x.join();
    pthread_barrier_destroy(&g_bar);
  }
}
#define hipLaunchKernelGGL(k, g, b, shm, st, ...) emu_launch(k, dim3(g), dim3(b), __VA_ARGS__)
