// Host stand-in for HIP used by the emulation tests, so that device code of
// diffpumas.jl_amd/csrc compiles unchanged with g++ and runs on the CPU.
//
// Kernels: every block runs as blockDim.x OS threads meeting at a shared
// barrier; blocks (x, then y, then z of the grid) run one after the other.  __shared__ variables are
// function statics (shared by the block's threads; blocks do not overlap).
// __ballot, __shfl_xor and wave barriers are exchanges between all threads
// of the block at a barrier, so they model 64-thread (one-wave) blocks and
// must be reached by every thread, which is what the kernels under test
// promise at those points.  Device memory is host memory; streams, events
// and copies are synchronous.  Graph capture is not emulated.
#pragma once
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <thread>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static
#define __constant__
#define HIP_DYNAMIC_SHARED(type, var) type* var = (type*)g_dyn_shared;
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3() {}
  dim3(unsigned a) : x(a) {}
  dim3(unsigned a, unsigned b, unsigned c = 1) : x(a), y(b), z(c) {}
};
inline thread_local dim3 blockIdx, threadIdx, blockDim, gridDim;
inline void* g_dyn_shared = nullptr;

// ---------------------------------------------------------- block sync
inline pthread_barrier_t g_bar;
inline unsigned long long g_acc;
inline double g_xch[1024];
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
inline double __shfl_xor(double v, int off, int width = 64) {
  (void)width;
  pthread_barrier_wait(&g_bar);
  g_xch[threadIdx.x] = v;
  pthread_barrier_wait(&g_bar);
  const double r = g_xch[threadIdx.x ^ off];
  pthread_barrier_wait(&g_bar);
  return r;
}
#define __builtin_amdgcn_wave_barrier() pthread_barrier_wait(&g_bar)
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
inline int __lane_id() { return threadIdx.x & 63; }
// value of lane k of the caller's wave (every thread of the block calls it)
inline int __builtin_amdgcn_readlane(int v, int k) {
  static int xch[1024];
  pthread_barrier_wait(&g_bar);
  xch[threadIdx.x] = v;
  pthread_barrier_wait(&g_bar);
  const int r = xch[(threadIdx.x & ~63u) + (unsigned)k];
  pthread_barrier_wait(&g_bar);
  return r;
}
// DPP with a quad_perm control (ctrl < 0x100): lane 4m + j takes src of lane
// 4m + ((ctrl >> 2j) & 3).  Other DPP controls are not emulated.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask,
                                       bool bound_ctrl) {
  (void)old; (void)row_mask; (void)bank_mask; (void)bound_ctrl;
  static int xch[1024];
  if (ctrl >= 0x100) abort();
  pthread_barrier_wait(&g_bar);
  xch[threadIdx.x] = src;
  pthread_barrier_wait(&g_bar);
  const unsigned j = threadIdx.x & 3u;
  const int r = xch[(threadIdx.x & ~3u) + ((unsigned)(ctrl >> (2 * j)) & 3u)];
  pthread_barrier_wait(&g_bar);
  return r;
}
inline int __double2hiint(double x) {
  long long b;
  memcpy(&b, &x, 8);
  return (int)(b >> 32);
}
inline int __double2loint(double x) {
  long long b;
  memcpy(&b, &x, 8);
  return (int)(unsigned)(b & 0xffffffffll);
}
inline double __hiloint2double(int hi, int lo) {
  const unsigned long long b = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  double x;
  memcpy(&x, &b, 8);
  return x;
}
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
// hardware reciprocal estimates: the host gives correctly rounded values (the
// device's refined estimates agree to about 1e-16)
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / sqrt(x); }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }

// ------------------------------------------------------ lane intrinsics
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned __umulhi(unsigned a, unsigned b) {
  return (unsigned)(((unsigned long long)a * b) >> 32);
}
inline void sincos(double x, double* s, double* c) { *s = sin(x); *c = cos(x); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline long max(long a, long b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline double atomicAdd(double* p, double v) {
  static_assert(sizeof(double) == sizeof(unsigned long long), "64-bit double");
  unsigned long long* q = (unsigned long long*)p;
  unsigned long long o = __atomic_load_n(q, __ATOMIC_SEQ_CST), n;
  double od;
  do {
    memcpy(&od, &o, 8);
    const double nd = od + v;
    memcpy(&n, &nd, 8);
  } while (!__atomic_compare_exchange_n(q, &o, n, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
  return od;
}
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) {
  unsigned long long o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return o;
}

// --------------------------------------------------------- kernel launch
template <class K, class... A>
void emu_launch_shm(K k, dim3 g, dim3 b, size_t shm, A... args) {
  std::vector<unsigned char> dyn(shm > 0 ? shm : 8);
  for (unsigned bz = 0; bz < g.z; ++bz)
    for (unsigned by = 0; by < g.y; ++by)
      for (unsigned blk = 0; blk < g.x; ++blk) {
        g_dyn_shared = dyn.data();
        pthread_barrier_init(&g_bar, nullptr, b.x);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < b.x; ++t)
          th.emplace_back([=]() {
            blockIdx.x = blk; blockIdx.y = by; blockIdx.z = bz;
            threadIdx.x = t; blockDim = b; gridDim = g;
            k(args...);
          });
        for (auto& x : th) x.join();
        pthread_barrier_destroy(&g_bar);
      }
}
template <class K, class... A>
void emu_launch(K k, dim3 g, dim3 b, A... args) {
  emu_launch_shm(k, g, b, 0, args...);
}
#define hipLaunchKernelGGL(k, g, b, shm, st, ...) emu_launch_shm(k, dim3(g), dim3(b), shm, __VA_ARGS__)

// ---------------------------------------------------------- runtime API
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorNotSupported = 801 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1 };
struct hipDeviceProp_t { int multiProcessorCount; };

inline const char* hipGetErrorString(hipError_t e) { return e ? "emulated HIP error" : "hipSuccess"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
// two "compute units": persistent grids stay at a couple of blocks
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  p->multiProcessorCount = 2;
  return hipSuccess;
}
template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) {
  *n = 1;
  return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t n) {
  *p = aligned_alloc(256, (n + 255) / 256 * 256);
  if (*p) memset(*p, 0xA5, n);  // uninitialised device memory is not zero
  return *p ? hipSuccess : 2;
}
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, int) {
  if (n) memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int k, hipStream_t = nullptr) {
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void* d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) {
  return hipMemset(d, v, n);
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  static int dummy;
  *s = &dummy;
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { static int dummy; *e = &dummy; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) {
  return hipErrorNotSupported;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
#define HIP_SYMBOL(x) (&x)
// device symbols are host objects: their address is the object's address
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
