// hip_runtime.h -- CPU stand-in for <hip/hip_runtime.h> with the subset of
// HIP that bf-new_amd/csrc/hip uses (test infrastructure; see emu_rt.cpp).
// Lanes are fibers on one host thread, switched at barriers and sleeps; a
// block's lanes share LDS.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#define __global__
#define __device__
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_load emu__hip_atomic_load
#define __hip_atomic_store emu__hip_atomic_store
#define __hip_atomic_fetch_add emu__hip_atomic_fetch_add
#define __hip_atomic_fetch_min emu__hip_atomic_fetch_min
#define __hip_atomic_fetch_max emu__hip_atomic_fetch_max
#define __hip_atomic_fetch_or emu__hip_atomic_fetch_or
#define __hip_atomic_exchange emu__hip_atomic_exchange
#define __HIP_MEMORY_SCOPE_WORKGROUP 0

struct dim3 {
  std::uint32_t x, y, z;
  dim3(std::uint32_t a = 1, std::uint32_t b = 1, std::uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct uint2 { std::uint32_t x, y; };
inline uint2 make_uint2(std::uint32_t a, std::uint32_t b) { return uint2{a, b}; }
struct alignas(16) uint4 { std::uint32_t x, y, z, w; };
inline uint4 make_uint4(std::uint32_t a, std::uint32_t b, std::uint32_t c, std::uint32_t d) {
  return uint4{a, b, c, d};
}

struct emu_graph;
struct emu_event;

namespace emu {
struct Block {
  std::uint32_t live = 0, count = 0;
  std::uint64_t gen = 0;
  int or_acc = 0;
  std::vector<std::uint32_t> lds;
  std::uint32_t wave_live[16] = {}, wave_count[16] = {};
  std::uint64_t wave_gen[16] = {};
  std::uint64_t wave_slot[1024] = {};
};
extern dim3 g_tid, g_bid, g_bdim, g_gdim;
extern Block* g_blk;
inline std::uint32_t* block_lds() { return g_blk->lds.data(); }
void yield();
void maybe_yield();
void barrier();
void wave_barrier();
void launch(dim3 grid, dim3 block, std::size_t lds, bool concurrent, std::function<void()> body);
bool capturing();
void capture(std::function<void()> f);
void begin_capture();
::emu_graph* end_capture();
}  // namespace emu

#define threadIdx (emu::g_tid)
#define blockIdx (emu::g_bid)
#define blockDim (emu::g_bdim)
#define gridDim (emu::g_gdim)

inline void __syncthreads() { emu::barrier(); }
inline int __syncthreads_or(int p) {
  emu::Block* b = emu::g_blk;
  if (p) b->or_acc = 1;
  emu::barrier();
  const int r = b->or_acc;
  emu::barrier();
  if (threadIdx.x == 0) b->or_acc = 0;
  emu::barrier();
  return r;
}
inline std::uint64_t emu_wave_exchange_read(std::uint64_t v, std::uint32_t src_lane) {
  emu::Block* b = emu::g_blk;
  const std::uint32_t w = threadIdx.x / 64, lane = threadIdx.x % 64;
  b->wave_slot[w * 64 + lane] = v;
  emu::wave_barrier();
  const std::uint64_t r = b->wave_slot[w * 64 + src_lane];
  emu::wave_barrier();
  return r;
}
inline unsigned long long __ballot(int p) {
  emu::Block* b = emu::g_blk;
  const std::uint32_t w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const std::uint32_t n = std::min<std::uint32_t>(64, blockDim.x - w * 64);
  b->wave_slot[w * 64 + lane] = p ? 1 : 0;
  emu::wave_barrier();
  unsigned long long m = 0;
  for (std::uint32_t i = 0; i < n; ++i) m |= (b->wave_slot[w * 64 + i] ? 1ull : 0ull) << i;
  emu::wave_barrier();
  return m;
}
inline std::uint32_t __shfl_xor(std::uint32_t v, int off, int = 64) {
  return static_cast<std::uint32_t>(emu_wave_exchange_read(v, (threadIdx.x % 64) ^ off));
}
inline std::uint32_t __shfl(std::uint32_t v, int src, int = 64) {
  return static_cast<std::uint32_t>(emu_wave_exchange_read(v, static_cast<std::uint32_t>(src) % 64));
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __ffs(unsigned x) { return __builtin_ffs(static_cast<int>(x)); }
inline float __uint_as_float(std::uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline std::uint32_t __float_as_uint(float f) { std::uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline void __threadfence() {}
inline void __threadfence_block() {}
inline long long clock64() {
  return std::chrono::steady_clock::now().time_since_epoch().count();
}
#define __builtin_amdgcn_fence(o, s) ((void)0)
#define __builtin_amdgcn_s_sleep(n) emu::yield()
#define __builtin_amdgcn_readfirstlane(x) (x)

template <typename T> inline T min(T a, T b) { return b < a ? b : a; }
template <typename T> inline T max(T a, T b) { return a < b ? b : a; }

template <typename T> inline T atomicAdd(T* p, T v) { emu::maybe_yield(); return std::atomic_ref<T>(*p).fetch_add(v); }
template <typename T> inline T atomicOr(T* p, T v) { return std::atomic_ref<T>(*p).fetch_or(v); }
template <typename T> inline T atomicMin(T* p, T v) {
  emu::maybe_yield();
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (v < cur && !a.compare_exchange_weak(cur, v)) {}
  return cur;
}
template <typename T> inline T atomicMax(T* p, T v) {
  emu::maybe_yield();
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (cur < v && !a.compare_exchange_weak(cur, v)) {}
  return cur;
}
template <typename T> inline T atomicCAS(T* p, T cmp, T v) {
  std::atomic_ref<T>(*p).compare_exchange_strong(cmp, v);
  return cmp;
}
template <typename T> inline T emu__hip_atomic_load(const T* p, int, int) {
  emu::maybe_yield();
  return std::atomic_ref<T>(*const_cast<T*>(p)).load();
}
template <typename T> inline void emu__hip_atomic_store(T* p, T v, int, int) { emu::maybe_yield(); std::atomic_ref<T>(*p).store(v); }
template <typename T> inline T emu__hip_atomic_fetch_add(T* p, T v, int, int) { return atomicAdd(p, v); }
template <typename T> inline T emu__hip_atomic_fetch_min(T* p, T v, int, int) { return atomicMin(p, v); }
template <typename T> inline T emu__hip_atomic_fetch_max(T* p, T v, int, int) { return atomicMax(p, v); }
template <typename T> inline T emu__hip_atomic_fetch_or(T* p, T v, int, int) { return atomicOr(p, v); }
template <typename T> inline T emu__hip_atomic_exchange(T* p, T v, int, int) { return std::atomic_ref<T>(*p).exchange(v); }

// ---- runtime API ----------------------------------------------------------
enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotReady = 600 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
typedef void* hipStream_t;
typedef struct emu_event* hipEvent_t;
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graph* hipGraphExec_t;
#define hipStreamNonBlocking 1u
#define hipHostMallocDefault 0u
struct hipDeviceProp_t {
  char name[256];
  char gcnArchName[256];
  std::size_t totalGlobalMem;
  std::size_t sharedMemPerBlock;
  std::size_t maxSharedMemoryPerMultiProcessor;
  int multiProcessorCount;
  int warpSize;
};
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipSetDevice(int);
hipError_t hipGetDeviceCount(int*);
hipError_t hipGetDeviceProperties(hipDeviceProp_t*, int);
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void**, std::size_t);
template <typename T> inline hipError_t hipMalloc(T** p, std::size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n); }
// Host memory stands in for device memory: a symbol's address is its own.
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
  *p = const_cast<void*>(symbol);
  return hipSuccess;
}
hipError_t hipFree(void*);
hipError_t hipHostMalloc(void**, std::size_t, unsigned = 0);
template <typename T> inline hipError_t hipHostMalloc(T** p, std::size_t n, unsigned f = 0) { return hipHostMalloc(reinterpret_cast<void**>(p), n, f); }
hipError_t hipHostFree(void*);
hipError_t hipMemcpyAsync(void*, const void*, std::size_t, hipMemcpyKind, hipStream_t = nullptr);
hipError_t hipMemsetAsync(void*, int, std::size_t, hipStream_t = nullptr);
hipError_t hipStreamCreateWithFlags(hipStream_t*, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipEventCreate(hipEvent_t*);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr);
hipError_t hipEventElapsedTime(float*, hipEvent_t, hipEvent_t);
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*);
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, char*, std::size_t);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
hipError_t hipGraphDestroy(hipGraph_t);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int*, const void*, int, std::size_t);

#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
  emu::launch(dim3(grid), dim3(block), (lds), false, [=]() { kernel(__VA_ARGS__); })

template <typename A0, typename A1, typename A2, typename A3>
inline hipError_t emu_launch_ptr(void (*f)(A0, A1, A2, A3), dim3 grid, dim3 block, void** args,
                                 std::size_t lds, hipStream_t, bool concurrent) {
  // arguments are copied now, as the runtime does at launch
  const A0 a0 = *static_cast<A0*>(args[0]);
  const A1 a1 = *static_cast<A1*>(args[1]);
  const A2 a2 = *static_cast<A2*>(args[2]);
  const A3 a3 = *static_cast<A3*>(args[3]);
  emu::launch(grid, block, lds, concurrent, [=]() { f(a0, a1, a2, a3); });
  return hipSuccess;
}
