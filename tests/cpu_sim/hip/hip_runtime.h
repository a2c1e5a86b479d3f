// CPU execution layer for the engine's device code (test infrastructure).
//
// Compiling gpugst_amd/csrc/gst_device.hip with g++ against this header runs
// the same item and plan functions on the host: every lane of a wave is a
// host thread, wave-wide operations (ballot, shuffles, the wave barrier)
// synchronise the 64 lanes of the wave, __syncthreads synchronises the block,
// and atomics are host atomics.  Shared (LDS) buffers are passed in by the
// driver.  Only what the engine's device code uses is provided.
#pragma once
#include <barrier>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __align__(x) __attribute__((aligned(x)))

#define __ATOMIC_RELAXED_HIP __ATOMIC_SEQ_CST
#define __HIP_MEMORY_SCOPE_WAVEFRONT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 1
#define __HIP_MEMORY_SCOPE_AGENT 2
#define __HIP_MEMORY_SCOPE_SYSTEM 3

namespace gsim {
struct Wave {
  std::barrier<> bar{64};
  uint64_t xch[64];
};
struct Block {
  std::barrier<> bar{256};
};
struct ThreadIdx {
  unsigned x;
};
extern thread_local int lane;
extern thread_local Wave* wave;
extern thread_local Block* block;
extern thread_local ThreadIdx thread_idx;
extern thread_local ThreadIdx block_idx;
extern thread_local ThreadIdx grid_dim;
extern thread_local uint64_t collectives;  // wave collectives executed by this lane
}  // namespace gsim

#define threadIdx gsim::thread_idx
#define blockIdx gsim::block_idx
#define gridDim gsim::grid_dim

inline int __lane_id() { return gsim::lane; }

struct alignas(16) uint4 {
  uint32_t x, y, z, w;
};

inline uint64_t __gsim_exchange(uint64_t v, int src) {
  gsim::Wave* w = gsim::wave;
  gsim::collectives++;
  w->xch[gsim::lane] = v;
  w->bar.arrive_and_wait();
  const uint64_t r = w->xch[src & 63];
  w->bar.arrive_and_wait();
  return r;
}

inline unsigned long long __ballot(int pred) {
  gsim::Wave* w = gsim::wave;
  gsim::collectives++;
  w->xch[gsim::lane] = pred ? 1 : 0;
  w->bar.arrive_and_wait();
  unsigned long long m = 0;
  for (int i = 0; i < 64; i++)
    if (w->xch[i]) m |= 1ull << i;
  w->bar.arrive_and_wait();
  return m;
}

template <class T>
inline T __shfl(T v, int src, int = 64) {
  return (T)__gsim_exchange((uint64_t)v, src);
}
template <class T>
inline T __shfl_up(T v, unsigned delta, int = 64) {
  const int src = gsim::lane - (int)delta;
  const T r = (T)__gsim_exchange((uint64_t)v, src < 0 ? gsim::lane : src);
  return src < 0 ? v : r;
}
template <class T>
inline T __shfl_xor(T v, int mask, int = 64) {
  return (T)__gsim_exchange((uint64_t)v, gsim::lane ^ mask);
}

// every emulated wave runs with all 64 lanes active: the first active lane is 0
inline uint32_t gsim_readfirstlane(uint32_t v) { return (uint32_t)__gsim_exchange(v, 0); }
#define __builtin_amdgcn_readfirstlane(x) gsim_readfirstlane(x)
inline uint32_t gsim_readlane(uint32_t v, int l) { return (uint32_t)__gsim_exchange(v, l); }
#define __builtin_amdgcn_readlane(x, l) gsim_readlane((x), (l))

inline unsigned long long __lanemask_lt() { return (1ull << gsim::lane) - 1ull; }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffs(unsigned x) { return __builtin_ffs((int)x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }

template <class T, class U>
inline T atomicAdd(T* p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicOr(T* p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicExch(T* p, U v) { return __atomic_exchange_n(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicMin(T* p, U vv) {
  const T v = (T)vv;
  T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}
#ifndef __clang__  // clang provides both as builtins
template <class T>
inline T __hip_atomic_load(const T* p, int, int) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline void __hip_atomic_store(T* p, U v, int, int) { __atomic_store_n(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T __hip_atomic_fetch_min(T* p, U v, int, int) { return atomicMin(p, v); }
#endif

inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline void __syncthreads() {
  __atomic_thread_fence(__ATOMIC_SEQ_CST);
  gsim::block->bar.arrive_and_wait();
}
#define __builtin_amdgcn_fence(a, b) __atomic_thread_fence(__ATOMIC_SEQ_CST)
#define __builtin_amdgcn_wave_barrier() (gsim::collectives++, gsim::wave->bar.arrive_and_wait())
#define __builtin_amdgcn_s_sleep(x) ((void)0)

#ifdef GST_CPU_SIM_HOST
// ---- host runtime stand-in for the CPU build of the whole library ----------
// Every operation completes when it is enqueued (the "device" is the host).
#include <cstddef>
typedef int hipError_t;
enum {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotReady = 600,
  hipErrorNotSupported = 801
};
typedef struct ihipStream* hipStream_t;
typedef struct ihipEvent* hipEvent_t;
typedef struct ihipGraph* hipGraph_t;
typedef struct ihipGraphExec* hipGraphExec_t;
typedef int hipMemcpyKind;
enum { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipStreamNonBlocking = 1, hipHostMallocDefault = 0, hipStreamCaptureModeThreadLocal = 1 };
enum { hipHostMallocMapped = 2, hipHostMallocCoherent = 0x40000000 };
struct hipDeviceProp_t {
  int multiProcessorCount;
};
struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
hipError_t hipGetDevice(int*);
hipError_t hipSetDevice(int);
hipError_t hipGetDeviceProperties(hipDeviceProp_t*, int);
hipError_t hipMalloc(void**, size_t);
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipFree(void*);
hipError_t hipHostMalloc(void**, size_t, unsigned);
template <class T>
inline hipError_t hipHostMalloc(T** p, size_t n, unsigned f) { return hipHostMalloc((void**)p, n, f); }
hipError_t hipHostFree(void*);
hipError_t hipHostGetDevicePointer(void**, void*, unsigned);
hipError_t hipMemcpy(void*, const void*, size_t, hipMemcpyKind);
hipError_t hipMemcpyAsync(void*, const void*, size_t, hipMemcpyKind, hipStream_t);
hipError_t hipMemset(void*, int, size_t);
hipError_t hipMemGetInfo(size_t*, size_t*);
hipError_t hipStreamCreateWithFlags(hipStream_t*, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipStreamQuery(hipStream_t);
hipError_t hipEventCreate(hipEvent_t*);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t);
hipError_t hipEventQuery(hipEvent_t);
hipError_t hipEventElapsedTime(float*, hipEvent_t, hipEvent_t);
hipError_t hipStreamBeginCapture(hipStream_t, int);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*);
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, char*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphDestroy(hipGraph_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t);
// a file-scope "device" array is host memory here: its address is the symbol's
#define HIP_SYMBOL(X) X
template <class T>
inline hipError_t hipGetSymbolAddress(void** p, const T& sym) {
  *p = (void*)&sym;
  return hipSuccess;
}
#endif
