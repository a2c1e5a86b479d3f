// Host runtime stand-in for the CPU build of libgst (test infrastructure).
#define GST_CPU_SIM_HOST 1
#include "hip/hip_runtime.h"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

struct ihipEvent {
  double t;
};

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) { p->multiProcessorCount = 1; return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
  // device memory is not zeroed: fill with a pattern like stale HBM contents
  if (posix_memalign(p, 256, n ? n : 1)) return hipErrorOutOfMemory;
  memset(*p, 0xA5, n);
  return hipSuccess;
}
hipError_t hipFree(void* p) { free(p); return hipSuccess; }
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned) { *d = h; return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
hipError_t hipMemset(void* p, int v, size_t n) { memset(p, v, n); return hipSuccess; }
hipError_t hipMemGetInfo(size_t* f, size_t* t) { *f = (size_t)2 << 30; *t = (size_t)4 << 30; return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (hipStream_t)malloc(8); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { free(s); return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
// Test hook: while set, every stream reports work still running (a device
// that does not finish in time); launches still complete when issued.
static int g_stream_busy = 0;
extern "C" void gsim_stream_busy(int on) { __atomic_store_n(&g_stream_busy, on, __ATOMIC_SEQ_CST); }
hipError_t hipStreamQuery(hipStream_t) {
  return __atomic_load_n(&g_stream_busy, __ATOMIC_SEQ_CST) ? hipErrorNotReady : hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) { *e = new ihipEvent{0}; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = now_ms(); return hipSuccess; }
// events complete when recorded (work completes when issued), unless the
// test hook holds every stream busy
hipError_t hipEventQuery(hipEvent_t) {
  return __atomic_load_n(&g_stream_busy, __ATOMIC_SEQ_CST) ? hipErrorNotReady : hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(b->t - a->t); return hipSuccess; }
// Stream capture: launches made while capturing are recorded (with their
// arguments copied) instead of run; a graph launch runs them in order.
struct ihipGraph {
  std::vector<std::function<void()>> ops;
};
struct ihipGraphExec {
  std::vector<std::function<void()>> ops;
};
static thread_local ihipGraph* g_capture = nullptr;
bool gsim_capturing() { return g_capture != nullptr; }
void gsim_record(std::function<void()> f) { g_capture->ops.push_back(std::move(f)); }
hipError_t hipStreamBeginCapture(hipStream_t, int) {
  if (g_capture) return hipErrorNotSupported;
  g_capture = new ihipGraph;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  if (!g_capture) return hipErrorNotSupported;
  *g = g_capture;
  g_capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t) {
  *e = new ihipGraphExec{g->ops};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& f : e->ops) f();
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e ? "mock runtime error" : "no error"; }
