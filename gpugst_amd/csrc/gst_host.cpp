// gst_host.cpp -- host side of libgst: graph residency, the slot engine that
// drives the step and plan kernels, group-distance jobs, memory admission,
// batch and server paths, and the C ABI declared in include/gst.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gst.h"
#include "gst_common.h"

extern "C" hipError_t gst_launch_step(int cls, int wpb, const gst::GraphDev* g, gst::Slot* slots, gst::Ctrl* ctrl,
                                      int grid, hipStream_t stream);
extern "C" hipError_t gst_launch_plan(gst::Slot* slots, gst::Ctrl* ctrl, uint64_t V, hipStream_t stream);
extern "C" hipError_t gst_step_occupancy(int cls, int wpb, int* blocks_per_cu);
extern "C" hipError_t gst_scratch_setup(int device, void** ptr);
extern "C" hipError_t gst_launch_count(gst::Slot* slot, uint64_t V, hipStream_t stream);
extern "C" hipError_t gst_launch_gather(const gst::GatherEnt* t, int n, gst::EdgeDev* dst, hipStream_t stream);
extern "C" hipError_t gst_bb_fill(uint32_t* p, uint64_t n, uint32_t val, hipStream_t stream);
extern "C" hipError_t gst_bb_seed(uint32_t* dist, uint8_t* flags, const int32_t* hubs, int H, hipStream_t stream);
extern "C" hipError_t gst_bb_pass(int what, const gst::BbArgs* a, hipStream_t stream);

namespace {

thread_local std::string g_last_error;

gst_status fail(gst_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail(GST_DEVICE_ERROR, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                         \
  } while (0)

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Engine steps per epoch before control returns to the host loop.
constexpr uint32_t kEpochMaxSteps = 1u << 20;
// (step, plan) pairs per captured hipGraph chunk.
constexpr int kGraphChunk = 16;

// First-fit allocator over one device region, with coalescing of free runs.
class Arena {
 public:
  void reset(size_t size) {
    free_.clear();
    size_ = size;
    if (size) free_[0] = size;
  }
  bool alloc(size_t len, size_t* off) {
    len = align256(len);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        *off = it->first;
        const size_t rest = it->second - len, noff = it->first + len;
        free_.erase(it);
        if (rest) free_[noff] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t len) {
    len = align256(len);
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t largest() const {
    size_t m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
  size_t size() const { return size_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0;
};

struct Options {
  int64_t max_slots = 512;  // queries and group-distance jobs admitted at once
  int64_t budget_mb = 0;  // 0: most of the free device memory
  int64_t blocks_per_cu = 0;  // 0: occupancy maximum
  int64_t chunk = 512;
  int64_t edge_cap = 65536;
  int64_t timeout_ms = 60000;  // host-side deadline for one chunk of engine steps
  int64_t graph = 0;  // replay chunks of engine steps as a captured hipGraph
  int64_t group_pass = 1;  // unbounded queries take level 1 from shared group-distance jobs
  int64_t time_steps = 0;  // time each step's lean and heavy kernels with events
  int64_t trace_slot = 0;  // with time_steps: slot whose phases the step times are binned by
  int64_t job_min_queries = 2;  // fewer unbounded queries admitted together run their own level 1
  int64_t deadline_ms = 0;  // a query still running this long after its submission is cancelled (0: none)
  int64_t cache_mb = 0;  // device memory for cached group-distance columns (0: no cache)
  int64_t approx_num = 1, approx_den = 1;  // progressive answers within num / den of a certified bound
  // schedule (the autotuner's knobs besides chunk / blocks_per_cu / graph / max_slots)
  int64_t block = 2;      // waves per block of the step kernels: 2, 4 or 8 (2: measured fastest)
  int64_t merge_tv = gst::kMergeTV;  // live vertices per MERGE / FINAL work item
  int64_t job_lanes = 0;  // lanes sharing an arc in job rounds (power of two; 0: the smallest with <= 8 chunks each)
  int64_t count_states = 0;  // count the finite DP states of every query (an extra pass)
  int64_t backbone = 1;  // unbounded queries and jobs admitted from now on run on the shortest-path backbone
  int64_t hub_bounds = 1;  // unbounded jobs admitted from now on start their rows at hub bounds
  int64_t hub16 = 1;  // ... computed in packed 16-bit arithmetic where the graph's hub distances allow
  // server (continuous stepping): engine steps per chunk between two host
  // visits, chunks kept queued on the device, and how long / for how many
  // pending requests admission waits while the device is busy, so that more
  // requests share a group-distance job
  int64_t serve_chunk = 4;
  int64_t serve_depth = 2;
  int64_t serve_wait_us = 10000;  // measured: 1.1M-vertex graph, 1000-3300 q/s offered
  int64_t serve_batch = 64;
};

}  // namespace

// index of a block shape (waves per block) in kWaveChoices
static int wpb_index(int64_t wpb) { return wpb == 2 ? 0 : (wpb == 8 ? 2 : 1); }

struct gst_graph {
  int device = 0;
  int num_cu = 0;
  int occ[3][gst::kNumCls] = {{1, 1}, {1, 1}, {1, 1}};  // blocks per CU of the step kernels, per block shape
  int64_t V = 0, n_arcs = 0;
  int64_t n_groups = 0;
  int64_t groups_version = 0;  // bumped by every gst_groups_set (group ids change meaning)
  std::vector<int64_t> h_gptr;
  std::vector<int32_t> h_gmem;
  int64_t* d_rowptr = nullptr;
  int32_t* d_col = nullptr;
  int32_t* d_wt = nullptr;
  int32_t* d_rev = nullptr;
  // shortest-path backbone: its arcs follow the graph's in d_col / d_wt /
  // d_rev and its row starts follow the graph's in d_rowptr (has_bb false
  // when the graph has no arcs or its build did not fit)
  bool has_bb = false;
  int64_t bb_arcs = 0, bb_rounds = 0;
  // exact distances from the backbone's hubs, V rows of kBbHubs, on the device
  // and on the host (the host forms each job's hub bounds); null / empty when
  // the build did not run
  uint32_t* d_hubd = nullptr;
  uint16_t* d_hubd16 = nullptr;  // the 16-bit copy (GraphDev::hubd16), or null
  std::vector<uint32_t> h_hubd;
  // vertices some hub reaches, then the others; n_hubr of the first kind
  uint32_t* d_hubv = nullptr;
  int64_t n_hubr = 0;
  double bb_ms = 0;
  int64_t* d_gptr = nullptr;
  int32_t* d_gmem = nullptr;
  Options opt;
  std::mutex dev_mu;  // one engine on the device at a time
  // Servers step continuously, so their chunks stay queued on the device
  // after they release dev_mu: cont_out counts those chunks (under dev_mu).
  // gst_groups_set raises quiesce, under which servers queue no new chunk,
  // and changes the device group arrays only once cont_out is 0.
  int cont_out = 0;
  std::atomic<int> quiesce{0};
  std::unique_ptr<class Engine> plain;
  std::unique_ptr<class Engine> batch;
  gst::GraphDev dev() const {
    gst::GraphDev g;
    g.rowptr = d_rowptr;
    g.col = d_col;
    g.wt = d_wt;
    g.rev = d_rev;
    g.gptr = d_gptr;
    g.gmem = d_gmem;
    g.hubd = d_hubd;
    g.hubd16 = d_hubd16;
    g.hubv = d_hubv;
    g.n_hubr = n_hubr;
    g.V = V;
    g.n_arcs = n_arcs;
    return g;
  }
};

namespace {

struct QueryReq {
  std::vector<int32_t> groups;
  int32_t D = -1;
  uint64_t tag = 0;
  double t_submit = 0;
  double t_deadline = 0;  // host time after which the query is cancelled (0: none)
};

struct Layout {
  size_t T, S1, live, lvb, act, mask, C, list, loff, istart, R, edges, rq, total;
  int edge_cap;
  uint32_t istride;
};

// entries per buffer of istart for ROUND items of `chunk` arcs
uint32_t istride_for(const gst_graph* g, int64_t chunk) {
  return (uint32_t)(g->n_arcs / std::max<int64_t>(chunk, 1) + 2);
}

// True when the query can take its level 1 from a group-distance job (k >= 2;
// k = 1 is answered at SEED).  A job serves queries of one hop bound.
bool uses_job(const gst_graph* g, int k, int D) { return g->opt.group_pass != 0 && k >= 2; }

// True when a query or job of hop bound D runs on the shortest-path backbone:
// unbounded only (a hop bound can need an arc that no shortest path uses).
bool uses_backbone(const gst_graph* g, int D) { return D < 0 && g->opt.backbone != 0 && g->has_bb; }

// Device memory of one query; `own` = it computes its own level 1 (S1 and C
// tables); a job-fed query reads both from its job's rows.
Layout layout_for(const gst_graph* g, int k, int D, bool own) {
  const size_t V = (size_t)g->V;
  const size_t W = size_t(1) << k;
  const size_t ND = D < 0 ? 1 : (size_t)D + 1;
  Layout L;
  int64_t cap = std::min<int64_t>((int64_t)(V * ND), g->opt.edge_cap);
  L.edge_cap = (int)std::max<int64_t>(cap, 64);
  L.istride = istride_for(g, g->opt.chunk);
  size_t off = 0;
  // states need heights 0..ceil(D/2): rooted at its center, a tree of diameter
  // <= D has height <= ceil(D/2)
  const size_t NS = D < 0 ? 1 : (size_t)(D + 1) / 2 + 1;
  L.T = off;    off += align256(V * NS * W * 8);
  L.S1 = off;   off += own ? align256(V * ND * k * 8) : 0;
  L.live = off; off += align256(V * 4);
  L.lvb = off;  off += align256(2 * ((V + 63) / 64) * 8);  // live bits, then hold bits
  L.act = off;  off += align256(V * 4);
  L.mask = off; off += align256(3 * V * 8);
  L.C = off;    off += own ? align256(V * ND * k * 4) : 0;
  L.list = off; off += align256(3 * V * 4);
  L.loff = off; off += align256(3 * V * 4);
  L.istart = off; off += align256(3 * (size_t)L.istride * 4);
  L.R = off;    off += align256(W * k * k * 4);
  L.edges = off; off += align256((size_t)L.edge_cap * sizeof(gst::EdgeDev));
  L.rq = off;   off += align256(3 * ((size_t)L.edge_cap + 2 * gst::kMaxK) * 4);
  L.total = off;
  return L;
}

Layout layout_for(const gst_graph* g, int k, int D) { return layout_for(g, k, D, !uses_job(g, k, D)); }

// Group-distance job over L labels.
struct JobLayout {
  size_t dist, mask, list, loff, istart, lgrp, jmem, ccol, hA, total;
  int Lp, G, chunk, cl;
  uint32_t istride;
};

JobLayout job_layout(const gst_graph* g, int L, int nmem, int D) {
  const size_t V = (size_t)g->V;
  JobLayout J;
  // rows of whole chunks, and of whole u64 words (FILL writes pairs of labels)
  const int unit = D < 0 ? gst::kJobChunkFlat : gst::kJobChunkHop;  // labels per dirty chunk
  J.Lp = (L + unit - 1) / unit * unit;
  // lanes per arc: the smallest power of two that gives each lane at most
  // kJobChunksPerLane chunks (64 chunks: 8 lanes, 8 arcs per wave iteration;
  // one or two chunks: 1 lane, 64 arcs)
  J.cl = unit;
  const int nch = J.Lp / unit;
  J.G = 1;
  while (J.G * gst::kJobChunksPerLane < nch) J.G <<= 1;
  if (g->opt.job_lanes > 0) J.G = (int)g->opt.job_lanes;
  J.chunk = (int)gst::kJobArcs;
  J.istride = istride_for(g, J.chunk);
  size_t off = 0;
  const size_t ND = D < 0 ? 1 : (size_t)D + 1;
  J.dist = off;   off += align256(V * ND * (size_t)J.Lp * 4);  // one row per vertex and height
  J.mask = off;   off += align256(3 * V * 8);
  J.list = off;   off += align256(3 * V * 4);
  J.loff = off;   off += align256(3 * V * 4);
  J.istart = off; off += align256(3 * (size_t)J.istride * 4);
  J.lgrp = off;   off += align256((size_t)L * 4);
  J.jmem = off;   off += align256((size_t)nmem * 4);
  J.ccol = off;   off += align256((size_t)L * 2 * sizeof(void*));
  J.hA = off;     off += align256((size_t)J.Lp * gst::kBbHubs * 4);
  J.total = off;
  return J;
}

gst_status validate_query(const gst_graph* g, const int32_t* groups, int32_t k, int32_t D) {
  if (!g) return fail(GST_INVALID_ARG, "null graph");
  if (g->n_groups <= 0) return fail(GST_INVALID_ARG, "no groups set (call gst_groups_set)");
  if (k < 1 || k > GST_MAX_K) return fail(GST_INVALID_ARG, "k=%d outside [1,%d]", k, GST_MAX_K);
  if (D < GST_UNBOUNDED || D > GST_MAX_D)
    return fail(GST_INVALID_ARG, "D=%d outside [-1,%d]", D, GST_MAX_D);
  if (!groups) return fail(GST_INVALID_ARG, "null group id list");
  for (int i = 0; i < k; i++)
    if (groups[i] < 0 || groups[i] >= g->n_groups)
      return fail(GST_INVALID_ARG, "group id %d outside [0,%lld)", groups[i], (long long)g->n_groups);
  return GST_OK;
}

void fill_infeasible(gst_result* r) {
  memset(r, 0, sizeof(*r));
  r->cost = GST_INFEASIBLE;
  r->root = -1;
  r->status = GST_OK;
}

struct SlotHost {
  bool used = false;
  bool job = false;      // a group-distance job (no result of its own)
  bool charged = false;  // job: its work has been charged to its members
  size_t off = 0, len = 0;
  uint64_t tag = 0;
  double t_submit = 0, t_admit = 0;
  double t_deadline = 0;  // 0: none
  bool flagged = false;   // its cancellation flag is set
  int64_t extra[5] = {0, 0, 0, 0, 0};  // level-1 work charged from the query's job
  uint32_t jmaxq = 0;     // largest frontier of its job
  int32_t cached = 0;     // its labels served from the group-distance cache
  // continuous stepping: admission number (a snapshot of the slot table
  // describes this slot only if taken after its entry was uploaded) and
  // whether a snapshot has shown it finished
  uint64_t aseq = 0;
  bool done_seen = false;
};

// One cached group-distance column (unbounded distances of one group to every
// vertex), in the engine's arena.
struct CacheEntry {
  size_t off = 0;
  uint64_t last = 0;   // LRU clock of the last admission that used it
  int refs = 0;        // jobs reading or writing it
  bool valid = false;  // written by a finished job
  bool stale = false;  // the groups were replaced while it was referenced
};

}  // namespace

class Engine {
 public:
  Engine(gst_graph* g, int max_slots) : g_(g), max_slots_(std::min(max_slots, gst::kMaxSlots)) {}
  ~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (h_cancel_) (void)hipHostFree(h_cancel_);
    if (borrowed_) {
      delete[] h_slots_;
      delete h_ctrl_;
    } else {
      if (arena_mem_) (void)hipFree(arena_mem_);
      if (d_slots_) (void)hipFree(d_slots_);
      if (d_ctrl_) (void)hipFree(d_ctrl_);
      if (h_slots_) (void)hipHostFree(h_slots_);
      if (h_ctrl_) (void)hipHostFree(h_ctrl_);
    }
    drop_graph();
    for (int b = 0; b < kContBuffers; b++) {
      if (cev_[b]) (void)hipEventDestroy(cev_[b]);
      if (cev0_[b]) (void)hipEventDestroy(cev0_[b]);
      if (h_snap_[b]) (void)hipHostFree(h_snap_[b]);
    }
    if (h_zero_) (void)hipHostFree(h_zero_);
    if (!borrowed_) {
      if (d_gtab_) (void)hipFree(d_gtab_);
      if (h_gtab_) (void)hipHostFree(h_gtab_);
      if (d_gstage_) (void)hipFree(d_gstage_);
      if (h_gstage_) (void)hipHostFree(h_gstage_);
    }
    if (xstream_) {
      (void)hipStreamSynchronize(xstream_);
      (void)hipStreamDestroy(xstream_);
    }
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    for (auto e : tev_) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  gst_status init(size_t arena_bytes) {
    HIP_TRY(hipSetDevice(g_->device));
    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&ev0_));
    HIP_TRY(hipEventCreate(&ev1_));
    HIP_TRY(hipMalloc(&d_slots_, sizeof(gst::Slot) * max_slots_));
    HIP_TRY(hipMalloc(&d_ctrl_, sizeof(gst::Ctrl)));
    HIP_TRY(hipHostMalloc(&h_slots_, sizeof(gst::Slot) * max_slots_, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&h_ctrl_, sizeof(gst::Ctrl), hipHostMallocDefault));
    memset(h_slots_, 0, sizeof(gst::Slot) * max_slots_);
    memset(h_ctrl_, 0, sizeof(gst::Ctrl));
    HIP_TRY(hipMemset(d_ctrl_, 0, sizeof(gst::Ctrl)));
    sh_.assign(max_slots_, SlotHost());
    // cancellation flags: host memory the plan kernel reads while it runs
    HIP_TRY(hipHostMalloc((void**)&h_cancel_, sizeof(uint32_t) * (max_slots_ + 1),
                          hipHostMallocMapped | hipHostMallocCoherent));
    memset(h_cancel_, 0, sizeof(uint32_t) * (max_slots_ + 1));
    HIP_TRY(hipHostGetDevicePointer((void**)&d_cancel_, h_cancel_, 0));
    return resize_arena(arena_bytes);
  }

  // An engine over a device region it does not own (the static scratch of
  // gst_solve_static): slot table, control block and arena are carved out of
  // [base, base + bytes); the host mirrors are plain host memory.  Nothing
  // here or later allocates device memory.  Every query it admits runs its own
  // level 1 (no group-distance job).
  gst_status init_static(char* base, size_t bytes) {
    borrowed_ = true;
    force_own_ = true;
    HIP_TRY(hipSetDevice(g_->device));
    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&ev0_));
    HIP_TRY(hipEventCreate(&ev1_));
    size_t off = 0;
    d_slots_ = (gst::Slot*)(base + off);
    off += align256(sizeof(gst::Slot) * max_slots_);
    d_ctrl_ = (gst::Ctrl*)(base + off);
    off += align256(sizeof(gst::Ctrl));
    if (off >= bytes) return fail(GST_OUT_OF_MEMORY, "static scratch too small for the engine tables");
    h_slots_ = new gst::Slot[max_slots_];
    h_ctrl_ = new gst::Ctrl;
    memset(h_slots_, 0, sizeof(gst::Slot) * max_slots_);
    memset(h_ctrl_, 0, sizeof(gst::Ctrl));
    HIP_TRY(hipMemcpyAsync(d_ctrl_, h_ctrl_, sizeof(gst::Ctrl), hipMemcpyHostToDevice, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    sh_.assign(max_slots_, SlotHost());
    arena_mem_ = base + off;
    arena_.reset((bytes - off) & ~size_t(255));
    return GST_OK;
  }

  // Points an idle engine at another graph of the same device (the static
  // engine serves every graph handle of its device, one call at a time).
  void rebind(gst_graph* g) {
    if (g_ != g) drop_graph();  // a captured chunk holds the old graph's arguments
    g_ = g;
  }

  gst_status resize_arena(size_t bytes) {
    if (borrowed_) return fail(GST_INVALID_ARG, "the static engine's arena is fixed");
    if (active_ > 0) return fail(GST_INVALID_ARG, "arena resize with active queries");
    if (arena_mem_) {
      HIP_TRY(hipFree(arena_mem_));
      arena_mem_ = nullptr;
    }
    bytes = align256(bytes);
    if (bytes) {
      hipError_t e = hipMalloc((void**)&arena_mem_, bytes);
      if (e != hipSuccess) {
        arena_mem_ = nullptr;
        arena_.reset(0);
        return fail(GST_OUT_OF_MEMORY, "hipMalloc(%zu) for the state arena failed: %s", bytes,
                    hipGetErrorString(e));
      }
    }
    arena_.reset(bytes);
    return GST_OK;
  }

  size_t arena_size() const { return arena_.size(); }
  // admitted queries (group-distance jobs are not counted)
  int active() const { return active_; }
  bool has_free_slot() const { return used_slots_ < max_slots_; }

  // Admits queries from the front of `pend` while slots and memory allow and
  // returns how many were admitted.  Consecutive unbounded queries (k >= 2)
  // share a group-distance job of up to kJobLabels labels; a job is sized for
  // the members that fit together with it.
  int admit_front(std::deque<QueryReq*>& pend) {
    int n = 0;
    while (!pend.empty()) {
      QueryReq* q0 = pend.front();
      const int k0 = (int)q0->groups.size();
      if (force_own_ || !uses_job(g_, k0, q0->D)) {
        if (used_slots_ >= max_slots_ || admit_query(*q0, -1, 0) < 0) break;
        pend.pop_front();
        n++;
        continue;
      }
      // the job takes the front query and the next pending queries of the
      // same bound (skipping others, which keep their order), within a window
      // of the queue
      std::vector<QueryReq*> grp;
      std::vector<size_t> at;  // their positions in pend
      int L = 0;
      const size_t window = std::min(pend.size(), (size_t)(4 * gst::kJobLabels));
      for (size_t i = 0; i < window; i++) {
        QueryReq* q = pend[i];
        const int k = (int)q->groups.size();
        if (!uses_job(g_, k, q->D) || q->D != q0->D) continue;
        if (L + k > gst::kJobLabels || used_slots_ + (int)grp.size() + 2 > max_slots_) break;
        grp.push_back(q);
        at.push_back(i);
        L += k;
      }
      if ((int64_t)grp.size() < g_->opt.job_min_queries) {
        // a job pays off from about two queries on: alone, the query's own
        // level 1 is faster
        if (used_slots_ >= max_slots_ || admit_query(*q0, -1, 0) < 0) break;
        pend.pop_front();
        n++;
        continue;
      }
      const size_t want = grp.size();
      while (!grp.empty() && !admit_job_group(grp, L)) {
        L -= (int)grp.back()->groups.size();
        grp.pop_back();
        at.pop_back();
      }
      for (size_t i = at.size(); i-- > 0;) pend.erase(pend.begin() + (std::ptrdiff_t)at[i]);
      n += (int)grp.size();
      if (grp.size() < want || grp.empty()) break;
    }
    return n;
  }

 private:
  int free_slot() const {
    for (int i = 0; i < max_slots_; i++)
      if (!sh_[i].used) return i;
    return -1;
  }

  // A query slot; job >= 0 names the slot of its group-distance job, lab its
  // first label there and jdist / jLp the job's distance rows (the query's
  // C).  Returns the slot index, -1 when memory or slots are exhausted.
  int admit_query(const QueryReq& q, int job, int lab, uint32_t* jdist = nullptr, int jLp = 0) {
    // jdist: the job's rows, ND rows of jLp labels per vertex
    const int k = (int)q.groups.size();
    if (k < 1 || k > gst::kMaxK) return -1;  // validated at submission; never admitted
    const bool own = job < 0;
    const Layout L = layout_for(g_, k, q.D, own);
    const int s = free_slot();
    if (s < 0) return -1;
    size_t off = 0;
    if (!alloc_or_evict(L.total, &off)) return -1;
    SlotHost& h = sh_[s];
    h.used = true;
    h.job = false;
    h.off = off;
    h.len = L.total;
    h.tag = q.tag;
    h.t_submit = q.t_submit;
    h.t_admit = now_ms();
    h.t_deadline = q.t_deadline;
    if (h.t_deadline > 0) n_deadline_++;
    h.flagged = false;
    h.jmaxq = 0;
    h.cached = 0;
    mark_admitted(s);
    if (h_cancel_) __atomic_store_n(&h_cancel_[1 + s], 0u, __ATOMIC_RELEASE);
    gst::Slot& d = h_slots_[s];
    memset(&d, 0, sizeof(d));
    d.kind = gst::KIND_QUERY;
    d.job = job;
    d.lab = lab;
    d.k = k;
    d.W = 1 << k;
    d.ND = q.D < 0 ? 1 : q.D + 1;
    d.D = q.D;
    d.bb = uses_backbone(g_, q.D) ? 1u : 0u;
    for (int i = 0; i < k; i++) d.qg[i] = q.groups[i];
    d.edge_cap = L.edge_cap;
    d.chunk = (int32_t)g_->opt.chunk;
    d.istride = L.istride;
    d.NS = q.D < 0 ? 1 : (q.D + 1) / 2 + 1;
    d.nT = (uint64_t)g_->V * d.NS * d.W;
    char* base = arena_mem_ + off;
    d.T = (uint64_t*)(base + L.T);
    d.S1 = own ? (uint64_t*)(base + L.S1) : nullptr;
    d.live = (uint32_t*)(base + L.live);
    d.lvb = (uint32_t*)(base + L.lvb);
    d.hold = (uint32_t*)(base + L.lvb + ((size_t)g_->V + 63) / 64 * 8);
    d.act = (uint32_t*)(base + L.act);
    d.mask = (uint64_t*)(base + L.mask);
    d.C = own ? (uint32_t*)(base + L.C) : jdist + lab;
    d.cstride = own ? (uint32_t)(k * d.ND) : (uint32_t)(d.ND * jLp);
    d.chs = own ? (uint32_t)k : (uint32_t)jLp;
    if (!own) {
      d.J = jdist + lab;
      d.jvs = (uint32_t)(d.ND * jLp);
      d.jds = q.D >= 0 ? (uint32_t)jLp : 0u;
    }
    d.list = (uint32_t*)(base + L.list);
    d.loff = (uint32_t*)(base + L.loff);
    d.istart = (uint32_t*)(base + L.istart);
    d.R = (uint32_t*)(base + L.R);
    d.edges = (gst::EdgeDev*)(base + L.edges);
    d.rq = (uint32_t*)(base + L.rq);
    d.phase = gst::PH_FILL;
    d.count = g_->opt.count_states ? 1u : 0u;
    d.mtv = (uint32_t)g_->opt.merge_tv;
    if (g_->opt.approx_num > g_->opt.approx_den) {
      d.anum = (uint32_t)g_->opt.approx_num;
      d.aden = (uint32_t)g_->opt.approx_den;
    }
    d.lbmin = gst::kInf;
    d.cert = gst::kInf;
    d.star_pk = ~0ull;
    d.best = gst::kInf;
    d.ans = ~0ull;
    d.out_cost = (int32_t)gst::kInf;
    d.out_root = -1;
    nslots_used_ = std::max(nslots_used_, s + 1);
    used_slots_++;
    active_++;
    return s;
  }

  void release_slot(int s) {
    SlotHost& h = sh_[s];
    if (h.flagged) {
      h.flagged = false;
      if (--n_flagged_ == 0 && h_cancel_) __atomic_store_n(&h_cancel_[0], 0u, __ATOMIC_RELEASE);
    }
    if (h.t_deadline > 0) {
      h.t_deadline = 0;
      n_deadline_--;
    }
    arena_.release(h.off, h.len);
    if (!h.job) active_--;
    if (!h.done_seen) n_inflight_--;  // never uploaded (a failed admission)
    h.used = false;
    h.job = false;
    h.charged = false;
    h.done_seen = false;
    used_slots_--;
    memset(&h_slots_[s], 0, sizeof(gst::Slot));  // PH_FREE
    while (nslots_used_ > 0 && !sh_[nslots_used_ - 1].used) nslots_used_--;
  }

  // One job over the labels of `grp` plus its member queries; all or nothing.
  bool admit_job_group(const std::vector<QueryReq*>& grp, int L) {
    const int D = grp.front()->D;
    const JobLayout J = job_layout(g_, L, (int)grp.size(), D);
    const int js = free_slot();
    if (js < 0) return false;
    size_t off = 0;
    if (!alloc_or_evict(J.total, &off)) return false;
    SlotHost& h = sh_[js];
    h.used = true;
    h.job = true;
    h.off = off;
    h.len = J.total;
    mark_admitted(js);
    used_slots_++;
    nslots_used_ = std::max(nslots_used_, js + 1);
    std::vector<int> mem;
    int lab = 0;
    uint32_t* jdist = (uint32_t*)(arena_mem_ + off + J.dist);
    for (QueryReq* q : grp) {
      const int s = admit_query(*q, js, lab, jdist, J.Lp);
      if (s < 0) {
        for (int m : mem) release_slot(m);
        release_slot(js);
        return false;
      }
      mem.push_back(s);
      lab += (int)q->groups.size();
    }
    job_members_[js] = mem;
    job_refs_[js] = (int)mem.size();
    for (int m : mem) member_job_[m] = js;
    gst::Slot& d = h_slots_[js];
    memset(&d, 0, sizeof(d));
    d.kind = gst::KIND_JOB;
    d.job = -1;
    d.D = D;
    d.ND = D < 0 ? 1 : D + 1;
    d.bb = uses_backbone(g_, D) ? 1u : 0u;
    d.L = L;
    d.Lp = J.Lp;
    d.jcl = (uint32_t)J.cl;
    d.G = J.G;
    d.chunk = J.chunk;
    d.istride = J.istride;
    d.nmem = (int32_t)mem.size();
    char* base = arena_mem_ + off;
    d.dist = (uint32_t*)(base + J.dist);
    d.mask = (uint64_t*)(base + J.mask);
    d.list = (uint32_t*)(base + J.list);
    d.loff = (uint32_t*)(base + J.loff);
    d.istart = (uint32_t*)(base + J.istart);
    d.lgrp = (int32_t*)(base + J.lgrp);
    d.jmem = (uint32_t*)(base + J.jmem);
    d.phase = gst::PH_FILL;
    // label -> group and member lists go up with the next epoch's uploads
    std::vector<int32_t> lg;
    for (QueryReq* q : grp) lg.insert(lg.end(), q->groups.begin(), q->groups.end());
    std::vector<uint32_t> jm(mem.begin(), mem.end());
    pending_uploads_.push_back({d.lgrp, std::vector<char>((char*)lg.data(), (char*)(lg.data() + lg.size()))});
    pending_uploads_.push_back({d.jmem, std::vector<char>((char*)jm.data(), (char*)(jm.data() + jm.size()))});
    if (D < 0 && g_->opt.hub_bounds != 0 && g_->d_hubd) {
      // hub bounds of the labels: hub h to the nearest member of the label's
      // group, label-major (padding labels: none)
      std::vector<uint32_t> hA((size_t)J.Lp * gst::kBbHubs, gst::kInf);
      const uint32_t* hd = g_->h_hubd.data();
      for (int l = 0; l < L; l++) {
        const int32_t gid = lg[l];
        uint32_t* x = hA.data() + (size_t)l * gst::kBbHubs;
        for (int64_t t = g_->h_gptr[gid]; t < g_->h_gptr[gid + 1]; t++) {
          const uint32_t* row = hd + (size_t)g_->h_gmem[t] * gst::kBbHubs;
          for (int hh = 0; hh < gst::kBbHubs; hh++) x[hh] = std::min(x[hh], row[hh]);
        }
      }
      // each member starts from its hub star, min over hubs h of the sum of
      // its groups' distances to h: shortest paths from h to every group form
      // a real walk-tree, and on power-law graphs a hub is often the best
      // star centre.  The job live test then runs before the true stars are
      // known (they are taken in the same pass).
      for (size_t m = 0; m < grp.size(); m++) {
        gst::Slot& q = h_slots_[mem[m]];
        uint32_t hs = gst::kInf;
        for (int hh = 0; hh < gst::kBbHubs; hh++) {
          uint32_t sum = 0;
          for (int i = 0; i < q.k; i++) sum = std::min(gst::kInf, sum + hA[(size_t)(q.lab + i) * gst::kBbHubs + hh]);
          hs = std::min(hs, sum);
        }
        q.best = hs;
      }
      d.hA = (const uint32_t*)(base + J.hA);
      d.nhubr = (uint32_t)g_->n_hubr;
      d.hub16 = g_->d_hubd16 != nullptr && g_->opt.hub16 != 0 ? 1u : 0u;
      pending_uploads_.push_back({(void*)d.hA, std::vector<char>((char*)hA.data(), (char*)(hA.data() + hA.size()))});
    }
    // group-distance cache (unbounded jobs): a label whose group has a valid
    // column is copied in instead of computed; a new group gets a column
    // its job fills at the end, when the budget allows
    if (D < 0 && g_->opt.cache_mb > 0) {
      sync_cache_version();
      std::vector<uint32_t*> cc(2 * (size_t)L, nullptr);
      std::vector<std::pair<int32_t, bool>>& refs = job_cache_refs_[js];
      std::unordered_set<int32_t> outs;
      int lab2 = 0;
      for (size_t m = 0; m < grp.size(); m++) {
        for (size_t i = 0; i < grp[m]->groups.size(); i++, lab2++) {
          const int32_t gid = grp[m]->groups[i];
          auto it = cache_.find(gid);
          if (it != cache_.end() && it->second.valid && !it->second.stale) {
            cc[lab2] = (uint32_t*)(arena_mem_ + it->second.off);
            it->second.refs++;
            it->second.last = ++cache_clock_;
            refs.push_back({gid, false});
            sh_[mem[m]].cached++;
            d.ncached++;
            cache_hits_++;
            continue;
          }
          cache_misses_++;
          if (it != cache_.end() || outs.count(gid)) continue;  // being written already
          // a group earns a column on its second request: one-off groups
          // (most of a uniform workload) cost no copy-out
          if (seen_.insert(gid).second) {
            if (seen_.size() > kSeenCap) seen_.clear();
            continue;
          }
          size_t coff = 0;
          if (!reserve_column(&coff)) continue;
          CacheEntry e;
          e.off = coff;
          e.refs = 1;
          e.last = ++cache_clock_;
          cache_[gid] = e;
          cache_bytes_ += column_bytes();
          cc[(size_t)L + lab2] = (uint32_t*)(arena_mem_ + coff);
          refs.push_back({gid, true});
          outs.insert(gid);
          d.nout++;
          cache_inserts_++;
        }
      }
      if (d.ncached || d.nout) {
        d.ccol = (uint32_t* const*)(base + J.ccol);
        pending_uploads_.push_back({base + J.ccol, std::vector<char>((char*)cc.data(), (char*)(cc.data() + cc.size()))});
      }
    }
    return true;
  }

  size_t column_bytes() const { return align256((size_t)g_->V * 4); }

  // Releases the cache entries no job references, least recently used first,
  // until `need` bytes are free in one run of the arena or `budget` bytes of
  // cache remain; false when not reached.
  bool evict_until(size_t need, size_t budget) {
    for (;;) {
      if (arena_.largest() >= need && cache_bytes_ <= budget) return true;
      auto victim = cache_.end();
      for (auto it = cache_.begin(); it != cache_.end(); ++it)
        if (it->second.refs == 0 && (victim == cache_.end() || it->second.last < victim->second.last)) victim = it;
      if (victim == cache_.end()) return false;
      arena_.release(victim->second.off, column_bytes());
      cache_bytes_ -= column_bytes();
      cache_.erase(victim);
      cache_evictions_++;
    }
  }

  // Arena allocation for admission: the cache gives way to queries.
  bool alloc_or_evict(size_t len, size_t* off) {
    if (arena_.alloc(len, off)) return true;
    if (cache_.empty() || !evict_until(align256(len), cache_bytes_)) return false;
    return arena_.alloc(len, off);
  }

  // A column for a new cache entry within the cache budget.
  bool reserve_column(size_t* off) {
    const size_t budget = (size_t)g_->opt.cache_mb << 20;
    const size_t cb = column_bytes();
    if (cb > budget) return false;
    if (cache_bytes_ + cb > budget && !evict_until(0, budget - cb)) return false;
    return arena_.alloc(cb, off);
  }

  // New group lists make every column meaningless: unreferenced ones go now,
  // referenced ones when their job ends.
  void sync_cache_version() {
    if (cache_version_ == g_->groups_version) return;
    cache_version_ = g_->groups_version;
    seen_.clear();
    for (auto it = cache_.begin(); it != cache_.end();) {
      if (it->second.refs == 0) {
        arena_.release(it->second.off, column_bytes());
        cache_bytes_ -= column_bytes();
        it = cache_.erase(it);
      } else {
        it->second.stale = true;
        ++it;
      }
    }
  }

  // A finished job hands back its cache references; the columns it wrote
  // become valid.
  void release_job_cache(int js) {
    auto jr = job_cache_refs_.find(js);
    if (jr == job_cache_refs_.end()) return;
    for (auto& pr : jr->second) {
      auto it = cache_.find(pr.first);
      if (it == cache_.end()) continue;
      it->second.refs--;
      if (pr.second) it->second.valid = true;
      if (it->second.stale && it->second.refs == 0) {
        arena_.release(it->second.off, column_bytes());
        cache_bytes_ -= column_bytes();
        cache_.erase(it);
      }
    }
    job_cache_refs_.erase(jr);
  }

  void clear_cache() {
    for (auto& kv : cache_) arena_.release(kv.second.off, column_bytes());
    cache_.clear();
    seen_.clear();
    cache_bytes_ = 0;
    job_cache_refs_.clear();
  }

 public:
  // Runs engine steps until `exit_after` queries finish (or max_steps steps
  // pass, or nothing is left to do); returns the finished query slots.
  // Steps are enqueued in chunks of (lean, heavy, plan) launches; the plan
  // kernel turns into a no-op once it has decided to stop, so a chunk may
  // overshoot safely.  The control header is read back after each chunk.
  // Finished group-distance jobs are released here.
  gst_status run_epoch(uint32_t exit_after, uint32_t max_steps, std::vector<int>* finished) {
    finished->clear();
    if (nslots_used_ == 0) return GST_OK;
    if (n_deadline_ > 0) flag_expired(now_ms());
    step_before_.resize(max_slots_);
    for (int i = 0; i < nslots_used_; i++) step_before_[i] = (uint32_t)h_slots_[i].ctr[0];
    const bool use_graph = g_->opt.graph != 0;
    const bool timed = g_->opt.time_steps != 0 && !use_graph;
    for (auto& u : pending_uploads_)
      HIP_TRY(hipMemcpyAsync(u.first, u.second.data(), u.second.size(), hipMemcpyHostToDevice, stream_));
    const size_t slot_bytes = sizeof(gst::Slot) * (use_graph ? max_slots_ : nslots_used_);
    HIP_TRY(hipMemcpyAsync(d_slots_, h_slots_, slot_bytes, hipMemcpyHostToDevice, stream_));
    gst::Ctrl& c = *h_ctrl_;
    memset(&c, 0, offsetof(gst::Ctrl, active));
    c.exit_after = std::max(1u, exit_after);
    c.max_steps = max_steps;
    c.nslots = (uint32_t)(use_graph ? max_slots_ : nslots_used_);
    c.trace_slot = timed ? (uint32_t)g_->opt.trace_slot : ~0u;
    c.cancel = d_cancel_;
    const size_t hdr = offsetof(gst::Ctrl, active);
    pair_ns_.clear();
    HIP_TRY(hipMemcpyAsync(d_ctrl_, h_ctrl_, hdr, hipMemcpyHostToDevice, stream_));
    const int grid_l = grid_size(gst::kClsLean), grid_h = grid_size(gst::kClsHeavy);
    const gst::GraphDev gd = g_->dev();
    if (timed && tev_.empty()) {
      tev_.resize(4 * 64);
      for (auto& e : tev_) HIP_TRY(hipEventCreate(&e));
    }
    HIP_TRY(hipEventRecord(ev0_, stream_));
    hipError_t le = gst_launch_plan(d_slots_, d_ctrl_, (uint64_t)g_->V, stream_);
    if (le != hipSuccess) return fail(GST_DEVICE_ERROR, "plan launch failed: %s", hipGetErrorString(le));
    int chunk = use_graph ? kGraphChunk : 4;
    for (;;) {
      if (use_graph) {
        if (gexec_ && (graph_grid_[0] != grid_l || graph_grid_[1] != grid_h || graph_wpb_ != wpb())) drop_graph();
        if (!gexec_) {
          gst_status st = capture_chunk(grid_l, grid_h);
          if (st != GST_OK) return st;
        }
        HIP_TRY(hipGraphLaunch(gexec_, stream_));
      } else {
        for (int i = 0; i < chunk; i++) {
          if (timed) HIP_TRY(hipEventRecord(tev_[4 * i], stream_));
          le = gst_launch_step(gst::kClsLean, wpb(), &gd, d_slots_, d_ctrl_, grid_l, stream_);
          if (timed && le == hipSuccess) le = hipEventRecord(tev_[4 * i + 1], stream_);
          if (le == hipSuccess) le = gst_launch_step(gst::kClsHeavy, wpb(), &gd, d_slots_, d_ctrl_, grid_h, stream_);
          if (timed && le == hipSuccess) le = hipEventRecord(tev_[4 * i + 2], stream_);
          if (le == hipSuccess) le = gst_launch_plan(d_slots_, d_ctrl_, (uint64_t)g_->V, stream_);
          if (le != hipSuccess) return fail(GST_DEVICE_ERROR, "step launch failed: %s", hipGetErrorString(le));
        }
      }
      HIP_TRY(hipMemcpyAsync(h_ctrl_, d_ctrl_, hdr, hipMemcpyDeviceToHost, stream_));
      gst_status st = wait_stream();
      if (st != GST_OK) return st;
      if (n_deadline_ > 0) flag_expired(now_ms());
      if (timed) {
        for (int i = 0; i < chunk; i++) {
          float a = 0, b = 0;
          HIP_TRY(hipEventElapsedTime(&a, tev_[4 * i], tev_[4 * i + 1]));
          HIP_TRY(hipEventElapsedTime(&b, tev_[4 * i + 1], tev_[4 * i + 2]));
          cls_ns_[0] += (int64_t)((double)a * 1e6);
          cls_ns_[1] += (int64_t)((double)b * 1e6);
          cls_launches_ += 1;
          pair_ns_.push_back((int64_t)((double)(a + b) * 1e6));
        }
      }
      if (h_ctrl_->exit_flag) break;
      if (!use_graph) chunk = std::min(chunk * 2, 64);
    }
    pending_uploads_.clear();
    HIP_TRY(hipEventRecord(ev1_, stream_));
    HIP_TRY(hipMemcpyAsync(h_slots_, d_slots_, sizeof(gst::Slot) * nslots_used_, hipMemcpyDeviceToHost, stream_));
    gst_status st = wait_stream();
    if (st != GST_OK) return st;
    if (timed && c.trace_slot != ~0u) {
      // step j of the epoch ran the plan that the trace logged at index j
      const size_t n = std::min<size_t>({pair_ns_.size(), (size_t)h_ctrl_->steps, (size_t)gst::kTraceSteps});
      std::vector<uint32_t> tr(n);
      if (n) {
        HIP_TRY(hipMemcpyAsync(tr.data(), d_ctrl_->trace, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
        st = wait_stream();
        if (st != GST_OK) return st;
      }
      for (size_t j = 0; j < n; j++) {
        if (!(tr[j] & 0x10000u)) continue;
        auto& e = phase_ns_[tr[j] & 0xFFFFu];
        e.first += pair_ns_[j];
        e.second += 1;
      }
    }
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
    kernel_ns_ += (int64_t)((double)ms * 1e6);
    launches_++;
    steps_total_ += h_ctrl_->steps;
    last_steps_ = h_ctrl_->steps;
    // progress guard: every epoch runs at least one step of some admitted slot
    bool any_step = false;
    for (int i = 0; i < nslots_used_; i++)
      if (sh_[i].used && ((uint32_t)h_slots_[i].ctr[0] != step_before_[i] || h_slots_[i].phase == gst::PH_DONE))
        any_step = true;
    if (!any_step)
      return fail(GST_DEVICE_ERROR, "engine epoch made no progress on %d admitted queries", active_);
    for (int i = 0; i < nslots_used_; i++) {
      if (!sh_[i].used || h_slots_[i].phase != gst::PH_DONE) continue;
      if (sh_[i].job) charge_job(i);
      else finished->push_back(i);
    }
    return GST_OK;
  }

  // A finished job's level-1 work is charged to its members, per label; its
  // distance rows stay until the last member has been harvested.
  void charge_job(int i) {
    if (sh_[i].charged) return;
    const gst::Slot& j = h_slots_[i];
    for (int m : job_members_.at(i)) {
      const gst::Slot& q = h_slots_[m];
      SlotHost& qh = sh_[m];
      for (int t = 1; t <= 4; t++) qh.extra[t] += (int64_t)(j.ctr[t] * (uint64_t)q.k / (uint64_t)std::max(1, j.L));
      qh.extra[0] += (int64_t)j.ctr[0];
      qh.jmaxq = std::max(qh.jmaxq, j.maxq);
    }
    jobs_done_++;
    for (int t = 0; t < 6; t++) job_ctr_[t] += (int64_t)j.ctr[t];
    job_ctr_[6] += (int64_t)j.nsrc;
    sh_[i].charged = true;
    release_job_cache(i);
  }

  // ---- continuous stepping (the server) ----
  // The device keeps stepping while the host harvests finished queries and
  // admits new ones between chunks of steps.  While it runs, the slot table on
  // the device is authoritative for admitted slots: the host uploads only the
  // entries of slots admitted since the last chunk (those were free or
  // finished on the device, so no step touches them, and the plan kernel takes
  // them up at its next run), and reads the table back after every chunk into
  // a snapshot buffer.  A snapshot describes a slot only if the slot's entry
  // was uploaded before that chunk (admission numbers).  Up to `depth` chunks
  // stay queued, so the device steps on while the host handles the oldest
  // one's snapshot.  Returns the queries the collected snapshot shows finished
  // (their final entries copied into the host mirror for harvest).
  gst_status cont_iter(int steps, int depth, std::vector<int>* finished, bool may_step = true) {
    finished->clear();
    if (poisoned_) return fail(GST_DEVICE_ERROR, "engine stopped by an earlier failure");
    depth = std::max(1, std::min(depth, kContBuffers));
    steps = std::max(1, steps);
    if (!cont_on_) {
      gst_status st = cont_start();
      if (st != GST_OK) return st;
    } else if (!fresh_.empty() || !pending_uploads_.empty()) {
      for (auto& u : pending_uploads_)
        HIP_TRY(hipMemcpyAsync(u.first, u.second.data(), u.second.size(), hipMemcpyHostToDevice, stream_));
      for (int s : fresh_)
        if (sh_[s].used && !sh_[s].done_seen)
          HIP_TRY(hipMemcpyAsync(d_slots_ + s, h_slots_ + s, sizeof(gst::Slot), hipMemcpyHostToDevice, stream_));
      fresh_.clear();
      // the plan kernel stands still once it has found no work: restart it
      HIP_TRY(hipMemcpyAsync(&d_ctrl_->exit_flag, h_zero_, sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
    }
    // host sources of the uploads stay alive until a chunk behind them is done
    std::vector<std::pair<void*, std::vector<char>>> keep;
    keep.swap(pending_uploads_);
    if (may_step && n_inflight_ > 0 && (int)chunks_.size() < depth) {
      int buf = 0;
      for (;; buf++) {
        bool busy = false;
        for (auto& c : chunks_) busy = busy || c.buf == buf;
        if (!busy) break;
      }
      const int grid_l = grid_size(gst::kClsLean), grid_h = grid_size(gst::kClsHeavy);
      const gst::GraphDev gd = g_->dev();
      HIP_TRY(hipEventRecord(cev0_[buf], stream_));
      for (int i = 0; i < steps; i++) {
        hipError_t le = gst_launch_step(gst::kClsLean, wpb(), &gd, d_slots_, d_ctrl_, grid_l, stream_);
        if (le == hipSuccess) le = gst_launch_step(gst::kClsHeavy, wpb(), &gd, d_slots_, d_ctrl_, grid_h, stream_);
        if (le == hipSuccess) le = gst_launch_plan(d_slots_, d_ctrl_, (uint64_t)g_->V, stream_);
        if (le != hipSuccess) {
          poisoned_ = true;
          return fail(GST_DEVICE_ERROR, "step launch failed: %s", hipGetErrorString(le));
        }
      }
      ContChunk c;
      c.buf = buf;
      c.aseq = aseq_;
      c.nslots = nslots_used_;
      c.keep.swap(keep);
      if (c.nslots > 0)
        HIP_TRY(hipMemcpyAsync(h_snap_[buf], d_slots_, sizeof(gst::Slot) * (size_t)c.nslots, hipMemcpyDeviceToHost,
                               stream_));
      HIP_TRY(hipEventRecord(cev_[buf], stream_));
      chunks_.push_back(std::move(c));
      steps_total_ += steps;
    } else if (!keep.empty()) {
      if (!chunks_.empty()) {
        for (auto& u : keep) chunks_.back().keep.push_back(std::move(u));
      } else {
        gst_status st = wait_stream();
        if (st != GST_OK) return st;
      }
    }
    if (chunks_.empty()) return GST_OK;
    if (may_step && (int)chunks_.size() < depth && n_inflight_ > 0) return GST_OK;  // fill the queue first
    return cont_collect(finished);
  }

  bool cont_outstanding() const { return !chunks_.empty(); }
  int cont_queued() const { return (int)chunks_.size(); }
  // continuous stepping: chunks run and their summed device time (between
  // a start and an end event of each chunk)
  int64_t cont_chunks() const { return cont_chunks_; }
  int64_t cont_busy_ns() const { return cont_busy_ns_; }

 private:
  static constexpr int kContBuffers = 2;
  struct ContChunk {
    int buf = 0;
    uint64_t aseq = 0;  // admissions uploaded before the chunk
    int nslots = 0;     // entries of its snapshot
    std::vector<std::pair<void*, std::vector<char>>> keep;
  };

  void mark_admitted(int s) {
    sh_[s].aseq = ++aseq_;
    sh_[s].done_seen = false;
    n_inflight_++;
    if (cont_on_) fresh_.push_back(s);
  }

  gst_status cont_start() {
    if (!xstream_) HIP_TRY(hipStreamCreateWithFlags(&xstream_, hipStreamNonBlocking));
    for (int b = 0; b < kContBuffers; b++) {
      if (!h_snap_[b]) HIP_TRY(hipHostMalloc((void**)&h_snap_[b], sizeof(gst::Slot) * max_slots_, hipHostMallocDefault));
      if (!cev_[b]) HIP_TRY(hipEventCreate(&cev_[b]));
      if (!cev0_[b]) HIP_TRY(hipEventCreate(&cev0_[b]));
    }
    if (!h_zero_) {
      HIP_TRY(hipHostMalloc((void**)&h_zero_, 64, hipHostMallocDefault));
      memset(h_zero_, 0, 64);
    }
    for (auto& u : pending_uploads_)
      HIP_TRY(hipMemcpyAsync(u.first, u.second.data(), u.second.size(), hipMemcpyHostToDevice, stream_));
    // the whole table: slots never admitted must read as free
    HIP_TRY(hipMemcpyAsync(d_slots_, h_slots_, sizeof(gst::Slot) * max_slots_, hipMemcpyHostToDevice, stream_));
    gst::Ctrl& c = *h_ctrl_;
    memset(&c, 0, offsetof(gst::Ctrl, active));
    c.exit_after = ~0u;  // the plan kernel stops only when no slot has work
    c.max_steps = ~0u;
    c.nslots = (uint32_t)max_slots_;
    c.trace_slot = ~0u;
    c.cancel = d_cancel_;
    HIP_TRY(hipMemcpyAsync(d_ctrl_, h_ctrl_, offsetof(gst::Ctrl, active), hipMemcpyHostToDevice, stream_));
    hipError_t le = gst_launch_plan(d_slots_, d_ctrl_, (uint64_t)g_->V, stream_);
    if (le != hipSuccess) return fail(GST_DEVICE_ERROR, "plan launch failed: %s", hipGetErrorString(le));
    fresh_.clear();
    cont_on_ = true;
    return GST_OK;
  }

  gst_status cont_collect(std::vector<int>* finished) {
    ContChunk c = std::move(chunks_.front());
    chunks_.pop_front();
    gst_status st = wait_event(cev_[c.buf]);
    if (st != GST_OK) return st;
    float ms = 0;
    if (hipEventElapsedTime(&ms, cev0_[c.buf], cev_[c.buf]) == hipSuccess) cont_busy_ns_ += (int64_t)((double)ms * 1e6);
    cont_chunks_++;
    const gst::Slot* snap = h_snap_[c.buf];
    uint64_t sum = 0;
    bool done_any = false;
    for (int s = 0; s < c.nslots; s++) {
      SlotHost& h = sh_[s];
      if (!h.used || h.done_seen || h.aseq > c.aseq) continue;
      const gst::Slot& x = snap[s];
      sum += (uint64_t)x.ctr[0];
      if (x.phase != gst::PH_DONE) continue;
      h.done_seen = true;
      n_inflight_--;
      done_any = true;
      memcpy((void*)&h_slots_[s], (const void*)&x, sizeof(gst::Slot));
      if (h.job) charge_job(s);
      else finished->push_back(s);
    }
    // progress guard: admitted slots that stop stepping are a device-side fault
    if (done_any || sum != cont_sum_ || n_inflight_ == 0) {
      cont_sum_ = sum;
      cont_stall_ = 0;
    } else if (++cont_stall_ > kContStallChunks) {
      poisoned_ = true;
      return fail(GST_DEVICE_ERROR, "engine made no progress on %d admitted slots", n_inflight_);
    }
    return GST_OK;
  }

  // wait_stream for one event: the end of a chunk
  gst_status wait_event(hipEvent_t ev) {
    const double t_end = now_ms() + (double)g_->opt.timeout_ms;
    for (int it = 0;; it++) {
      const hipError_t q = hipEventQuery(ev);
      if (q == hipSuccess) return GST_OK;
      if (q != hipErrorNotReady) {
        poisoned_ = true;
        return fail(GST_DEVICE_ERROR, "engine step failed: %s", hipGetErrorString(q));
      }
      const double t = now_ms();
      if (t > t_end) {
        poisoned_ = true;
        return fail(GST_DEVICE_ERROR, "engine step did not finish within %lld ms", (long long)g_->opt.timeout_ms);
      }
      if (n_deadline_ > 0 && (it & 15) == 0) flag_expired(t);
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }

  // the stream harvest copies run on: beside the queued chunks while stepping
  // continuously (a finished slot is not touched by any step)
  hipStream_t xs() const { return cont_on_ ? xstream_ : stream_; }

 public:

  // A harvested member releases its hold on its job; the last one frees it.
  void drop_job_ref(int member) {
    auto it = member_job_.find(member);
    if (it == member_job_.end()) return;
    const int js = it->second;
    member_job_.erase(it);
    if (--job_refs_[js] == 0) {
      job_refs_.erase(js);
      job_members_.erase(js);
      release_slot(js);
    }
  }

  // Waits for the stream with a deadline, so a launch that never finishes is
  // reported instead of blocking the caller forever.
  // While it waits, it raises the cancellation flag of every query past its
  // deadline; the plan kernel reads the flags at each step.  After a timeout
  // or a device error the engine is poisoned: steps may still be running and
  // copies into the host mirrors may still be queued, so nothing of the
  // engine is touched again until its stream has drained (recover()).
  gst_status wait_stream() {
    const double t_end = now_ms() + (double)g_->opt.timeout_ms;
    for (int it = 0;; it++) {
      const hipError_t q = hipStreamQuery(stream_);
      if (q == hipSuccess) return GST_OK;
      if (q != hipErrorNotReady) {
        poisoned_ = true;
        return fail(GST_DEVICE_ERROR, "engine step failed: %s", hipGetErrorString(q));
      }
      const double t = now_ms();
      if (t > t_end) {
        poisoned_ = true;
        return fail(GST_DEVICE_ERROR, "engine step did not finish within %lld ms", (long long)g_->opt.timeout_ms);
      }
      if (n_deadline_ > 0 && (it & 15) == 0) flag_expired(t);
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }

  // Raises the cancellation flag of slot s (a query); returns false when the
  // engine has no flags (the static engine).
  bool flag_slot(int s) {
    if (!h_cancel_) return false;
    SlotHost& h = sh_[s];
    if (!h.used || h.job || h.flagged) return true;
    h.flagged = true;
    n_flagged_++;
    __atomic_store_n(&h_cancel_[1 + s], 1u, __ATOMIC_RELEASE);
    __atomic_store_n(&h_cancel_[0], 1u, __ATOMIC_RELEASE);
    return true;
  }
  void flag_expired(double t) {
    for (int i = 0; i < nslots_used_; i++)
      if (sh_[i].used && !sh_[i].job && !sh_[i].flagged && sh_[i].t_deadline > 0 && t > sh_[i].t_deadline)
        flag_slot(i);
  }
  // Cancels the admitted query with this tag; false if none.
  bool cancel_tag(uint64_t tag) {
    for (int i = 0; i < nslots_used_; i++)
      if (sh_[i].used && !sh_[i].job && sh_[i].tag == tag) return flag_slot(i);
    return false;
  }
  void cancel_all() {
    for (int i = 0; i < nslots_used_; i++)
      if (sh_[i].used && !sh_[i].job) flag_slot(i);
  }

  bool poisoned() const { return poisoned_; }
  void poison() { poisoned_ = true; }
  // Makes a poisoned engine usable again once its stream has drained: the
  // admitted queries are dropped (their callers already got the error).
  gst_status recover() {
    if (!poisoned_) return GST_OK;
    const hipError_t q = hipStreamQuery(stream_);
    if (q == hipErrorNotReady)
      return fail(GST_DEVICE_ERROR, "an earlier call timed out and its engine steps are still running on the device");
    if (q != hipSuccess) return fail(GST_DEVICE_ERROR, "engine stream failed: %s", hipGetErrorString(q));
    poisoned_ = false;
    reset_slots();
    return GST_OK;
  }

  // A chunk of kGraphChunk (lean, heavy, plan) launches captured once and
  // replayed.
  gst_status capture_chunk(int grid_l, int grid_h) {
    const gst::GraphDev gd = g_->dev();
    HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int i = 0; i < kGraphChunk && le == hipSuccess; i++) {
      le = gst_launch_step(gst::kClsLean, wpb(), &gd, d_slots_, d_ctrl_, grid_l, stream_);
      if (le == hipSuccess) le = gst_launch_step(gst::kClsHeavy, wpb(), &gd, d_slots_, d_ctrl_, grid_h, stream_);
      if (le == hipSuccess) le = gst_launch_plan(d_slots_, d_ctrl_, (uint64_t)g_->V, stream_);
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(stream_, &graph);
    if (le != hipSuccess) return fail(GST_DEVICE_ERROR, "capture launch failed: %s", hipGetErrorString(le));
    if (e != hipSuccess) return fail(GST_DEVICE_ERROR, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&gexec_, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      gexec_ = nullptr;
      return fail(GST_DEVICE_ERROR, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    graph_grid_[0] = grid_l;
    graph_grid_[1] = grid_h;
    graph_wpb_ = wpb();
    return GST_OK;
  }

  void drop_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    gexec_ = nullptr;
  }

  // Copies the trees of the finished queries `slots` to the host: one gather
  // kernel puts them side by side in a device staging buffer, one transfer
  // brings that to pinned host memory.  harvest() then takes each tree from
  // there (instead of a device-to-host copy of its own, each of which waits
  // for its transfer into unpinned memory).
  gst_status gather_trees(const std::vector<int>& slots) {
    gents_.clear();
    uint64_t tot = 0;
    for (int s : slots) {
      const gst::Slot& d = h_slots_[s];
      staged_[s] = -1;
      if (d.status != 0 || d.out_cost >= (int32_t)gst::kInf || d.out_nedges <= 0) continue;
      gst::GatherEnt e;
      e.src = d.edges;
      e.off = tot;
      e.n = (uint32_t)d.out_nedges;
      e.pad = 0;
      gents_.push_back(e);
      staged_[s] = (int64_t)tot;
      tot += (uint64_t)d.out_nedges;
    }
    if (gents_.empty()) return GST_OK;
    if (gents_.size() > gtab_cap_) {
      if (d_gtab_) (void)hipFree(d_gtab_);
      if (h_gtab_) (void)hipHostFree(h_gtab_);
      d_gtab_ = nullptr;
      h_gtab_ = nullptr;
      gtab_cap_ = 0;
      const size_t cap = std::max<size_t>(gents_.size(), 1024);
      HIP_TRY(hipMalloc((void**)&d_gtab_, cap * sizeof(gst::GatherEnt)));
      HIP_TRY(hipHostMalloc((void**)&h_gtab_, cap * sizeof(gst::GatherEnt), hipHostMallocDefault));
      gtab_cap_ = cap;
    }
    if (tot > gstage_cap_) {
      if (d_gstage_) (void)hipFree(d_gstage_);
      if (h_gstage_) (void)hipHostFree(h_gstage_);
      d_gstage_ = nullptr;
      h_gstage_ = nullptr;
      gstage_cap_ = 0;
      const size_t cap = std::max<size_t>((size_t)tot * 2, 65536);
      HIP_TRY(hipMalloc((void**)&d_gstage_, cap * sizeof(gst::EdgeDev)));
      HIP_TRY(hipHostMalloc((void**)&h_gstage_, cap * sizeof(gst::EdgeDev), hipHostMallocDefault));
      gstage_cap_ = cap;
    }
    memcpy(h_gtab_, gents_.data(), gents_.size() * sizeof(gst::GatherEnt));
    HIP_TRY(hipMemcpyAsync(d_gtab_, h_gtab_, gents_.size() * sizeof(gst::GatherEnt), hipMemcpyHostToDevice, xs()));
    const hipError_t le = gst_launch_gather(d_gtab_, (int)gents_.size(), d_gstage_, xs());
    if (le != hipSuccess) return fail(GST_DEVICE_ERROR, "tree gather launch failed: %s", hipGetErrorString(le));
    HIP_TRY(hipMemcpyAsync(h_gstage_, d_gstage_, tot * sizeof(gst::EdgeDev), hipMemcpyDeviceToHost, xs()));
    const hipError_t e = hipStreamSynchronize(xs());
    if (e != hipSuccess) return fail(GST_DEVICE_ERROR, "tree copy failed: %s", hipGetErrorString(e));
    return GST_OK;
  }

  // Copies the result of a finished query out and releases its memory.  The
  // tree goes to a malloc'd buffer owned by `out`, or, with ext, into the
  // caller's buffer of ext_cap edges (out->edges stays null then).  A tree
  // that gather_trees staged is taken from the staging buffer; any other is
  // copied from the device on its own.
  gst_status harvest(int s, gst_result* out, uint64_t* tag, gst_edge* ext = nullptr, int32_t ext_cap = 0) {
    const gst::Slot& d = h_slots_[s];
    SlotHost& h = sh_[s];
    memset(out, 0, sizeof(*out));
    const double t = now_ms();
    out->latency_ms = t - h.t_submit;
    out->device_ms = t - h.t_admit;
    out->cost = d.out_cost;
    out->root = d.out_root;
    out->status = GST_OK;
    out->counters.steps = (int64_t)d.ctr[0] + h.extra[0];
    out->counters.frontier_entries = (int64_t)d.ctr[1];
    out->counters.arcs = (int64_t)d.ctr[2];
    out->counters.relaxations = (int64_t)d.ctr[3];
    out->counters.improvements = (int64_t)d.ctr[4];
    out->counters.merge_evals = (int64_t)d.ctr[5];
    out->counters.live_vertices = (int64_t)d.nlive;
    out->counters.shared_entries = h.extra[1];
    out->counters.shared_arcs = h.extra[2];
    out->counters.shared_relaxations = h.extra[3];
    out->counters.active_vertices = (int64_t)d.nact;
    out->counters.max_frontier = (int64_t)std::max(d.maxq, h.jmaxq);
    out->counters.cached_labels = h.cached;
    out->counters.states_finite = -1;
    out->counters.lower_bound = d.cert >= gst::kInf ? GST_INFEASIBLE : (int64_t)d.cert;
    if (d.count && d.status == 0 && d.phase == gst::PH_DONE) {
      // the reference counts finite states after its main loop (count_set);
      // here a pass of its own over the finished query's tables
      unsigned long long nf = 0;
      gst::Slot* ds = d_slots_ + s;
      hipError_t e = gst_launch_count(ds, (uint64_t)g_->V, xs());  // nfinite starts at 0 (admission)
      if (e == hipSuccess) e = hipMemcpyAsync(&nf, &ds->nfinite, sizeof(nf), hipMemcpyDeviceToHost, xs());
      if (e == hipSuccess) e = hipStreamSynchronize(xs());
      if (e == hipSuccess) out->counters.states_finite = (int64_t)nf;
    }
    gst_status st = GST_OK;
    if (d.status == gst::kStatusCancelled) {
      out->status = GST_CANCELLED;
      out->cost = GST_INFEASIBLE;
      out->root = -1;
      out->n_edges = 0;
      st = fail(GST_CANCELLED, "query cancelled before it finished (request or deadline)");
    } else if (d.status == gst::kStatusEdgeCap) {
      out->status = GST_DEVICE_ERROR;
      st = fail(GST_DEVICE_ERROR, "tree of the query exceeds the edge capacity %d", d.edge_cap);
      out->n_edges = 0;
    } else if (d.status != 0) {
      out->status = GST_DEVICE_ERROR;
      st = fail(GST_DEVICE_ERROR, "tree reconstruction found no shortest-path arc (status %u)", d.status);
      out->n_edges = 0;
    } else if (d.out_cost >= (int32_t)gst::kInf) {
      out->cost = GST_INFEASIBLE;
      out->root = -1;
      out->n_edges = 0;
    } else {
      out->n_edges = d.out_nedges;
      if (ext) {
        if (d.out_nedges > ext_cap) {
          out->status = GST_INVALID_ARG;
          st = fail(GST_INVALID_ARG, "the tree has %d edges; the edge buffer holds %d", d.out_nedges, ext_cap);
        } else if (d.out_nedges > 0) {
          hipError_t e = hipMemcpyAsync(ext, d.edges, sizeof(gst_edge) * (size_t)d.out_nedges,
                                        hipMemcpyDeviceToHost, xs());
          if (e == hipSuccess) e = hipStreamSynchronize(xs());
          if (e != hipSuccess) {
            out->status = GST_DEVICE_ERROR;
            st = fail(GST_DEVICE_ERROR, "edge copy failed: %s", hipGetErrorString(e));
          }
        }
      } else if (d.out_nedges > 0) {
        out->edges = (gst_edge*)malloc(sizeof(gst_edge) * (size_t)d.out_nedges);
        if (!out->edges) return fail(GST_OUT_OF_MEMORY, "host edge buffer");
        auto sg = staged_.find(s);
        if (sg != staged_.end() && sg->second >= 0) {
          static_assert(sizeof(gst_edge) == sizeof(gst::EdgeDev), "edge layouts differ");
          memcpy(out->edges, h_gstage_ + sg->second, sizeof(gst_edge) * (size_t)d.out_nedges);
        } else {
          hipError_t e = hipMemcpyAsync(out->edges, d.edges, sizeof(gst_edge) * (size_t)d.out_nedges,
                                        hipMemcpyDeviceToHost, xs());
          if (e == hipSuccess) e = hipStreamSynchronize(xs());
          if (e != hipSuccess) {
            out->status = GST_DEVICE_ERROR;
            st = fail(GST_DEVICE_ERROR, "edge copy failed: %s", hipGetErrorString(e));
          }
        }
      }
    }
    staged_.erase(s);
    *tag = h.tag;
    for (auto& x : h.extra) x = 0;
    release_slot(s);
    drop_job_ref(s);
    return st;
  }

  // Drops every admitted query and job (after a device error) so the engine
  // is reusable.
  void reset_slots() {
    clear_cache();
    for (int i = 0; i < max_slots_; i++) {
      if (sh_[i].used) arena_.release(sh_[i].off, sh_[i].len);
      sh_[i] = SlotHost();
      memset(&h_slots_[i], 0, sizeof(gst::Slot));
    }
    if (h_cancel_) memset(h_cancel_, 0, sizeof(uint32_t) * (max_slots_ + 1));
    n_flagged_ = 0;
    n_deadline_ = 0;
    active_ = 0;
    used_slots_ = 0;
    nslots_used_ = 0;
    n_inflight_ = 0;
    fresh_.clear();
    chunks_.clear();
    staged_.clear();
    cont_on_ = false;  // the next continuous chunk uploads the whole table again
    pending_uploads_.clear();
    job_members_.clear();
    job_refs_.clear();
    member_job_.clear();
  }

  uint32_t last_steps() const { return last_steps_; }
  int64_t kernel_ns() const { return kernel_ns_; }
  int64_t launches() const { return launches_; }
  int64_t steps_total() const { return steps_total_; }
  int64_t cls_ns(int c) const { return cls_ns_[c]; }
  int64_t cls_launches() const { return cls_launches_; }
  int64_t jobs_done() const { return jobs_done_; }
  int64_t job_ctr(int t) const { return job_ctr_[t]; }
  // summed step time (lean + heavy) and step count of the traced slot's
  // steps in `phase` at `level` (options time_steps and trace_slot)
  int64_t phase_stat(int phase, int level, bool count) const {
    auto it = phase_ns_.find((uint32_t)phase | ((uint32_t)level << 8));
    if (it == phase_ns_.end()) return 0;
    return count ? it->second.second : it->second.first;
  }
  void reset_stats() {
    phase_ns_.clear();
    kernel_ns_ = launches_ = steps_total_ = 0;
    cls_ns_[0] = cls_ns_[1] = cls_launches_ = jobs_done_ = 0;
    for (auto& x : job_ctr_) x = 0;
  }
  size_t largest_free() const { return arena_.largest(); }

  int wpb() const { return gst::kWaveChoices[wpb_index(g_->opt.block)]; }

  int grid_size(int cls) const {
    const int occ = g_->occ[wpb_index(g_->opt.block)][cls];
    int bpc = occ;
    if (g_->opt.blocks_per_cu > 0) bpc = std::min<int>((int)g_->opt.blocks_per_cu, occ);
    return std::max(1, bpc) * g_->num_cu;
  }

 private:
  gst_graph* g_;
  int max_slots_;
  bool borrowed_ = false;   // device tables live in memory the engine does not own
  bool force_own_ = false;  // every query runs its own level 1
  bool poisoned_ = false;   // a wait timed out or the device failed: see recover()
  uint32_t* h_cancel_ = nullptr;  // cancellation flags (pinned host memory)
  const uint32_t* d_cancel_ = nullptr;  // the same words as the device reads them
  int n_flagged_ = 0;       // slots whose flag is set
  int n_deadline_ = 0;      // admitted queries with a deadline
  // group-distance cache: group id -> column in the arena
  std::unordered_map<int32_t, CacheEntry> cache_;
  std::map<int, std::vector<std::pair<int32_t, bool>>> job_cache_refs_;  // job slot -> (group, writes it)
  std::unordered_set<int32_t> seen_;  // groups requested once since they last left the cache
  static constexpr size_t kSeenCap = size_t(1) << 22;
  size_t cache_bytes_ = 0;
  uint64_t cache_clock_ = 0;
  int64_t cache_version_ = -1;
 public:
  int64_t cache_hits_ = 0, cache_misses_ = 0, cache_inserts_ = 0, cache_evictions_ = 0;
  size_t cache_entries() const { return cache_.size(); }
 private:
  hipStream_t stream_ = nullptr;
  gst::Slot* d_slots_ = nullptr;
  gst::Slot* h_slots_ = nullptr;
  gst::Ctrl* d_ctrl_ = nullptr;
  gst::Ctrl* h_ctrl_ = nullptr;
  char* arena_mem_ = nullptr;
  Arena arena_;
  std::vector<SlotHost> sh_;
  std::vector<uint32_t> step_before_;
  int nslots_used_ = 0;
  int used_slots_ = 0;  // queries and jobs
  int active_ = 0;      // queries
  uint32_t last_steps_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  std::vector<hipEvent_t> tev_;  // per-step class timing (option time_steps)
  hipGraphExec_t gexec_ = nullptr;
  int graph_grid_[2] = {0, 0};
  int graph_wpb_ = 0;
  int64_t kernel_ns_ = 0, launches_ = 0, steps_total_ = 0;
  int64_t cls_ns_[2] = {0, 0}, cls_launches_ = 0, jobs_done_ = 0;
  int64_t job_ctr_[7] = {0, 0, 0, 0, 0, 0, 0};  // the job slots' ctr[0..5], then their nsrc
  std::vector<int64_t> pair_ns_;  // per step of the current epoch (time_steps)
  std::map<uint32_t, std::pair<int64_t, int64_t>> phase_ns_;
  // host copies of job tables, uploaded at the start of the next epoch (or
  // before the next chunk, stepping continuously)
  std::vector<std::pair<void*, std::vector<char>>> pending_uploads_;
  // continuous stepping
  static constexpr int kContStallChunks = 4096;
  bool cont_on_ = false;
  hipStream_t xstream_ = nullptr;
  gst::Slot* h_snap_[kContBuffers] = {nullptr, nullptr};
  hipEvent_t cev_[kContBuffers] = {nullptr, nullptr};   // end of a chunk (after its snapshot)
  hipEvent_t cev0_[kContBuffers] = {nullptr, nullptr};  // start of a chunk
  int64_t cont_chunks_ = 0, cont_busy_ns_ = 0;
  // tree gather (gather_trees): entry table and staging buffers, and the
  // staging offset of each harvested slot's tree
  std::vector<gst::GatherEnt> gents_;
  gst::GatherEnt* d_gtab_ = nullptr;
  gst::GatherEnt* h_gtab_ = nullptr;
  size_t gtab_cap_ = 0;
  gst::EdgeDev* d_gstage_ = nullptr;
  gst::EdgeDev* h_gstage_ = nullptr;
  size_t gstage_cap_ = 0;
  std::unordered_map<int, int64_t> staged_;
  uint32_t* h_zero_ = nullptr;
  std::deque<ContChunk> chunks_;
  uint64_t aseq_ = 0;
  std::vector<int> fresh_;  // admitted since the last upload
  int n_inflight_ = 0;      // admitted slots no snapshot has shown finished
  uint64_t cont_sum_ = 0;
  int cont_stall_ = 0;
  std::map<int, std::vector<int>> job_members_;
  std::map<int, int> job_refs_;    // job slot -> members not yet harvested
  std::map<int, int> member_job_;  // member slot -> its job slot
};

namespace {

// Parallel loop over [0, n) on up to hardware_concurrency threads.
template <class F>
void parallel_for(int64_t n, F f) {
  int nt = (int)std::min<int64_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
  if (n < 100000) nt = 1;
  std::vector<std::thread> th;
  const int64_t per = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int64_t b = t * per, e = std::min(n, b + per);
    if (b >= e) break;
    th.emplace_back([=]() { f(b, e); });
  }
  for (auto& x : th) x.join();
}

// Parallel loop over the vertices [0, V), in ranges of about equal arc counts
// (power-law graphs put most arcs on a few low-numbered hubs).
template <class F>
void parallel_for_rows(const int64_t* beg_pos, int64_t V, F f) {
  int nt = (int)std::min<int64_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
  const int64_t arcs = beg_pos[V];
  if (V < 100000 && arcs < 1000000) nt = 1;
  std::vector<std::thread> th;
  int64_t b = 0;
  for (int t = 0; t < nt && b < V; t++) {
    const int64_t target = arcs / nt * (t + 1);
    int64_t e = t + 1 == nt ? V : (int64_t)(std::upper_bound(beg_pos, beg_pos + V + 1, target) - beg_pos) - 1;
    e = std::max(e, b + 1);
    e = std::min(e, V);
    th.emplace_back([=]() { f(b, e); });
    b = e;
  }
  if (b < V) f(b, V);
  for (auto& x : th) x.join();
}

gst_status engine_for(gst_graph* g, bool batch, size_t need, Engine** out) {
  std::unique_ptr<Engine>& e = batch ? g->batch : g->plain;
  if (e) {
    gst_status st = e->recover();
    if (st != GST_OK) return st;
  }
  if (!e) {
    int slots = batch ? (int)g->opt.max_slots : 2;  // plain: the query and its job
    std::unique_ptr<Engine> ne(new Engine(g, slots));
    size_t bytes = need;
    if (batch) {
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(GST_DEVICE_ERROR, "hipMemGetInfo failed");
      bytes = g->opt.budget_mb > 0 ? (size_t)g->opt.budget_mb << 20 : (size_t)(fr * 0.85);
      bytes = std::min(bytes, (size_t)(fr * 0.95));
    }
    gst_status st = ne->init(bytes);
    if (st != GST_OK) return st;
    e = std::move(ne);
  } else if (!batch && e->arena_size() < need) {
    gst_status st = e->resize_arena(need);
    if (st != GST_OK) return st;
  }
  *out = e.get();
  return GST_OK;
}

// Device bytes one query needs while admitted, with the share of its
// group-distance job when it has one (a job of its own when alone).
size_t single_query_bytes(const gst_graph* g, int k, int D) {
  if (!uses_job(g, k, D) || g->opt.job_min_queries > 1) return layout_for(g, k, D, true).total;
  return layout_for(g, k, D, false).total + job_layout(g, k, 1, D).total;
}

// Drives an engine over a list of queries: admits while memory and slots
// allow, runs epochs, harvests finished queries, refills.
gst_status run_queries(gst_graph* g, Engine* eng, std::vector<QueryReq>& qs, gst_result* out,
                       gst_edge* ext = nullptr, int32_t ext_cap = 0) {
  std::deque<QueryReq*> pend;
  for (size_t i = 0; i < qs.size(); i++) {
    qs[i].tag = i;
    pend.push_back(&qs[i]);
  }
  std::vector<int> fin;
  gst_status first_err = eng->recover();
  if (first_err != GST_OK) return first_err;
  while (!pend.empty() || eng->active() > 0) {
    if (!pend.empty()) {
      const int n = eng->admit_front(pend);
      if (n == 0 && eng->active() == 0)
        return fail(GST_OUT_OF_MEMORY, "query of k=%zu D=%d does not fit the state arena",
                    pend.front()->groups.size(), pend.front()->D);
    }
    const int act = eng->active();
    const uint32_t exit_after = pend.empty() ? (uint32_t)act : std::max(1u, (uint32_t)act / 8);
    gst_status st = eng->run_epoch(exit_after, kEpochMaxSteps, &fin);
    if (st != GST_OK) {
      eng->poison();  // reset once the stream has drained (next call)
      return st;
    }
    // the trees of the finished queries come out together (one gather, one
    // transfer), before anything is admitted into their released memory
    if (!fin.empty() && ext == nullptr) {
      gst_status gs = eng->gather_trees(fin);
      if (gs != GST_OK) {
        eng->poison();
        return gs;
      }
    }
    for (int s : fin) {
      uint64_t tag = 0;
      gst_result r;
      gst_status hs = eng->harvest(s, &r, &tag, ext, ext_cap);
      out[tag] = r;
      if (hs != GST_OK && first_err == GST_OK) first_err = hs;
    }
  }
  return first_err;
}

// The engine of gst_solve_static on each device: built over the static
// scratch on the first call there and kept until the process ends (like the
// scratch itself); calls on one device take its lock, so they never overlap.
struct StaticDev {
  std::mutex mu;
  Engine* eng = nullptr;
};
StaticDev g_static[gst::kMaxDevices];

// arena bytes of the static engine (2 slots: the scratch minus its tables)
size_t static_arena_bytes() {
  return gst::kScratchBytes - align256(2 * sizeof(gst::Slot)) - align256(sizeof(gst::Ctrl));
}

// Shortest-path backbone of a new graph: the arcs that the hub distances do
// not prove redundant (see the backbone kernels), in the same row order, with
// their own reverse-arc table.  Built once; unbounded queries and jobs run on
// it (option backbone).  Device buffers of the build are released before it
// returns; when they do not fit, the graph simply has no backbone.
gst_status build_backbone(gst_graph* g, const int64_t* beg_pos, const std::vector<int32_t>& col,
                          const std::vector<int32_t>& wt, const std::vector<int32_t>& rev) {
  const double t0 = now_ms();
  const int64_t V = g->V, E = g->n_arcs;
  if (E == 0) return GST_OK;
  // hubs: the vertices of highest degree (ties: lower id)
  const int H = (int)std::min<int64_t>(gst::kBbHubs, V);
  std::vector<int32_t> hubs(V);
  for (int64_t v = 0; v < V; v++) hubs[v] = (int32_t)v;
  std::partial_sort(hubs.begin(), hubs.begin() + H, hubs.end(), [&](int32_t a, int32_t b) {
    const int64_t da = beg_pos[a + 1] - beg_pos[a], db = beg_pos[b + 1] - beg_pos[b];
    return da != db ? da > db : a < b;
  });
  hubs.resize(H);
  const size_t Vp = (size_t)(V + 3) & ~size_t(3), Ep = (size_t)(E + 3) & ~size_t(3);
  uint32_t* dist = nullptr;
  uint8_t* flags = nullptr;
  uint8_t* keep = nullptr;
  uint32_t* changed = nullptr;
  int32_t* dhubs = nullptr;
  hipStream_t st = nullptr;
  auto release = [&]() {
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(dist);
    (void)hipFree(flags);
    (void)hipFree(keep);
    (void)hipFree(changed);
    (void)hipFree(dhubs);
  };
  if (hipMalloc(&dist, (size_t)V * gst::kBbHubs * 4) != hipSuccess || hipMalloc(&flags, 2 * Vp) != hipSuccess ||
      hipMalloc(&keep, Ep) != hipSuccess || hipMalloc(&changed, 256) != hipSuccess ||
      hipMalloc(&dhubs, sizeof(int32_t) * gst::kBbHubs) != hipSuccess) {
    release();
    (void)hipGetLastError();
    return GST_OK;  // no room for the build: the graph runs without a backbone
  }
  std::vector<uint8_t> hk(E);
  int64_t rounds = 0;
  gst_status err = [&]() -> gst_status {
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_TRY(gst_bb_fill(dist, (uint64_t)V * gst::kBbHubs, gst::kInf, st));
    HIP_TRY(gst_bb_fill((uint32_t*)flags, 2 * Vp / 4, 0u, st));
    HIP_TRY(hipMemcpyAsync(dhubs, hubs.data(), sizeof(int32_t) * H, hipMemcpyHostToDevice, st));
    uint8_t* cur = flags;
    uint8_t* nxt = flags + Vp;
    HIP_TRY(gst_bb_seed(dist, cur, dhubs, H, st));
    gst::BbArgs a;
    a.rowptr = g->d_rowptr;
    a.col = g->d_col;
    a.wt = g->d_wt;
    a.V = (uint64_t)V;
    a.n_arcs = (uint64_t)E;
    a.dist = dist;
    a.changed = changed;
    a.keep = keep;
    // label-correcting rounds until no row changes.  A vertex whose shortest
    // path from a hub has i arcs holds its distance after round i (its
    // predecessor held its own after round i - 1, changed then and relaxed in
    // round i), so at most V + 1 rounds: the hop depth of the hubs' trees + 1
    for (;;) {
      if (rounds > V + 1) return fail(GST_DEVICE_ERROR, "backbone hub rounds did not converge");
      HIP_TRY(gst_bb_fill(changed, 1, 0u, st));
      a.dcur = cur;
      a.dnxt = nxt;
      HIP_TRY(gst_bb_pass(0, &a, st));
      uint32_t ch = 0;
      HIP_TRY(hipMemcpyAsync(&ch, changed, sizeof(ch), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      rounds++;
      if (!ch) break;
      HIP_TRY(gst_bb_fill((uint32_t*)cur, Vp / 4, 0u, st));
      std::swap(cur, nxt);
    }
    HIP_TRY(gst_bb_pass(1, &a, st));
    HIP_TRY(hipMemcpyAsync(hk.data(), keep, (size_t)E, hipMemcpyDeviceToHost, st));
    g->h_hubd.resize((size_t)V * gst::kBbHubs);
    HIP_TRY(hipMemcpyAsync(g->h_hubd.data(), dist, g->h_hubd.size() * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return GST_OK;
  }();
  if (err == GST_OK) {
    // the hub distances stay: jobs start their rows at hub bounds, computed
    // for the vertices some hub reaches
    std::vector<uint32_t> hv;
    hv.reserve(V);
    for (int pass = 0; pass < 2; pass++)
      for (int64_t v = 0; v < V; v++) {
        const uint32_t* r = g->h_hubd.data() + (size_t)v * gst::kBbHubs;
        bool reached = false;
        for (int h = 0; h < gst::kBbHubs && !reached; h++) reached = r[h] < gst::kInf;
        if (reached == (pass == 0)) hv.push_back((uint32_t)v);
        if (pass == 0 && reached) g->n_hubr++;
      }
    if (hipMalloc(&g->d_hubv, sizeof(uint32_t) * V) == hipSuccess &&
        hipMemcpy(g->d_hubv, hv.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice) == hipSuccess) {
      g->d_hubd = dist;
      dist = nullptr;
      // a 16-bit copy when every finite hub distance is below half of
      // kHub16Inf: then a bound plus a distance never reaches it, and the
      // jobs' hub FILL works on two labels per packed instruction
      std::atomic<uint32_t> dmax{0};
      const size_t nh = g->h_hubd.size();
      parallel_for((int64_t)nh, [&](int64_t b, int64_t e) {
        uint32_t m = 0;
        for (int64_t i = b; i < e; i++)
          if (g->h_hubd[i] < gst::kInf) m = std::max(m, g->h_hubd[i]);
        uint32_t cur = dmax.load();
        while (m > cur && !dmax.compare_exchange_weak(cur, m)) {}
      });
      if (dmax.load() < gst::kHub16Inf / 2) {
        std::vector<uint16_t> h16(nh);
        parallel_for((int64_t)nh, [&](int64_t b, int64_t e) {
          for (int64_t i = b; i < e; i++)
            h16[i] = (uint16_t)(g->h_hubd[i] < gst::kInf ? g->h_hubd[i] : gst::kHub16Inf);
        });
        if (hipMalloc(&g->d_hubd16, nh * 2) != hipSuccess ||
            hipMemcpy(g->d_hubd16, h16.data(), nh * 2, hipMemcpyHostToDevice) != hipSuccess) {
          (void)hipFree(g->d_hubd16);
          (void)hipGetLastError();
          g->d_hubd16 = nullptr;
        }
      }
    } else {
      (void)hipFree(g->d_hubv);
      (void)hipGetLastError();
      g->d_hubv = nullptr;
      g->n_hubr = 0;
      g->h_hubd.clear();
    }
  } else {
    g->h_hubd.clear();
  }
  release();
  if (err != GST_OK) return err;
  // the test is symmetric (d(h,u) + d(h,v), and reverse arcs have equal
  // weights), so an arc and its reverse are kept together
  std::atomic<int64_t> asym{-1};
  parallel_for(E, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++)
      if (hk[i] != hk[rev[i]]) asym = i;
  });
  if (asym >= 0)
    return fail(GST_DEVICE_ERROR, "backbone test kept arc %lld but not its reverse", (long long)asym.load());
  std::vector<int64_t> brow(V + 1, 0);
  parallel_for_rows(beg_pos, V, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; u++) {
      int64_t c = 0;
      for (int64_t i = beg_pos[u]; i < beg_pos[u + 1]; i++) c += hk[i];
      brow[u + 1] = c;
    }
  });
  for (int64_t u = 0; u < V; u++) brow[u + 1] += brow[u];
  const int64_t nb = brow[V];
  // arc indices of the backbone run on from E in the shared arrays, and a
  // grow predecessor holds one in 30 bits
  if (E + nb >= (int64_t(1) << 30)) return GST_OK;
  std::vector<int32_t> nidx(E, -1), acol(E + nb), awt(E + nb), arev(E + nb);
  std::vector<int64_t> arow(2 * (V + 1));
  std::copy(col.begin(), col.end(), acol.begin());
  std::copy(wt.begin(), wt.end(), awt.begin());
  std::copy(rev.begin(), rev.end(), arev.begin());
  for (int64_t v = 0; v <= V; v++) {
    arow[v] = beg_pos[v];
    arow[V + 1 + v] = E + brow[v];
  }
  parallel_for_rows(beg_pos, V, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; u++) {
      int64_t j = E + brow[u];
      for (int64_t i = beg_pos[u]; i < beg_pos[u + 1]; i++) {
        if (!hk[i]) continue;
        nidx[i] = (int32_t)j;
        acol[j] = col[i];
        awt[j] = wt[i];
        j++;
      }
    }
  });
  parallel_for(E, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++)
      if (hk[i]) arev[nidx[i]] = nidx[rev[i]];
  });
  int64_t* d_rowptr = nullptr;
  int32_t *d_col = nullptr, *d_wt = nullptr, *d_rev = nullptr;
  auto drop = [&]() {
    (void)hipFree(d_rowptr);
    (void)hipFree(d_col);
    (void)hipFree(d_wt);
    (void)hipFree(d_rev);
    (void)hipGetLastError();
  };
  if (hipMalloc(&d_rowptr, sizeof(int64_t) * arow.size()) != hipSuccess ||
      hipMalloc(&d_col, sizeof(int32_t) * acol.size()) != hipSuccess ||
      hipMalloc(&d_wt, sizeof(int32_t) * awt.size()) != hipSuccess ||
      hipMalloc(&d_rev, sizeof(int32_t) * arev.size()) != hipSuccess) {
    drop();
    return GST_OK;  // no room: the graph runs without a backbone
  }
  gst_status cst = [&]() -> gst_status {
    HIP_TRY(hipMemcpy(d_rowptr, arow.data(), sizeof(int64_t) * arow.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_col, acol.data(), sizeof(int32_t) * acol.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_wt, awt.data(), sizeof(int32_t) * awt.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_rev, arev.data(), sizeof(int32_t) * arev.size(), hipMemcpyHostToDevice));
    return GST_OK;
  }();
  if (cst != GST_OK) {
    drop();
    return cst;
  }
  for (void* p : {(void*)g->d_rowptr, (void*)g->d_col, (void*)g->d_wt, (void*)g->d_rev}) (void)hipFree(p);
  g->d_rowptr = d_rowptr;
  g->d_col = d_col;
  g->d_wt = d_wt;
  g->d_rev = d_rev;
  g->has_bb = true;
  g->bb_arcs = nb;
  g->bb_rounds = rounds;
  g->bb_ms = now_ms() - t0;
  return GST_OK;
}

// Device memory of a graph (its engines are gone already).
void free_graph_memory(gst_graph* g) {
  for (void* p : {(void*)g->d_rowptr, (void*)g->d_col, (void*)g->d_wt, (void*)g->d_rev, (void*)g->d_gptr,
                  (void*)g->d_gmem, (void*)g->d_hubd, (void*)g->d_hubd16, (void*)g->d_hubv})
    if (p) (void)hipFree(p);
}

}  // namespace

// ------------------------------------------------------------------- C ABI

extern "C" {

const char* gst_last_error(void) { return g_last_error.c_str(); }
const char* gst_version(void) { return "gpugst-amd 0.1 (gfx950)"; }

gst_status gst_graph_create(const int64_t* beg_pos, int64_t V, const int32_t* adj, const int32_t* w,
                            int64_t n_arcs, gst_graph** out) {
  if (!out) return fail(GST_INVALID_ARG, "null output handle");
  *out = nullptr;
  if (!beg_pos || V < 1 || V >= (int64_t(1) << 31) - 1) return fail(GST_INVALID_ARG, "bad V=%lld", (long long)V);
  if (n_arcs < 0 || n_arcs >= (int64_t(1) << 30)) return fail(GST_INVALID_ARG, "bad n_arcs=%lld", (long long)n_arcs);
  if (n_arcs > 0 && (!adj || !w)) return fail(GST_INVALID_ARG, "null adjacency or weights");
  if (beg_pos[0] != 0 || beg_pos[V] != n_arcs)
    return fail(GST_INVALID_ARG, "beg_pos must start at 0 and end at n_arcs");
  for (int64_t v = 0; v < V; v++)
    if (beg_pos[v + 1] < beg_pos[v]) return fail(GST_INVALID_ARG, "beg_pos decreases at %lld", (long long)v);
  std::atomic<int64_t> bad_col{-1}, bad_w{-1};
  parallel_for(n_arcs, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      if (adj[i] < 0 || adj[i] >= V) bad_col = i;
      if (w[i] < 1 || w[i] > (1 << 20)) bad_w = i;
    }
  });
  if (bad_col >= 0) return fail(GST_INVALID_ARG, "adjacency entry %lld out of range", (long long)bad_col.load());
  if (bad_w >= 0) return fail(GST_INVALID_ARG, "weight of arc %lld outside [1, 2^20]", (long long)bad_w.load());

  // rows sorted by (neighbour, weight), so the reverse of every arc is found
  // by binary search
  std::vector<int32_t> col(adj, adj + n_arcs), wt(w, w + n_arcs);
  parallel_for_rows(beg_pos, V, [&](int64_t b, int64_t e) {
    std::vector<std::pair<int32_t, int32_t>> tmp;
    for (int64_t v = b; v < e; v++) {
      const int64_t r0 = beg_pos[v], r1 = beg_pos[v + 1];
      bool sorted = true;
      for (int64_t i = r0 + 1; i < r1 && sorted; i++)
        if (col[i] < col[i - 1] || (col[i] == col[i - 1] && wt[i] < wt[i - 1])) sorted = false;
      if (sorted) continue;
      tmp.clear();
      for (int64_t i = r0; i < r1; i++) tmp.emplace_back(col[i], wt[i]);
      std::sort(tmp.begin(), tmp.end());
      for (int64_t i = r0; i < r1; i++) {
        col[i] = tmp[i - r0].first;
        wt[i] = tmp[i - r0].second;
      }
    }
  });
  std::vector<int32_t> rev(n_arcs);
  std::atomic<int64_t> asym{-1};
  parallel_for_rows(beg_pos, V, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; u++) {
      for (int64_t i = beg_pos[u]; i < beg_pos[u + 1]; i++) {
        const int32_t v = col[i], ww = wt[i];
        const int32_t* lo = col.data() + beg_pos[v];
        const int32_t* hi = col.data() + beg_pos[v + 1];
        const int32_t* it = std::lower_bound(lo, hi, (int32_t)u);
        int64_t j = it - col.data();
        while (j < beg_pos[v + 1] && col[j] == u && wt[j] < ww) j++;
        if (j >= beg_pos[v + 1] || col[j] != u || wt[j] != ww) {
          asym = i;
          continue;
        }
        rev[i] = (int32_t)j;
      }
    }
  });
  if (asym >= 0)
    return fail(GST_INVALID_ARG, "arc %lld has no reverse arc of equal weight (graph must be undirected)",
                (long long)asym.load());

  std::unique_ptr<gst_graph> g(new gst_graph());
  HIP_TRY(hipGetDevice(&g->device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, g->device));
  g->num_cu = prop.multiProcessorCount;
  for (int b = 0; b < 3; b++)
    for (int c = 0; c < gst::kNumCls; c++) {
      int occ = 1;
      if (gst_step_occupancy(c, gst::kWaveChoices[b], &occ) == hipSuccess && occ > 0) g->occ[b][c] = occ;
    }
  g->V = V;
  g->n_arcs = n_arcs;
  {
    // resolve this device's static scratch now (gst_solve_static reports a
    // failure on its own call)
    void* sp = nullptr;
    (void)gst_scratch_setup(g->device, &sp);
  }
  HIP_TRY(hipMalloc(&g->d_rowptr, sizeof(int64_t) * (V + 1)));
  HIP_TRY(hipMalloc(&g->d_col, sizeof(int32_t) * std::max<int64_t>(n_arcs, 1)));
  HIP_TRY(hipMalloc(&g->d_wt, sizeof(int32_t) * std::max<int64_t>(n_arcs, 1)));
  HIP_TRY(hipMalloc(&g->d_rev, sizeof(int32_t) * std::max<int64_t>(n_arcs, 1)));
  HIP_TRY(hipMemcpy(g->d_rowptr, beg_pos, sizeof(int64_t) * (V + 1), hipMemcpyHostToDevice));
  if (n_arcs) {
    HIP_TRY(hipMemcpy(g->d_col, col.data(), sizeof(int32_t) * n_arcs, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->d_wt, wt.data(), sizeof(int32_t) * n_arcs, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->d_rev, rev.data(), sizeof(int32_t) * n_arcs, hipMemcpyHostToDevice));
  }
  gst_status bst = build_backbone(g.get(), beg_pos, col, wt, rev);
  if (bst != GST_OK) {
    free_graph_memory(g.get());
    return bst;
  }
  *out = g.release();
  return GST_OK;
}

gst_status gst_backbone_get(gst_graph* g, int64_t* beg_pos_out, int32_t* adj_out, int32_t* w_out,
                            int64_t* n_arcs_out) {
  if (!g || !beg_pos_out || !n_arcs_out) return fail(GST_INVALID_ARG, "null argument");
  const bool bb = g->has_bb;
  const int64_t n = bb ? g->bb_arcs : g->n_arcs, first = bb ? g->n_arcs : 0;
  std::lock_guard<std::mutex> lk(g->dev_mu);
  HIP_TRY(hipMemcpy(beg_pos_out, g->d_rowptr + (bb ? g->V + 1 : 0), sizeof(int64_t) * (g->V + 1),
                    hipMemcpyDeviceToHost));
  for (int64_t v = 0; v <= g->V; v++) beg_pos_out[v] -= first;
  if (adj_out && n) HIP_TRY(hipMemcpy(adj_out, g->d_col + first, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  if (w_out && n) HIP_TRY(hipMemcpy(w_out, g->d_wt + first, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  *n_arcs_out = n;
  return GST_OK;
}

gst_status gst_graph_destroy(gst_graph* g) {
  if (!g) return GST_OK;
  {
    std::lock_guard<std::mutex> lk(g->dev_mu);
    g->plain.reset();
    g->batch.reset();
  }
  free_graph_memory(g);
  delete g;
  return GST_OK;
}

gst_status gst_groups_set(gst_graph* g, const int64_t* grp_ptr, const int32_t* members, int64_t n_groups) {
  if (!g || !grp_ptr || n_groups < 1) return fail(GST_INVALID_ARG, "bad group arguments");
  if (grp_ptr[0] != 0) return fail(GST_INVALID_ARG, "grp_ptr must start at 0");
  for (int64_t i = 0; i < n_groups; i++)
    if (grp_ptr[i + 1] < grp_ptr[i]) return fail(GST_INVALID_ARG, "grp_ptr decreases at %lld", (long long)i);
  const int64_t nm = grp_ptr[n_groups];
  if (nm > 0 && !members) return fail(GST_INVALID_ARG, "null members");
  for (int64_t i = 0; i < nm; i++)
    if (members[i] < 0 || members[i] >= g->V)
      return fail(GST_INVALID_ARG, "member %lld = %d out of range", (long long)i, members[i]);
  // servers of this graph queue no new chunk from now on; the arrays change
  // once the chunks they queued have finished
  struct Quiesce {
    std::atomic<int>& q;
    explicit Quiesce(std::atomic<int>& x) : q(x) { q++; }
    ~Quiesce() { q--; }
  } quiesce(g->quiesce);
  std::unique_lock<std::mutex> lk(g->dev_mu);
  while (g->cont_out > 0) {
    lk.unlock();
    std::this_thread::sleep_for(std::chrono::microseconds(100));
    lk.lock();
  }
  if (g->d_gptr) (void)hipFree(g->d_gptr);
  if (g->d_gmem) (void)hipFree(g->d_gmem);
  g->d_gptr = nullptr;
  g->d_gmem = nullptr;
  HIP_TRY(hipMalloc(&g->d_gptr, sizeof(int64_t) * (n_groups + 1)));
  HIP_TRY(hipMalloc(&g->d_gmem, sizeof(int32_t) * std::max<int64_t>(nm, 1)));
  HIP_TRY(hipMemcpy(g->d_gptr, grp_ptr, sizeof(int64_t) * (n_groups + 1), hipMemcpyHostToDevice));
  if (nm) HIP_TRY(hipMemcpy(g->d_gmem, members, sizeof(int32_t) * nm, hipMemcpyHostToDevice));
  g->h_gptr.assign(grp_ptr, grp_ptr + n_groups + 1);
  g->h_gmem.assign(members, members + nm);
  g->n_groups = n_groups;
  g->groups_version++;
  return GST_OK;
}

static int64_t* option_ref(gst_graph* g, const std::string& k) {
  if (k == "max_slots") return &g->opt.max_slots;
  if (k == "budget_mb") return &g->opt.budget_mb;
  if (k == "blocks_per_cu") return &g->opt.blocks_per_cu;
  if (k == "chunk") return &g->opt.chunk;
  if (k == "edge_cap") return &g->opt.edge_cap;
  if (k == "timeout_ms") return &g->opt.timeout_ms;
  if (k == "graph") return &g->opt.graph;
  if (k == "group_pass") return &g->opt.group_pass;
  if (k == "time_steps") return &g->opt.time_steps;
  if (k == "trace_slot") return &g->opt.trace_slot;
  if (k == "job_min_queries") return &g->opt.job_min_queries;
  if (k == "deadline_ms") return &g->opt.deadline_ms;
  if (k == "count_states") return &g->opt.count_states;
  if (k == "cache_mb") return &g->opt.cache_mb;
  if (k == "block") return &g->opt.block;
  if (k == "merge_tv") return &g->opt.merge_tv;
  if (k == "job_lanes") return &g->opt.job_lanes;
  if (k == "approx_num") return &g->opt.approx_num;
  if (k == "approx_den") return &g->opt.approx_den;
  if (k == "backbone") return &g->opt.backbone;
  if (k == "hub_bounds") return &g->opt.hub_bounds;
  if (k == "hub16") return &g->opt.hub16;
  if (k == "serve_chunk") return &g->opt.serve_chunk;
  if (k == "serve_depth") return &g->opt.serve_depth;
  if (k == "serve_wait_us") return &g->opt.serve_wait_us;
  if (k == "serve_batch") return &g->opt.serve_batch;
  return nullptr;
}

gst_status gst_set_option(gst_graph* g, const char* key, int64_t value) {
  if (!g || !key) return fail(GST_INVALID_ARG, "null argument");
  if (std::string(key) == "reset_stats") {
    std::lock_guard<std::mutex> lk(g->dev_mu);
    if (g->plain) g->plain->reset_stats();
    if (g->batch) g->batch->reset_stats();
    return GST_OK;
  }
  int64_t* r = option_ref(g, key);
  if (!r) return fail(GST_INVALID_ARG, "unknown option '%s'", key);
  const std::string k(key);
  if (k == "max_slots" && (value < 1 || value > gst::kMaxSlots))
    return fail(GST_INVALID_ARG, "max_slots outside [1,%d]", gst::kMaxSlots);
  if (k == "chunk" && (value < 64 || value > (1 << 20))) return fail(GST_INVALID_ARG, "chunk outside [64, 2^20]");
  if (k == "edge_cap" && (value < 64 || value > (1 << 26))) return fail(GST_INVALID_ARG, "edge_cap outside [64, 2^26]");
  if ((k == "budget_mb" || k == "blocks_per_cu" || k == "timeout_ms" || k == "trace_slot" || k == "deadline_ms" ||
       k == "count_states" || k == "cache_mb" || k == "backbone" || k == "hub_bounds" || k == "hub16") && value < 0)
    return fail(GST_INVALID_ARG, "negative value");
  if (k == "block" && value != 2 && value != 4 && value != 8)
    return fail(GST_INVALID_ARG, "block (waves per block) must be 2, 4 or 8");
  if (k == "merge_tv" && (value < 1 || value > 4096)) return fail(GST_INVALID_ARG, "merge_tv outside [1, 4096]");
  if (k == "serve_chunk" && (value < 1 || value > 64)) return fail(GST_INVALID_ARG, "serve_chunk outside [1, 64]");
  if (k == "serve_depth" && (value < 1 || value > 2)) return fail(GST_INVALID_ARG, "serve_depth must be 1 or 2");
  if (k == "serve_wait_us" && (value < 0 || value > 10000000))
    return fail(GST_INVALID_ARG, "serve_wait_us outside [0, 10^7]");
  if (k == "serve_batch" && (value < 1 || value > gst::kMaxSlots))
    return fail(GST_INVALID_ARG, "serve_batch outside [1, %d]", gst::kMaxSlots);
  if (k == "job_lanes" && (value < 0 || value > 64 || (value & (value - 1))))
    return fail(GST_INVALID_ARG, "job_lanes must be 0 or a power of two up to 64");
  if ((k == "approx_num" || k == "approx_den") && (value < 1 || value > (1 << 20)))
    return fail(GST_INVALID_ARG, "%s outside [1, 2^20]", key);
  if (k == "job_min_queries" && (value < 1 || value > gst::kJobLabels))
    return fail(GST_INVALID_ARG, "job_min_queries outside [1,%d]", gst::kJobLabels);
  std::lock_guard<std::mutex> lk(g->dev_mu);
  *r = value;
  if (k == "max_slots" || k == "budget_mb") g->batch.reset();  // rebuilt on next use
  return GST_OK;
}

gst_status gst_get_option(gst_graph* g, const char* key, int64_t* value) {
  if (!g || !key || !value) return fail(GST_INVALID_ARG, "null argument");
  const std::string k(key);
  if (k == "num_cu") { *value = g->num_cu; return GST_OK; }
  if (k == "backbone_arcs") { *value = g->has_bb ? g->bb_arcs : g->n_arcs; return GST_OK; }
  if (k == "backbone_rounds") { *value = g->bb_rounds; return GST_OK; }
  if (k == "backbone_us") { *value = (int64_t)(g->bb_ms * 1e3); return GST_OK; }
  if (k == "hub_bytes") { *value = g->d_hubd16 && g->opt.hub16 ? 2 : 4; return GST_OK; }
  if (k.rfind("plain_", 0) == 0 || k.rfind("batch_", 0) == 0) {
    Engine* e = (k[0] == 'p') ? g->plain.get() : g->batch.get();
    const std::string f = k.substr(6);
    if (f == "kernel_ns") { *value = e ? e->kernel_ns() : 0; return GST_OK; }
    if (f == "launches") { *value = e ? e->launches() : 0; return GST_OK; }
    if (f == "steps") { *value = e ? e->steps_total() : 0; return GST_OK; }
    if (f == "lean_ns") { *value = e ? e->cls_ns(gst::kClsLean) : 0; return GST_OK; }
    if (f == "heavy_ns") { *value = e ? e->cls_ns(gst::kClsHeavy) : 0; return GST_OK; }
    if (f == "timed_steps") { *value = e ? e->cls_launches() : 0; return GST_OK; }
    if (f == "jobs") { *value = e ? e->jobs_done() : 0; return GST_OK; }
    if (f == "job_entries") { *value = e ? e->job_ctr(1) : 0; return GST_OK; }
    if (f == "job_arcs") { *value = e ? e->job_ctr(2) : 0; return GST_OK; }
    if (f == "job_relaxations") { *value = e ? e->job_ctr(3) : 0; return GST_OK; }
    if (f == "job_improvements") { *value = e ? e->job_ctr(4) : 0; return GST_OK; }
    if (f == "job_atomics") { *value = e ? e->job_ctr(5) : 0; return GST_OK; }
    if (f == "job_src_chunks") { *value = e ? e->job_ctr(6) : 0; return GST_OK; }
    if (f == "cache_hits") { *value = e ? e->cache_hits_ : 0; return GST_OK; }
    if (f == "cache_misses") { *value = e ? e->cache_misses_ : 0; return GST_OK; }
    if (f == "cache_inserts") { *value = e ? e->cache_inserts_ : 0; return GST_OK; }
    if (f == "cache_evictions") { *value = e ? e->cache_evictions_ : 0; return GST_OK; }
    if (f == "cache_entries") { *value = e ? (int64_t)e->cache_entries() : 0; return GST_OK; }
    // phase_ns:<phase>:<level>, phase_steps:<phase>:<level>
    int ph = 0, lv = 0;
    if (sscanf(f.c_str(), "phase_ns:%d:%d", &ph, &lv) == 2) {
      *value = e ? e->phase_stat(ph, lv, false) : 0;
      return GST_OK;
    }
    if (sscanf(f.c_str(), "phase_steps:%d:%d", &ph, &lv) == 2) {
      *value = e ? e->phase_stat(ph, lv, true) : 0;
      return GST_OK;
    }
  }
  auto grid_of = [&](int c) {
    const int occ = g->occ[wpb_index(g->opt.block)][c];
    const int64_t b = g->opt.blocks_per_cu > 0 ? std::min<int64_t>(g->opt.blocks_per_cu, occ) : occ;
    return std::max<int64_t>(1, b) * g->num_cu;
  };
  if (k == "occupancy") { *value = g->occ[wpb_index(g->opt.block)][gst::kClsLean]; return GST_OK; }
  if (k == "occupancy_heavy") { *value = g->occ[wpb_index(g->opt.block)][gst::kClsHeavy]; return GST_OK; }
  if (k == "grid") { *value = grid_of(gst::kClsLean); return GST_OK; }
  if (k == "grid_heavy") { *value = grid_of(gst::kClsHeavy); return GST_OK; }
  int64_t* r = option_ref(g, key);
  if (!r) return fail(GST_INVALID_ARG, "unknown option '%s'", key);
  *value = *r;
  return GST_OK;
}

gst_status gst_query_bytes(gst_graph* g, int32_t k, int32_t D, int64_t* bytes) {
  if (!g || !bytes || k < 1 || k > GST_MAX_K || D < -1 || D > GST_MAX_D) return fail(GST_INVALID_ARG, "bad arguments");
  *bytes = (int64_t)single_query_bytes(g, k, D);
  return GST_OK;
}

static bool has_empty_group(const gst_graph* g, const int32_t* ids, int k) {
  for (int i = 0; i < k; i++)
    if (g->h_gptr[ids[i] + 1] == g->h_gptr[ids[i]]) return true;
  return false;
}

gst_status gst_solve(gst_graph* g, const int32_t* group_ids, int32_t k, int32_t D, gst_result* out) {
  if (!out) return fail(GST_INVALID_ARG, "null result");
  memset(out, 0, sizeof(*out));
  gst_status st = validate_query(g, group_ids, k, D);
  if (st != GST_OK) return st;
  if (has_empty_group(g, group_ids, k)) {
    fill_infeasible(out);
    return GST_OK;
  }
  std::lock_guard<std::mutex> lk(g->dev_mu);
  Engine* eng = nullptr;
  st = engine_for(g, false, single_query_bytes(g, k, D), &eng);
  if (st != GST_OK) return st;
  std::vector<QueryReq> qs(1);
  qs[0].groups.assign(group_ids, group_ids + k);
  qs[0].D = D;
  qs[0].t_submit = now_ms();
  if (g->opt.deadline_ms > 0) qs[0].t_deadline = qs[0].t_submit + (double)g->opt.deadline_ms;
  return run_queries(g, eng, qs, out);
}

gst_status gst_solve_static(gst_graph* g, const int32_t* group_ids, int32_t k, int32_t D, gst_edge* edges,
                            int32_t edge_cap, gst_result* out) {
  if (!out) return fail(GST_INVALID_ARG, "null result");
  memset(out, 0, sizeof(*out));
  gst_status st = validate_query(g, group_ids, k, D);
  if (st != GST_OK) return st;
  if (edge_cap < 0 || (edge_cap > 0 && !edges)) return fail(GST_INVALID_ARG, "bad edge buffer");
  if (g->device < 0 || g->device >= gst::kMaxDevices) return fail(GST_INVALID_ARG, "device %d", g->device);
  if (has_empty_group(g, group_ids, k)) {
    fill_infeasible(out);
    return GST_OK;
  }
  const size_t need = layout_for(g, k, D, true).total;
  if (need > static_arena_bytes())
    return fail(GST_OUT_OF_MEMORY,
                "a query of k=%d D=%d on %lld vertices needs %zu bytes of state; the static scratch holds "
                "%zu (gst_solve allocates what it needs instead)",
                k, D, (long long)g->V, need, static_arena_bytes());
  std::lock_guard<std::mutex> lk(g->dev_mu);
  StaticDev& sd = g_static[g->device];
  std::lock_guard<std::mutex> sl(sd.mu);
  if (!sd.eng) {
    void* base = nullptr;
    const hipError_t e = gst_scratch_setup(g->device, &base);
    if (e != hipSuccess) return fail(GST_DEVICE_ERROR, "static scratch address: %s", hipGetErrorString(e));
    std::unique_ptr<Engine> ne(new Engine(g, 2));
    st = ne->init_static((char*)base, gst::kScratchBytes);
    if (st != GST_OK) return st;
    sd.eng = ne.release();
  }
  sd.eng->rebind(g);
  std::vector<QueryReq> qs(1);
  qs[0].groups.assign(group_ids, group_ids + k);
  qs[0].D = D;
  qs[0].t_submit = now_ms();
  return run_queries(g, sd.eng, qs, out, edges, edge_cap);
}

gst_status gst_solve_batch(gst_graph* g, int64_t n, const gst_query* q, gst_result* out) {
  if (!g || n < 0 || (n > 0 && (!q || !out))) return fail(GST_INVALID_ARG, "bad batch arguments");
  std::vector<QueryReq> qs;
  std::vector<int64_t> map;
  const double t0 = now_ms();
  for (int64_t i = 0; i < n; i++) {
    memset(&out[i], 0, sizeof(gst_result));
    gst_status st = validate_query(g, q[i].group_ids, q[i].k, q[i].D);
    if (st != GST_OK) return st;
    if (has_empty_group(g, q[i].group_ids, q[i].k)) {
      fill_infeasible(&out[i]);
      continue;
    }
    QueryReq r;
    r.groups.assign(q[i].group_ids, q[i].group_ids + q[i].k);
    r.D = q[i].D;
    r.t_submit = t0;
    if (g->opt.deadline_ms > 0) r.t_deadline = t0 + (double)g->opt.deadline_ms;
    qs.push_back(std::move(r));
    map.push_back(i);
  }
  if (qs.empty()) return GST_OK;
  std::lock_guard<std::mutex> lk(g->dev_mu);
  Engine* eng = nullptr;
  gst_status st = engine_for(g, true, 0, &eng);
  if (st != GST_OK) return st;
  std::vector<gst_result> tmp(qs.size());
  st = run_queries(g, eng, qs, tmp.data());
  for (size_t i = 0; i < qs.size(); i++) out[map[i]] = tmp[i];
  return st;
}

void gst_result_free(gst_result* r) {
  if (!r) return;
  free(r->edges);
  r->edges = nullptr;
  r->n_edges = 0;
}

}  // extern "C"

// ------------------------------------------------------------------ server

struct gst_server {
  gst_graph* g = nullptr;
  std::unique_ptr<Engine> eng;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_new, cv_done;
  std::deque<std::pair<uint64_t, QueryReq>> pending;
  std::unordered_map<uint64_t, gst_result> done;
  std::unordered_map<uint64_t, int> known;  // 0 queued, 1 admitted, 2 done
  uint64_t next_ticket = 1;
  bool stop = false;
  gst_status error = GST_OK;
  std::string error_msg;
  size_t max_pending = 1 << 20;
  int waiters = 0;  // threads inside gst_wait
  // device statistics (gst_server_stat), written by the server thread
  std::atomic<int64_t> steps{0}, chunks{0}, busy_ns{0};
};

namespace {

gst_result cancelled_result() {
  gst_result r;
  memset(&r, 0, sizeof(r));
  r.status = GST_CANCELLED;
  r.cost = GST_INFEASIBLE;
  r.root = -1;
  return r;
}

// Answers the pending requests whose deadline has passed (caller holds s->mu).
void drop_expired(gst_server* s) {
  const double t = now_ms();
  auto expired = [t](const std::pair<uint64_t, QueryReq>& pr) {
    return pr.second.t_deadline > 0 && t > pr.second.t_deadline;
  };
  if (std::none_of(s->pending.begin(), s->pending.end(), expired)) return;
  std::deque<std::pair<uint64_t, QueryReq>> rest;
  for (auto& pr : s->pending) {
    if (expired(pr)) {
      s->done[pr.first] = cancelled_result();
      s->known[pr.first] = 2;
    } else {
      rest.push_back(std::move(pr));
    }
  }
  s->pending.swap(rest);
  s->cv_done.notify_all();
}

void server_loop(gst_server* s) {
  std::vector<int> fin;
  Engine* e = s->eng.get();
  bool quiesced = false;
  for (;;) {
    // while the group arrays are being replaced the device gets no new
    // chunk: wait a little instead of spinning
    if (quiesced) std::this_thread::sleep_for(std::chrono::microseconds(200));
    quiesced = false;
    int64_t chunk, depth, wait_us, batch;
    {
      std::lock_guard<std::mutex> dk(s->g->dev_mu);
      chunk = s->g->opt.serve_chunk;
      depth = s->g->opt.serve_depth;
      wait_us = s->g->opt.serve_wait_us;
      batch = s->g->opt.serve_batch;
    }
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // sleep only while the device holds nothing of this server's
      s->cv_new.wait(lk, [&] {
        return s->stop || !s->pending.empty() || e->active() > 0 || e->cont_outstanding();
      });
      if (s->stop && s->pending.empty() && e->active() == 0 && !e->cont_outstanding()) break;
      drop_expired(s);
      // While the device is busy, admission may hold requests back a little
      // so that more of them share a group-distance job; an idle device
      // takes a request at once.
      bool admit = !s->pending.empty();
      if (admit && e->active() > 0 && (int64_t)s->pending.size() < batch)
        admit = now_ms() - s->pending.front().second.t_submit >= (double)wait_us * 1e-3;
      while (admit && !s->pending.empty()) {
        std::deque<QueryReq*> pend;
        for (auto& pr : s->pending) {
          if (pend.size() >= (size_t)gst::kMaxSlots) break;
          pr.second.tag = pr.first;
          pend.push_back(&pr.second);
        }
        const std::vector<QueryReq*> offered(pend.begin(), pend.end());
        int n;
        {
          std::lock_guard<std::mutex> dk(s->g->dev_mu);
          n = e->admit_front(pend);
        }
        if (n > 0) {
          // the admitted requests are the offered ones no longer pending (not
          // always a prefix: a job gathers requests of one bound from further
          // back in the queue)
          std::unordered_set<QueryReq*> left(pend.begin(), pend.end());
          std::unordered_set<uint64_t> taken;
          for (QueryReq* q : offered)
            if (!left.count(q)) taken.insert(q->tag);
          std::deque<std::pair<uint64_t, QueryReq>> rest;
          for (auto& pr : s->pending) {
            if (taken.count(pr.first)) s->known[pr.first] = 1;
            else rest.push_back(std::move(pr));
          }
          s->pending.swap(rest);
        }
        if (n > 0 || e->active() > 0) break;
        // the first pending query can never fit: answer it with an error result
        gst_result r;
        memset(&r, 0, sizeof(r));
        r.status = GST_OUT_OF_MEMORY;
        r.cost = GST_INFEASIBLE;
        r.root = -1;
        s->done[s->pending.front().first] = r;
        s->known[s->pending.front().first] = 2;
        s->pending.pop_front();
        s->cv_done.notify_all();
      }
    }
    if (e->active() == 0 && !e->cont_outstanding()) continue;
    // One chunk of steps goes on the device; the snapshot of the chunk before
    // it (or of this one, with nothing queued behind) is read back.
    gst_status st;
    std::vector<std::pair<uint64_t, gst_result>> out;
    bool held = false;
    {
      std::lock_guard<std::mutex> dk(s->g->dev_mu);
      // while the group arrays are being replaced, queue no new chunk
      held = s->g->quiesce.load() > 0;
      const int before = e->cont_queued();
      st = e->cont_iter((int)chunk, (int)depth, &fin, !held);
      s->g->cont_out += e->cont_queued() - before;
      if (st != GST_OK) {
        s->g->cont_out -= e->cont_queued();  // poisoned: its chunks are never collected
        e->poison();  // the server stops; nothing runs on this engine again
      }
      if (st == GST_OK && !fin.empty()) {
        st = e->gather_trees(fin);  // the finished trees, in one transfer
        if (st != GST_OK) e->poison();
      }
      for (int slot : fin) {
        if (st != GST_OK) break;
        uint64_t tag = 0;
        gst_result r;
        e->harvest(slot, &r, &tag);
        out.push_back({tag, r});
      }
      s->steps = e->steps_total();
      s->chunks = e->cont_chunks();
      s->busy_ns = e->cont_busy_ns();
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (st != GST_OK) {
      s->error = st;
      s->error_msg = g_last_error;
      s->stop = true;
      s->cv_done.notify_all();
      break;
    }
    for (auto& pr : out) {
      s->done[pr.first] = pr.second;
      s->known[pr.first] = 2;
    }
    if (!out.empty()) s->cv_done.notify_all();
    if (held) quiesced = true;
  }
}


}  // namespace

extern "C" {

gst_status gst_server_create(gst_graph* g, int64_t budget_bytes, int32_t max_slots, gst_server** out) {
  if (!g || !out) return fail(GST_INVALID_ARG, "null argument");
  if (g->n_groups <= 0) return fail(GST_INVALID_ARG, "no groups set");
  std::unique_ptr<gst_server> s(new gst_server());
  s->g = g;
  int slots = max_slots > 0 ? std::min(max_slots, gst::kMaxSlots) : (int)g->opt.max_slots;
  s->eng.reset(new Engine(g, slots));
  size_t bytes = (size_t)budget_bytes;
  if (budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(GST_DEVICE_ERROR, "hipMemGetInfo failed");
    bytes = (size_t)(fr * 0.85);
  }
  {
    std::lock_guard<std::mutex> dk(g->dev_mu);
    gst_status st = s->eng->init(bytes);
    if (st != GST_OK) return st;
  }
  gst_server* sp = s.get();
  s->worker = std::thread(server_loop, sp);
  *out = s.release();
  return GST_OK;
}

gst_status gst_submit(gst_server* s, const gst_query* q, uint64_t* ticket) {
  if (!s || !q || !ticket) return fail(GST_INVALID_ARG, "null argument");
  gst_status st = validate_query(s->g, q->group_ids, q->k, q->D);
  if (st != GST_OK) return st;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error != GST_OK) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
  if (s->stop) return fail(GST_CANCELLED, "server is shutting down");
  if (s->pending.size() >= s->max_pending) return fail(GST_QUEUE_FULL, "server queue full");
  const uint64_t t = s->next_ticket++;
  *ticket = t;
  if (has_empty_group(s->g, q->group_ids, q->k)) {
    gst_result r;
    fill_infeasible(&r);
    s->done[t] = r;
    s->known[t] = 2;
    s->cv_done.notify_all();
    return GST_OK;
  }
  QueryReq r;
  r.groups.assign(q->group_ids, q->group_ids + q->k);
  r.D = q->D;
  r.t_submit = now_ms();
  if (s->g->opt.deadline_ms > 0) r.t_deadline = r.t_submit + (double)s->g->opt.deadline_ms;
  s->pending.emplace_back(t, std::move(r));
  s->known[t] = 0;
  s->cv_new.notify_one();
  return GST_OK;
}

gst_status gst_server_stat(gst_server* s, const char* key, int64_t* value) {
  if (!s || !key || !value) return fail(GST_INVALID_ARG, "null argument");
  const std::string k(key);
  if (k == "steps") *value = s->steps.load();
  else if (k == "chunks") *value = s->chunks.load();
  else if (k == "busy_us") *value = s->busy_ns.load() / 1000;
  else return fail(GST_INVALID_ARG, "unknown server statistic '%s'", key);
  return GST_OK;
}

gst_status gst_poll(gst_server* s, uint64_t ticket, gst_result* out) {
  if (!s || !out) return fail(GST_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->done.find(ticket);
  if (it != s->done.end()) {
    *out = it->second;
    s->done.erase(it);
    s->known.erase(ticket);
    return GST_OK;
  }
  if (s->known.count(ticket)) {
    if (s->error != GST_OK) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
    return GST_NOT_READY;
  }
  return fail(GST_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
}

gst_status gst_wait(gst_server* s, uint64_t ticket, gst_result* out) {
  if (!s || !out) return fail(GST_INVALID_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->known.count(ticket)) return fail(GST_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
  s->waiters++;
  s->cv_done.wait(lk, [&] { return s->done.count(ticket) || s->error != GST_OK; });
  s->waiters--;
  if (s->waiters == 0) s->cv_done.notify_all();  // gst_server_destroy waits for the last waiter
  auto it = s->done.find(ticket);
  if (it == s->done.end()) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
  *out = it->second;
  s->done.erase(it);
  s->known.erase(ticket);
  return GST_OK;
}

gst_status gst_cancel(gst_server* s, uint64_t ticket) {
  if (!s) return fail(GST_INVALID_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto kn = s->known.find(ticket);
  if (kn == s->known.end()) return fail(GST_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
  if (kn->second == 0) {
    // still queued: answered at once
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
      if (it->first == ticket) {
        s->pending.erase(it);
        break;
      }
    s->done[ticket] = cancelled_result();
    kn->second = 2;
    s->cv_done.notify_all();
  } else if (kn->second == 1) {
    // on the device: the plan kernel stops it at its next step
    std::lock_guard<std::mutex> dk(s->g->dev_mu);
    s->eng->cancel_tag(ticket);
  }
  return GST_OK;
}

gst_status gst_server_destroy(gst_server* s) {
  if (!s) return GST_OK;
  {
    // queued requests are answered GST_CANCELLED, running ones are cancelled
    // on the device, and every waiter wakes before the server goes away
    std::unique_lock<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& pr : s->pending) {
      s->done[pr.first] = cancelled_result();
      s->known[pr.first] = 2;
    }
    s->pending.clear();
    std::lock_guard<std::mutex> dk(s->g->dev_mu);
    s->eng->cancel_all();
    s->cv_done.notify_all();
  }
  s->cv_new.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  for (auto& kv : s->done) gst_result_free(&kv.second);
  {
    std::lock_guard<std::mutex> dk(s->g->dev_mu);
    s->eng.reset();
  }
  delete s;
  return GST_OK;
}

}  // extern "C"
