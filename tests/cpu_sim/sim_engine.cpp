// CPU driver for the engine's device code (test infrastructure): runs the
// plan / step cycle of one query with host threads standing in for waves.
#define GST_CPU_SIM 1
#include "gst_device.hip"

#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

namespace gsim {
thread_local int lane = 0;
thread_local Wave* wave = nullptr;
thread_local Block* block = nullptr;
thread_local ThreadIdx thread_idx = {0};
thread_local uint64_t collectives = 0;
}  // namespace gsim

namespace {

int g_fill = 0;  // byte pattern of freshly allocated device-like memory

template <class T>
T* zalloc(size_t n) {
  void* p = nullptr;
  if (posix_memalign(&p, 256, std::max<size_t>(n, 1) * sizeof(T))) return nullptr;
  memset(p, g_fill, std::max<size_t>(n, 1) * sizeof(T));
  return (T*)p;
}

void run_plan(gst::Slot* slots, gst::Ctrl* ctrl, uint64_t V) {
  if (ctrl->exit_flag) return;
  static uint32_t sh_items[gst::kMaxSlots];
  static uint32_t sh_scan[2 * gst::kBlock];
  gsim::Block blk;
  std::vector<gsim::Wave> waves(gst::kWaves);
  std::vector<std::thread> th;
  for (int t = 0; t < gst::kBlock; t++)
    th.emplace_back([&, t]() {
      gsim::lane = t & 63;
      gsim::wave = &waves[t >> 6];
      gsim::block = &blk;
      gsim::thread_idx.x = (unsigned)t;
      gst::phase_b(slots, ctrl, V, sh_items, sh_scan);
    });
  for (auto& x : th) x.join();
}

void run_step(const gst::GraphDev& g, gst::Slot* slots, gst::Ctrl* ctrl, int n_waves, gst::WaveLDS* lds) {
  std::vector<gsim::Wave> waves(n_waves);
  std::vector<std::thread> th;
  for (int w = 0; w < n_waves; w++)
    for (int l = 0; l < 64; l++)
      th.emplace_back([&, w, l]() {
        gsim::lane = l;
        gsim::wave = &waves[w];
        gsim::thread_idx.x = (unsigned)((w % gst::kWaves) * 64 + l);
        gst::step_body(g, slots, ctrl, &lds[w]);
      });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" int gst_sim_solve(const int64_t* rowptr, int64_t V, const int32_t* col, const int32_t* wt,
                             int64_t n_arcs, const int32_t* rev, const int64_t* gptr, const int32_t* gmem,
                             const int32_t* qgroups, int32_t k, int32_t D, int32_t edge_cap, int32_t chunk,
                             int32_t n_waves, int32_t max_steps, int32_t verbose, int32_t* out /* cost, root, n_edges, steps */,
                             int32_t* edges /* edge_cap * 4 */, int64_t* counters /* 6 */) {
  const size_t W = size_t(1) << k, ND = D < 0 ? 1 : (size_t)D + 1;
  g_fill = 0;
  gst::Slot* slot = zalloc<gst::Slot>(1);
  gst::Ctrl* ctrl = zalloc<gst::Ctrl>(1);
  g_fill = (verbose & 2) ? 0xA5 : 0;  // the per-query arrays start as garbage when asked
  slot->k = k;
  slot->W = (int)W;
  slot->ND = (int)ND;
  slot->D = D;
  for (int i = 0; i < k; i++) slot->qg[i] = qgroups[i];
  slot->edge_cap = edge_cap;
  slot->chunk = chunk;
  slot->nT = V * ND * W;
  slot->T = zalloc<uint64_t>(V * ND * W);
  slot->S1 = zalloc<uint64_t>(V * ND * k);
  slot->live = zalloc<uint32_t>(V);
  slot->lvf = zalloc<uint8_t>(((V + 7) / 8) * 8);
  slot->mask = zalloc<uint64_t>(3 * V);
  slot->C = zalloc<uint32_t>(((V * k + 1) / 2) * 2);
  slot->list = zalloc<uint32_t>(3 * V);
  slot->loff = zalloc<uint32_t>(3 * V);
  slot->R = zalloc<uint32_t>(W * k * k);
  slot->edges = zalloc<gst::EdgeDev>(edge_cap);
  slot->rq = zalloc<uint32_t>(3 * ((size_t)edge_cap + 2 * gst::kMaxK));
  // garbage in the never-initialised T segments, as device memory would hold
  for (size_t i = 0; i < V * ND * W; i++) slot->T[i] = 0x5A5A5A5A5A5A5A5Aull;
  slot->phase = gst::PH_FILL;
  slot->best = gst::kInf;
  slot->ans = ~0ull;
  slot->out_cost = (int32_t)gst::kInf;
  slot->out_root = -1;
  ctrl->exit_after = 1;
  ctrl->max_steps = 0xFFFFFFFFu;
  ctrl->nslots = 1;
  gst::GraphDev g;
  g.rowptr = rowptr;
  g.col = col;
  g.wt = wt;
  g.rev = rev;
  g.gptr = gptr;
  g.gmem = gmem;
  g.V = V;
  g.n_arcs = n_arcs;
  std::vector<gst::WaveLDS> lds(n_waves);
  run_plan(slot, ctrl, (uint64_t)V);
  int steps = 0;
  while (!ctrl->exit_flag && steps < max_steps) {
    if (verbose)
      fprintf(stderr, "step %d: phase %d level %d items %u (work %u) best %u nlive %u\n", steps, slot->phase,
              slot->level, ctrl->item_total, slot->n_work, slot->best, slot->nlive);
    run_step(g, slot, ctrl, n_waves, lds.data());
    run_plan(slot, ctrl, (uint64_t)V);
    steps++;
  }
  out[0] = slot->out_cost;
  out[1] = slot->out_root;
  out[2] = slot->out_nedges;
  out[3] = steps;
  const int ne = std::min(slot->out_nedges, edge_cap);
  for (int i = 0; i < ne; i++) {
    edges[4 * i + 0] = slot->edges[i].u;
    edges[4 * i + 1] = slot->edges[i].v;
    edges[4 * i + 2] = slot->edges[i].w;
    edges[4 * i + 3] = slot->edges[i].parent;
  }
  for (int i = 0; i < 6; i++) counters[i] = (int64_t)slot->ctr[i];
  const int rc = ctrl->exit_flag ? (slot->status ? 2 : 0) : 1;
  free(slot->T); free(slot->S1); free(slot->live); free(slot->lvf); free(slot->mask); free(slot->C);
  free(slot->list); free(slot->loff); free(slot->R); free(slot->edges); free(slot->rq);
  free(slot);
  free(ctrl);
  return rc;
}
