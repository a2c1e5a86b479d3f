// Kernel launches of the CPU build of libgst: the engine's step and plan
// kernels run as host threads, one per lane (test infrastructure).  A step
// kernel's waves only claim items, so the launcher starts no more waves than
// the class has items (none when it has none): the surplus waves of a GPU
// launch would claim nothing and exit.
#define GST_CPU_SIM 1
#define GST_CPU_SIM_HOST 1
#include "gst_device.hip"

#include <functional>
#include <thread>
#include <vector>

namespace gsim {
thread_local int lane = 0;
thread_local Wave* wave = nullptr;
thread_local Block* block = nullptr;
thread_local ThreadIdx thread_idx = {0};
thread_local ThreadIdx block_idx = {0};
thread_local ThreadIdx grid_dim = {1};
thread_local uint64_t collectives = 0;
}  // namespace gsim

// stream capture (mock runtime): record the launch instead of running it
bool gsim_capturing();
void gsim_record(std::function<void()> f);

extern "C" hipError_t gst_launch_plan(gst::Slot* slots, gst::Ctrl* ctrl, uint64_t V, hipStream_t s) {
  if (gsim_capturing()) {
    gsim_record([=]() { gst_launch_plan(slots, ctrl, V, s); });
    return hipSuccess;
  }
  if (ctrl->exit_flag) return hipSuccess;
  static uint32_t sh_items[gst::kNumCls][gst::kMaxSlots];
  static uint32_t sh_scan[gst::kNumCls][2][gst::kBlock];
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
  return hipSuccess;
}

template <int CLS, int WPB, class LDS>
static void run_class(const gst::GraphDev* g, gst::Slot* slots, gst::Ctrl* ctrl, int grid) {
  if (ctrl->exit_flag || ctrl->item_total[CLS] == 0) return;
  int n_waves = grid * WPB;
  if ((uint32_t)n_waves > ctrl->item_total[CLS]) n_waves = (int)ctrl->item_total[CLS];
  std::vector<gsim::Wave> waves(n_waves);
  std::vector<LDS> lds(n_waves);
  std::vector<std::thread> th;
  const gst::GraphDev gd = *g;
  for (int w = 0; w < n_waves; w++)
    for (int l = 0; l < 64; l++)
      th.emplace_back([&, w, l]() {
        gsim::lane = l;
        gsim::wave = &waves[w];
        gsim::thread_idx.x = (unsigned)((w % WPB) * 64 + l);
        gsim::block_idx.x = (unsigned)(w / WPB);
        gsim::grid_dim.x = (unsigned)((n_waves + WPB - 1) / WPB);
        gst::step_body<CLS, WPB>(gd, slots, ctrl, &lds[w]);
      });
  for (auto& x : th) x.join();
}

template <int WPB>
static void run_step(int cls, const gst::GraphDev* g, gst::Slot* slots, gst::Ctrl* ctrl, int grid) {
  if (cls == gst::kClsLean)
    run_class<gst::kClsLean, WPB, gst::FrontLDS>(g, slots, ctrl, grid);
  else
    run_class<gst::kClsHeavy, WPB, gst::WaveLDS>(g, slots, ctrl, grid);
}

extern "C" hipError_t gst_launch_step(int cls, int wpb, const gst::GraphDev* g, gst::Slot* slots, gst::Ctrl* ctrl,
                                      int grid, hipStream_t s) {
  if (gsim_capturing()) {
    const gst::GraphDev gd = *g;  // kernel arguments are copied at capture
    gsim_record([=]() { gst_launch_step(cls, wpb, &gd, slots, ctrl, grid, s); });
    return hipSuccess;
  }
  switch (wpb) {
    case 2: run_step<2>(cls, g, slots, ctrl, grid); break;
    case 4: run_step<4>(cls, g, slots, ctrl, grid); break;
    case 8: run_step<8>(cls, g, slots, ctrl, grid); break;
    default: return hipErrorInvalidValue;
  }
  return hipSuccess;
}

extern "C" hipError_t gst_launch_gather(const gst::GatherEnt* t, int n, gst::EdgeDev* dst, hipStream_t) {
  for (int b = 0; b < n; b++)
    for (uint32_t i = 0; i < t[b].n; i++) dst[t[b].off + i] = t[b].src[i];
  return hipSuccess;
}

extern "C" hipError_t gst_launch_count(gst::Slot* slot, uint64_t V, hipStream_t) {
  const uint32_t tiles = (uint32_t)((V + gst::kDistTV - 1) / gst::kDistTV);
  for (uint32_t t = 0; t < tiles; t++) {
    gsim::Wave w;
    std::vector<std::thread> th;
    for (int l = 0; l < 64; l++)
      th.emplace_back([&, l]() {
        gsim::lane = l;
        gsim::wave = &w;
        gst::item_count(slot, t, V);
      });
    for (auto& x : th) x.join();
  }
  return hipSuccess;
}

// Backbone construction: the fill and seed as host loops, the hub rounds and
// the arc test as emulated waves taking items in a grid-stride loop.
extern "C" hipError_t gst_bb_fill(uint32_t* p, uint64_t n, uint32_t val, hipStream_t) {
  for (uint64_t i = 0; i < n; i++) p[i] = val;
  return hipSuccess;
}

extern "C" hipError_t gst_bb_seed(uint32_t* dist, uint8_t* flags, const int32_t* hubs, int H, hipStream_t) {
  if (H < 1 || H > gst::kBbHubs) return hipErrorInvalidValue;
  for (int t = 0; t < H; t++) {
    dist[(uint64_t)hubs[t] * gst::kBbHubs + t] = 0u;
    flags[hubs[t]] = 1;
  }
  return hipSuccess;
}

extern "C" hipError_t gst_bb_pass(int what, const gst::BbArgs* a, hipStream_t) {
  if (what != 0 && what != 1) return hipErrorInvalidValue;
  const uint64_t items = (a->n_arcs + gst::kBbArcs - 1) / gst::kBbArcs;
  if (items == 0) return hipSuccess;
  const int nw = (int)(items < 4 ? items : 4);
  std::vector<gsim::Wave> waves(nw);
  std::vector<std::thread> th;
  const gst::BbArgs args = *a;
  for (int w = 0; w < nw; w++)
    for (int l = 0; l < 64; l++)
      th.emplace_back([&, w, l]() {
        gsim::lane = l;
        gsim::wave = &waves[w];
        for (uint64_t it = (uint64_t)w; it < items; it += (uint64_t)nw) {
          if (what == 0) gst::bb_round_item(args, it);
          else gst::bb_test_item(args, it);
        }
      });
  for (auto& x : th) x.join();
  return hipSuccess;
}

extern "C" hipError_t gst_step_occupancy(int, int wpb, int* blocks_per_cu) {
  *blocks_per_cu = wpb >= 4 ? 1 : 4 / wpb;  // at least four emulated waves per launch
  return hipSuccess;
}
