// gst_device.hip -- the DPBF engine kernels for gfx950 (CDNA4, wave64).
//
// Many queries ("slots") advance together.  One engine step is one launch of
// each work-class kernel (gst_lean_kernel: fill, seed and the group-distance
// jobs' items; gst_heavy_kernel: the per-query DP items): waves pull the
// step's items (a prefix over slots)
// with one atomic each.  Between steps the one-block gst_plan_kernel advances
// every slot's state machine and publishes the next step's items.  Blocks never
// wait for each other; every hand-off between blocks happens at a kernel
// boundary.
//
// Per query the DP is evaluated level by level in |p| (Dreyfus-Wagner order):
//   level 1  : group->vertex distances, either as SEED + frontier rounds of the
//              query itself (hop-bounded DP: T(v,{i},d) per hop depth d; the
//              reference's dis0_init_1 SSSP, code/GPUGST/src/DPQ.cu:241-271)
//              or, for unbounded queries, from a group-distance job: one
//              multi-label SSSP shared by a batch of queries, whose distance
//              rows the queries then read as their C; the job then builds
//              every member's vv and R, and one pass over its rows gives each
//              member its star bound and live vertices (JLIVE)
//   DIST     : compact distance table C[v][i], inter-group distances vv,
//              route table R (host loops of code/GPUGST/src/DPQ.cu:409-483
//              moved on the device) -- queries with their own level 1
//   level L  : MERGE pass (every vertex segment staged in LDS; all sub-mask
//              splits of every |p| = L subset, hop pairs d1 + d2 <= D, new
//              depth max(d1,d2) -- code/D-GPUGST/src/DPQ.cu:100-124) followed by
//              frontier rounds of grow (code/GPUGST/src/DPQ.cu:113-136,
//              code/D-GPUGST/src/DPQ.cu:80-98) with lower-bound pruning
//   FINAL    : full-set merge, argmin over (v, d) (DPQ.cu:558-575)
//   RECON    : predecessor chase into tree edges (the reference's commented
//              out reconstruction, DPQ.cu:577-606, and CPUHOP.h:234-259)
// Merges only read levels < L, which are final, so every state is exact when
// its level ends; grow within a level is a label-correcting closure.
//
// Pruning: a candidate for T(u, p, d) is dropped when cand + lb(u, ~p) > best,
// best being the cost of a real walk-tree (star or complement merge) and lb a
// proven lower bound on the cheapest tree rooted at u covering ~p (one-label
// bound and tour bound, GPUGST+ one_label_lb_1, code/GPUGST+/src/
// GSTnonHop.cu:88-138).  Using '>' keeps every state of an optimal witness.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "gst_common.h"

namespace gst {

// per-wave LDS of the arc-parallel items: 64 staged frontier entries
struct FrontLDS {
  uint32_t wv[64];
  uint32_t woff[64];
  uint32_t wrs[64];
  uint32_t pad[64];
  uint64_t wm[64];
  uint32_t av[128];  // vertices appended by this wave, not yet published
};

// per-wave LDS of the vertex-segment items
struct WaveLDS {
  FrontLDS f;
  uint64_t stage[kStageWords];
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x) {
  const int lane = lane_id();
  uint32_t incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  return incl - x;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t y = (uint64_t)__shfl_xor((unsigned long long)x, o, 64);
    x = y < x ? y : x;
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t y = __shfl_xor(x, o, 64);
    x = y < x ? y : x;
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += (uint64_t)__shfl_xor((unsigned long long)x, o, 64);
  return x;
}

__device__ __forceinline__ uint32_t sat(uint32_t a) { return a < kInf ? a : kInf; }

// Device tables are reached through pointers held in the slot table, which
// the compiler cannot place in an address space.  Cast to the global one,
// their loads and atomics are global_* instructions instead of flat ones
// (whose waits also cover LDS traffic).
#ifdef GST_CPU_SIM
#define GST_AS_GLOBAL
#define GST_AS_CONST
#else
#define GST_AS_GLOBAL __attribute__((address_space(1)))
// memory no kernel writes while it runs: wave-uniform loads become scalar ones
#define GST_AS_CONST __attribute__((address_space(4)))
#endif
// a job chunk of CL labels, and the 16-byte granule of copies and fills
template <int CL>
struct JChunk;
template <>
struct JChunk<4> {
  typedef uint32_t type __attribute__((vector_size(16)));
};
template <>
struct JChunk<16> {
  typedef uint32_t type __attribute__((vector_size(64)));
};
typedef uint32_t bchunk __attribute__((vector_size(16)));  // four labels (hub or job)
typedef GST_AS_GLOBAL const bchunk g_cgran;
typedef GST_AS_GLOBAL bchunk g_gran;
typedef GST_AS_GLOBAL uint32_t g_u32;
typedef GST_AS_GLOBAL const uint32_t g_cu32;
typedef GST_AS_GLOBAL unsigned long long g_u64;

// A store that is not read again soon (non-temporal: kept out of the caches).
__device__ __forceinline__ void store_stream(uint64_t* p, uint64_t v) {
#ifdef GST_CPU_SIM
  *p = v;
#else
  __builtin_nontemporal_store(v, (GST_AS_GLOBAL uint64_t*)p);
#endif
}
__device__ __forceinline__ void store_stream(g_u32* p, uint32_t v) {
#ifdef GST_CPU_SIM
  *p = v;
#else
  __builtin_nontemporal_store(v, p);
#endif
}
__device__ __forceinline__ void store_stream(g_gran* p, bchunk v) {
#ifdef GST_CPU_SIM
  *p = v;
#else
  __builtin_nontemporal_store(v, p);
#endif
}

// State (v, p, d): singletons in the compact S1 table (heights 0..D) or,
// when the slot's distances come from a group-distance job, synthesised from
// the job's rows with the kPredPath predecessor; every other subset in the
// vertex-major T segments (heights 0..NS-1).
struct States {
  g_u64* T;
  g_u64* S1;
  g_cu32* J;
  uint32_t jvs, jds;
  int k, ND, NS, W;
  // address of a state that is written (never a singleton of a C-fed slot)
  __device__ __forceinline__ g_u64* at(uint32_t v, uint32_t p, uint32_t d) const {
    if ((p & (p - 1u)) == 0u) return S1 + ((uint64_t)v * ND + d) * (uint64_t)k + (uint32_t)(__ffs(p) - 1);
    return T + ((uint64_t)v * NS + d) * (uint64_t)W + p;
  }
  __device__ __forceinline__ uint64_t get(uint32_t v, uint32_t p, uint32_t d) const {
    if ((p & (p - 1u)) == 0u) {
      const uint32_t i = (uint32_t)(__ffs(p) - 1);
      if (S1) return S1[((uint64_t)v * ND + d) * (uint64_t)k + i];
      return ((uint64_t)jcost(v, i, d) << 32) | kPredPath;
    }
    return T[((uint64_t)v * NS + d) * (uint64_t)W + p];
  }
  // cost of the singleton T(v, {i}, d) of a job-fed query
  __device__ __forceinline__ uint32_t jcost(uint64_t v, uint32_t i, uint32_t d) const {
    return J[v * jvs + (uint64_t)d * jds + i];
  }
};

__device__ __forceinline__ States states_of(const Slot* sp) {
  States st;
  st.T = (g_u64*)sp->T;
  st.S1 = (g_u64*)sp->S1;
  st.J = (g_cu32*)sp->J;
  st.jvs = sp->jvs;
  st.jds = sp->jds;
  st.k = sp->k;
  st.ND = sp->ND;
  st.NS = sp->NS;
  st.W = sp->W;
  return st;
}

// Arcs of a slot: the graph's shortest-path backbone (Slot::bb) for
// unbounded queries and jobs, every arc otherwise.  Only the row starts
// differ: the backbone's arcs follow the graph's in col / wt / rev.
struct Arcs {
  const int64_t* rowptr;
  const int32_t* col;
  const int32_t* wt;
  const int32_t* rev;
};

__device__ __forceinline__ Arcs arcs_of(const GraphDev& g, const Slot* sp) {
  Arcs a;
  a.rowptr = g.rowptr + (sp->bb ? (uint64_t)g.V + 1u : 0u);
  a.col = g.col;
  a.wt = g.wt;
  a.rev = g.rev;
  return a;
}

// Row of vertex v in the distance table C.  A hop query's table holds D + 1
// cumulative rows per vertex, row h the group distances within h hops (its own
// table, or its job's cumulative rows); an unbounded query has
// one row (the job's row for a job-fed query).
__device__ __forceinline__ g_cu32* crow(const Slot* sp, uint64_t v, int h) {
  return (g_cu32*)sp->C + v * sp->cstride + (sp->D >= 0 ? (uint64_t)h * (uint64_t)sp->chs : 0);
}

// Lower bound on the cost of the cheapest tree containing u that covers the
// groups in x (x non-empty): max( max_{i in x} c_i(u),
//   ceil( min_{i != j in x} (c_i(u) + R[x][i][j] + c_j(u)) / 2 ) ).
// The second term halves a closed walk from u through every group of x; R is
// the cheapest group-to-group route visiting all of x from i to j.
// lower_bound_c takes the vertex's group distances c[0..k) already read,
// lower_bound reads them from its C row.
__device__ __forceinline__ uint32_t lower_bound_c(const uint32_t* c, int k, uint32_t x, g_cu32* __restrict__ R) {
  uint32_t lb1 = 0;
#pragma unroll
  for (int i = 0; i < kMaxK; i++)
    if (i < k && ((x >> i) & 1u)) lb1 = c[i] > lb1 ? c[i] : lb1;
  if (lb1 >= kInf) return kInf;
  if (__popc(x) >= 2) {
    g_cu32* Rx = R + (size_t)x * (uint32_t)(k * k);
    uint32_t tour = 3u * kInf;
#pragma unroll
    for (int i = 0; i < kMaxK; i++) {
      if (i >= k || !((x >> i) & 1u)) continue;
#pragma unroll
      for (int j = 0; j < kMaxK; j++) {
        if (j >= k || j == i || !((x >> j) & 1u)) continue;
        uint32_t t = c[i] + Rx[i * k + j] + c[j];
        tour = t < tour ? t : tour;
      }
    }
    uint32_t lb2 = (tour + 1u) >> 1;
    lb1 = lb2 > lb1 ? lb2 : lb1;
  }
  return sat(lb1);
}

__device__ __forceinline__ uint32_t lower_bound(g_cu32* __restrict__ Crow, int k, uint32_t x,
                                                g_cu32* __restrict__ R) {
  uint32_t c[kMaxK];
#pragma unroll
  for (int i = 0; i < kMaxK; i++) c[i] = (i < k) ? Crow[i] : kInf;
  return lower_bound_c(c, k, x, R);
}

// lower_bound(Crow, k, x, R) <= lim, decided with early exits: false as soon
// as one group of x is farther than lim, true at the first pair (i, j) whose
// closed walk c_i + R[x][i][j] + c_j is at most 2 lim (ceil(t / 2) <= lim
// exactly when t <= 2 lim).  Used where only the comparison matters.
__device__ __forceinline__ bool lb_le(g_cu32* __restrict__ Crow, int k, uint32_t x, g_cu32* __restrict__ R,
                                      uint32_t lim) {
  if (lim >= kInf) return true;  // lower_bound saturates at kInf
  uint32_t c[kMaxK];
#pragma unroll
  for (int i = 0; i < kMaxK; i++) {
    c[i] = (i < k) ? Crow[i] : kInf;
    if (i < k && ((x >> i) & 1u) && c[i] > lim) return false;
  }
  if (__popc(x) < 2) return true;
  g_cu32* Rx = R + (size_t)x * (uint32_t)(k * k);
  const uint32_t lim2 = 2u * lim;
#pragma unroll
  for (int i = 0; i < kMaxK; i++) {
    if (i >= k || !((x >> i) & 1u)) continue;
#pragma unroll
    for (int j = 0; j < kMaxK; j++) {
      if (j >= k || j == i || !((x >> j) & 1u)) continue;
      if (c[i] + Rx[i * k + j] + c[j] <= lim2) return true;
    }
  }
  return false;
}

// Wave-aggregated append of vertex u (when flag) to frontier buffer `buf`:
// one 64-bit atomicAdd of (count << 32 | degree sum) per wave gives every
// appended vertex both its list position and its arc offset.  Each ROUND item
// of the next round covers `chunk` arcs of that arc space; the vertex that
// holds an item's first arc records its list position in istart, so the item
// starts without a search.
__device__ __forceinline__ void wave_append(bool flag, uint32_t u, Slot* sp, int buf,
                                            const int64_t* rowptr, uint64_t V) {
  const uint64_t ball = __ballot(flag);
  if (ball == 0) return;
  const int lane = lane_id();
  uint32_t deg = 0;
  if (flag) deg = (uint32_t)(rowptr[u + 1] - rowptr[u]);
  const uint32_t excl = wave_excl_scan(deg);
  const uint32_t total = __shfl(excl + deg, 63, 64);
  const int leader = __ffsll((long long)ball) - 1;
  unsigned long long old = 0;
  if (lane == leader) {
    old = atomicAdd((unsigned long long*)&sp->app[buf],
                    ((unsigned long long)__popcll(ball) << 32) | (unsigned long long)total);
  }
  old = __shfl(old, leader, 64);
  if (flag) {
    const uint32_t rank = __popcll(ball & __lanemask_lt());
    const uint32_t pos = (uint32_t)(old >> 32) + rank;
    const uint32_t off = (uint32_t)old + excl;
    sp->list[(uint64_t)buf * V + pos] = u;
    sp->loff[(uint64_t)buf * V + pos] = off;
    if (deg) {
      const uint32_t ch = (uint32_t)sp->chunk;
      uint32_t* ist = sp->istart + (uint64_t)buf * sp->istride;
      for (uint32_t c = (off + ch - 1) / ch; c * ch < off + deg; c++) ist[c] = pos;
    }
  }
}

// Staged append for the ROUND items: appended vertices collect in the wave's
// LDS and are published in batches of 64 or more with one atomicAdd, so the
// waves of a dense round do not queue on the frontier counter once per
// iteration.  flush_append must run before the item ends.
__device__ void flush_append(FrontLDS* L, uint32_t& apn, Slot* sp, int buf, const int64_t* rowptr,
                             uint64_t V) {
  if (apn == 0) return;
  wave_sync();
  const int lane = lane_id();
  uint32_t u0 = 0, u1 = 0, d0 = 0, d1 = 0;
  if ((uint32_t)lane < apn) {
    u0 = L->av[lane];
    d0 = (uint32_t)(rowptr[u0 + 1] - rowptr[u0]);
  }
  if ((uint32_t)lane + 64 < apn) {
    u1 = L->av[lane + 64];
    d1 = (uint32_t)(rowptr[u1 + 1] - rowptr[u1]);
  }
  const uint32_t e0 = wave_excl_scan(d0);
  const uint32_t t0 = __shfl(e0 + d0, 63, 64);
  const uint32_t e1 = wave_excl_scan(d1);
  const uint32_t t1 = __shfl(e1 + d1, 63, 64);
  // the publishing lane comes from a ballot, so no other lane-0 test of the
  // caller can be merged with this one
  const int leader = __ffsll((long long)__ballot(1)) - 1;
  unsigned long long old = 0;
  if (lane == leader)
    old = atomicAdd((unsigned long long*)&sp->app[buf], ((unsigned long long)apn << 32) | (t0 + t1));
  old = __shfl(old, leader, 64);
  const uint32_t pos0 = (uint32_t)(old >> 32), off0 = (uint32_t)old;
  const uint32_t ch = (uint32_t)sp->chunk;
  uint32_t* ist = sp->istart + (uint64_t)buf * sp->istride;
  for (int h = 0; h < 2; h++) {
    const uint32_t idx = (uint32_t)lane + 64u * (uint32_t)h;
    if (idx >= apn) continue;
    const uint32_t u = h ? u1 : u0, deg = h ? d1 : d0;
    const uint32_t pos = pos0 + idx;
    const uint32_t off = off0 + (h ? t0 + e1 : e0);
    sp->list[(uint64_t)buf * V + pos] = u;
    sp->loff[(uint64_t)buf * V + pos] = off;
    for (uint32_t c = (off + ch - 1) / ch; c * ch < off + deg; c++) ist[c] = pos;
  }
  apn = 0;
  wave_sync();
}

__device__ __forceinline__ void stage_append(bool flag, uint32_t u, FrontLDS* L, uint32_t& apn, Slot* sp,
                                             int buf, const int64_t* rowptr, uint64_t V) {
  const uint64_t ball = __ballot(flag);
  if (ball == 0) return;
  if (flag) L->av[apn + __popcll(ball & __lanemask_lt())] = u;
  apn += (uint32_t)__popcll(ball);
  if (apn >= 64) flush_append(L, apn, sp, buf, rowptr, V);
}

// Sets the bits `imp` of the next round's dirty mask of u; true when u had
// none (u then joins the next frontier).  A plain read first skips the atomic
// when the bits are already set, as they soon are on a hub that many arcs
// improve in one round.
__device__ __forceinline__ bool mark_dirty(uint64_t* mnxt, uint32_t u, uint64_t imp) {
  const uint64_t seen = __hip_atomic_load(&mnxt[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if ((seen & imp) == imp) return false;
  const unsigned long long om = atomicOr((unsigned long long*)&mnxt[u], (unsigned long long)imp);
  return om == 0;
}

// Sets v's hold bit (flag: v received a finite state with |p| >= 2) and
// appends v to the active list when the bit is new.  Wave-collective: every
// lane calls it, with flag false where it has nothing.  A merge of a subset of
// three or more groups at v needs a part of two or more groups at v, so only
// vertices in the active list can hold such a merge.
__device__ __forceinline__ void mark_hold(bool flag, uint32_t v, Slot* sp) {
  bool fresh = false;
  if (flag) {
    uint32_t* w = sp->hold + (v >> 5);
    const uint32_t bit = 1u << (v & 31u);
    if (!(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit))
      fresh = (atomicOr(w, bit) & bit) == 0u;
  }
  const uint64_t ball = __ballot(fresh);
  if (ball == 0) return;
  const int leader = __ffsll((long long)ball) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(&sp->nact, (uint32_t)__popcll(ball));
  base = __shfl(base, leader, 64);
  if (fresh) sp->act[base + __popcll(ball & __lanemask_lt())] = v;
}

// ---------------------------------------------------------------- items ----

// u64 words that FILL initialises: job [distances][masks]; query [S1][masks]
// [live bits][hold bits] (S1 only when the query computes its own level 1;
// its C table needs no fill, DIST writes every entry)
__device__ __forceinline__ void fill_regions(const Slot* sp, uint64_t V, uint64_t* nA, uint64_t* nM,
                                             uint64_t* nF) {
  if (sp->kind == KIND_JOB) {
    // height 0 of every row (COPY writes the others); rows of hub bounds are
    // written by their own items
    *nA = sp->hA ? 0 : V * (uint64_t)sp->Lp / 2;
    *nM = 3 * V;
    *nF = 0;
    return;
  }
  const bool own = sp->S1 != nullptr;
  *nA = own ? V * (uint64_t)sp->ND * (uint64_t)sp->k : 0;
  // A query fed by a job starts at level 2, where only live vertices are ever
  // marked dirty (merges run over live and active vertices, grows skip
  // targets that are not live): its LIVE clears the masks of its live
  // vertices instead.
  *nM = sp->job >= 0 ? 0 : 3 * V;
  *nF = 2 * ((V + 63) / 64);  // live bits, then hold bits
}

// Job FILL of an unbounded job with hub bounds (Slot::hA): every label l of
// row v starts at min_h hA[l][h] + d(h, v), the length of a real walk from a
// member of l's group through hub h.  These values obey the triangle
// inequality along every arc (d(h, .) are exact distances), so no arc improves
// one of them: the rounds seeded at the members only lower the labels whose
// bound is not their distance, and the result is the same exact distances.
// An item takes 64 labels (one per lane, its kBbHubs bounds in registers) and
// kUbTV of the vertices some hub reaches, one at a time: the vertex's hub
// distances are the same for the whole wave and arrive as scalar loads.
constexpr uint32_t kUbTV = 256;

__device__ __forceinline__ uint32_t ub_groups(const Slot* sp) { return ((uint32_t)sp->Lp + 63u) / 64u; }

__device__ void item_ubfill(const GraphDev& g, Slot* sp, uint32_t tile) {
  const int lane = lane_id();
  const uint32_t ngrp = ub_groups(sp);
  const uint32_t l = (tile % ngrp) * 64u + (uint32_t)lane;
  const uint64_t i0 = (uint64_t)(tile / ngrp) * kUbTV;
  const uint64_t i1 = i0 + kUbTV < (uint64_t)g.n_hubr ? i0 + kUbTV : (uint64_t)g.n_hubr;
  const uint32_t Lp = (uint32_t)sp->Lp;
  const bool lv = l < Lp;
  uint32_t a[kBbHubs];
  const bchunk* arow = (const bchunk*)(sp->hA + (uint64_t)(lv ? l : 0u) * kBbHubs);
#pragma unroll
  for (int j = 0; j < kBbHubs / 4; j++) {
    const bchunk x = arow[j];
#pragma unroll
    for (int q = 0; q < 4; q++) a[4 * j + q] = x[q];
  }
  const GST_AS_CONST uint32_t* hubd = (const GST_AS_CONST uint32_t*)g.hubd;
  const GST_AS_CONST uint32_t* hubv = (const GST_AS_CONST uint32_t*)g.hubv;
  g_u32* dist = (g_u32*)sp->dist;
  for (uint64_t i = i0; i < i1; i++) {
    const uint64_t v = (uint64_t)hubv[i];
    const GST_AS_CONST uint32_t* hr = hubd + v * kBbHubs;
    uint32_t u = kInf;
#pragma unroll
    for (int h = 0; h < kBbHubs; h++) {
      const uint32_t t = a[h] + hr[h];  // both <= kInf: no overflow
      u = t < u ? t : u;
    }
    if (lv) store_stream(&dist[v * Lp + l], u);
  }
}

// item_ubfill in packed 16-bit arithmetic (graphs whose hub distances are all
// below kHub16Inf / 2, GraphDev::hubd16): a lane takes labels l and l + 64 of
// a group of 128, one per half of a packed word, so each packed add and min
// serves two labels.  Bounds and distances are clamped to kHub16Inf; two
// finite values add below it and a sum at or above it means "no bound" (the
// row starts at infinity, which is always a valid start).
typedef unsigned short us2 __attribute__((vector_size(4)));

__device__ __forceinline__ us2 us2_min(us2 a, us2 b) {
#ifdef GST_CPU_SIM
  us2 r;
  r[0] = a[0] < b[0] ? a[0] : b[0];
  r[1] = a[1] < b[1] ? a[1] : b[1];
  return r;
#else
  return __builtin_elementwise_min(a, b);
#endif
}

__device__ __forceinline__ uint32_t ub16_groups(const Slot* sp) { return ((uint32_t)sp->Lp + 127u) / 128u; }

// One vertex's packed bounds: the least of a[h] + d(h, v) over the hubs, for
// the lane's two labels (hr: the vertex's kBbHubs / 2 hub-distance pairs).
template <class HR>
__device__ __forceinline__ us2 ub16_vertex(const us2* a, HR hr) {
  us2 u = us2{(unsigned short)kHub16Inf, (unsigned short)kHub16Inf};
#pragma unroll
  for (int h = 0; h < kBbHubs / 2; h++) {
    const uint32_t pr = hr(h);
    const unsigned short d0 = (unsigned short)(pr & 0xFFFFu), d1 = (unsigned short)(pr >> 16);
    u = us2_min(u, a[2 * h] + us2{d0, d0});
    u = us2_min(u, a[2 * h + 1] + us2{d1, d1});
  }
  return u;
}

// A group of 128 labels takes 64 lanes per vertex, one vertex at a time (its
// hub distances arrive as scalar loads).  The last group of a row whose
// length is not a multiple of 128 has rem labels: it takes the smallest P of
// 8, 16, 32, 64 lanes with 2P >= rem per vertex and 64 / P vertices at a
// time, so that a job of, say, 160 labels does not pay a second full group.
__device__ void item_ubfill16(const GraphDev& g, Slot* sp, uint32_t tile) {
  const int lane = lane_id();
  const uint32_t ngrp = ub16_groups(sp);
  const uint32_t Lp = (uint32_t)sp->Lp;
  const uint32_t base = (tile % ngrp) * 128u;
  const uint32_t rem = Lp - base < 128u ? Lp - base : 128u;
  uint32_t sh = 6;  // log2 of the lanes per vertex
  while (sh > 3 && (1u << sh) >= rem) sh--;
  const uint32_t P = 1u << sh;
  const uint32_t sub = (uint32_t)lane >> sh, r = (uint32_t)lane & (P - 1u);
  const uint32_t l0 = base + r, l1 = base + P + r;
  const uint64_t i0 = (uint64_t)(tile / ngrp) * kUbTV;
  const uint64_t i1 = i0 + kUbTV < (uint64_t)g.n_hubr ? i0 + kUbTV : (uint64_t)g.n_hubr;
  const bool lv0 = l0 < Lp, lv1 = l1 < Lp;
  us2 a[kBbHubs];
  const bchunk* r0 = (const bchunk*)(sp->hA + (uint64_t)(lv0 ? l0 : 0u) * kBbHubs);
  const bchunk* r1 = (const bchunk*)(sp->hA + (uint64_t)(lv1 ? l1 : 0u) * kBbHubs);
#pragma unroll
  for (int j = 0; j < kBbHubs / 4; j++) {
    const bchunk x0 = r0[j], x1 = r1[j];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t b0 = lv0 && x0[q] < kHub16Inf ? x0[q] : kHub16Inf;
      const uint32_t b1 = lv1 && x1[q] < kHub16Inf ? x1[q] : kHub16Inf;
      a[4 * j + q] = us2{(unsigned short)b0, (unsigned short)b1};
    }
  }
  g_u32* dist = (g_u32*)sp->dist;
  if (P == 64u) {
    const GST_AS_CONST uint32_t* hubd = (const GST_AS_CONST uint32_t*)g.hubd16;  // hub pairs
    const GST_AS_CONST uint32_t* hubv = (const GST_AS_CONST uint32_t*)g.hubv;
    for (uint64_t i = i0; i < i1; i++) {
      const uint64_t v = (uint64_t)hubv[i];
      const GST_AS_CONST uint32_t* hr = hubd + v * (kBbHubs / 2);
      const us2 u = ub16_vertex(a, [hr](int h) { return hr[h]; });
      if (lv0) store_stream(&dist[v * Lp + l0], u[0] < kHub16Inf ? (uint32_t)u[0] : kInf);
      if (lv1) store_stream(&dist[v * Lp + l1], u[1] < kHub16Inf ? (uint32_t)u[1] : kInf);
    }
  } else {
    const uint32_t vpi = 64u >> sh;
    for (uint64_t i = i0 + sub; i < i1; i += vpi) {
      const uint64_t v = (uint64_t)g.hubv[i];
      g_cgran* hq = (g_cgran*)(g.hubd16 + v * kBbHubs);  // kBbHubs / 8 granules of four pairs
      bchunk q[kBbHubs / 8];
#pragma unroll
      for (int j = 0; j < kBbHubs / 8; j++) q[j] = hq[j];
      const us2 u = ub16_vertex(a, [&q](int h) { return q[h >> 2][h & 3]; });
      if (lv0) store_stream(&dist[v * Lp + l0], u[0] < kHub16Inf ? (uint32_t)u[0] : kInf);
      if (lv1) store_stream(&dist[v * Lp + l1], u[1] < kHub16Inf ? (uint32_t)u[1] : kInf);
    }
  }
}

// The rows of 64 vertices no hub reaches: infinite for every label (one lane
// per chunk, a row at a time).
__device__ void item_ubfill_inf(const GraphDev& g, Slot* sp, uint32_t tile) {
  const uint64_t n_un = (uint64_t)g.V - (uint64_t)g.n_hubr;
  const uint64_t i0 = (uint64_t)tile * 64u;
  const uint64_t i1 = i0 + 64u < n_un ? i0 + 64u : n_un;
  const int ng = sp->Lp / 4;  // 16-byte granules per row
  bchunk inf;
  for (int q = 0; q < 4; q++) inf[q] = kInf;
  for (uint64_t i = i0; i < i1; i++) {
    const uint64_t v = (uint64_t)g.hubv[(uint64_t)g.n_hubr + i];
    g_gran* row = (g_gran*)(sp->dist + v * (uint64_t)sp->Lp);
    for (int c = lane_id(); c < ng; c += 64) store_stream(&row[c], inf);
  }
}

__device__ void item_fill(Slot* sp, uint32_t idx, uint64_t V) {
  uint64_t nA, nM, nF;
  fill_regions(sp, V, &nA, &nM, &nF);
  const uint64_t total = nA + nM + nF;
  const uint64_t base = (uint64_t)idx * kFillWords;
  const uint64_t end = base + kFillWords < total ? base + kFillWords : total;
  const bool job = sp->kind == KIND_JOB;
  uint64_t* A = job ? (uint64_t*)sp->dist : sp->S1;
  uint64_t* M = sp->mask;
  uint64_t* F64 = (uint64_t*)sp->lvb;
  const uint64_t cinf = ((uint64_t)kInf << 32) | kInf;
  const uint64_t aval = job ? cinf : kInfState;
  // a hop job's height-0 rows: Lp / 2 words per vertex, ND * Lp / 2 words apart
  const bool rows = job && sp->ND > 1;
  const uint64_t hw = (uint64_t)sp->Lp / 2, rw = (uint64_t)sp->ND * hw;
  for (uint64_t t = base + lane_id(); t < end; t += 64) {
    if (t < nA)
      store_stream(&A[rows ? (t / hw) * rw + t % hw : t], aval);  // read back row by row, much later
    else if (t < nA + nM)
      M[t - nA] = 0;
    else
      F64[t - nA - nM] = 0;
  }
}

// FILL of an unbounded job with hub bounds (heavy kernel): the rows' items
// first, then the masks'.
__device__ void item_fill_hub(const GraphDev& g, Slot* sp, uint32_t idx, uint64_t V) {
  const uint32_t nr =
      (uint32_t)(((uint64_t)g.n_hubr + kUbTV - 1u) / kUbTV) * (sp->hub16 ? ub16_groups(sp) : ub_groups(sp));
  const uint32_t nu = (uint32_t)((V - (uint64_t)g.n_hubr + 63u) / 64u);
  if (idx < nr && sp->hub16) item_ubfill16(g, sp, idx);
  else if (idx < nr) item_ubfill(g, sp, idx);
  else if (idx < nr + nu) item_ubfill_inf(g, sp, idx - nr);
  else item_fill(sp, idx - nr - nu, V);
}

__device__ void item_seed(const GraphDev& g, Slot* sp, uint32_t i) {
  const Arcs A = arcs_of(g, sp);
  const int ND = sp->ND;
  const int nxt = (sp->step + 1) % 3;
  const uint64_t V = (uint64_t)g.V;
  const int32_t grp = sp->qg[i];
  const int64_t beg = g.gptr[grp], end = g.gptr[grp + 1];
  uint64_t* mn = sp->mask + (uint64_t)nxt * V;
  const int k = sp->k;
  uint32_t vmin = 0xFFFFFFFFu;
  for (int64_t base = beg; base < end; base += 64) {
    const int64_t t = base + lane_id();
    bool app = false;
    uint32_t v = 0;
    if (t < end) {
      v = (uint32_t)g.gmem[t];
      sp->S1[((uint64_t)v * ND) * k + i] = (uint64_t)kPredSeed;  // cost 0, d = 0
      unsigned long long old = atomicOr((unsigned long long*)&mn[v], 1ull << (1u << i));
      app = (old == 0);
      vmin = v < vmin ? v : vmin;
    }
    wave_append(app, v, sp, nxt, A.rowptr, V);
  }
  if (k == 1) {
    // a member alone is an optimal tree; the smallest member id is the root
    vmin = wave_min_u32(vmin);
    if (lane_id() == 0 && vmin != 0xFFFFFFFFu) atomicMin((unsigned long long*)&sp->ans, (unsigned long long)vmin);
  }
}

__device__ void item_clear(Slot* sp, uint32_t idx, uint64_t V) {
  const int buf = (sp->step + 2) % 3;
  const uint32_t n = (uint32_t)(sp->app[buf] >> 32);
  const uint32_t* lst = sp->list + (uint64_t)buf * V;
  uint64_t* m = sp->mask + (uint64_t)buf * V;
  const uint32_t base = idx * kClearPer;
  const uint32_t end = base + kClearPer < n ? base + kClearPer : n;
  for (uint32_t t = base + lane_id(); t < end; t += 64) m[lst[t]] = 0;
}

// Stages the 64 frontier entries from list index i on: vertex, arc offset,
// row start and dirty mask; entries past the end get offset arcs_total.
__device__ __forceinline__ void stage_front(const int64_t* rowptr, FrontLDS* L, uint32_t i, uint32_t n,
                                            uint32_t arcs_total, const uint32_t* lst, const uint32_t* lof,
                                            const uint64_t* mcur) {
  const int lane = lane_id();
  const uint32_t idx = i + lane;
  const bool valid = idx < n;
  const uint32_t v = valid ? lst[idx] : 0u;
  L->wv[lane] = v;
  L->woff[lane] = valid ? lof[idx] : arcs_total;
  L->wrs[lane] = valid ? (uint32_t)rowptr[v] : 0u;
  L->wm[lane] = valid ? mcur[v] : 0ull;
}

// staged entry whose arc range holds arc position pos
__device__ __forceinline__ int find_entry(const FrontLDS* L, uint32_t pos) {
  int o = 0;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1)
    if (L->woff[o + s] <= pos) o += s;
  return o;
}

// One chunk of arcs of the current frontier: grow T(v,p,d) -> T(u,p,d') for
// every arc v->u in the chunk and every dirty subset p of v.
__device__ void item_round(const GraphDev& g, Slot* sp, uint32_t chunk, FrontLDS* L) {
  const Arcs A = arcs_of(g, sp);
  const int lane = lane_id();
  const uint64_t V = (uint64_t)g.V;
  const int k = sp->k, W = sp->W, D = sp->D;
  const bool hop = D >= 0;
  const int level = sp->level;
  // source heights that may grow: singletons up to D (the C rows need every
  // height), states of two or more groups up to NS - 1
  const int dmax = hop ? (level >= 2 ? sp->NS - 1 : D) : 1;
  const uint32_t full = (uint32_t)W - 1;
  const uint32_t r = sp->step;
  const int cur = r % 3, nxt = (r + 1) % 3;
  const uint64_t appc = sp->app[cur];
  const uint32_t n = (uint32_t)(appc >> 32);
  const uint32_t arcs_total = (uint32_t)appc;
  const uint32_t CH = (uint32_t)sp->chunk;
  const uint32_t a0 = chunk * CH;
  const uint32_t a1 = a0 + CH < arcs_total ? a0 + CH : arcs_total;
  const uint32_t* lst = sp->list + (uint64_t)cur * V;
  const uint32_t* lof = sp->loff + (uint64_t)cur * V;
  const uint64_t* mcur = sp->mask + (uint64_t)cur * V;
  uint64_t* mnxt = sp->mask + (uint64_t)nxt * V;
  const States S = states_of(sp);
  g_cu32* C = (g_cu32*)sp->C;
  g_cu32* R = (g_cu32*)sp->R;

  uint64_t n_relax = 0, n_impr = 0;
  uint32_t apn = 0;
  uint32_t a = a0;
  for (uint32_t i = sp->istart[(uint64_t)cur * sp->istride + chunk]; a < a1; i += 64) {
    stage_front(A.rowptr, L, i, n, arcs_total, lst, lof, mcur);
    const uint32_t wend = (i + 64 < n) ? lof[i + 64] : arcs_total;
    wave_sync();
    const uint32_t hia = a1 < wend ? a1 : wend;
    // the arc of the lane's next iteration is read while this one is grown
    int no = 0;
    uint32_t ne = 0, nu = 0, nw = 0;
    if (a + lane < hia) {
      no = find_entry(L, a + lane);
      ne = L->wrs[no] + (a + lane - L->woff[no]);
      nu = (uint32_t)A.col[ne];
      nw = (uint32_t)A.wt[ne];
    }
    for (uint32_t base = a; base < hia; base += 64) {
      const uint32_t pos = base + lane;
      bool app = false, held = false;
      const int o = no;
      const uint32_t e = ne, w = nw;
      uint32_t u = pos < hia ? nu : 0u;
      if (pos + 64 < hia) {
        no = find_entry(L, pos + 64);
        ne = L->wrs[no] + (pos + 64 - L->woff[no]);
        nu = (uint32_t)A.col[ne];
        nw = (uint32_t)A.wt[ne];
      }
      if (pos < hia) {
        const uint32_t vs = L->wv[o];
        const uint64_t m = L->wm[o];
        uint32_t revarc = 0xFFFFFFFFu;
        uint64_t improved = 0;
        // above level 1 only live targets can hold useful states
        uint64_t mm = (level >= 2 && !((sp->lvb[u >> 5] >> (u & 31u)) & 1u)) ? 0ull : m;
        while (mm) {
          const uint32_t p = (uint32_t)(__ffsll((long long)mm) - 1);
          mm &= mm - 1;
          const uint32_t x = full ^ p;
          uint32_t lbx = 0xFFFFFFFFu;
          for (int d = 0; d < dmax; d++) {
            const uint32_t sc = (uint32_t)(S.get(vs, p, d) >> 32);
            if (sc >= kInf) continue;
            const uint32_t cand = sc + w;
            if (cand >= kInf) continue;
            n_relax++;
            const int td = hop ? d + 1 : 0;
            if (level >= 2) {
              // bound first: a target that fails it is never read, so only the
              // segments of live vertices are ever touched.  A hop state of
              // height td reaches the rest of its tree within D - td hops.
              uint32_t lb;
              if (hop) {
                lb = lower_bound(crow(sp, u, D - td), k, x, R);
              } else {
                if (lbx == 0xFFFFFFFFu) lbx = lower_bound(C + (uint64_t)u * sp->cstride, k, x, R);
                lb = lbx;
              }
              const uint32_t bst = __hip_atomic_load(&sp->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              if (cand + lb > bst) continue;
            }
            g_u64* tp = S.at(u, p, td);
            const uint32_t tc = (uint32_t)(*tp >> 32);
            if (cand >= tc) continue;
            if (revarc == 0xFFFFFFFFu) revarc = (uint32_t)A.rev[e];
            const unsigned long long nv = ((unsigned long long)cand << 32) | (kPredGrow | revarc);
            const unsigned long long old = __hip_atomic_fetch_min(tp, nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(old >> 32) > cand) {
              improved |= 1ull << p;
              n_impr++;
              if (__popc(x) <= level) {
                // a complete walk-tree: this state plus the complement state at u
                uint32_t cbest = kInf;
                const int d2max = hop ? ((D - td) < sp->NS - 1 ? (D - td) : sp->NS - 1) : 0;
                for (int d2 = 0; d2 <= d2max; d2++) {
                  const uint32_t c2 = (uint32_t)(S.get(u, x, d2) >> 32);
                  cbest = c2 < cbest ? c2 : cbest;
                }
                if (cbest < kInf) {
                  const uint32_t ub = sat(cand + cbest);
                  if (ub < __hip_atomic_load(&sp->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicMin(&sp->best, ub);
                }
              }
            }
          }
        }
        if (improved) app = mark_dirty(mnxt, u, improved);
        held = improved != 0;
      }
      if (level >= 2) mark_hold(held, u, sp);
      stage_append(app, u, L, apn, sp, nxt, A.rowptr, V);
    }
    a = hia;
    wave_sync();
  }
  flush_append(L, apn, sp, nxt, A.rowptr, V);
  n_relax = wave_sum_u64(n_relax);
  n_impr = wave_sum_u64(n_impr);
  if (lane == 0) {
    atomicAdd(&sp->ctr[3], (unsigned long long)n_relax);
    atomicAdd(&sp->ctr[4], (unsigned long long)n_impr);
  }
}

// ------------------------------------------------- group-distance jobs ----

// Job SEED: distance 0 at every member of label l's group.  A label whose
// distances came from the cache is final already: it is not seeded, so its
// chunk bit is only ever set by the other labels of its chunk (relaxing its
// exact distances again never lowers them).
__device__ void item_jseed(const GraphDev& g, Slot* sp, uint32_t l) {
  const Arcs A = arcs_of(g, sp);
  if (sp->ccol && sp->ccol[l]) return;
  const int nxt = (sp->step + 1) % 3;
  const uint64_t V = (uint64_t)g.V;
  const int32_t grp = sp->lgrp[l];
  const int64_t beg = g.gptr[grp], end = g.gptr[grp + 1];
  uint64_t* mn = sp->mask + (uint64_t)nxt * V;
  const uint64_t rs = (uint64_t)sp->ND * (uint64_t)sp->Lp;  // row stride per vertex
  for (int64_t base = beg; base < end; base += 64) {
    const int64_t t = base + lane_id();
    bool app = false;
    uint32_t v = 0;
    if (t < end) {
      v = (uint32_t)g.gmem[t];
      sp->dist[(uint64_t)v * rs + l] = 0u;  // height 0
      const unsigned long long old = atomicOr((unsigned long long*)&mn[v], 1ull << (l / sp->jcl));
      app = (old == 0);
    }
    wave_append(app, v, sp, nxt, A.rowptr, V);
  }
}


// One label of a job relaxation: target word tw held t when read, the
// candidate is s + w.  The atomicMin returns nothing, so the wave never waits
// on it (arcs into a hub all hit the same row); the chunk counts as improved
// when the candidate beat the value read.  That may mark a chunk whose value
// another arc lowered further in the meantime: it is relaxed once more in the
// next round, with its current value.  Every mark follows a strict decrease
// of an integer, so the rounds still end.
__device__ __forceinline__ bool jrelax(uint32_t s, uint32_t t, uint32_t w, g_u32* tw, uint32_t& n_at) {
  if (s >= kInf) return false;
  const uint32_t cand = s + w;
  if (cand >= t || cand >= kInf) return false;
  n_at++;
  (void)__hip_atomic_fetch_min(tw, cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Job ROUND: one chunk of arcs of the job's frontier.  G lanes share an arc
// (64 / G arcs per wave iteration); lane r of the group relaxes the dirty
// CL-label chunks r, r + G, ... of the source row into the target
// row (one vector load per chunk, coalesced over the group).  Unbounded jobs
// relax within the one row per vertex until no chunk is dirty; hop jobs relax
// the rows of height level - 1 into those of height level, one round per
// height, after COPY has set height level to height level - 1: the rows hold
// the distances within level hops, and a round's frontier is only the chunks
// that the previous height improved.
template <int CL>
__device__ void item_jround_t(const GraphDev& g, Slot* sp, uint32_t item, FrontLDS* L) {
  typedef typename JChunk<CL>::type jchunk;
  typedef GST_AS_GLOBAL const jchunk g_cchunk;
  const Arcs A = arcs_of(g, sp);
  const int lane = lane_id();
  const uint64_t V = (uint64_t)g.V;
  const uint32_t r = sp->step;
  const int cur = r % 3, nxt = (r + 1) % 3;
  const uint64_t appc = sp->app[cur];
  const uint32_t n = (uint32_t)(appc >> 32);
  const uint32_t arcs_total = (uint32_t)appc;
  const uint32_t CH = (uint32_t)sp->chunk;
  const uint32_t a0 = item * CH;
  const uint32_t a1 = a0 + CH < arcs_total ? a0 + CH : arcs_total;
  const uint32_t* lst = sp->list + (uint64_t)cur * V;
  const uint32_t* lof = sp->loff + (uint64_t)cur * V;
  const uint64_t* mcur = sp->mask + (uint64_t)cur * V;
  uint64_t* mnxt = sp->mask + (uint64_t)nxt * V;
  const uint64_t Lp = (uint64_t)sp->Lp;
  const uint64_t rs = (uint64_t)sp->ND * Lp;  // row stride per vertex
  const bool hop = sp->D >= 0;
  const uint32_t* sdist = sp->dist + (hop ? (uint64_t)(sp->level - 1) * Lp : 0);
  uint32_t* tdist = sp->dist + (hop ? (uint64_t)sp->level * Lp : 0);
  const int nch = sp->Lp / CL;
  const int G = sp->G;
  const int APW = 64 / G;
  const int sub = lane / G, rr = lane - sub * G;
  uint64_t n_relax = 0, n_impr = 0;
  uint32_t n_at = 0, apn = 0, n_src = 0;
  uint32_t a = a0;
  for (uint32_t i = sp->istart[(uint64_t)cur * sp->istride + item]; a < a1; i += 64) {
    stage_front(A.rowptr, L, i, n, arcs_total, lst, lof, mcur);
    const uint32_t wend = (i + 64 < n) ? lof[i + 64] : arcs_total;
    wave_sync();
    const uint32_t hia = a1 < wend ? a1 : wend;
    // the arc of the group's next iteration is read while this one is relaxed
    int no = 0;
    uint32_t nu = 0, nw = 0;
    if (a + (uint32_t)sub < hia) {
      no = find_entry(L, a + (uint32_t)sub);
      const uint32_t e = L->wrs[no] + (a + (uint32_t)sub - L->woff[no]);
      nu = (uint32_t)A.col[e];
      nw = (uint32_t)A.wt[e];
    }
    for (uint32_t base = a; base < hia; base += (uint32_t)APW) {
      const uint32_t pos = base + (uint32_t)sub;
      uint64_t imp = 0;
      const int o = no;
      const uint32_t w = nw;
      uint32_t u = pos < hia ? nu : 0u;
      if (pos + (uint32_t)APW < hia) {
        no = find_entry(L, pos + (uint32_t)APW);
        const uint32_t e = L->wrs[no] + (pos + (uint32_t)APW - L->woff[no]);
        nu = (uint32_t)A.col[e];
        nw = (uint32_t)A.wt[e];
      }
      if (pos < hia) {
        const uint32_t vs = L->wv[o];
        const uint64_t m = L->wm[o];
        g_cchunk* srow = (g_cchunk*)(sdist + (uint64_t)vs * rs);
        g_cchunk* trow = (g_cchunk*)(tdist + (uint64_t)u * rs);
        g_u32* tw = (g_u32*)(tdist + (uint64_t)u * rs);
        // the source's chunks are counted once per entry, at its first arc
        const uint32_t first = (pos == L->woff[o]) ? 1u : 0u;
        for (int c = rr; c < nch; c += G) {
          if (!((m >> c) & 1ull)) continue;
          const jchunk sv = srow[c];
          const jchunk tv = trow[c];
          n_src += first;
          bool im = false;
#pragma unroll
          for (int q = 0; q < CL; q++) im |= jrelax(sv[q], tv[q], w, tw + CL * c + q, n_at);
          n_relax += CL;
          if (im) {
            imp |= 1ull << c;
            n_impr++;
          }
        }
      }
      // the improved chunks of one arc, gathered on the group's first lane
      for (int o2 = 1; o2 < G; o2 <<= 1) imp |= (uint64_t)__shfl_xor((unsigned long long)imp, o2, 64);
      const bool app = (pos < hia && rr == 0 && imp) ? mark_dirty(mnxt, u, imp) : false;
      stage_append(app, u, L, apn, sp, nxt, A.rowptr, V);
    }
    a = hia;
    wave_sync();
  }
  flush_append(L, apn, sp, nxt, A.rowptr, V);
  n_relax = wave_sum_u64(n_relax);
  n_impr = wave_sum_u64(n_impr);
  const uint64_t n_atom = wave_sum_u64(n_at);
  const uint64_t n_srcw = wave_sum_u64(n_src);
  if (lane == 0) {
    atomicAdd(&sp->ctr[3], (unsigned long long)n_relax);
    atomicAdd(&sp->ctr[4], (unsigned long long)n_impr);
    atomicAdd(&sp->ctr[5], (unsigned long long)n_atom);  // job: atomicMin issued
    atomicAdd(&sp->nsrc, (unsigned long long)n_srcw);
  }
}

__device__ __forceinline__ void item_jround(const GraphDev& g, Slot* sp, uint32_t item, FrontLDS* L) {
  if (sp->jcl == (uint32_t)kJobChunkFlat) item_jround_t<kJobChunkFlat>(g, sp, item, L);
  else item_jround_t<kJobChunkHop>(g, sp, item, L);
}

// Job CACHEIN / CACHEOUT over a tile of kStarTV vertices: one lane per label,
// walking the tile's vertices, so each lane reads (or writes) one contiguous
// run of its cache column and the lanes together touch a row's labels side by
// side.  CACHEIN copies cached columns into the rows (after FILL set them to
// INF); CACHEOUT copies a finished unbounded job's new labels out to their
// columns.
__device__ void item_jcache(Slot* sp, uint32_t tile, uint64_t V, bool out) {
  const uint64_t v0 = (uint64_t)tile * kStarTV;
  const uint64_t v1 = v0 + kStarTV < V ? v0 + kStarTV : V;
  const uint64_t rs = (uint64_t)sp->ND * (uint64_t)sp->Lp;
  const uint32_t L = (uint32_t)sp->L;
  for (uint32_t l = (uint32_t)lane_id(); l < L; l += 64) {
    uint32_t* col = sp->ccol[out ? L + l : l];
    if (!col) continue;
    g_u32* c = (g_u32*)col;
    g_u32* d = (g_u32*)sp->dist + l;
    if (out) {
      for (uint64_t v = v0; v < v1; v++) c[v] = d[v * rs];
    } else {
      for (uint64_t v = v0; v < v1; v++) d[v * rs] = c[v];
    }
  }
}

// Hop job COPY: height `level` of the rows of kStarTV vertices starts as a copy
// of height level - 1 (one lane per chunk, coalesced).
__device__ void item_jcopy(Slot* sp, uint32_t tile, uint64_t V) {
  const uint64_t v0 = (uint64_t)tile * kStarTV;
  const uint64_t v1 = v0 + kStarTV < V ? v0 + kStarTV : V;
  const uint64_t Lp = (uint64_t)sp->Lp;
  const uint64_t rs = (uint64_t)sp->ND * Lp;
  const int ng = (int)(Lp / 4);  // 16-byte granules per row
  const uint64_t from = (uint64_t)(sp->level - 1) * Lp, to = (uint64_t)sp->level * Lp;
  for (uint64_t v = v0; v < v1; v++) {
    g_u32* row = (g_u32*)(sp->dist + v * rs);
    for (int c = lane_id(); c < ng; c += 64)
      store_stream((g_gran*)(row + to + (uint64_t)c * 4u), *(g_cgran*)(row + from + (uint64_t)c * 4u));
  }
}

// Job STAR: the star bound min_v sum_i c_i(v) of every member query, from
// its k labels of each distance row.  One lane per member (64 members at a
// time): the row base is the same for the lanes of a member group and each
// lane adds its labels' offset, so a row is fetched once for the job; the
// labels of kStarBatch vertices are read before they are summed.  A last
// group of fewer than 64 members takes the smallest power of two P of lanes
// that holds it, and the 64 / P groups of P lanes take turns over the
// vertices.  A hop job's height d
// holds the distance within d hops, so height floor(D/2) gives the star arms.
constexpr int kStarBatch = 2;

__device__ void item_jstar(Slot* slots, Slot* sp, uint32_t tile, uint64_t V) {
  const int lane = lane_id();
  const uint64_t v0 = (uint64_t)tile * kStarTV;
  const uint64_t v1 = v0 + kStarTV < V ? v0 + kStarTV : V;
  const uint64_t Lp = (uint64_t)sp->Lp, rs = (uint64_t)sp->ND * Lp;
  const uint64_t hs = sp->D >= 0 ? (uint64_t)(sp->D / 2) : 0u;
  const int nmem = sp->nmem;
  for (int mg = 0; mg < nmem; mg += 64) {
    const int rem = nmem - mg < 64 ? nmem - mg : 64;
    int sh = 6;  // log2 of the lanes per member group
    while (sh > 0 && (1 << (sh - 1)) >= rem) sh--;
    const int m = mg + (lane & ((1 << sh) - 1));
    const uint64_t sub = (uint64_t)(lane >> sh), ngr = (uint64_t)(64 >> sh);
    const bool has = m < nmem;
    Slot* q = has ? &slots[sp->jmem[m]] : nullptr;
    const int k = has ? q->k : 0;
    const uint32_t off = has ? (uint32_t)q->lab + (uint32_t)(hs * Lp) : 0u;
    uint32_t star = kInf;
    for (uint64_t vb = v0 + kStarBatch * sub; vb < v1; vb += kStarBatch * ngr) {
      uint32_t c[kStarBatch][kMaxK];
#pragma unroll
      for (int b = 0; b < kStarBatch; b++) {
        const bool ok = has && vb + b < v1;
        g_cu32* row = (g_cu32*)sp->dist + (vb + b) * rs;
#pragma unroll
        for (int i = 0; i < kMaxK; i++) c[b][i] = (ok && i < k) ? row[off + (uint32_t)i] : 0u;
      }
#pragma unroll
      for (int b = 0; b < kStarBatch; b++) {
        if (!has || vb + b >= v1) continue;
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < kMaxK; i++) s = sat(s + c[b][i]);
        star = s < star ? s : star;
      }
    }
    // shortest paths from one vertex to every group form a connected cover,
    // so their sum bounds the optimum from above (for a hop query, paths of
    // at most floor(D/2) hops: every two of them are within D hops of each
    // other); one atomic per member and tile, only when it lowers best
    for (int o = 1 << sh; o < 64; o <<= 1) {
      const uint32_t a = __shfl_xor(star, o, 64);
      star = a < star ? a : star;
    }
    if (has && sub == 0 && star < kInf &&
        star < __hip_atomic_load(&q->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(&q->best, star);
  }
}

// ------------------------------------------------------ per-query DIST ----

// Distance table, inter-group distances and route table after level 1.
// The query's own C rows are as long as its S1 rows (cstride == ND * k), so a
// tile of vertices is one contiguous run of both: lanes walk it element by
// element (vertex, height, group), coalesced, each taking the minimum over the
// heights up to its own.  Job-fed queries have no DIST tiles: their C is the
// job's rows.
__device__ void item_dist(const GraphDev& g, Slot* sp, uint32_t tile) {
  const int lane = lane_id();
  const uint64_t V = (uint64_t)g.V;
  const int k = sp->k, ND = sp->ND;
  const bool hop = sp->D >= 0;
  const uint64_t v0 = (uint64_t)tile * kDistTV;
  const uint64_t v1 = v0 + kDistTV < V ? v0 + kDistTV : V;
  const uint32_t rl = (uint32_t)(ND * k);
  uint32_t* Ct = sp->C + v0 * rl;
  const uint32_t nE = (uint32_t)(v1 - v0) * rl;
  const uint64_t* S1t = sp->S1 + v0 * rl;
  for (uint32_t e = (uint32_t)lane; e < nE; e += 64) {
    const uint32_t vr = e / rl, rem = e - vr * rl;
    const uint32_t d = rem / (uint32_t)k, i = rem - d * (uint32_t)k;
    const uint64_t* col = S1t + vr * rl + i;
    uint32_t c = kInf;
    for (uint32_t dd = 0; dd <= d; dd++) {
      const uint32_t t = (uint32_t)(col[dd * (uint32_t)k] >> 32);
      c = t < c ? t : c;
    }
    Ct[e] = c;  // within d hops
  }
  // star upper bound; the arms of a hop star are pairwise within D hops when
  // each has at most floor(D/2)
  const int hs = hop ? sp->D / 2 : 0;
  uint32_t star = kInf;
  for (uint64_t v = v0 + (uint64_t)lane; v < v1; v += 64) {
    const uint64_t* row = sp->S1 + v * rl;
    uint32_t s = 0;
    for (int i = 0; i < k; i++) {
      uint32_t c = kInf;
      for (int d = 0; d <= hs; d++) {
        const uint32_t t = (uint32_t)(row[d * k + i] >> 32);
        c = t < c ? t : c;
      }
      s = sat(s + c);
    }
    star = s < star ? s : star;
  }
  star = wave_min_u32(star);
  // one word for every tile of the query: lower it only when the tile beats it
  if (lane == 0 && star < kInf && star < __hip_atomic_load(&sp->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMin(&sp->best, star);
}

__device__ void item_vvr(const GraphDev& g, Slot* sp, WaveLDS* L) {
  const int lane = lane_id();
  const int k = sp->k, W = sp->W, ND = sp->ND;
  const States S = states_of(sp);
  uint32_t* vvl = (uint32_t*)L->f.wv;  // k*k <= 36 words
  for (int i = 0; i < k; i++) {
    const int32_t grp = sp->qg[i];
    const int64_t beg = g.gptr[grp], end = g.gptr[grp + 1];
    uint32_t loc[kMaxK];
#pragma unroll
    for (int j = 0; j < kMaxK; j++) loc[j] = kInf;
    for (int64_t t = beg + lane; t < end; t += 64) {
      const uint32_t v = (uint32_t)g.gmem[t];
#pragma unroll
      for (int j = 0; j < kMaxK; j++) {
        if (j >= k) continue;
        uint32_t c = kInf;
        if (sp->S1) {
          // C is being written by this step's DIST tiles: read S1
          for (int d = 0; d < ND; d++) {
            const uint32_t t2 = (uint32_t)(sp->S1[((uint64_t)v * ND + d) * k + j] >> 32);
            c = t2 < c ? t2 : c;
          }
        } else {
          // the job's rows (for an unbounded query they are its C)
          for (int d = 0; d < ND; d++) {
            const uint32_t t2 = S.jcost(v, (uint32_t)j, (uint32_t)d);
            c = t2 < c ? t2 : c;
          }
        }
        loc[j] = c < loc[j] ? c : loc[j];
      }
    }
#pragma unroll
    for (int j = 0; j < kMaxK; j++) {
      uint32_t m = wave_min_u32(loc[j]);
      if (lane == 0 && j < k) vvl[i * k + j] = (i == j) ? 0u : m;
    }
  }
  wave_sync();
  // route table R[x][i][j]: cheapest sequence of inter-group hops that starts
  // in group i, visits every group of x, and ends in group j (i != j)
  uint32_t* Rl = (uint32_t*)L->stage;
  const int kk = k * k;
  const int nR = W * kk;
  for (int t = lane; t < nR; t += 64) Rl[t] = kInf;
  wave_sync();
  for (int c = 2; c <= k; c++) {
    for (int t = lane; t < nR; t += 64) {
      const uint32_t x = (uint32_t)(t / kk);
      const int i = (t / k) % k, j = t % k;
      if (__popc(x) != c || i == j || !((x >> i) & 1u) || !((x >> j) & 1u)) continue;
      uint32_t best;
      if (c == 2) {
        best = vvl[i * k + j];
      } else {
        const uint32_t y = x ^ (1u << j);
        best = kInf;
        for (int m = 0; m < k; m++) {
          if (m == i || !((y >> m) & 1u)) continue;
          const uint32_t cand = sat(Rl[y * kk + i * k + m] + vvl[m * k + j]);
          best = cand < best ? cand : best;
        }
      }
      Rl[t] = best;
    }
    wave_sync();
  }
  for (int t = lane; t < nR; t += 64) sp->R[t] = Rl[t];
  if (lane < kk) sp->vv[lane] = vvl[lane];
}

// Stages the W * NS segments of nv live vertices into LDS (row-major like T);
// singleton entries are taken from S1 or the job rows, the unused p = 0
// column is INF.
__device__ __forceinline__ void stage_rows(const Slot* sp, const uint32_t* ids, int nv, int rowlen,
                                           int W, uint64_t* stage) {
  const States S = states_of(sp);
  for (int t = lane_id(); t < nv * rowlen; t += 64) {
    const int jj = t / rowlen, w = t - jj * rowlen;
    const uint32_t v = ids[jj];
    const int d = w / W;
    const uint32_t p = (uint32_t)(w - d * W);
    stage[t] = (p == 0) ? kInfState : S.get(v, p, (uint32_t)d);
  }
}

// Progressive answers need, per candidate vertex, the full-set bound itself
// (its minimum over the candidates bounds the optimum, whose tree has its
// vertices among them) and the cheapest star (arms within floor(D/2) hops).
// Wave-collective (lanes without a vertex pass has = false); returns whether
// the lane's vertex is live.  The minima go straight to the slot, so LIVE's
// loop carries nothing extra.
__device__ bool prog_bounds(Slot* sp, bool has, uint64_t v, uint32_t lim) {
  uint32_t lb = kInf;
  uint64_t pk = ~0ull;
  if (has) {
    const int k = sp->k;
    g_cu32* cr = crow(sp, v, sp->D);
    lb = lower_bound(cr, k, (uint32_t)sp->W - 1, (g_cu32*)sp->R);
    g_cu32* cs = sp->D >= 0 ? crow(sp, v, sp->D / 2) : cr;
    uint32_t st = 0;
    for (int i = 0; i < k; i++) st = sat(st + cs[i]);
    if (st < kInf) pk = ((uint64_t)st << 32) | v;
  }
  const uint32_t lbw = wave_min_u32(lb);
  const uint64_t pkw = wave_min_u64(pk);
  if (lane_id() == 0) {
    if (lbw < kInf) atomicMin(&sp->lbmin, lbw);
    if (pkw != ~0ull) atomicMin((unsigned long long*)&sp->star_pk, (unsigned long long)pkw);
  }
  return has && lb <= lim;
}

// Live list after level 1: a vertex whose lower bound for covering every
// group (the tour bound over the full set) exceeds the current upper bound
// cannot lie on a tree cheaper than that bound, so it never holds a useful
// state of level >= 2.  Live vertices get their T segments initialised here
// and a live flag; only they are ever read or written later.
__device__ void item_live(Slot* sp, uint32_t tile, uint64_t V, WaveLDS* L) {
  const int lane = lane_id();
  const int k = sp->k, W = sp->W;
  const int rowlen = W * sp->NS;
  const uint32_t bst = __hip_atomic_load(&sp->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t v0 = (uint64_t)tile * kDistTV;
  const uint64_t v1 = v0 + kDistTV < V ? v0 + kDistTV : V;
  const bool prog = sp->anum != 0;
  for (uint64_t vb = v0; vb < v1; vb += 64) {
    const uint64_t idx = vb + lane;
    uint64_t v = 0;
    bool lv = false;
    if (idx < v1) v = idx;
    if (prog) {
      lv = prog_bounds(sp, idx < v1, v, bst < kInf ? bst : kInf - 1);
    } else if (idx < v1) {
      // lower bound on any tree through v that covers every group
      lv = lb_le(crow(sp, v, sp->D), k, (uint32_t)W - 1, (g_cu32*)sp->R, bst < kInf ? bst : kInf - 1);
    }
    if (lv) atomicOr(&sp->lvb[v >> 5], 1u << (uint32_t)(v & 31u));
    const uint64_t ball = __ballot(lv);
    if (ball == 0) continue;
    const int nl = __popcll(ball);
    const uint32_t rank = __popcll(ball & __lanemask_lt());
    uint32_t base = 0;
    const int leader = __ffsll((long long)ball) - 1;
    if (lane == leader) base = atomicAdd(&sp->nlive, (uint32_t)nl);
    base = __shfl(base, leader, 64);
    if (lv) {
      sp->live[base + rank] = (uint32_t)v;
      L->f.wv[rank] = (uint32_t)v;
    }
    wave_sync();
    for (int t = lane; t < nl * rowlen; t += 64) {
      const int jj = t / rowlen;
      // not read again before the merges above level 2
      store_stream(&sp->T[(uint64_t)L->f.wv[jj] * rowlen + (t - jj * rowlen)], kInfState);
    }
    wave_sync();
  }
}

// lower_bound(Crow, k, W - 1, R) with the group distances c[0..k) and the full
// set's route entries Rf[i * kMaxK + j] already in registers.
__device__ __forceinline__ uint32_t lb_full_regs(const uint32_t* c, int k, const uint32_t* Rf) {
  uint32_t lb1 = 0;
#pragma unroll
  for (int i = 0; i < kMaxK; i++)
    if (i < k) lb1 = c[i] > lb1 ? c[i] : lb1;
  if (lb1 >= kInf) return kInf;
  if (k >= 2) {
    uint32_t tour = 3u * kInf;
#pragma unroll
    for (int i = 0; i < kMaxK; i++) {
#pragma unroll
      for (int j = 0; j < kMaxK; j++) {
        if (i >= k || j >= k || i == j) continue;
        const uint32_t t = c[i] + Rf[i * kMaxK + j] + c[j];
        tour = t < tour ? t : tour;
      }
    }
    const uint32_t lb2 = (tour + 1u) >> 1;
    lb1 = lb2 > lb1 ? lb2 : lb1;
  }
  return sat(lb1);
}

// Job LIVE: the live test of item_live for every member query of a job in one
// pass over the job's rows.  A tile of kJLiveTV vertices; one lane per member
// (64 members at a time) holds its query's full-set route entries, bound and
// row offset, and the wave walks the tile's vertices, so the members read
// their labels side by side from the same row (each row is fetched once for
// the whole job instead of a line per member and vertex).  A last group of
// fewer than 64 members takes the smallest power of two P of lanes that holds
// it, and the 64 / P groups of P lanes walk different 32-vertex words of the
// tile.  A lane gathers the live bits of 32 vertices into a word and stores it
// into its query's live bits (the tile owns those words; the query's FILL
// cleared them).  The
// query's LIVE then lists its live vertices and initialises their T segments
// (item_live2).  Progressive members also get the minima prog_bounds keeps.
constexpr uint32_t kJLiveTV = 256;
constexpr int kJLiveBatch = 2;  // vertices whose labels are read before the tests

__device__ void item_jlive(Slot* slots, Slot* sp, uint32_t tile, uint64_t V) {
  const int lane = lane_id();
  const uint64_t v0 = (uint64_t)tile * kJLiveTV;
  const uint64_t v1 = v0 + kJLiveTV < V ? v0 + kJLiveTV : V;
  const int nmem = sp->nmem;
  for (int mg = 0; mg < nmem; mg += 64) {
    const int rem = nmem - mg < 64 ? nmem - mg : 64;
    int sh = 6;  // log2 of the lanes per member group
    while (sh > 0 && (1 << (sh - 1)) >= rem) sh--;
    const int m = mg + (lane & ((1 << sh) - 1));
    const uint64_t sub = (uint64_t)(lane >> sh), ngr = (uint64_t)(64 >> sh);
    const bool has = m < nmem;
    Slot* q = has ? &slots[sp->jmem[m]] : nullptr;
    const int k = has ? q->k : 0;
    uint32_t Rf[kMaxK * kMaxK];
    g_cu32* Rq = has ? (g_cu32*)q->R + (uint64_t)((1u << k) - 1u) * (uint64_t)(k * k) : nullptr;
#pragma unroll
    for (int i = 0; i < kMaxK; i++)
#pragma unroll
      for (int j = 0; j < kMaxK; j++) Rf[i * kMaxK + j] = (i < k && j < k) ? Rq[i * k + j] : kInf;
    const uint32_t bst = has ? __hip_atomic_load(&q->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kInf;
    const uint32_t lim = bst < kInf ? bst : kInf - 1;
    const bool prog = has && q->anum != 0;
    // an unbounded job's live test also takes each member's star (its labels
    // are the bound row itself); a hop job's STAR step has done that
    const bool flat = sp->D < 0;
    // A member's labels sit at offset lab in every row of the job (rows of
    // rs words per vertex, heights Lp words apart): the row base is the same
    // for the whole wave, the lane adds its offset for height D (bound) and
    // floor(D / 2) (star).
    const uint64_t Lp = (uint64_t)sp->Lp, rs = (uint64_t)sp->ND * Lp;
    const uint32_t co = has ? (uint32_t)q->lab + (uint32_t)(sp->D >= 0 ? (uint64_t)sp->D * Lp : 0) : 0u;
    const uint32_t so = has ? (uint32_t)q->lab + (uint32_t)(sp->D >= 0 ? (uint64_t)(sp->D / 2) * Lp : 0) : 0u;
    uint32_t lbmn = kInf, smin = kInf;
    uint64_t pkmn = ~0ull;
    for (uint64_t w0 = v0 + 32u * sub; w0 < v1; w0 += 32u * ngr) {
      const uint64_t w1 = w0 + 32 < v1 ? w0 + 32 : v1;
      uint32_t word = 0;
      for (uint64_t vb = w0; vb < w1; vb += kJLiveBatch) {
        uint32_t c[kJLiveBatch][kMaxK];
#pragma unroll
        for (int b = 0; b < kJLiveBatch; b++) {
          const bool ok = has && vb + b < w1;
          g_cu32* row = (g_cu32*)sp->dist + (vb + b) * rs;
#pragma unroll
          for (int i = 0; i < kMaxK; i++) c[b][i] = (ok && i < k) ? row[co + (uint32_t)i] : kInf;
        }
#pragma unroll
        for (int b = 0; b < kJLiveBatch; b++) {
          if (!has || vb + b >= w1) continue;
          const uint32_t lb = lb_full_regs(c[b], k, Rf);
          if (lb <= lim) word |= 1u << (uint32_t)(vb + b - w0);
          uint32_t st = 0;
          if (flat) {
#pragma unroll
            for (int i = 0; i < kMaxK; i++)
              if (i < k) st = sat(st + c[b][i]);
            smin = st < smin ? st : smin;
          }
          if (prog) {
            lbmn = lb < lbmn ? lb : lbmn;
            if (!flat) {
              g_cu32* row = (g_cu32*)sp->dist + (vb + b) * rs;
              for (int i = 0; i < k; i++) st = sat(st + row[so + (uint32_t)i]);
            }
            const uint64_t pk = ((uint64_t)st << 32) | (vb + b);
            if (st < kInf && pk < pkmn) pkmn = pk;
          }
        }
      }
      if (word) q->lvb[w0 >> 5] = word;
    }
    // the star is the cost of a real walk-tree: shortest paths from one
    // vertex to every group.  A member's lane groups combine their stars,
    // and its first lane lowers best only when the tile's star beats it:
    // every tile of every member on the same word would otherwise queue at
    // its L2 channel (progressive members, rare, keep one atomic per lane)
    for (uint32_t o = 64u / (uint32_t)ngr; o < 64u; o <<= 1) {
      const uint32_t a = __shfl_xor(smin, (int)o, 64);
      smin = a < smin ? a : smin;
    }
    if (has && sub == 0 && smin <= lim) atomicMin(&q->best, smin);  // lim: best when the tile began
    if (prog) {
      if (lbmn < kInf) atomicMin(&q->lbmin, lbmn);
      if (pkmn != ~0ull) atomicMin((unsigned long long*)&q->star_pk, (unsigned long long)pkmn);
    }
  }
}

// LIVE of a query fed by a group-distance job, whose job has set its live
// bits (item_jlive): a tile of kLive2TV vertices, one 32-bit word per lane,
// lists the live vertices (one atomic per tile), initialises their T
// segments and clears their dirty masks.
constexpr uint32_t kLive2TV = 2048;

__device__ void item_live2(Slot* sp, uint32_t tile, uint64_t V, WaveLDS* L) {
  const int lane = lane_id();
  const uint64_t nw = (V + 31) / 32;
  const uint64_t wi = (uint64_t)tile * (kLive2TV / 32) + (uint64_t)lane;
  uint32_t word = wi < nw ? sp->lvb[wi] : 0u;
  const uint32_t cnt = (uint32_t)__popc(word);
  const uint32_t excl = wave_excl_scan(cnt);
  const uint32_t total = __shfl(excl + cnt, 63, 64);
  if (total == 0) return;
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(&sp->nlive, total);
  base = __shfl(base, 0, 64);
  uint32_t* ids = (uint32_t*)L->stage;  // kLive2TV ids at most
  uint32_t o = excl;
  while (word) {
    const uint32_t b = (uint32_t)(__ffs(word) - 1);
    word &= word - 1u;
    const uint32_t v = (uint32_t)(wi * 32u + b);
    ids[o] = v;
    sp->live[base + o] = v;
    o++;
  }
  wave_sync();
  const uint32_t rowlen = (uint32_t)(sp->W * sp->NS);
  for (uint32_t t = (uint32_t)lane; t < total * rowlen; t += 64) {
    const uint32_t jj = t / rowlen;
    // not read again before the merges above level 2
    store_stream(&sp->T[(uint64_t)ids[jj] * rowlen + (t - jj * rowlen)], kInfState);
  }
  // the three dirty masks of each live vertex (FILL left them alone)
  for (uint32_t t = (uint32_t)lane; t < total * 3u; t += 64) {
    const uint32_t jj = t / 3u;
    sp->mask[(uint64_t)(t - jj * 3u) * V + ids[jj]] = 0ull;
  }
  wave_sync();
}

// Level-L merge: T(v,p,nd) for every |p| = L subset p and height nd of every
// live vertex, from the final lower levels.  The W * NS segments of up to VPW
// live vertices are staged in LDS; the wave's lanes then take the (height,
// vertex, subset) tuples of the staged vertices in turn, height-major so that
// neighbouring lanes run the same height-pair loops.
__device__ void item_merge(const GraphDev& g, Slot* sp, uint32_t tile, WaveLDS* L) {
  const Arcs A = arcs_of(g, sp);
  const int lane = lane_id();
  const uint64_t V = (uint64_t)g.V;
  const int k = sp->k, W = sp->W, ND = sp->ND, NS = sp->NS, D = sp->D;
  const bool hop = D >= 0;
  const int Dlim = hop ? D : 0;
  const int level = sp->level;
  const uint32_t full = (uint32_t)W - 1;
  const int nxt = (sp->step + 1) % 3;
  uint64_t* mnxt = sp->mask + (uint64_t)nxt * V;
  // Level 2 reads only singletons (every state of two or more groups is still
  // INF), so it stages the NS * k singletons of a vertex instead of its whole
  // segment: row[d * k + i] instead of row[d * W + p].
  const bool l2 = level == 2;
  const int rowlen = l2 ? NS * k : W * NS;
  // each staged vertex also brings its C rows (ND * k words, in pairs)
  const int crl = (ND * k + 1) / 2;
  const int VPW = (kStageWords / (rowlen + crl)) < 64 ? (kStageWords / (rowlen + crl)) : 64;
  const States S = states_of(sp);
  // this level's subsets, listed in LDS
  uint32_t* subs = L->f.pad;
  const bool mine = lane < W && __popc((uint32_t)lane) == level;
  const uint64_t sball = __ballot(mine);
  if (mine) subs[__popcll(sball & __lanemask_lt())] = (uint32_t)lane;
  const int nL = __popcll(sball);
  // level 2 merges singletons of every live vertex; a higher level needs a part
  // of two or more groups at the vertex, so only active vertices can hold one
  const uint32_t* vl = level >= 3 ? sp->act : sp->live;
  const uint64_t nl = level >= 3 ? sp->nact : sp->nlive;
  const uint64_t mtv = sp->mtv;
  const uint64_t v0 = (uint64_t)tile * mtv;
  const uint64_t v1 = v0 + mtv < nl ? v0 + mtv : nl;
  const uint32_t bst0 = __hip_atomic_load(&sp->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t ubmin = kInf;
  uint64_t n_eval = 0;
  for (uint64_t vb = v0; vb < v1; vb += VPW) {
    const int nv = (int)((v1 - vb) < (uint64_t)VPW ? (v1 - vb) : (uint64_t)VPW);
    if (l2) {
      for (int t = lane; t < nv * rowlen; t += 64) {
        const int jj = t / rowlen, rem = t - jj * rowlen;
        const int d = rem / k;
        L->stage[t] = S.get(vl[vb + jj], 1u << (rem - d * k), (uint32_t)d);
      }
    } else {
      stage_rows(sp, vl + vb, nv, rowlen, W, L->stage);
    }
    uint32_t* cst = (uint32_t*)(L->stage + nv * rowlen);  // C rows of the staged vertices
    if (l2 && !sp->S1 && NS == ND) {
      // a job-fed query's singletons are its C rows: take them from the stage
      wave_sync();
      for (int t = lane; t < nv * rowlen; t += 64) cst[t] = (uint32_t)(L->stage[t] >> 32);
    } else {
      for (int t = lane; t < nv * ND * k; t += 64) {
        const int jj = t / (ND * k), rem = t - jj * ND * k;
        const int d = rem / k;
        cst[t] = crow(sp, vl[vb + jj], d)[rem - d * k];
      }
    }
    if (lane < nv) {
      L->f.wv[lane] = vl[vb + lane];
      L->f.wm[lane] = 0;
    }
    wave_sync();
    const int per_h = nv * nL;  // tuples of one height
    for (int t = lane; t < per_h * NS; t += 64) {
      const int nd = t / per_h;
      const int r = t - nd * per_h;
      const int jj = r / nL;
      const uint32_t p = subs[r - jj * nL];
      const uint32_t v = L->f.wv[jj];
      const uint64_t* row = L->stage + jj * rowlen;
      const uint32_t x = full ^ p;
      // The subtree part reaches every group of p within nd hops and the rest
      // every group of ~p within D - nd hops, on disjoint edges: when the two
      // farthest groups already cost more than the upper bound, the state
      // cannot pass the prune below and its walk-trees cannot lower the bound.
      {
        const uint32_t* ct = cst + jj * ND * k + (hop ? nd * k : 0);
        const uint32_t* cr = cst + jj * ND * k + (hop ? (D - nd) * k : 0);
        uint32_t mt = 0, mr = 0;
        for (int i = 0; i < k; i++) {
          if ((p >> i) & 1u) mt = ct[i] > mt ? ct[i] : mt;
          else mr = cr[i] > mr ? cr[i] : mr;
        }
        if (sat(mt + mr) > bst0) continue;
      }
      // state (d, q) of the staged vertex, in either layout
      auto st = [&](int d, uint32_t q) -> uint64_t {
        return l2 ? row[d * k + (__ffs(q) - 1)] : row[d * W + q];
      };
      const uint32_t low = p & (0u - p);
      uint64_t bestpk = l2 ? kInfState : row[nd * W + p];
      for (uint32_t sub = (p - 1) & p; sub; sub = (sub - 1) & p) {
        if (!(sub & low)) continue;
        const uint32_t sub2 = p ^ sub;
        // pairs (d1, d2) with max(d1, d2) == nd and d1 + d2 <= Dlim
        const int lim = Dlim - nd;
        if (lim < 0) continue;
        const int dm = nd < lim ? nd : lim;
        for (int d2 = 0; d2 <= dm; d2++) {  // d1 = nd
          const uint64_t a = st(nd, sub), b = st(d2, sub2);
          const uint32_t c = sat((uint32_t)(a >> 32) + (uint32_t)(b >> 32));
          n_eval++;
          if (c >= kInf) continue;
          const uint64_t pk = ((uint64_t)c << 32) | (kPredMerge | sub | ((uint32_t)nd << 16) | ((uint32_t)d2 << 23));
          bestpk = pk < bestpk ? pk : bestpk;
        }
        const int dm1 = (nd - 1) < lim ? (nd - 1) : lim;
        for (int d1 = 0; d1 <= dm1; d1++) {  // d2 = nd, d1 < nd
          const uint64_t a = st(d1, sub), b = st(nd, sub2);
          const uint32_t c = sat((uint32_t)(a >> 32) + (uint32_t)(b >> 32));
          n_eval++;
          if (c >= kInf) continue;
          const uint64_t pk = ((uint64_t)c << 32) | (kPredMerge | sub | ((uint32_t)d1 << 16) | ((uint32_t)nd << 23));
          bestpk = pk < bestpk ? pk : bestpk;
        }
      }
      const uint32_t c = (uint32_t)(bestpk >> 32);
      if (c >= kInf) continue;
      // a state of height nd reaches the rest of its tree within D - nd hops
      // (the staged C rows hold every height of the vertex)
      {
        const uint32_t* cb = cst + jj * ND * k + (hop ? (D - nd) * k : 0);
        uint32_t cv[kMaxK];
#pragma unroll
        for (int i = 0; i < kMaxK; i++) cv[i] = i < k ? cb[i] : kInf;
        if (c + lower_bound_c(cv, k, x, (g_cu32*)sp->R) > bst0) continue;
      }
      sp->T[((uint64_t)v * NS + nd) * W + p] = bestpk;
      atomicOr((unsigned long long*)&L->f.wm[jj], 1ull << p);
      // upper bounds from complete walk-trees at v: this state plus one
      // shortest path per missing group (hop: paths of at most h hops, with
      // nd + h <= D and, for two or more paths, 2h <= D) ...
      const int h = hop ? ((__popc(x) >= 2 && D / 2 < D - nd) ? D / 2 : D - nd) : 0;
      const uint32_t* cr = cst + jj * ND * k + h * k;
      uint32_t psum = 0;
      for (int i = 0; i < k; i++)
        if ((x >> i) & 1u) psum = sat(psum + cr[i]);
      uint32_t ub = sat(c + psum);
      // ... or the complement state at v, when its level is final
      if (__popc(x) < level) {
        const int d2m = (Dlim - nd) < NS - 1 ? (Dlim - nd) : NS - 1;
        for (int d2 = 0; d2 <= d2m; d2++) {
          const uint32_t u2 = sat(c + (uint32_t)(st(d2, x) >> 32));
          ub = u2 < ub ? u2 : ub;
        }
      }
      ubmin = ub < ubmin ? ub : ubmin;
    }
    wave_sync();
    bool app = false;
    uint32_t v = 0;
    if (lane < nv) {
      v = L->f.wv[lane];
      const uint64_t m = L->f.wm[lane];
      if (m) {
        mnxt[v] = m;
        app = true;
      }
    }
    if (level == 2) mark_hold(app, v, sp);  // above level 2 the vertex is active already
    wave_append(app, v, sp, nxt, A.rowptr, V);
    wave_sync();
  }
  ubmin = wave_min_u32(ubmin);
  n_eval = wave_sum_u64(n_eval);
  if (lane == 0) {
    if (ubmin < kInf && ubmin < __hip_atomic_load(&sp->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(&sp->best, ubmin);
    atomicAdd(&sp->ctr[5], (unsigned long long)n_eval);
  }
}

// Full-set merge: W/2 lanes per vertex, one lane per split (sub contains bit 0).
__device__ void item_final(const GraphDev& g, Slot* sp, uint32_t tile, WaveLDS* L) {
  const int lane = lane_id();
  const int k = sp->k, W = sp->W, ND = sp->ND, NS = sp->NS, D = sp->D;
  const bool hop = D >= 0;
  const int Dlim = hop ? D : 0;
  const uint32_t full = (uint32_t)W - 1;
  const int rowlen = W * NS;
  const int half = W / 2;  // W >= 2
  const int crl = (ND * k + 1) / 2;  // C rows of a staged vertex, in u64 words
  int VPF = 64 / half;
  const int cap = kStageWords / (rowlen + crl);
  if (VPF > cap) VPF = cap;
  const uint32_t bst0 = __hip_atomic_load(&sp->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int j = lane / half;
  const int s = lane % half;
  const uint32_t sub = 2u * (uint32_t)s + 1u;
  const uint32_t sub2 = full ^ sub;
  // with k >= 3 one part of every split has two or more groups: active vertices
  const uint32_t* vl = k >= 3 ? sp->act : sp->live;
  const uint64_t nl = k >= 3 ? sp->nact : sp->nlive;
  const uint64_t mtv = sp->mtv;
  const uint64_t v0 = (uint64_t)tile * mtv;
  const uint64_t v1 = v0 + mtv < nl ? v0 + mtv : nl;
  uint64_t ansmin = ~0ull;
  uint64_t n_eval = 0;
  for (uint64_t vb = v0; vb < v1; vb += VPF) {
    const int nv = (int)((v1 - vb) < (uint64_t)VPF ? (v1 - vb) : (uint64_t)VPF);
    stage_rows(sp, vl + vb, nv, rowlen, W, L->stage);
    uint32_t* cst = (uint32_t*)(L->stage + nv * rowlen);
    for (int t = lane; t < nv * ND * k; t += 64) {
      const int jj = t / (ND * k), rem = t - jj * ND * k;
      const int d = rem / k;
      cst[t] = crow(sp, vl[vb + jj], d)[rem - d * k];
    }
    wave_sync();
    const bool vact = j < nv;
    const uint32_t v = vact ? vl[vb + j] : 0u;
    const uint64_t* row = L->stage + (vact ? j : 0) * rowlen;
    for (int nd = 0; nd < NS; nd++) {
      // a tree of height nd at v reaches every group within nd hops: when the
      // farthest one costs more than the upper bound it cannot be the answer
      bool vok = vact;
      if (vok) {
        const uint32_t* ct = cst + j * ND * k + (hop ? nd * k : 0);
        uint32_t mt = 0;
        for (int i = 0; i < k; i++) mt = ct[i] > mt ? ct[i] : mt;
        vok = mt <= bst0;
      }
      const bool sact = vok && sub != full;
      uint64_t bestpk = (vok && s == 0) ? row[nd * W + full] : kInfState;
      if (sact) {
        const int lim = Dlim - nd;
        if (lim >= 0) {
          const int dm = nd < lim ? nd : lim;
          for (int d2 = 0; d2 <= dm; d2++) {
            const uint64_t a = row[nd * W + sub], b = row[d2 * W + sub2];
            const uint32_t c = sat((uint32_t)(a >> 32) + (uint32_t)(b >> 32));
            n_eval++;
            if (c >= kInf) continue;
            const uint64_t pk = ((uint64_t)c << 32) | (kPredMerge | sub | ((uint32_t)nd << 16) | ((uint32_t)d2 << 23));
            bestpk = pk < bestpk ? pk : bestpk;
          }
          const int dm1 = (nd - 1) < lim ? (nd - 1) : lim;
          for (int d1 = 0; d1 <= dm1; d1++) {
            const uint64_t a = row[d1 * W + sub], b = row[nd * W + sub2];
            const uint32_t c = sat((uint32_t)(a >> 32) + (uint32_t)(b >> 32));
            n_eval++;
            if (c >= kInf) continue;
            const uint64_t pk = ((uint64_t)c << 32) | (kPredMerge | sub | ((uint32_t)d1 << 16) | ((uint32_t)nd << 23));
            bestpk = pk < bestpk ? pk : bestpk;
          }
        }
      }
      // min over the half lanes of this vertex (segments are aligned powers of two)
      for (int o = half / 2; o > 0; o >>= 1) {
        const uint64_t y = (uint64_t)__shfl_xor((unsigned long long)bestpk, o, 64);
        bestpk = y < bestpk ? y : bestpk;
      }
      if (vact && s == 0 && (uint32_t)(bestpk >> 32) < kInf) {
        sp->T[((uint64_t)v * NS + nd) * W + full] = bestpk;
        const uint64_t a = (bestpk & 0xFFFFFFFF00000000ull) | v;
        ansmin = a < ansmin ? a : ansmin;
      }
    }
    wave_sync();
  }
  ansmin = wave_min_u64(ansmin);
  n_eval = wave_sum_u64(n_eval);
  if (lane == 0) {
    if (ansmin != ~0ull) atomicMin((unsigned long long*)&sp->ans, (unsigned long long)ansmin);
    atomicAdd(&sp->ctr[5], (unsigned long long)n_eval);
  }
}

// Predecessor chase from the optimal root: breadth-first over the witness,
// one lane per state; edges get their parent edge index so that a tree-like
// walk (hop DP) keeps its shape.  A singleton read from a job's rows
// (kPredPath) has no stored predecessor: its next edge is an arc v->u with
// J[u][i] + w == J[v][i] (for a hop job, J at height d - 1 against height d),
// found by a wave-wide scan of v's row (it exists because the rows are exact).
__device__ void item_recon(const GraphDev& g, Slot* sp) {
  const Arcs A = arcs_of(g, sp);
  const int lane = lane_id();
  const int W = sp->W, NS = sp->NS, D = sp->D;
  const States S = states_of(sp);
  const bool hop = D >= 0;
  const uint32_t full = (uint32_t)W - 1;
  // the answer: the DP's best full-set state, or (progressive answers) the
  // cheapest star, rebuilt as one shortest path per group from its centre
  const bool star = sp->recon_star != 0;
  const uint64_t ans = star ? (uint64_t)sp->star_pk : sp->ans;
  const uint32_t cost = (uint32_t)(ans >> 32);
  if (cost >= kInf) {
    if (lane == 0) {
      sp->out_cost = (int32_t)kInf;
      sp->out_root = -1;
      sp->out_nedges = 0;
      sp->out_d = -1;
    }
    return;
  }
  const uint32_t root = (uint32_t)ans;
  // smallest depth holding the optimum at the root
  uint32_t cd = kInf;
  if (lane < NS) cd = (uint32_t)(S.get(root, full, (uint32_t)lane) >> 32);
  const uint64_t hit = __ballot(lane < NS && cd == cost);
  const int dstar = hit ? (__ffsll((long long)hit) - 1) : 0;

  const uint32_t qcap = (uint32_t)sp->edge_cap + 2 * kMaxK;
  uint32_t* qv = sp->rq;
  uint32_t* qpd = sp->rq + qcap;
  int32_t* qat = (int32_t*)(sp->rq + 2 * qcap);
  if (!star && lane == 0) {
    qv[0] = root;
    qpd[0] = full | ((uint32_t)dstar << 16);
    qat[0] = -1;
  }
  if (star && lane < sp->k) {
    // one arm per group: its singleton at the lowest height h <= floor(D/2)
    // that holds the distance within floor(D/2) hops
    uint32_t sd = 0;
    if (hop) {
      const uint32_t want = crow(sp, root, D / 2)[lane];
      sd = (uint32_t)(D / 2);
      for (int h = D / 2; h >= 0; h--)
        if ((uint32_t)(S.get(root, 1u << lane, (uint32_t)h) >> 32) == want) sd = (uint32_t)h;
    }
    qv[lane] = root;
    qpd[lane] = (1u << lane) | (sd << 16);
    qat[lane] = -1;
  }
  __threadfence();
  wave_sync();
  uint32_t head = 0, tail = star ? (uint32_t)sp->k : 1u, ne = 0;
  bool overflow = false, broken = false;
  while (head < tail) {
    const uint32_t cnt = (tail - head) < 64u ? (tail - head) : 64u;
    const uint32_t t = head + lane;
    const bool act = (uint32_t)lane < cnt;
    uint32_t v = 0, p = 0, d = 0, pred = 0, sc = 0;
    int32_t att = -1;
    if (act) {
      v = qv[t];
      const uint32_t pd = qpd[t];
      p = pd & 0xFFFFu;
      d = pd >> 16;
      att = qat[t];
      const uint64_t st = S.get(v, p, d);
      pred = (uint32_t)st;
      sc = (uint32_t)(st >> 32);
    }
    uint32_t type = pred >> 30;
    // shortest-path singletons: resolve the next arc with the whole wave
    uint32_t parc = 0xFFFFFFFFu;
    uint64_t pb = __ballot(act && type == 3u && sc > 0u);
    while (pb) {
      const int l = __ffsll((long long)pb) - 1;
      pb &= pb - 1;
      const uint32_t lv = __shfl(v, l, 64);
      const uint32_t li = (uint32_t)(__ffs(__shfl(p, l, 64)) - 1);
      const uint32_t lc = __shfl(sc, l, 64);
      const uint32_t ld = __shfl(d, l, 64);
      const uint32_t e0 = (uint32_t)A.rowptr[lv], e1 = (uint32_t)A.rowptr[lv + 1];
      uint32_t found = 0xFFFFFFFFu;
      for (uint32_t b = e0; b < e1; b += 64) {
        const uint32_t e = b + lane;
        bool ok = false;
        if (e < e1) {
          const uint32_t u = (uint32_t)A.col[e];
          ok = S.jcost(u, li, hop ? ld - 1u : 0u) + (uint32_t)A.wt[e] == lc;
        }
        const uint64_t hb = __ballot(ok);
        if (hb) {
          found = b + (uint32_t)(__ffsll((long long)hb) - 1);
          break;
        }
      }
      if (lane == l) parc = found;
    }
    if (act && type == 3u) {
      if (sc == 0u) {
        type = 0u;  // a member of the group: leaf
      } else if (parc == 0xFFFFFFFFu) {
        type = 0u;
        broken = true;
      } else {
        type = 1u;
        pred = kPredGrow | parc;
      }
    }
    const bool isg = act && type == 1u;
    const bool ism = act && type == 2u;
    const uint64_t gb = __ballot(isg);
    const uint32_t eidx = ne + __popcll(gb & __lanemask_lt());
    uint32_t c1v = 0, c1pd = 0, c2v = 0, c2pd = 0;
    int32_t c1a = att;
    if (isg) {
      const uint32_t ea = pred & kArcMask;
      const uint32_t u = (uint32_t)A.col[ea];
      const int32_t w = A.wt[ea];
      if (eidx < (uint32_t)sp->edge_cap) {
        EdgeDev ed;
        ed.u = (int32_t)v;
        ed.v = (int32_t)u;
        ed.w = w;
        ed.parent = att;
        sp->edges[eidx] = ed;
      }
      c1v = u;
      c1pd = p | ((hop ? d - 1 : 0u) << 16);
      c1a = (int32_t)eidx;
    } else if (ism) {
      const uint32_t s1 = pred & 0xFFFFu;
      const uint32_t d1 = (pred >> 16) & 0x7Fu, d2 = (pred >> 23) & 0x7Fu;
      c1v = v;
      c1pd = s1 | (d1 << 16);
      c2v = v;
      c2pd = (p ^ s1) | (d2 << 16);
    }
    const uint32_t nch = isg ? 1u : (ism ? 2u : 0u);
    const uint32_t off = wave_excl_scan(nch);
    const uint32_t tot = __shfl(off + nch, 63, 64);
    if (tail + tot > qcap) {
      overflow = true;
      break;
    }
    if (nch >= 1) {
      qv[tail + off] = c1v;
      qpd[tail + off] = c1pd;
      qat[tail + off] = c1a;
    }
    if (nch == 2) {
      qv[tail + off + 1] = c2v;
      qpd[tail + off + 1] = c2pd;
      qat[tail + off + 1] = att;
    }
    ne += __popcll(gb);
    head += cnt;
    tail += tot;
    __threadfence();
    wave_sync();
  }
  const bool bad = __ballot(broken) != 0;
  if (lane == 0) {
    sp->out_cost = (int32_t)cost;
    sp->out_root = (int32_t)root;
    sp->out_nedges = (int32_t)ne;
    sp->out_d = dstar;
    if (overflow || ne > (uint32_t)sp->edge_cap) sp->status = kStatusEdgeCap;
    if (bad) sp->status = kStatusNoPath;
  }
}

// Finite-state count (option count_states; the reference's count_set,
// code/GPUGST/src/DPQ.cu:225-240, D-GPUGST :128-148): per vertex of a tile,
// the finite singleton states over every height (from S1 or the job's rows)
// and, at live vertices (the only ones whose T segment is initialised), the
// finite states of two or more groups.  Run by its own kernel on a finished
// query before its memory is released, so the step kernels carry none of it.
__device__ void item_count(Slot* sp, uint32_t tile, uint64_t V) {
  const int lane = lane_id();
  const uint32_t k = (uint32_t)sp->k, W = (uint32_t)sp->W, ND = (uint32_t)sp->ND, NS = (uint32_t)sp->NS;
  const uint64_t v0 = (uint64_t)tile * kDistTV;
  const uint64_t v1 = v0 + kDistTV < V ? v0 + kDistTV : V;
  // singletons: S1 words (cost in the high half) or the job's distance rows
  const uint32_t* S1 = (const uint32_t*)sp->S1;
  const uint32_t* J = sp->J;
  const uint32_t* T = (const uint32_t*)sp->T;
  uint32_t n = 0;
  for (uint64_t v = v0 + (uint64_t)lane; v < v1; v += 64) {
    for (uint32_t d = 0; d < ND; d++)
      for (uint32_t i = 0; i < k; i++) {
        const uint32_t c = S1 ? S1[2 * ((v * ND + d) * k + i) + 1] : J[v * sp->jvs + (uint64_t)d * sp->jds + i];
        n += c < kInf ? 1u : 0u;
      }
    if (!((sp->lvb[v >> 5] >> (v & 31u)) & 1u)) continue;
    for (uint32_t d = 0; d < NS; d++)
      for (uint32_t p = 3; p < W; p++)
        if (p & (p - 1u)) n += T[2 * ((v * NS + d) * W + p) + 1] < kInf ? 1u : 0u;
  }
  const uint64_t t = wave_sum_u64(n);
  if (lane == 0 && t) atomicAdd(&sp->nfinite, (unsigned long long)t);
}

// -------------------------------------------------------- state machine ----

__device__ uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Progressive answers: the candidate cost is within anum / aden of the
// certified lower bound (the candidates' full-set bound, or the distance
// between two groups, both at most the optimum), in exact integers.
__device__ uint32_t certified_lb(const Slot* s) {
  uint32_t lb = s->lbmin;
  for (int i = 0; i < s->k; i++)
    for (int j = 0; j < s->k; j++)
      if (i != j && s->vv[i * s->k + j] < kInf && s->vv[i * s->k + j] > lb) lb = s->vv[i * s->k + j];
  return lb;
}

__device__ bool approx_met(Slot* s, uint32_t cand) {
  if (!s->anum || cand >= kInf) return false;
  const uint32_t lb = certified_lb(s);
  if (lb >= kInf || (uint64_t)cand * s->aden > (uint64_t)s->anum * lb) return false;
  s->cert = lb;
  return true;
}

__device__ void advance_level(Slot* s) {
  if (s->level == 1) {
    s->phase = PH_DIST;
  } else if (s->level + 1 < s->k) {
    if (s->anum) {
      // progressive answers: a FINAL over the levels done so far first
      s->phase = PH_FINAL;
      s->probe = 1;
    } else {
      s->phase = PH_MERGE;
      s->level++;
    }
  } else {
    s->phase = PH_FINAL;
  }
}

__device__ void complete_step(Slot* slots, Slot* s, Ctrl* ctrl) {
  const uint32_t r = s->step;
  bool keep_frontier = false;
  if (s->kind == KIND_JOB) {
    switch (s->phase) {
      case PH_FILL:
        s->phase = s->ncached ? PH_CACHEIN : PH_SEED;
        break;
      case PH_CACHEIN:
        s->phase = PH_SEED;
        break;
      case PH_CACHEOUT:
        s->phase = PH_STAR;
        break;
      case PH_SEED:
        // a hop job computes heights 1..D, each a COPY and one round
        s->level = 1;
        s->phase = (s->D < 0) ? PH_ROUND : (s->D == 0 ? PH_STAR : PH_COPY);
        break;
      case PH_COPY:
        // the round's frontier is the one the previous round appended: the
        // step index (which selects the frontier buffers) stays
        s->phase = PH_ROUND;
        keep_frontier = true;
        break;
      case PH_ROUND:
        s->ctr[1] += (unsigned long long)(s->app[r % 3] >> 32);
        s->ctr[2] += (unsigned long long)(uint32_t)s->app[r % 3];
        if ((uint32_t)(s->app[r % 3] >> 32) > s->maxq) s->maxq = (uint32_t)(s->app[r % 3] >> 32);
        if (s->D < 0) {
          if ((s->app[(r + 1) % 3] >> 32) == 0) s->phase = s->nout ? PH_CACHEOUT : PH_STAR;
        } else if (s->level >= s->D) {
          s->phase = PH_STAR;
        } else {
          s->level++;
          s->phase = PH_COPY;  // every height is written, also after an empty round
        }
        break;
      case PH_STAR:
        // the members' route tables were built in this step (beside the star
        // tiles of a hop job)
        s->phase = PH_JLIVE;
        break;
      case PH_JLIVE:
        for (int m = 0; m < s->nmem; m++) slots[s->jmem[m]].jdone = 1u;
        s->phase = PH_DONE;
        break;
      default:
        break;
    }
  } else {
    switch (s->phase) {
      case PH_FILL:
        s->phase = s->S1 ? PH_SEED : PH_WAITJ;
        break;
      case PH_SEED:
        if (s->k == 1) {
          s->phase = PH_RECON;  // the answer was set while seeding
        } else {
          s->phase = PH_ROUND;
          s->level = 1;
        }
        break;
      case PH_ROUND:
        s->ctr[1] += (unsigned long long)(s->app[r % 3] >> 32);
        s->ctr[2] += (unsigned long long)(uint32_t)s->app[r % 3];
        if ((uint32_t)(s->app[r % 3] >> 32) > s->maxq) s->maxq = (uint32_t)(s->app[r % 3] >> 32);
        if ((s->app[(r + 1) % 3] >> 32) == 0) advance_level(s);
        break;
      case PH_DIST:
        s->phase = PH_LIVE;
        break;
      case PH_LIVE:
        if (approx_met(s, (uint32_t)(s->star_pk >> 32))) {
          s->recon_star = 1;  // the star is close enough to the bound
          s->phase = PH_RECON;
        } else if (s->k == 2) {
          s->phase = PH_FINAL;
        } else {
          s->phase = PH_MERGE;
          s->level = 2;
        }
        break;
      case PH_MERGE:
        if ((s->app[(r + 1) % 3] >> 32) == 0) advance_level(s);
        else s->phase = PH_ROUND;
        break;
      case PH_FINAL:
        if (s->probe) {
          s->probe = 0;
          const uint32_t a = (uint32_t)(s->ans >> 32), st = (uint32_t)(s->star_pk >> 32);
          if (approx_met(s, a < st ? a : st)) {
            s->recon_star = st < a;
            s->phase = PH_RECON;
          } else {
            s->phase = PH_MERGE;
            s->level++;
          }
        } else {
          s->cert = (uint32_t)(s->ans >> 32);  // the DP is complete: the answer is the optimum
          s->phase = PH_RECON;
        }
        break;
      case PH_RECON:
        s->phase = PH_DONE;
        atomicAdd(&ctrl->finished, 1u);
        break;
      default:
        break;
    }
  }
  s->app[(r + 2) % 3] = 0;  // cleared by this step's CLEAR items
  if (!keep_frontier) s->step = r + 1;
  s->ctr[0]++;
}

// Sets the slot's work for the next step: n_work items of class cls for its
// phase and n_clr CLEAR items (lean); returns their sum.
__device__ uint32_t plan_step(Slot* s, uint64_t V) {
  const uint32_t r = s->step;
  uint32_t work = 0;
  int cls = kClsLean;
  const bool job = s->kind == KIND_JOB;
  s->n_hv = 0;
  switch (s->phase) {
    case PH_FILL: {
      uint64_t nA, nM, nF;
      fill_regions(s, V, &nA, &nM, &nF);
      work = ceil_div(nA + nM + nF, kFillWords);
      // hub bounds keep 64 distances per lane in registers: the heavy kernel
      // has them, the job rounds' kernel does not
      if (job && s->hA) {
        work += ceil_div(s->nhubr, kUbTV) * (s->hub16 ? ub16_groups(s) : ub_groups(s)) + ceil_div(V - s->nhubr, 64u);
        cls = kClsHeavy;
      }
      break;
    }
    case PH_SEED: work = job ? (uint32_t)s->L : (uint32_t)s->k; break;
    case PH_ROUND:
      work = ceil_div((uint32_t)s->app[r % 3], (uint32_t)s->chunk);
      if (!job) cls = kClsHeavy;  // the per-query grow carries the lower-bound code
      break;
    case PH_STAR:
      if (s->hA) {
        // hub bounds: the members start from their hub stars and the job live
        // test takes the true stars, so only the route tables are built here
        work = (uint32_t)s->nmem;
        cls = kClsHeavy;
      } else {
        work = ceil_div(V, kStarTV);
        s->n_hv = (uint32_t)s->nmem;  // heavy: the route table of each member
      }
      break;
    case PH_COPY: work = ceil_div(V, kStarTV); break;
    // the route entries of 64 members stay in registers: the heavy kernel
    // has them, the job rounds' kernel does not
    case PH_JLIVE: work = ceil_div(V, kJLiveTV); cls = kClsHeavy; break;
    // cache copies run in the heavy kernel: they would push the lean one
    // (job rounds) over its register budget
    case PH_CACHEIN:
    case PH_CACHEOUT: work = ceil_div(V, kStarTV); cls = kClsHeavy; break;
    case PH_DIST: work = (s->S1 ? ceil_div(V, kDistTV) : 0u) + 1u; cls = kClsHeavy; break;
    case PH_LIVE:
      // a job-fed query lists the live bits its job set
      work = s->job >= 0 ? ceil_div(V, kLive2TV) : ceil_div(V, kDistTV);
      cls = kClsHeavy;
      break;
    case PH_MERGE: work = ceil_div(s->level >= 3 ? s->nact : s->nlive, s->mtv); cls = kClsHeavy; break;
    case PH_FINAL: work = ceil_div(s->k >= 3 ? s->nact : s->nlive, s->mtv); cls = kClsHeavy; break;
    case PH_RECON: work = 1; cls = kClsHeavy; break;
    default: return (s->n_work = s->n_clr = 0);
  }
  // every active step runs at least one (possibly empty) item so that it completes
  if (work == 0) work = 1;
  s->n_work = work;
  s->cls = cls;
  s->n_clr = ceil_div((uint32_t)(s->app[(r + 2) % 3] >> 32), kClearPer);
  return work + s->n_clr + s->n_hv;
}

// Run by the one-block plan kernel between steps.
__device__ void phase_b(Slot* slots, Ctrl* ctrl, uint64_t V, uint32_t (*sh_items)[kMaxSlots],
                        uint32_t (*sh_scan)[2][kBlock]) {
  const int tid = threadIdx.x;
  const int nslots = (int)ctrl->nslots;
  for (int i = tid; i < nslots; i += kBlock) {
    Slot* s = &slots[i];
    if (s->in_step) {
      complete_step(slots, s, ctrl);
      s->in_step = 0;
    }
  }
  // Host-set cancellation: a query asked to stop ends here, without an
  // answer.  A query fed by a group-distance job stops only once its job has
  // finished, because the job writes into its members' slots until then.
  const uint32_t* cf = ctrl->cancel;
  if (cf && __hip_atomic_load(&cf[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
    for (int i = tid; i < nslots; i += kBlock) {
      Slot* s = &slots[i];
      if (s->kind != KIND_QUERY || s->phase == PH_FREE || s->phase == PH_DONE) continue;
      if (s->job >= 0 && !s->jdone) continue;
      if (__hip_atomic_load(&cf[1 + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == 0u) continue;
      s->phase = PH_DONE;
      s->status = kStatusCancelled;
      s->out_cost = (int32_t)kInf;
      s->out_root = -1;
      s->out_nedges = 0;
      atomicAdd(&ctrl->finished, 1u);
    }
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) {
    const uint32_t steps = ++ctrl->steps;  // only this block touches it now
    const uint32_t fin = __hip_atomic_load(&ctrl->finished, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh_scan[0][0][0] = (fin >= ctrl->exit_after || steps >= ctrl->max_steps) ? 1u : 0u;
  }
  __syncthreads();
  const bool leave = sh_scan[0][0][0] != 0;
  __syncthreads();
  for (int i = tid; i < nslots; i += kBlock) {
    Slot* s = &slots[i];
    // a query whose job has finished leaves the wait (also on the last step
    // of an epoch, so that the job can be released by the host)
    // (its route table and live bits come from the job: straight to LIVE)
    if (s->phase == PH_WAITJ && s->jdone) s->phase = PH_LIVE;
    uint32_t lean = 0, heavy = 0;
    if (!leave && s->phase != PH_FREE && s->phase != PH_DONE) {
      const uint32_t items = plan_step(s, V);
      s->in_step = items ? 1u : 0u;
      if (items) {
        lean = s->n_clr + (s->cls == kClsLean ? s->n_work : 0u);
        heavy = (s->cls == kClsHeavy ? s->n_work : 0u) + s->n_hv;
      }
    }
    sh_items[kClsLean][i] = lean;
    sh_items[kClsHeavy][i] = heavy;
    if ((uint32_t)i == ctrl->trace_slot && !leave)
      ctrl->trace[(ctrl->steps - 1u) & (uint32_t)(kTraceSteps - 1)] =
          (uint32_t)s->phase | ((uint32_t)s->level << 8) | ((lean + heavy) ? 0x10000u : 0u);
  }
  __syncthreads();
  // per class: exclusive scan over slots with work, compacting them into the
  // active list.  Thread t owns slots [b0, b1); a scan within each wave, then
  // the waves' totals (kBlock / 64 of them) give every thread its offsets.
  const int per = (nslots + kBlock - 1) / kBlock;
  const int b0 = tid * per;
  const int b1 = (b0 + per) < nslots ? (b0 + per) : nslots;
  const int wid = tid >> 6;
  uint32_t xcnt[kNumCls], xact[kNumCls];
  for (int c = 0; c < kNumCls; c++) {
    uint32_t cnt = 0, act = 0;
    for (int i = b0; i < b1; i++) {
      cnt += sh_items[c][i];
      act += sh_items[c][i] ? 1u : 0u;
    }
    xcnt[c] = wave_excl_scan(cnt);
    xact[c] = wave_excl_scan(act);
    if ((tid & 63) == 63) {
      sh_scan[c][0][wid] = xcnt[c] + cnt;
      sh_scan[c][1][wid] = xact[c] + act;
    }
  }
  __syncthreads();
  uint32_t any = 0;
  for (int c = 0; c < kNumCls; c++) {
    uint32_t tc = 0, ta = 0;
    for (int w = 0; w < kBlock / 64; w++) {
      const uint32_t x = sh_scan[c][0][w], y = sh_scan[c][1][w];
      if (w < wid) {
        xcnt[c] += x;
        xact[c] += y;
      }
      tc += x;
      ta += y;
    }
    if (tid == 0) {
      ctrl->item_total[c] = tc;
      ctrl->n_active[c] = ta;
      ctrl->prefix[c][ta] = tc;
      ctrl->item_next[c] = 0;
    }
    any += ta;
  }
  if (tid == 0) ctrl->exit_flag = (leave || any == 0) ? 1u : 0u;
  for (int c = 0; c < kNumCls; c++) {
    uint32_t cs = xcnt[c], as = xact[c];
    for (int i = b0; i < b1; i++) {
      if (sh_items[c][i]) {
        ctrl->active[c][as] = (uint32_t)i;
        ctrl->prefix[c][as] = cs;
        as++;
        cs += sh_items[c][i];
      }
    }
  }
}

// One engine step of one work class for one wave: pull items of the current
// plan until they are exhausted.
template <int CLS, int WPB, class LDS>
__device__ void step_body(const GraphDev& g, Slot* slots, Ctrl* ctrl, LDS* L) {
  const int lane = lane_id();
  const uint64_t V = (uint64_t)g.V;
  if (ctrl->exit_flag) return;
  const uint32_t total = ctrl->item_total[CLS];
  // waves beyond the item count have nothing to claim: leave before touching
  // the shared counter (a small step then costs one atomic per item, not one
  // per wave of the grid)
  // (the same for the whole wave: read into a scalar register, so that the
  // loop below stays uniform)
  const uint32_t wid = __builtin_amdgcn_readfirstlane(blockIdx.x * (uint32_t)WPB + (threadIdx.x >> 6));
  if (wid >= total) return;
  const uint32_t nact = ctrl->n_active[CLS];
  const uint32_t* prefix = ctrl->prefix[CLS];
  const uint32_t* active = ctrl->active[CLS];
  // A wave claims a run of `run` consecutive items per atomic: all waves
  // claim from one counter, whose atomics serialise at its L2 bank, so a step
  // with many small items (LIVE: one per 256 vertices of every query) would
  // otherwise queue on it.  The run leaves about 8 claims per wave.
  const uint32_t nwaves = gridDim.x * (uint32_t)WPB;
  uint32_t run = total / (8u * nwaves);
  run = run < 1u ? 1u : (run > 16u ? 16u : run);
  // A wave's first run of items is its own (items wid * run ...), so a wave
  // starts without waiting on the counter; the counter hands out the runs
  // after the first nwaves ones.
  uint32_t it0 = wid * run;
  for (bool first = true;; first = false) {
    // Claim items for the whole wave.  Every lane executes the atomic (only
    // lane 0 adds) and the index is read from lane 0 into a scalar register,
    // so the claim sits under no lane-dependent branch and the loop test is
    // uniform.  With the claim under `if (lane == 0)`, the optimizer threads
    // the lane != 0 edge of the `if (lane == 0)` blocks that end several items
    // straight to the broadcast: the loop splits in two and lanes 1..63 re-run
    // the same item without lane 0, never finishing.
    if (!first) {
      const uint32_t mine = atomicAdd(&ctrl->item_next[CLS], lane == 0 ? run : 0u);
      it0 = nwaves * run + __builtin_amdgcn_readfirstlane(mine);
    }
    if (it0 >= total) break;
    const uint32_t it1 = it0 + run < total ? it0 + run : total;
    for (uint32_t it = it0; it < it1; it++) {
      uint32_t lo = 0, hi = nact;  // last active slot whose prefix <= it
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= it) lo = mid; else hi = mid;
      }
      Slot* sp = &slots[active[lo]];
      const uint32_t local = it - prefix[lo];
      if constexpr (CLS == kClsLean) {
        const uint32_t nwork = sp->cls == kClsLean ? sp->n_work : 0u;
        if (local >= nwork) {
          item_clear(sp, local - nwork, V);
          continue;
        }
        if (sp->kind == KIND_JOB) {
          switch (sp->phase) {
            case PH_FILL: item_fill(sp, local, V); break;
            case PH_SEED: item_jseed(g, sp, local); break;
            case PH_ROUND: item_jround(g, sp, local, (FrontLDS*)L); break;
            case PH_STAR: item_jstar(slots, sp, local, V); break;
            case PH_COPY: item_jcopy(sp, local, V); break;
            default: break;
          }
        } else {
          switch (sp->phase) {
            case PH_FILL: item_fill(sp, local, V); break;
            case PH_SEED: item_seed(g, sp, local); break;
            default: break;
          }
        }
      } else {
        const uint32_t nwork = sp->n_work;
        switch (sp->phase) {
          case PH_DIST:
            if (local + 1 == nwork) item_vvr(g, sp, (WaveLDS*)L);
            else item_dist(g, sp, local);
            break;
          case PH_ROUND: item_round(g, sp, local, &((WaveLDS*)L)->f); break;
          case PH_LIVE:
            if (sp->job >= 0) item_live2(sp, local, V, (WaveLDS*)L);
            else item_live(sp, local, V, (WaveLDS*)L);
            break;
          case PH_STAR: item_vvr(g, &slots[sp->jmem[local]], (WaveLDS*)L); break;  // job: a member's tables
          case PH_JLIVE: item_jlive(slots, sp, local, V); break;
          case PH_MERGE: item_merge(g, sp, local, (WaveLDS*)L); break;
          case PH_FINAL: item_final(g, sp, local, (WaveLDS*)L); break;
          case PH_RECON: item_recon(g, sp); break;
          case PH_CACHEIN: item_jcache(sp, local, V, false); break;
          case PH_FILL: item_fill_hub(g, sp, local, V); break;  // jobs with hub bounds
          case PH_CACHEOUT: item_jcache(sp, local, V, true); break;
          default: break;
        }
      }
    }
  }
}

// ------------------------------------------------- shortest-path backbone ----
// An arc (u, v, w) with a u-v path shorter than w lies on no shortest path and
// on no optimal group Steiner tree: replacing it by the shorter path, then
// breaking the cycles that closes, would leave a cheaper connected subgraph
// over the same vertices.  So the unbounded DP and the group->vertex distances
// give the same answers on the remaining arcs (the backbone).  A hop bound can
// need such an arc (the only route within D hops may be a heavy direct one),
// so hop-bounded queries keep the whole graph.
// Redundancy is proved through the exact distances d(h, .) from kBbHubs hub
// vertices: d(h, u) + d(h, v) is the length of a real u-v walk, so
// min_h d(h, u) + d(h, v) < w proves the arc redundant; an arc no hub proves
// redundant is kept.  The hub distances come from label-correcting rounds over
// all arcs, where a source whose row changed in the round before relaxes its
// row into its targets.
// the row holding arc e < n_arcs: rowptr[r] <= e < rowptr[r + 1]
__device__ __forceinline__ uint64_t bb_row_of(const int64_t* rowptr, uint64_t V, uint64_t e) {
  uint64_t lo = 0, hi = V;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((uint64_t)rowptr[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

// One hub round over the arcs of one item (kBbArcs consecutive arcs): 16
// lanes per arc, four arcs at a time; lane r of a group relaxes hub labels
// 4r .. 4r+3 of the source row into the target row.
__device__ void bb_round_item(const BbArgs& a, uint64_t item) {
  const int r = lane_id() & 15;
  const uint64_t a0 = item * kBbArcs;
  const uint64_t a1 = a0 + kBbArcs < a.n_arcs ? a0 + kBbArcs : a.n_arcs;
  uint64_t e = a0 + (uint64_t)(lane_id() >> 4);
  uint64_t s = e < a1 ? bb_row_of(a.rowptr, a.V, e) : 0;
  bool any = false;
  for (; e < a1; e += 4) {
    while ((uint64_t)a.rowptr[s + 1] <= e) s++;
    if (!a.dcur[s]) continue;
    const uint64_t v = (uint64_t)a.col[e];
    const uint32_t w = (uint32_t)a.wt[e];
    const bchunk su = ((const bchunk*)(a.dist + s * kBbHubs))[r];
    const bchunk tv = ((const bchunk*)(a.dist + v * kBbHubs))[r];
    bool im = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (su[q] >= kInf) continue;
      const uint32_t cand = su[q] + w;
      if (cand >= kInf || cand >= tv[q]) continue;
      atomicMin(a.dist + v * kBbHubs + 4 * r + q, cand);
      im = true;
    }
    if (im) {
      a.dnxt[v] = 1;
      any = true;
    }
  }
  if (any) *a.changed = 1;
}

// Arc test of one item: 16 lanes per arc, each the least d(h, u) + d(h, v)
// over its four hubs, then the least over the group against the weight.
__device__ void bb_test_item(const BbArgs& a, uint64_t item) {
  const int r = lane_id() & 15;
  const uint64_t a0 = item * kBbArcs;
  const uint64_t a1 = a0 + kBbArcs < a.n_arcs ? a0 + kBbArcs : a.n_arcs;
  const uint64_t e0 = a0 + (uint64_t)(lane_id() >> 4);
  uint64_t s = e0 < a1 ? bb_row_of(a.rowptr, a.V, e0) : 0;
  for (uint64_t base = a0; base < a1; base += 4) {
    const uint64_t e = base + (uint64_t)(lane_id() >> 4);
    const bool valid = e < a1;
    uint32_t m = 0xFFFFFFFFu, w = 0;
    if (valid) {
      while ((uint64_t)a.rowptr[s + 1] <= e) s++;
      const uint64_t v = (uint64_t)a.col[e];
      w = (uint32_t)a.wt[e];
      const bchunk su = ((const bchunk*)(a.dist + s * kBbHubs))[r];
      const bchunk tv = ((const bchunk*)(a.dist + v * kBbHubs))[r];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t x = su[q] + tv[q];  // both <= kInf: no overflow
        m = x < m ? x : m;
      }
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
      const uint32_t y = __shfl_xor(m, o, 64);
      m = y < m ? y : m;
    }
    if (valid && r == 0) a.keep[e] = m < w ? 0 : 1;
  }
}

#ifndef GST_CPU_SIM
// Backbone construction (once per graph): a fill, the hub seeds, then the
// hub rounds and the arc test as grid-stride loops over items, one wave per item.
__global__ __launch_bounds__(256) void gst_bb_fill_kernel(uint32_t* p, uint64_t n, uint32_t val) {
  const uint64_t stride = (uint64_t)gridDim.x * 256u;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) p[i] = val;
}

__global__ __launch_bounds__(64) void gst_bb_seed_kernel(uint32_t* dist, uint8_t* flags, const int32_t* hubs,
                                                         int H) {
  const int t = (int)threadIdx.x;
  if (t < H) {
    dist[(uint64_t)hubs[t] * kBbHubs + t] = 0u;
    flags[hubs[t]] = 1;
  }
}

__global__ __launch_bounds__(256) void gst_bb_round_kernel(BbArgs a, uint64_t n_items) {
  const uint64_t nw = (uint64_t)gridDim.x * 4u;
  for (uint64_t it = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6); it < n_items; it += nw) bb_round_item(a, it);
}

__global__ __launch_bounds__(256) void gst_bb_test_kernel(BbArgs a, uint64_t n_items) {
  const uint64_t nw = (uint64_t)gridDim.x * 4u;
  for (uint64_t it = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6); it < n_items; it += nw) bb_test_item(a, it);
}

// One engine step of each class: every wave pulls items until the class's
// plan is exhausted.  There is no waiting between blocks; the step ends when
// the kernels end, and the next plan kernel sees everything they wrote.
// Five waves per SIMD (96 VGPRs) without spills: held there so that code
// added to rare items cannot lower the job rounds' occupancy.
// Blocks of WPB waves (a schedule option); the claim loop is per wave, so the
// block shape sets only how waves share a CU's LDS and how the grid fills it.
template <int WPB>
__global__ __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(4))) void gst_lean_kernel(GraphDev g,
                                                                                                Slot* slots,
                                                                                                Ctrl* ctrl) {
  __shared__ FrontLDS lds[WPB];
  step_body<kClsLean, WPB>(g, slots, ctrl, &lds[threadIdx.x >> 6]);
}

// Three waves per SIMD is what its LDS allows (3 blocks of 4 waves per CU);
// the register budget is held to match (one loop-invariant 16-byte value is
// kept in scratch).
template <int WPB>
__global__ __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(3))) void gst_heavy_kernel(GraphDev g,
                                                                                                 Slot* slots,
                                                                                                 Ctrl* ctrl) {
  __shared__ WaveLDS lds[WPB];
  step_body<kClsHeavy, WPB>(g, slots, ctrl, &lds[threadIdx.x >> 6]);
}

// Between two steps, one block advances every slot's state machine and
// publishes the next step's work items.  Once it has set exit_flag it leaves
// everything untouched, so the host may enqueue more steps than are needed.
// Finite-state count of one finished query: one wave per tile of kDistTV
// vertices.
__global__ __launch_bounds__(64) void gst_count_kernel(Slot* slot, uint64_t V) {
  item_count(slot, blockIdx.x, V);
}

// Trees of finished queries gathered into one staging buffer (one block per
// tree), so the host copies them out with one transfer.
__global__ __launch_bounds__(256) void gst_gather_kernel(const GatherEnt* t, EdgeDev* dst) {
  const GatherEnt e = t[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < e.n; i += blockDim.x) dst[e.off + i] = e.src[i];
}

__global__ __launch_bounds__(kBlock) void gst_plan_kernel(Slot* slots, Ctrl* ctrl, uint64_t V) {
  __shared__ uint32_t sh_items[kNumCls][kMaxSlots];
  __shared__ uint32_t sh_scan[kNumCls][2][kBlock];
  if (ctrl->exit_flag) return;
  phase_b(slots, ctrl, V, sh_items, sh_scan);
}
#endif

}  // namespace gst

#ifndef GST_CPU_SIM
namespace {
template <int WPB>
hipError_t launch_step_wpb(int cls, const gst::GraphDev* g, gst::Slot* slots, gst::Ctrl* ctrl, int grid,
                           hipStream_t stream) {
  if (cls == gst::kClsLean)
    hipLaunchKernelGGL(gst::gst_lean_kernel<WPB>, dim3(grid), dim3(WPB * 64), 0, stream, *g, slots, ctrl);
  else
    hipLaunchKernelGGL(gst::gst_heavy_kernel<WPB>, dim3(grid), dim3(WPB * 64), 0, stream, *g, slots, ctrl);
  return hipGetLastError();
}
}  // namespace

// wpb: waves per block, one of kWaveChoices
extern "C" hipError_t gst_launch_step(int cls, int wpb, const gst::GraphDev* g, gst::Slot* slots, gst::Ctrl* ctrl,
                                      int grid, hipStream_t stream) {
  switch (wpb) {
    case 2: return launch_step_wpb<2>(cls, g, slots, ctrl, grid, stream);
    case 4: return launch_step_wpb<4>(cls, g, slots, ctrl, grid, stream);
    case 8: return launch_step_wpb<8>(cls, g, slots, ctrl, grid, stream);
    default: return hipErrorInvalidValue;
  }
}

extern "C" hipError_t gst_launch_plan(gst::Slot* slots, gst::Ctrl* ctrl, uint64_t V, hipStream_t stream) {
  hipLaunchKernelGGL(gst::gst_plan_kernel, dim3(1), dim3(gst::kBlock), 0, stream, slots, ctrl, V);
  return hipGetLastError();
}

extern "C" hipError_t gst_launch_gather(const gst::GatherEnt* t, int n, gst::EdgeDev* dst, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gst::gst_gather_kernel, dim3((unsigned)n), dim3(256), 0, stream, t, dst);
  return hipGetLastError();
}

extern "C" hipError_t gst_launch_count(gst::Slot* slot, uint64_t V, hipStream_t stream) {
  const uint32_t tiles = (uint32_t)((V + gst::kDistTV - 1) / gst::kDistTV);
  hipLaunchKernelGGL(gst::gst_count_kernel, dim3(tiles), dim3(64), 0, stream, slot, V);
  return hipGetLastError();
}

extern "C" hipError_t gst_step_occupancy(int cls, int wpb, int* blocks_per_cu) {
  const void* f = nullptr;
  switch (wpb) {
    case 2: f = cls == gst::kClsLean ? (const void*)gst::gst_lean_kernel<2> : (const void*)gst::gst_heavy_kernel<2>; break;
    case 4: f = cls == gst::kClsLean ? (const void*)gst::gst_lean_kernel<4> : (const void*)gst::gst_heavy_kernel<4>; break;
    case 8: f = cls == gst::kClsLean ? (const void*)gst::gst_lean_kernel<8> : (const void*)gst::gst_heavy_kernel<8>; break;
    default: return hipErrorInvalidValue;
  }
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, wpb * 64, 0);
}

namespace {
int bb_grid(uint64_t waves) {
  const uint64_t b = (waves + 3) / 4;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}
}  // namespace

// Backbone construction: n words of p set to val.
extern "C" hipError_t gst_bb_fill(uint32_t* p, uint64_t n, uint32_t val, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint64_t b = (n + 255) / 256;
  hipLaunchKernelGGL(gst::gst_bb_fill_kernel, dim3((unsigned)(b > 8192 ? 8192 : b)), dim3(256), 0, stream, p, n, val);
  return hipGetLastError();
}

// Hub l (l < H <= kBbHubs) at distance 0 from itself, its row marked changed.
extern "C" hipError_t gst_bb_seed(uint32_t* dist, uint8_t* flags, const int32_t* hubs, int H, hipStream_t stream) {
  if (H < 1 || H > gst::kBbHubs) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gst::gst_bb_seed_kernel, dim3(1), dim3(64), 0, stream, dist, flags, hubs, H);
  return hipGetLastError();
}

// what: 0 = one hub round, 1 = the arc test
extern "C" hipError_t gst_bb_pass(int what, const gst::BbArgs* a, hipStream_t stream) {
  const uint64_t items = (a->n_arcs + gst::kBbArcs - 1) / gst::kBbArcs;
  if (items == 0) return hipSuccess;
  if (what == 0)
    hipLaunchKernelGGL(gst::gst_bb_round_kernel, dim3(bb_grid(items)), dim3(256), 0, stream, *a, items);
  else if (what == 1)
    hipLaunchKernelGGL(gst::gst_bb_test_kernel, dim3(bb_grid(items)), dim3(256), 0, stream, *a, items);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
#endif  // GST_CPU_SIM

// ------------------------------------------------------- static scratch ----
// Device memory of the allocation-free entry point (gst_solve_static): the
// slot table, the control block and one query's state arena are carved out of
// this file-scope array, so that path calls no allocator.  Each GPU holds its
// own copy at its own address.  In host code the array's name is not a device
// address; gst_scratch_setup resolves the address once per device (with that
// device current) and every host use goes through the pointer it returns.
namespace gst {
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
}  // namespace gst

extern "C" hipError_t gst_scratch_setup(int device, void** ptr) {
  static std::mutex mu;
  static void* addr[gst::kMaxDevices] = {};
  if (device < 0 || device >= gst::kMaxDevices || !ptr) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(gst::scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidValue;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}
