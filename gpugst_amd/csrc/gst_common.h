// gst_common.h -- layouts shared by the host engine and the gfx950 kernels.
//
// State layout (one query, "slot"): T[((v * ND) + d) * W + p] is the packed
// 64-bit state (cost << 32 | pred) of the DP entry T(v, p, d) for |p| >= 2 --
// vertex-major, so the W * ND costs of one vertex form one contiguous,
// coalesced segment.  Singleton states |p| = 1 (the group->vertex distances,
// needed at every vertex) live in the compact S1[((v * ND) + d) * k + i].
// ND = D + 1 for the hop-bounded DP and 1 for the unbounded one.  The reference
// keeps 28 B / 40 B AoS nodes per state (code/GPUGST/include/DPQ.cuh:51-58,
// code/D-GPUGST/include/DPQ.cuh:51-58) and writes cost and predecessor with two
// separate, racing stores; here both live in one word updated by one atomicMin.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace gst {

constexpr uint32_t kInf = 0x3FFFFFFFu;  // cost sentinel, saturating arithmetic below 2^31
constexpr uint64_t kInfState = (uint64_t(kInf) << 32) | 0xFFFFFFFFull;
constexpr int kMaxK = 6;                // W = 2^k <= 64: one wave lane per subset
constexpr int kMaxW = 1 << kMaxK;
constexpr int kMaxND = 16;              // D <= 15
constexpr int kMaxSlots = 1024;         // concurrently admitted queries per engine
// Static device scratch of the allocation-free entry point gst_solve_static:
// its slot table, control block and one query's state arena.  Every GPU the
// library runs on holds one copy of it until the process ends.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

// pred word: bits 31..30 type; grow: bits 29..0 = index of the arc v->u in the
// row of the state's own vertex v (so u and w(u,v) are one load away during
// reconstruction); merge: bits 0..15 = sub-mask p1, 16..22 = d1, 23..29 = d2.
constexpr uint32_t kPredSeed = 0u;
constexpr uint32_t kPredGrow = 1u << 30;
constexpr uint32_t kPredMerge = 2u << 30;
// singleton state read from the distance table C (no stored predecessor):
// reconstruction finds the next edge of a shortest path by scanning the row
constexpr uint32_t kPredPath = 3u << 30;
constexpr uint32_t kArcMask = (1u << 30) - 1;

// block shape and work-item granularities of the engine kernel
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kFillWords = 16384;  // u64 words per FILL item
constexpr uint32_t kMergeTV = 64;       // default vertices per MERGE / FINAL item (option merge_tv)
// waves per block of the step kernels: a schedule option ("block"); the plan
// kernel always runs kBlock threads
constexpr int kWaveChoices[3] = {2, 4, 8};
constexpr uint32_t kDistTV = 256;       // vertices per DIST item
constexpr uint32_t kClearPer = 1024;    // mask clears per CLEAR item
constexpr int kStageWords = 1152;       // per-wave LDS staging (u64): 64*ND rows or the W*k*k route table

// Group-distance jobs: one multi-label SSSP per batch of unbounded queries.
// Labels are (query, group) pairs; the distance row of a vertex holds all of
// them (u32, padded to whole chunks), so an arc relaxation reads and writes
// coalesced chunks shared by every query of the job.  A chunk is the unit of
// the dirty bits: unbounded jobs use 16 labels (64 bytes, one HBM line per
// relaxation), hop jobs 4 (their rounds relax only the chunks the height below
// improved, and coarser bits relax more labels there).  Swept 4 / 8 / 16 for
// both kinds: 16 / 4 were fastest (batch of 256 at k = 5: 53.0 -> 48.6 ms;
// D = 4, batch of 128: 58.1 ms at 4 labels, 62 at 8, 74 at 16).
constexpr int kJobChunkFlat = 16;      // labels per chunk of an unbounded job's row
constexpr int kJobChunkHop = 4;        // ... of a hop job's row
constexpr int kJobLabels = 256;        // labels per job (at most 64 chunks)
constexpr uint32_t kJobArcs = 512;      // arcs per job ROUND item (swept: 128 / 256 / 512 / 1024)
constexpr int kJobChunksPerLane = 8;    // chunks one lane relaxes per arc (swept: 2 / 4 / 8)
constexpr uint32_t kStarTV = 64;        // vertices per job STAR item

// Shortest-path backbone, built once per graph: exact distances from kBbHubs
// hub vertices (one 256-byte row per vertex) prove arcs redundant.
constexpr int kBbHubs = 64;
constexpr uint32_t kBbArcs = 256;       // arcs per backbone work item
constexpr uint32_t kHub16Inf = 0x7FFF;  // "no bound" in the 16-bit hub distances

// work classes: each has its own kernel, so the arc-parallel items do not run
// at the register budget (and occupancy) of the merge items
constexpr int kClsLean = 0;   // FILL, SEED, job ROUND, STAR, CLEAR
constexpr int kClsHeavy = 1;  // query ROUND, DIST, LIVE, MERGE, FINAL, RECON
constexpr int kNumCls = 2;
constexpr int kTraceSteps = 4096;  // steps of one epoch whose phase can be logged

enum Phase : int32_t {
  PH_FREE = 0,
  PH_FILL = 1,   // initialise T, masks, distance table
  PH_SEED = 2,   // T(v, {i}, 0) = 0 for members of query group i
  PH_ROUND = 3,  // one frontier round of grow (label-correcting) at level L
  PH_DIST = 4,   // group->vertex distance table + inter-group distances + route table
  PH_MERGE = 5,  // level-L merge pass over every vertex segment
  PH_FINAL = 6,  // full-set merge + argmin
  PH_RECON = 7,  // predecessor chase -> tree edges
  PH_DONE = 8,
  PH_LIVE = 9,   // live-vertex list + initialisation of their T segments
  PH_WAITJ = 10, // query waiting for its group-distance job (no work items)
  PH_STAR = 11,    // job: every member's route table (and, for a hop job, its star bound)
  PH_COPY = 12,    // hop job: height level starts as a copy of height level - 1
  PH_JLIVE = 13,   // job: live bits of every member query (one pass over its rows)
  PH_CACHEIN = 14, // job: copy cached group-distance columns into its rows
  PH_CACHEOUT = 15 // job: copy its new labels' distances out into cache columns
};

// Slot::status values set on the device
constexpr uint32_t kStatusEdgeCap = 1;    // tree larger than the edge buffer
constexpr uint32_t kStatusNoPath = 2;     // reconstruction found no shortest-path arc
constexpr uint32_t kStatusCancelled = 3;  // the host cancelled the query (request or deadline)

enum SlotKind : int32_t {
  KIND_QUERY = 0,
  KIND_JOB = 1
};

// The graph's arcs are followed by those of its shortest-path backbone (the
// arcs (u, v, w) with no u-v path shorter than w, same row order): col / wt /
// rev hold n_arcs + backbone arcs, rowptr holds the graph's V + 1 row starts,
// then the backbone's (into the same arrays, so from n_arcs on).  A slot with
// Slot::bb set (unbounded queries and jobs) reads rows through the second
// half of rowptr; its arc indices (grow predecessors, frontier rows) are
// indices into the same col / wt / rev.
struct GraphDev {
  const int64_t* rowptr;  // V + 1 (+ V + 1 of the backbone)
  const int32_t* col;     // n_arcs (+ backbone arcs)
  const int32_t* wt;      // as col
  const int32_t* rev;     // as col: index of the reverse arc
  const int64_t* gptr;    // n_groups + 1
  const int32_t* gmem;    // members
  // exact distances d(h, v) from the kBbHubs hub vertices of the backbone
  // build, V rows of kBbHubs (null: none)
  const uint32_t* hubd;
  // the same distances as 16-bit words, kHub16Inf for unreached, when every
  // finite one is below kHub16Inf / 2 (so two of them add without overflow);
  // null otherwise
  const uint16_t* hubd16;
  // vertices by hub reach: the n_hubr vertices some hub reaches, then the rest
  // (jobs compute hub bounds for the first ones only; the others' are
  // infinite)
  const uint32_t* hubv;
  int64_t n_hubr;
  int64_t V;
  int64_t n_arcs;
};

// Arguments of the backbone construction passes (hub rounds, arc test).
struct BbArgs {
  const int64_t* rowptr;
  const int32_t* col;
  const int32_t* wt;
  uint64_t V, n_arcs;
  uint32_t* dist;       // V rows of kBbHubs labels (256 bytes)
  const uint8_t* dcur;  // per vertex: its row changed in the round before
  uint8_t* dnxt;        // per vertex: its row changed in this round
  uint32_t* changed;    // set when any row changed in this round
  uint8_t* keep;        // per arc: 0 when proved redundant
};

struct EdgeDev {
  int32_t u, v, w, parent;
};

// One finished query's tree for the gather kernel: n edges from src to
// position off of the staging buffer.
struct GatherEnt {
  const EdgeDev* src;
  uint64_t off;
  uint32_t n, pad;
};

struct alignas(64) Slot {
  // ---- query or job (set at admission) ----
  int32_t kind;         // KIND_QUERY or KIND_JOB
  int32_t job;          // query fed by a job: its slot index, else -1
  int32_t lab;          // query fed by a job: its first label in the job
  uint32_t jdone;       // set by the job when this query's C is complete
  int32_t k, W, ND, D;  // D < 0: unbounded (ND == 1); ND = D + 1 heights of S1 / C / job rows
  int32_t NS, pad5;     // heights of the states T: floor(D/2) + 1 rounded up (1 unbounded)
  int32_t qg[kMaxK];
  int32_t edge_cap;
  int32_t chunk;        // arcs per ROUND item
  uint32_t istride;     // entries per buffer of istart
  int32_t L, Lp, G;     // job: labels, padded row length, lanes per arc (power of two)
  int32_t nmem;         // job: member queries
  uint32_t cstride;     // words per vertex of C (ND * k for an own table, the job's row stride)
  uint32_t chs;         // words between the heights of one vertex in C (k, or the job's Lp)
  uint64_t nT;          // V * ND * W
  uint64_t* T;          // states with |p| >= 2; only live vertices' segments are touched
  uint64_t* S1;         // V * ND * k singleton states T(v, {i}, d); null when C comes from a job
  uint32_t* istart;     // 3 * istride: frontier index holding the first arc of each ROUND item
  uint32_t* dist;       // job: V * ND * Lp distances (per vertex, one row of Lp labels per height)
  int32_t* lgrp;        // job: group id of each label
  uint32_t* jmem;       // job: member slot indices
  unsigned long long nsrc;  // job ROUND: source chunks read (the dirty chunks of its frontier entries)
  uint32_t jcl, pad3;    // job: labels per dirty chunk (kJobChunkFlat or kJobChunkHop)
  uint32_t* live;       // V: vertices that can lie on a tree cheaper than the bound
  uint32_t* lvb;        // V bits: live flag per vertex
  uint32_t* hold;       // V bits: the vertex holds a finite state with |p| >= 2
  uint32_t* act;        // V: vertices with their hold bit set, in order of setting
  uint64_t* mask;       // 3 * V (triple-buffered dirty sub-masks)
  uint32_t* C;          // group->vertex distances within h hops, C[v * cstride + h * chs + i]:
                        // the query's own V * ND * k table, or its k columns of its job's
                        // rows (a hop job turns them cumulative at STAR)
  uint32_t* J;          // job-fed query: its first label in the job's rows (null otherwise);
                        // singleton T(v, {i}, d) costs J[v * jvs + d * jds + i]
  uint32_t jvs, jds;    // job-fed query: row stride per vertex and per height
  uint32_t* list;       // 3 * V frontier vertices
  uint32_t* loff;       // 3 * V arc offsets of frontier entries
  uint32_t* R;          // W * k * k route table
  EdgeDev* edges;       // edge_cap
  uint32_t* rq;         // 3 * (edge_cap + 2 * kMaxK) reconstruction queue
  // ---- engine state ----
  int32_t phase, level;
  uint32_t step, in_step;
  uint32_t n_work, n_clr;   // this step's phase items and CLEAR items
  int32_t cls;              // work class of the phase items
  uint32_t n_hv;            // job STAR: heavy items beside the lean ones (a route table per member)
  uint64_t app[3];      // (count << 32) | arcs of the three frontier buffers
  uint32_t best;        // upper bound on the optimum (cost of a real walk-tree)
  uint32_t status;
  uint32_t nlive;
  uint32_t nact;        // entries of act
  uint64_t ans;         // (cost << 32) | root vertex
  uint32_t vv[kMaxK * kMaxK];
  int32_t out_cost, out_root, out_nedges, out_d;
  unsigned long long ctr[6];
  // ---- fields read off the hot paths (kept after the others) ----
  uint32_t maxq;        // largest frontier (vertices) of one of its rounds
  uint32_t count;       // 1: count its finite states when it has finished
  unsigned long long nfinite;  // finite states (singletons over all heights + T)
  // job: 2 * L column pointers; entry l (l < L) is the cache column of label l
  // to copy in (null: computed by the job), entry L + l the column its
  // distances are copied out to (null: not cached)
  uint32_t* const* ccol;
  uint32_t ncached;     // job: labels whose distances come from cache columns (not seeded)
  uint32_t nout;        // job: labels copied out to the cache
  // progressive answers: stop once a walk-tree within anum / aden of a
  // certified lower bound is known (anum == 0: exact)
  uint32_t anum, aden;
  uint32_t lbmin;       // LIVE: min over candidate vertices of the full-set lower bound
  uint32_t cert;        // certified lower bound on the optimum of the answer given
  unsigned long long star_pk;  // LIVE: cheapest star, (cost << 32) | centre vertex
  uint32_t probe;       // the FINAL under way is a checkpoint between levels
  uint32_t recon_star;  // RECON rebuilds the star of star_pk (shortest paths from its centre)
  uint32_t mtv;         // live vertices per MERGE / FINAL item (schedule option merge_tv)
  uint32_t bb;          // 1: the slot's arcs are the graph's shortest-path backbone
  uint32_t nhubr;       // job with hub bounds: vertices some hub reaches (GraphDev::n_hubr)
  // unbounded job: hub bounds of its labels, hA[l * kBbHubs + h] = distance
  // from hub h to the nearest member of label l's group (null: its rows start
  // at infinity)
  const uint32_t* hA;
  // 1: the job's hub FILL uses GraphDev::hubd16 (two labels per lane in
  // packed 16-bit arithmetic)
  uint32_t hub16;
};

struct alignas(64) Ctrl {
  uint32_t item_next[kNumCls];
  uint32_t item_total[kNumCls];
  uint32_t n_active[kNumCls];
  uint32_t exit_flag;
  uint32_t finished;
  uint32_t steps;
  uint32_t exit_after;
  uint32_t max_steps;
  uint32_t nslots;
  uint32_t trace_slot;  // slot whose planned phase is logged per step (~0u: none)
  uint32_t pad0;
  // Host-set cancellation, read by the plan kernel every step: host memory
  // the device can read (null: none).  Word 0 is nonzero once any flag is
  // set; word 1 + i asks slot i to stop.
  const uint32_t* cancel;
  uint32_t active[kNumCls][kMaxSlots];
  uint32_t prefix[kNumCls][kMaxSlots + 1];
  uint32_t trace[kTraceSteps];  // phase | level << 8 | planned << 16, by step of the epoch
};

}  // namespace gst
