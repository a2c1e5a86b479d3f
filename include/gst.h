/*
 * gst.h -- C ABI of the MI355X group Steiner tree engine (libgst.so).
 *
 * Drop-in boundary for the reference's DPBF query path.  The reference has
 * no library boundary; its per-query entry points are C++ functions taking
 * host graph containers by reference:
 *
 *   DPBF_GPU(node**, node*, CSR_graph&, std::vector<int>& query_groups,
 *            graph_v_of_v_idealID& group_graph, graph_v_of_v_idealID&,
 *            int& real_cost, non_overlapped_group_sets, double& rt, int& RAM,
 *            records&)                        code/GPUGST/include/DPQ.cuh:63
 *   DPBF_gpu(CSR_graph&, std::vector<int>&, graph_v_of_v_idealID&,
 *            graph_v_of_v_idealID&, int D, double*, int& real_cost, int& RAM,
 *            records&)                        code/D-GPUGST/include/DPQ.cuh:63
 *   main() of GPUGST+ / D-GPUGST+ driving balanced_push per query
 *                                             code/GPUGST+/src/GSTnonHop.cu:161-510
 *                                             code/D-GPUGST+/src/GPUHop.cu:95-268
 *
 * Every function returns a gst_status.  An infeasible query is not an error:
 * it is reported with cost == GST_INFEASIBLE (the reference returns 0 or
 * 100000 instead, code/GPUGST/src/DPQ.cu:571-575, code/GPUGST+/src/
 * GSTnonHop.cu:435-440).  No C++ types cross this boundary.
 *
 * Threading: a graph handle is immutable after gst_groups_set and may be
 * shared by threads; device work from one graph handle is serialised (one
 * engine at a time), so concurrency comes from gst_solve_batch and the
 * server, which advance many queries in every engine step.
 */
#ifndef GST_H_
#define GST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GST_OK = 0,
  GST_INVALID_ARG = 1,
  GST_OUT_OF_MEMORY = 2,
  GST_DEVICE_ERROR = 3,
  GST_CANCELLED = 4,
  GST_NOT_READY = 5,
  GST_QUEUE_FULL = 6
} gst_status;

/* cost sentinel of an infeasible query (no tree covers every group) */
#define GST_INFEASIBLE 0x3FFFFFFF
/* D value meaning "no diameter bound" */
#define GST_UNBOUNDED (-1)
/* largest supported number of query groups and hop bound */
#define GST_MAX_K 6
#define GST_MAX_D 15

typedef struct gst_graph gst_graph;
typedef struct gst_server gst_server;

/* One tree edge.  Edges are listed so that each edge's parent copy is known:
 * `parent` is the index of the edge whose child end is this edge's parent
 * end (u), or -1 when u is the root.  For the hop-bounded DP the optimum is
 * a tree-like walk; a vertex may then appear in two branches, and this
 * structure still identifies the walk's tree shape. */
typedef struct {
  int32_t u;      /* parent vertex (closer to the root) */
  int32_t v;      /* child vertex */
  int32_t w;      /* edge weight */
  int32_t parent; /* index of the parent edge, -1 if u is the root */
} gst_edge;

/* Per-query device counters.  The reference's `records{process_queue_num,
 * counts}` (code/GPUGST/include/DPQ.cuh:59-62) map to frontier_entries and
 * states_finite (the reference counts the finite states of its table with
 * count_set after the main loop, code/GPUGST/src/DPQ.cu:545-551; here the
 * same count is an optional pass, option "count_states").  Its RAM figure,
 * finite states + largest queue + V * 2^k (DPQ.cu:549), is
 * states_finite + max_frontier + V << k.  The first five count the query's
 * own work; a query whose level 1 came from a shared group-distance job
 * (unbounded, k >= 2) also gets its share of that job (per label) in the
 * shared_ fields. */
typedef struct {
  int64_t steps;            /* engine steps (rounds + merges + phases), incl. its job's */
  int64_t frontier_entries; /* frontier vertices processed */
  int64_t arcs;             /* CSR arcs scanned */
  int64_t relaxations;      /* (arc, subset, hop) candidates evaluated */
  int64_t improvements;     /* successful packed atomicMin updates */
  int64_t merge_evals;      /* (split, hop pair) merge candidates evaluated */
  int64_t live_vertices;    /* vertices whose full-set lower bound admits a tree <= the bound */
  int64_t shared_entries;   /* share of its job's frontier entries */
  int64_t shared_arcs;      /* share of its job's arcs */
  int64_t shared_relaxations; /* share of its job's label relaxations */
  int64_t active_vertices;  /* live vertices that received a state of two or more groups */
  int64_t max_frontier;     /* largest frontier of one of its rounds (its job's included), vertices */
  int64_t states_finite;    /* finite DP states (v, p[, d]) at the end: singletons over every height
                               plus the stored states of two or more groups; -1 unless the option
                               "count_states" was set when the query was admitted */
  int64_t cached_labels;    /* its groups whose distances came from the group-distance cache */
  int64_t lower_bound;      /* certified lower bound on the optimum: the cost itself when the
                               answer is optimal; with a progressive ratio num / den the answer
                               satisfies cost * den <= num * lower_bound (GST_INFEASIBLE when
                               no tree exists) */
} gst_counters;

typedef struct {
  int32_t cost;     /* optimal cost, or GST_INFEASIBLE (also when the query was cancelled) */
  int32_t root;     /* root vertex of the returned tree, -1 if infeasible */
  int32_t n_edges;  /* number of records in edges */
  int32_t status;   /* gst_status of this query (GST_CANCELLED: stopped by gst_cancel, a
                       deadline or gst_server_destroy, no answer) */
  gst_edge *edges;  /* owned by the library until gst_result_free */
  gst_counters counters;
  double latency_ms; /* submit -> result available (batch and server) */
  double device_ms;  /* admission -> completion on the device */
} gst_result;

typedef struct {
  const int32_t *group_ids; /* k distinct-or-not group ids (indices into the groups set) */
  int32_t k;                /* 1 .. GST_MAX_K */
  int32_t D;                /* GST_UNBOUNDED or 0 .. GST_MAX_D */
} gst_query;

/* Graph: CSR with int64 row pointers beg_pos[V+1] and int32 adjacency and
 * weights of n_arcs arcs; undirected, both arcs present with equal weights
 * (code/GPUGST+/include/graph.hpp:12-118 loads the same three arrays from
 * X_beg_pos.bin / X_csr.bin / X_weight.bin).  Weights must be in [1, 2^20].
 * The library copies the arrays to the device and keeps them resident.
 * It also builds the graph's shortest-path backbone once, on the device: the
 * arcs (u, v, w) that no u-v path shorter than w is known to replace (exact
 * distances from the 64 vertices of highest degree prove the others
 * redundant).  No shortest path and no optimal group Steiner tree uses a
 * redundant arc, so queries without a hop bound run on the backbone with the
 * same costs (option "backbone"); hop-bounded queries use every arc.  The
 * reference runs every query on the full CSR (code/GPUGST+/src/
 * GSTnonHop.cu:168-180). */
gst_status gst_graph_create(const int64_t *beg_pos, int64_t V,
                            const int32_t *adj, const int32_t *w,
                            int64_t n_arcs, gst_graph **out);
gst_status gst_graph_destroy(gst_graph *g);

/* Copies out the backbone: beg_pos_out[V+1] and, when non-null, adj_out /
 * w_out of *n_arcs_out arcs (the count alone when they are null).  A graph
 * without a backbone (no arcs, or no device memory for its build) reports
 * every arc.  No counterpart in the reference (it has no backbone). */
gst_status gst_backbone_get(gst_graph *g, int64_t *beg_pos_out, int32_t *adj_out,
                            int32_t *w_out, int64_t *n_arcs_out);

/* Groups: group g's members are members[grp_ptr[g] .. grp_ptr[g+1]), as read
 * from the .g file line "g{g+1}:" (code/GPUGST/include/
 * graph_hash_of_mixed_weighted_read_for_GSTP.h:190-239).  Replaces any
 * earlier set. */
gst_status gst_groups_set(gst_graph *g, const int64_t *grp_ptr,
                          const int32_t *members, int64_t n_groups);

/* One query, one call: the plain one-query-at-a-time path (the role of
 * DPBF_GPU / DPBF_gpu).  D = GST_UNBOUNDED for the unbounded DP. */
gst_status gst_solve(gst_graph *g, const int32_t *group_ids, int32_t k,
                     int32_t D, gst_result *out);

/* gst_solve without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * query's state lives in a static device array of the library (64 MiB per
 * GPU, held until the process ends; a query whose state needs more -- about
 * V * 2^k * 8 B plus ~80 B per vertex, times the state heights of a hop
 * bound -- is rejected with GST_OUT_OF_MEMORY).  The tree is written to the
 * caller's `edges` (edge_cap records; GST_INVALID_ARG if the tree has more,
 * with out->n_edges set to its size); out->edges stays NULL.  Calls on one
 * GPU share that array and are serialised by the library; it keeps the last
 * call's contents, which the next call re-initialises.  Every query runs its
 * own group-distance pass (no shared job). */
gst_status gst_solve_static(gst_graph *g, const int32_t *group_ids, int32_t k,
                            int32_t D, gst_edge *edges, int32_t edge_cap,
                            gst_result *out);

/* Many queries advanced together in shared engine steps, with memory-based
 * admission (continuous batching).  out[i] answers q[i]. */
gst_status gst_solve_batch(gst_graph *g, int64_t n, const gst_query *q,
                           gst_result *out);

void gst_result_free(gst_result *r);

/* Tunables: "max_slots", "budget_mb", "blocks_per_cu", "chunk",
 * "edge_cap", "timeout_ms" (host-side deadline for one chunk of engine steps;
 * on expiry the call returns GST_DEVICE_ERROR and the engine refuses work
 * until the device has finished those steps), "deadline_ms" (0: none; a
 * query of gst_solve / gst_solve_batch / gst_submit still running this long
 * after its submission is stopped on the device at its next step and
 * answered with status GST_CANCELLED), "count_states" (1: fill
 * counters.states_finite with an extra pass), "approx_num" and
 * "approx_den" (progressive answers, default 1 / 1 = exact: a query stops
 * as soon as the cheapest tree it has found costs at most num / den times a
 * certified lower bound -- checked in exact integers after its group
 * distances and after each level of the DP -- and returns that tree with the
 * bound in counters.lower_bound; the reference's maximum_return_app_ratio
 * early exit, code/PrunedDP++/include/CPUNONHOP.h:730-755, whose int
 * division is replaced by this exact test), "cache_mb" (device memory of
 * an engine's group-distance cache: columns of exact distances of one group
 * to every vertex, kept across batches and server epochs, least recently used
 * first out, evicted before admission fails; 0: no cache; used by queries
 * without a hop bound that share a group-distance job), "graph" (replay engine steps as a hipGraph),
 * "group_pass" (1: unbounded queries share group-distance jobs; 0: each
 * query runs its own level 1), "job_min_queries" (default 2: fewer
 * unbounded queries admitted together run their own level 1 instead of a
 * job), "time_steps" (time every step's kernels),
 * "trace_slot" (with time_steps: the engine slot whose phase each step's
 * time is binned by; slot 0 is the query of a one-query call),
 * "backbone" (default 1: queries without a hop bound admitted from then on
 * run on the graph's shortest-path backbone; 0: on every arc), "hub_bounds"
 * (default 1: group-distance jobs without a hop bound start their rows at
 * bounds through the backbone's hub vertices; 0: at infinity), "hub16"
 * (default 1: those bounds in packed 16-bit arithmetic when the graph's hub
 * distances allow; 0: 32-bit), "serve_chunk", "serve_depth",
 * "serve_wait_us", "serve_batch" (the server, see gst_server_create).
 * Read-only: "num_cu", "backbone_arcs", "backbone_rounds" (hub rounds of
 * its build), "backbone_us" (build time), "hub_bytes" (bytes per hub distance
 * a job's hub FILL reads: 2 in packed 16-bit arithmetic, else 4), "occupancy", "occupancy_heavy", "grid",
 * "grid_heavy", "{plain,batch}_{kernel_ns,launches,steps,lean_ns,heavy_ns,
 * timed_steps,jobs,job_entries,job_arcs,job_relaxations,job_improvements,
 * job_atomics,job_src_chunks,cache_hits,cache_misses,cache_inserts,cache_evictions,
 * cache_entries}", "{plain,batch}_phase_ns:<phase>:<level>" and
 * "{plain,batch}_phase_steps:<phase>:<level>" (step time and step count of
 * the traced slot per engine phase).
 * "reset_stats" (set) clears the statistics.  Unknown keys return
 * GST_INVALID_ARG. */
gst_status gst_set_option(gst_graph *g, const char *key, int64_t value);
gst_status gst_get_option(gst_graph *g, const char *key, int64_t *value);

/* Device bytes one query of (k, D) needs while admitted. */
gst_status gst_query_bytes(gst_graph *g, int32_t k, int32_t D, int64_t *bytes);

/* Server: a background thread keeps the graph resident and admits submitted
 * queries into the running engine as memory and slots free up.  The engine
 * steps continuously: between chunks of "serve_chunk" engine steps (option,
 * default 4; "serve_depth" chunks, default 2, stay queued on the device) the
 * thread uploads newly admitted queries, which the running step plan takes up
 * at its next step, and copies finished answers out.  While the device is
 * busy, admission may wait up to "serve_wait_us" microseconds for
 * "serve_batch" pending requests, so that they share a group-distance job
 * (defaults 10000 and 64; 0 and 1 admit every request at once). */
gst_status gst_server_create(gst_graph *g, int64_t budget_bytes,
                             int32_t max_slots, gst_server **out);
gst_status gst_submit(gst_server *s, const gst_query *q, uint64_t *ticket);
/* GST_NOT_READY while pending; on GST_OK *out owns the result. */
gst_status gst_poll(gst_server *s, uint64_t ticket, gst_result *out);
gst_status gst_wait(gst_server *s, uint64_t ticket, gst_result *out);
/* Cancels a request: a queued one is answered at once, a running one is
 * stopped by the engine at its next step (a request fed by a shared
 * group-distance job once that job is done); its result has status
 * GST_CANCELLED unless it finished first.  No reference counterpart (the
 * reference has no request loop). */
gst_status gst_cancel(gst_server *s, uint64_t ticket);
/* Device statistics of the server: "steps" (engine steps run), "chunks"
 * (chunks collected), "busy_us" (their summed device time).  No reference
 * counterpart (monitoring of the request loop). */
gst_status gst_server_stat(gst_server *s, const char *key, int64_t *value);
/* Answers every queued request with GST_CANCELLED, cancels the running ones,
 * wakes every gst_wait and returns once they have returned; no call on s may
 * start after it. */
gst_status gst_server_destroy(gst_server *s);

/* Thread-local text of the last error. */
const char *gst_last_error(void);
const char *gst_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GST_H_ */
