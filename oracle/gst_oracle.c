/*
 * gst_oracle.c -- CPU restatement of the reference's exact group-Steiner-tree
 * dynamic programs.  TEST INFRASTRUCTURE ONLY: this file is compiled into
 * oracle/_build/libgst_oracle.so and is loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, always as
 * the checker.  The product path (gpugst_amd/) never links or calls it.
 *
 * What it restates (reference paths relative to the GPUGST repository):
 *
 *  orc_dpbf(D < 0)  -- unpruned DPBF over states (v, p): best-first (Dijkstra)
 *      order, every subset of a member's group mask seeded at cost 0
 *      (code/PrunedDP++/include/CPUNONHOP.h:594-624), grow over every arc
 *      (:820-894), merge with every disjoint partner (:898-983), first popped
 *      full-set state is the optimum (:658-678), tree rebuilt from
 *      back-pointers (:390-432).  PrunedDP++'s pruning rules and its
 *      approximation-ratio early exit (:730-755) are deliberately absent: the
 *      oracle is the exact optimum.
 *  orc_dpbf(D >= 0) -- the hop-bounded DP of code/D-PrunedDP++/include/
 *      CPUHOP.h:107-288: states (v, p, d), singleton seeds at d = 0
 *      (:158-170), lazy-deletion Dijkstra (:174-185), grow to d+1 only while
 *      d < D (:192-206), merge (v,p1,d1)+(v,p2,d2) -> (v, p1|p2, max(d1,d2))
 *      for d2 <= D - d1 (:208-227), first popped full state wins (:186-190),
 *      BFS back-tracking of (type, u | p1, p2, d1, d2) (:234-259).
 *  orc_sssp         -- multi-source Dijkstra, the group-to-vertex distance of
 *      code/PrunedDP++/include/CPUNONHOP.h:21-80 (dummy source joined to every
 *      member by 0-weight arcs, then graph_v_of_v_idealID_shortest_paths.h:
 *      24-81).
 *  orc_steiner_dw   -- Dreyfus-Wagner Steiner tree (textbook; not in the
 *      reference), used only as an independent cross-check through the
 *      GST -> Steiner reduction.
 *
 * Conventions shared with the product C-ABI: CSR with int64 row pointers and
 * int32 columns / weights, both arc directions present; group g's members are
 * gmem[gptr[g] .. gptr[g+1]); bit i of a subset mask is the i-th query group.
 * Costs are int32; GST_ORC_INF marks "no tree".  Tree edges are returned as
 * (parent vertex, child vertex, weight, parent edge index) records, parent
 * edge index -1 meaning the child hangs off the root copy.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int64_t orc_now_ns(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (int64_t)ts.tv_sec * 1000000000 + ts.tv_nsec;
}

#define GST_ORC_INF 0x3FFFFFFF

typedef struct {
  int32_t u, v, w, parent;
} orc_edge;

/* ---------------- binary min-heap of packed (cost << 34 | state) keys ----- */
typedef struct {
  uint64_t *a;
  size_t n, cap;
} heap_t;

static int heap_push(heap_t *h, uint64_t key) {
  if (h->n == h->cap) {
    size_t nc = h->cap ? h->cap * 2 : 1024;
    uint64_t *na = (uint64_t *)realloc(h->a, nc * sizeof(uint64_t));
    if (!na) return -1;
    h->a = na;
    h->cap = nc;
  }
  size_t i = h->n++;
  while (i > 0) {
    size_t par = (i - 1) >> 1;
    if (h->a[par] <= key) break;
    h->a[i] = h->a[par];
    i = par;
  }
  h->a[i] = key;
  return 0;
}

static uint64_t heap_pop(heap_t *h) {
  uint64_t top = h->a[0];
  uint64_t last = h->a[--h->n];
  size_t i = 0, n = h->n;
  for (;;) {
    size_t l = 2 * i + 1;
    if (l >= n) break;
    size_t c = (l + 1 < n && h->a[l + 1] < h->a[l]) ? l + 1 : l;
    if (h->a[c] >= last) break;
    h->a[i] = h->a[c];
    i = c;
  }
  if (n) h->a[i] = last;
  return top;
}

#define STATE_BITS 34
#define STATE_MASK ((1ULL << STATE_BITS) - 1)

/* ---------------- multi-source Dijkstra ------------------------------------ */
int orc_sssp(int64_t V, const int64_t *rowptr, const int32_t *col,
             const int32_t *wt, int64_t nsrc, const int32_t *src,
             int32_t *dist) {
  heap_t h = {0, 0, 0};
  for (int64_t v = 0; v < V; v++) dist[v] = GST_ORC_INF;
  for (int64_t i = 0; i < nsrc; i++) {
    int32_t s = src[i];
    if (s < 0 || s >= V) return -2;
    if (dist[s] != 0) {
      dist[s] = 0;
      if (heap_push(&h, (uint64_t)s)) return -1;
    }
  }
  while (h.n) {
    uint64_t key = heap_pop(&h);
    int64_t v = (int64_t)(key & STATE_MASK);
    int32_t c = (int32_t)(key >> STATE_BITS);
    if (c != dist[v]) continue;
    for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
      int32_t u = col[e];
      int64_t nc = (int64_t)c + wt[e];
      if (nc < dist[u]) {
        dist[u] = (int32_t)nc;
        if (heap_push(&h, ((uint64_t)nc << STATE_BITS) | (uint64_t)u)) return -1;
      }
    }
  }
  free(h.a);
  return 0;
}

/* ---------------- exact DPBF (unbounded and hop-bounded) ------------------- */
/* pred encoding: type 0 seed; type 1 grow from vertex a over arc b (arc a->v
 * in a's row); type 2 merge of (v, a, d1) and (v, p ^ a, d2) with b = d1<<16|d2. */
typedef struct {
  int32_t a, b;
  int8_t type;
} pred_t;

int orc_dpbf(int64_t V, const int64_t *rowptr, const int32_t *col,
             const int32_t *wt, int32_t k, const int64_t *gptr,
             const int32_t *gmem, int32_t D, int32_t *cost_out,
             int32_t *root_out, orc_edge *edges, int32_t edge_cap,
             int32_t *n_edges, int64_t *stats) {
  if (k < 1 || k > 16 || V <= 0) return -2;
  const int hop = D >= 0;
  const int64_t ND = hop ? (int64_t)D + 1 : 1;
  const int64_t W = (int64_t)1 << k;
  const int32_t full = (int32_t)(W - 1);
  const int64_t NS = V * ND * W;
  if ((uint64_t)NS > STATE_MASK) return -2;
  *cost_out = GST_ORC_INF;
  *root_out = -1;
  *n_edges = 0;

  int32_t *cost = (int32_t *)malloc((size_t)NS * sizeof(int32_t));
  pred_t *pred = (pred_t *)malloc((size_t)NS * sizeof(pred_t));
  uint32_t *gmask = (uint32_t *)calloc((size_t)V, sizeof(uint32_t));
  if (!cost || !pred || !gmask) {
    free(cost); free(pred); free(gmask);
    return -1;
  }
  for (int64_t s = 0; s < NS; s++) cost[s] = GST_ORC_INF;
#define SIDX(v, p, d) ((((int64_t)(v) * ND) + (d)) * W + (p))

  for (int i = 0; i < k; i++)
    for (int64_t j = gptr[i]; j < gptr[i + 1]; j++) {
      int32_t v = gmem[j];
      if (v < 0 || v >= V) {
        free(cost); free(pred); free(gmask);
        return -2;
      }
      gmask[v] |= 1u << i;
    }

  heap_t h = {0, 0, 0};
  int64_t pops = 0, pushes = 0;
  for (int64_t v = 0; v < V; v++) {
    uint32_t m = gmask[v];
    if (!m) continue;
    if (!hop) {
      /* every non-empty subset of the member's mask (CPUNONHOP.h:594-624) */
      for (uint32_t p = m; p; p = (p - 1) & m) {
        int64_t s = SIDX(v, p, 0);
        cost[s] = 0;
        pred[s].type = 0;
        heap_push(&h, (uint64_t)s);
        pushes++;
      }
    } else {
      /* singletons at d = 0 (CPUHOP.h:158-170) */
      for (int i = 0; i < k; i++)
        if (m & (1u << i)) {
          int64_t s = SIDX(v, 1u << i, 0);
          cost[s] = 0;
          pred[s].type = 0;
          heap_push(&h, (uint64_t)s);
          pushes++;
        }
    }
  }

  int64_t root_state = -1;
  while (h.n) {
    uint64_t key = heap_pop(&h);
    int64_t s = (int64_t)(key & STATE_MASK);
    int32_t c = (int32_t)(key >> STATE_BITS);
    if (c != cost[s]) continue; /* lazy deletion */
    pops++;
    int32_t p = (int32_t)(s % W);
    int64_t vd = s / W;
    int32_t d = (int32_t)(vd % ND);
    int64_t v = vd / ND;
    if (p == full) {
      root_state = s;
      break;
    }
    /* grow */
    if (!hop || d < D) {
      int32_t nd = hop ? d + 1 : 0;
      for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
        int32_t u = col[e];
        int64_t nc = (int64_t)c + wt[e];
        int64_t t = SIDX(u, p, nd);
        if (nc < cost[t]) {
          cost[t] = (int32_t)nc;
          pred[t].type = 1;
          pred[t].a = (int32_t)v;
          pred[t].b = (int32_t)e;
          heap_push(&h, ((uint64_t)nc << STATE_BITS) | (uint64_t)t);
          pushes++;
        }
      }
    }
    /* merge with every disjoint non-empty partner */
    int32_t comp = full ^ p;
    for (int32_t p2 = comp; p2; p2 = (p2 - 1) & comp) {
      int32_t q = p | p2;
      for (int32_t d2 = 0; d2 < ND; d2++) {
        if (hop && d + d2 > D) break;
        int32_t c2 = cost[SIDX(v, p2, d2)];
        if (c2 >= GST_ORC_INF) continue;
        int32_t nd = d > d2 ? d : d2;
        int64_t nc = (int64_t)c + c2;
        int64_t t = SIDX(v, q, nd);
        if (nc < cost[t]) {
          cost[t] = (int32_t)nc;
          pred[t].type = 2;
          pred[t].a = p;
          pred[t].b = (d << 16) | d2;
          heap_push(&h, ((uint64_t)nc << STATE_BITS) | (uint64_t)t);
          pushes++;
        }
      }
    }
  }
  free(h.a);
  if (stats) {
    stats[0] = pops;
    stats[1] = pushes;
  }

  int rc = 0;
  if (root_state >= 0) {
    *cost_out = cost[root_state];
    *root_out = (int32_t)((root_state / W) / ND);
    /* breadth-first back-tracking; each queue entry carries the index of the
     * edge whose child copy it hangs from */
    int64_t qcap = 1024, qh = 0, qt = 0;
    int64_t *qs = (int64_t *)malloc((size_t)qcap * sizeof(int64_t));
    int32_t *qa = (int32_t *)malloc((size_t)qcap * sizeof(int32_t));
    int32_t ne = 0;
    qs[qt] = root_state;
    qa[qt++] = -1;
    while (qh < qt) {
      int64_t s = qs[qh];
      int32_t att = qa[qh++];
      pred_t pr = pred[s];
      int32_t p = (int32_t)(s % W);
      int64_t vd = s / W;
      int32_t d = (int32_t)(vd % ND);
      int32_t v = (int32_t)(vd / ND);
      int64_t c1 = -1, c2 = -1;
      int32_t a1 = att, a2 = att;
      if (pr.type == 1) {
        if (ne < edge_cap) {
          edges[ne].u = v;
          edges[ne].v = pr.a;
          edges[ne].w = wt[pr.b];
          edges[ne].parent = att;
        } else {
          rc = 1; /* edge buffer too small; count continues */
        }
        a1 = ne++;
        c1 = SIDX(pr.a, p, hop ? d - 1 : 0);
      } else if (pr.type == 2) {
        int32_t d1 = pr.b >> 16, d2 = pr.b & 0xFFFF;
        c1 = SIDX(v, pr.a, d1);
        c2 = SIDX(v, p ^ pr.a, d2);
      }
      if (qt + 2 > qcap) {
        qcap *= 2;
        qs = (int64_t *)realloc(qs, (size_t)qcap * sizeof(int64_t));
        qa = (int32_t *)realloc(qa, (size_t)qcap * sizeof(int32_t));
      }
      if (c1 >= 0) { qs[qt] = c1; qa[qt++] = a1; }
      if (c2 >= 0) { qs[qt] = c2; qa[qt++] = a2; }
    }
    *n_edges = ne;
    free(qs);
    free(qa);
  }
  free(cost);
  free(pred);
  free(gmask);
  return rc;
#undef SIDX
}

/* ---------------- Dreyfus-Wagner Steiner tree ------------------------------ */
/* dp[S][v] = cheapest tree containing terminals S and vertex v.  Textbook
 * O(3^t V + 2^t (E + V log V)); used only to cross-check orc_dpbf through the
 * reduction GST -> Steiner (one dummy terminal per group). */
int orc_steiner_dw(int64_t V, const int64_t *rowptr, const int32_t *col,
                   const int32_t *wt, int32_t t, const int32_t *term,
                   int64_t *cost_out) {
  if (t < 1 || t > 12) return -2;
  const int64_t W = (int64_t)1 << t;
  int64_t *dp = (int64_t *)malloc((size_t)(W * V) * sizeof(int64_t));
  if (!dp) return -1;
  const int64_t BIG = (int64_t)1 << 60;
  for (int64_t i = 0; i < W * V; i++) dp[i] = BIG;
  heap_t h = {0, 0, 0};
  for (int64_t S = 1; S < W; S++) {
    int64_t *row = dp + S * V;
    if ((S & (S - 1)) == 0) {
      int i = __builtin_ctzll((unsigned long long)S);
      row[term[i]] = 0;
    } else {
      int64_t low = S & -S;
      for (int64_t A = (S - 1) & S; A; A = (A - 1) & S) {
        if (!(A & low)) continue;
        const int64_t *ra = dp + A * V, *rb = dp + (S ^ A) * V;
        for (int64_t v = 0; v < V; v++) {
          int64_t c = ra[v] + rb[v];
          if (c < row[v]) row[v] = c;
        }
      }
    }
    /* Dijkstra closure of row over the graph; costs fit in 30 bits here */
    h.n = 0;
    for (int64_t v = 0; v < V; v++)
      if (row[v] < BIG) heap_push(&h, ((uint64_t)row[v] << STATE_BITS) | (uint64_t)v);
    while (h.n) {
      uint64_t key = heap_pop(&h);
      int64_t v = (int64_t)(key & STATE_MASK);
      int64_t c = (int64_t)(key >> STATE_BITS);
      if (c != row[v]) continue;
      for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
        int32_t u = col[e];
        int64_t nc = c + wt[e];
        if (nc < row[u]) {
          row[u] = nc;
          heap_push(&h, ((uint64_t)nc << STATE_BITS) | (uint64_t)u);
        }
      }
    }
  }
  int64_t best = BIG;
  for (int64_t v = 0; v < V; v++)
    if (dp[(W - 1) * V + v] < best) best = dp[(W - 1) * V + v];
  *cost_out = best >= BIG ? -1 : best;
  free(dp);
  free(h.a);
  return 0;
}

/* ---------------- exact pruned best-first DPBF (CPU baseline) -------------- */
/* Exact-mode restatement of PrunedDP++ (code/PrunedDP++/include/CPUNONHOP.h:
 * 499-992) for the unbounded DP: k multi-source Dijkstra runs for the
 * group->vertex distances (:21-80), inter-group distances and route tables
 * (:175-210, :257-386), a best-first loop over (v, p) popping states by cost
 * (:635-985) with lower-bound pruning of grow and merge candidates (:436-496)
 * and an upper bound from tree + shortest paths (:684-726) and complement
 * merges.  Deviations that keep it exact: no approximation-ratio exit
 * (:730-755), no 1/2 and 2/3 decomposition rules, and a candidate is dropped
 * only when cost + lb > best (strict), which never removes a state of an
 * optimal witness.  Returns the optimal cost only.  stats (4 entries): pops,
 * heap capacity, ns of the precompute (distances, route tables, bounds) and
 * ns of the main loop -- the reference times its main loop alone
 * (CPUNONHOP.h:634-665). */
static uint32_t orc_lb(const int32_t *c, int k, uint32_t x, const int32_t *R) {
  int64_t lb1 = 0;
  for (int i = 0; i < k; i++)
    if ((x >> i) & 1u) lb1 = c[i] > lb1 ? c[i] : lb1;
  if (lb1 >= GST_ORC_INF) return GST_ORC_INF;
  if (__builtin_popcount(x) >= 2) {
    int64_t tour = (int64_t)3 * GST_ORC_INF;
    const int32_t *Rx = R + (size_t)x * k * k;
    for (int i = 0; i < k; i++) {
      if (!((x >> i) & 1u)) continue;
      for (int j = 0; j < k; j++) {
        if (j == i || !((x >> j) & 1u)) continue;
        int64_t t = (int64_t)c[i] + Rx[i * k + j] + c[j];
        if (t < tour) tour = t;
      }
    }
    int64_t lb2 = (tour + 1) / 2;
    if (lb2 > lb1) lb1 = lb2;
  }
  return lb1 >= GST_ORC_INF ? GST_ORC_INF : (uint32_t)lb1;
}

int orc_dpbf_pruned(int64_t V, const int64_t *rowptr, const int32_t *col, const int32_t *wt,
                    int32_t k, const int64_t *gptr, const int32_t *gmem, int32_t *cost_out,
                    int64_t *stats) {
  if (k < 1 || k > 16 || V <= 0) return -2;
  const int64_t t_begin = orc_now_ns();
  const int64_t W = (int64_t)1 << k;
  const int32_t full = (int32_t)(W - 1);
  *cost_out = GST_ORC_INF;
  if (stats) stats[0] = stats[1] = stats[2] = stats[3] = 0;
  for (int i = 0; i < k; i++)
    if (gptr[i + 1] == gptr[i]) return 0; /* empty group: infeasible */
  if (k == 1) {
    *cost_out = 0;
    return 0;
  }
  int32_t *C = (int32_t *)malloc((size_t)(V * k) * sizeof(int32_t));
  int32_t *dist = (int32_t *)malloc((size_t)V * sizeof(int32_t));
  int32_t *R = (int32_t *)malloc((size_t)(W * k * k) * sizeof(int32_t));
  int32_t *cost = (int32_t *)malloc((size_t)(V * W) * sizeof(int32_t));
  uint8_t *done = (uint8_t *)calloc((size_t)(V * W), 1);
  uint32_t *gmask = (uint32_t *)calloc((size_t)V, sizeof(uint32_t));
  if (!C || !dist || !R || !cost || !done || !gmask) {
    free(C); free(dist); free(R); free(cost); free(done); free(gmask);
    return -1;
  }
  for (int i = 0; i < k; i++) {
    orc_sssp(V, rowptr, col, wt, gptr[i + 1] - gptr[i], gmem + gptr[i], dist);
    for (int64_t v = 0; v < V; v++) C[v * k + i] = dist[v];
    for (int64_t j = gptr[i]; j < gptr[i + 1]; j++) gmask[gmem[j]] |= 1u << i;
  }
  int32_t vv[16 * 16];
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) {
      int32_t m = GST_ORC_INF;
      for (int64_t t = gptr[i]; t < gptr[i + 1]; t++) {
        int32_t c = C[(int64_t)gmem[t] * k + j];
        if (c < m) m = c;
      }
      vv[i * k + j] = (i == j) ? 0 : m;
    }
  for (int64_t t = 0; t < W * k * k; t++) R[t] = GST_ORC_INF;
  for (int c = 2; c <= k; c++)
    for (int64_t x = 1; x < W; x++) {
      if (__builtin_popcountll(x) != c) continue;
      for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
          if (i == j || !((x >> i) & 1) || !((x >> j) & 1)) continue;
          int64_t best = GST_ORC_INF;
          if (c == 2) {
            best = vv[i * k + j];
          } else {
            int64_t y = x ^ (1LL << j);
            for (int m = 0; m < k; m++) {
              if (m == i || !((y >> m) & 1)) continue;
              int64_t cand = (int64_t)R[y * k * k + i * k + m] + vv[m * k + j];
              if (cand < best) best = cand;
            }
          }
          R[x * k * k + i * k + j] = best >= GST_ORC_INF ? GST_ORC_INF : (int32_t)best;
        }
    }
  /* star upper bound */
  int64_t best = GST_ORC_INF;
  for (int64_t v = 0; v < V; v++) {
    int64_t s = 0;
    for (int i = 0; i < k; i++) s += C[v * k + i];
    if (s < best) best = s;
  }
  for (int64_t s = 0; s < V * W; s++) cost[s] = GST_ORC_INF;
  heap_t h = {0, 0, 0};
  int64_t pops = 0;
  for (int64_t v = 0; v < V; v++) {
    uint32_t m = gmask[v];
    if (!m) continue;
    for (uint32_t p = m; p; p = (p - 1) & m) {
      cost[v * W + p] = 0;
      heap_push(&h, (uint64_t)(v * W + p));
    }
  }
  int32_t answer = GST_ORC_INF;
  const int64_t t_loop = orc_now_ns();
  while (h.n) {
    uint64_t key = heap_pop(&h);
    int64_t s = (int64_t)(key & STATE_MASK);
    int32_t c = (int32_t)(key >> STATE_BITS);
    if (c != cost[s] || done[s]) continue;
    done[s] = 1;
    pops++;
    int32_t p = (int32_t)(s % W);
    int64_t v = s / W;
    if (p == full) {
      answer = c;
      break;
    }
    const int32_t *cv = C + v * k;
    int32_t x = full ^ p;
    /* upper bounds: tree + shortest paths, complement merge */
    int64_t ub = c;
    for (int i = 0; i < k; i++)
      if ((x >> i) & 1) ub += cv[i];
    if (cost[v * W + x] < GST_ORC_INF && (int64_t)c + cost[v * W + x] < ub) ub = (int64_t)c + cost[v * W + x];
    if (ub < best) best = ub;
    /* grow */
    for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
      int32_t u = col[e];
      int64_t nc = (int64_t)c + wt[e];
      int64_t t = (int64_t)u * W + p;
      if (done[t] || nc >= cost[t]) continue;
      if (nc + orc_lb(C + (int64_t)u * k, k, (uint32_t)x, R) > best) continue;
      cost[t] = (int32_t)nc;
      heap_push(&h, ((uint64_t)nc << STATE_BITS) | (uint64_t)t);
    }
    /* merge with settled disjoint partners */
    for (int32_t p2 = x; p2; p2 = (p2 - 1) & x) {
      int64_t s2 = v * W + p2;
      if (!done[s2]) continue;
      int32_t q = p | p2;
      int64_t nc = (int64_t)c + cost[s2];
      int64_t t = v * W + q;
      if (done[t] || nc >= cost[t]) continue;
      if (q != full && nc + orc_lb(cv, k, (uint32_t)(full ^ q), R) > best) continue;
      cost[t] = (int32_t)nc;
      heap_push(&h, ((uint64_t)nc << STATE_BITS) | (uint64_t)t);
    }
  }
  *cost_out = answer;
  if (stats) {
    const int64_t t_end = orc_now_ns();
    stats[0] = pops;
    stats[1] = (int64_t)h.cap;
    stats[2] = t_loop - t_begin;
    stats[3] = t_end - t_loop;
  }
  free(h.a);
  free(C); free(dist); free(R); free(cost); free(done); free(gmask);
  return 0;
}
