// gst_data.cpp -- deterministic synthetic inputs for the benchmark configs and
// CSR construction.  Host-only C++ (no device code), exported with C linkage
// and prefixed gstd_.
//
// Random numbers: SplitMix64 seeds xoshiro256** (Blackman & Vigna); bounded
// integers by rejection, so every stream is identical across compilers and
// standard libraries.  Generators:
//  * Chung-Lu: expected degree w_i proportional to (i+1)^(-1/(gamma-1));
//    endpoints drawn proportional to w through a cumulative table; self loops
//    and duplicates are dropped until m distinct undirected edges exist.
//  * R-MAT: 2^scale x 2^scale recursive quadrants with probabilities
//    (a, b, c, 1-a-b-c) -- the Graph500 initiator a=.57, b=c=.19 that the
//    reference vendors (code/D-GPUGST+/include/generator/graph_generator.c:
//    23-29) -- endpoints >= V, self loops and duplicates rejected, then a
//    seeded random relabelling.
//  * groups: size ~ Geometric(1/mean) on {1,2,...} by Bernoulli trials
//    (resampled above max), members uniform without replacement.
//  * queries: k distinct group ids uniform over an eligible id list.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform integer in [0, n), n >= 1, by rejection of the biased tail
  uint64_t below(uint64_t n) {
    const uint64_t lim = UINT64_MAX - (UINT64_MAX % n);
    for (;;) {
      const uint64_t x = next();
      if (x < lim) return x % n;
    }
  }
  // uniform double in [0, 1) with 53 random bits
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// open-addressing set of undirected edge keys (min << 32 | max)
class EdgeSet {
 public:
  explicit EdgeSet(size_t expected) {
    size_t cap = 16;
    while (cap < expected * 2) cap <<= 1;
    tab_.assign(cap, ~0ull);
    mask_ = cap - 1;
  }
  bool insert(uint64_t key) {
    size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 17) & mask_;
    for (;;) {
      if (tab_[h] == ~0ull) {
        tab_[h] = key;
        return true;
      }
      if (tab_[h] == key) return false;
      h = (h + 1) & mask_;
    }
  }

 private:
  std::vector<uint64_t> tab_;
  size_t mask_;
};

inline uint64_t ekey(uint32_t a, uint32_t b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)a << 32) | b;
}

int worker_count() { return (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency())); }

// f(i) for i in [0, n) on up to worker_count() threads, in contiguous ranges
template <class F>
void parallel_range(int64_t n, F f) {
  const int nt = (int)std::min<int64_t>(worker_count(), std::max<int64_t>(1, n / 4096));
  if (nt <= 1) {
    for (int64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  const int64_t per = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      const int64_t a = t * per, b = std::min(n, (t + 1) * per);
      for (int64_t i = a; i < b; i++) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int64_t gstd_chung_lu(int64_t V, int64_t m, double gamma, uint64_t seed, int32_t* eu, int32_t* ev) {
  if (V < 2 || m < 0 || gamma <= 1.0 || m > V * (V - 1) / 2) return -1;
  std::vector<double> cum(V);
  const double ex = -1.0 / (gamma - 1.0);
  double acc = 0;
  for (int64_t i = 0; i < V; i++) {
    acc += std::pow((double)(i + 1), ex);
    cum[i] = acc;
  }
  Rng rng(seed);
  EdgeSet set((size_t)m);
  int64_t n = 0;
  uint64_t tries = 0;
  const uint64_t max_tries = (uint64_t)m * 200 + 1000;
  // Candidates in blocks: the draws stay in one sequential stream, the
  // endpoint searches run in parallel, and the candidates are accepted in
  // stream order, so the edge list is the same as drawing one at a time.
  const int64_t B = (int64_t)1 << 20;
  std::vector<double> xs(2 * B);
  std::vector<int64_t> ends(2 * B);
  while (n < m && tries < max_tries) {
    const int64_t nb = (int64_t)std::min<uint64_t>((uint64_t)B, max_tries - tries);
    for (int64_t i = 0; i < 2 * nb; i++) xs[i] = rng.unit() * acc;
    parallel_range(2 * nb, [&](int64_t i) {
      ends[i] = std::upper_bound(cum.begin(), cum.end(), xs[i]) - cum.begin();
    });
    for (int64_t i = 0; i < nb && n < m; i++) {
      tries++;
      const int64_t a = ends[2 * i], b = ends[2 * i + 1];
      if (a >= V || b >= V || a == b) continue;
      if (!set.insert(ekey((uint32_t)a, (uint32_t)b))) continue;
      eu[n] = (int32_t)std::min(a, b);
      ev[n] = (int32_t)std::max(a, b);
      n++;
    }
  }
  return n;
}

int64_t gstd_rmat(int64_t V, int64_t m, int32_t scale, double a, double b, double c, uint64_t seed,
                  int32_t* eu, int32_t* ev) {
  if (V < 2 || scale < 1 || scale > 31 || ((int64_t)1 << scale) < V || m < 0) return -1;
  Rng rng(seed);
  EdgeSet set((size_t)m);
  int64_t n = 0;
  uint64_t tries = 0;
  const uint64_t max_tries = (uint64_t)m * 400 + 1000;
  const double ab = a + b, abc = a + b + c;
  while (n < m && tries < max_tries) {
    tries++;
    uint64_t u = 0, v = 0;
    for (int l = 0; l < scale; l++) {
      const double r = rng.unit();
      u <<= 1;
      v <<= 1;
      if (r < a) {
      } else if (r < ab) {
        v |= 1;
      } else if (r < abc) {
        u |= 1;
      } else {
        u |= 1;
        v |= 1;
      }
    }
    if ((int64_t)u >= V || (int64_t)v >= V || u == v) continue;
    if (!set.insert(ekey((uint32_t)u, (uint32_t)v))) continue;
    eu[n] = (int32_t)u;
    ev[n] = (int32_t)v;
    n++;
  }
  // seeded relabelling (Fisher-Yates over the vertex ids)
  std::vector<int32_t> perm(V);
  std::iota(perm.begin(), perm.end(), 0);
  Rng pr(seed ^ 0x5DEECE66Dull);
  for (int64_t i = V - 1; i > 0; i--) std::swap(perm[i], perm[pr.below((uint64_t)i + 1)]);
  for (int64_t i = 0; i < n; i++) {
    const int32_t x = perm[eu[i]], y = perm[ev[i]];
    eu[i] = std::min(x, y);
    ev[i] = std::max(x, y);
  }
  return n;
}

// Sorts undirected edges (u < v) by (u, v) so weights can be assigned in a
// canonical order.
void gstd_sort_edges(int64_t m, int32_t* eu, int32_t* ev) {
  std::vector<uint64_t> k(m);
  for (int64_t i = 0; i < m; i++) k[i] = ekey((uint32_t)eu[i], (uint32_t)ev[i]);
  // sorted runs in parallel, then pairwise merges (the keys are distinct, so
  // the result is the one order)
  const int nt = m > (1 << 20) ? worker_count() : 1;
  std::vector<int64_t> cut(nt + 1);
  for (int t = 0; t <= nt; t++) cut[t] = m * t / nt;
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t]() { std::sort(k.begin() + cut[t], k.begin() + cut[t + 1]); });
    for (auto& x : th) x.join();
  }
  for (int w = 1; w < nt; w *= 2) {
    std::vector<std::thread> th;
    for (int t = 0; t + w < nt; t += 2 * w) {
      const int64_t a = cut[t], mid = cut[t + w], b = cut[std::min(nt, t + 2 * w)];
      th.emplace_back([&, a, mid, b]() { std::inplace_merge(k.begin() + a, k.begin() + mid, k.begin() + b); });
    }
    for (auto& x : th) x.join();
  }
  for (int64_t i = 0; i < m; i++) {
    eu[i] = (int32_t)(k[i] >> 32);
    ev[i] = (int32_t)(k[i] & 0xFFFFFFFFu);
  }
}

void gstd_weights(int64_t m, int32_t lo, int32_t hi, uint64_t seed, int32_t* w) {
  Rng rng(seed);
  const uint64_t span = (uint64_t)(hi - lo + 1);
  for (int64_t i = 0; i < m; i++) w[i] = lo + (int32_t)rng.below(span);
}

// CSR of the undirected edge list, both arcs, rows sorted by neighbour.
int64_t gstd_csr(int64_t V, int64_t m, const int32_t* eu, const int32_t* ev, const int32_t* w,
                 int64_t* rowptr, int32_t* col, int32_t* wt) {
  std::vector<int64_t> deg(V + 1, 0);
  for (int64_t i = 0; i < m; i++) {
    if (eu[i] < 0 || eu[i] >= V || ev[i] < 0 || ev[i] >= V) return -1;
    deg[eu[i]]++;
    deg[ev[i]]++;
  }
  rowptr[0] = 0;
  for (int64_t v = 0; v < V; v++) rowptr[v + 1] = rowptr[v] + deg[v];
  std::vector<int64_t> pos(rowptr, rowptr + V);
  for (int64_t i = 0; i < m; i++) {
    int64_t p = pos[eu[i]]++;
    col[p] = ev[i];
    wt[p] = w[i];
    p = pos[ev[i]]++;
    col[p] = eu[i];
    wt[p] = w[i];
  }
  const int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  const int64_t per = (V + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    th.emplace_back([&, t]() {
      std::vector<std::pair<int32_t, int32_t>> tmp;
      for (int64_t v = t * per; v < std::min(V, (t + 1) * per); v++) {
        const int64_t r0 = rowptr[v], r1 = rowptr[v + 1];
        tmp.clear();
        for (int64_t i = r0; i < r1; i++) tmp.emplace_back(col[i], wt[i]);
        std::sort(tmp.begin(), tmp.end());
        for (int64_t i = r0; i < r1; i++) {
          col[i] = tmp[i - r0].first;
          wt[i] = tmp[i - r0].second;
        }
      }
    });
  }
  for (auto& x : th) x.join();
  return rowptr[V];
}

// Connected components by BFS; comp[v] = component id (ids by increasing
// smallest vertex).  Returns the id of the largest component (smallest id on
// ties); *lcc_size receives its size.
int64_t gstd_components(int64_t V, const int64_t* rowptr, const int32_t* col, int32_t* comp,
                        int64_t* lcc_size) {
  std::fill(comp, comp + V, -1);
  std::vector<int32_t> q;
  q.reserve(V);
  int32_t nc = 0;
  int64_t best = -1, bsz = -1;
  for (int64_t s = 0; s < V; s++) {
    if (comp[s] >= 0) continue;
    q.clear();
    q.push_back((int32_t)s);
    comp[s] = nc;
    for (size_t h = 0; h < q.size(); h++) {
      const int32_t v = q[h];
      for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
        const int32_t u = col[e];
        if (comp[u] < 0) {
          comp[u] = nc;
          q.push_back(u);
        }
      }
    }
    if ((int64_t)q.size() > bsz) {
      bsz = (int64_t)q.size();
      best = nc;
    }
    nc++;
  }
  if (lcc_size) *lcc_size = bsz;
  return best;
}

void gstd_group_sizes(int64_t G, int64_t mean, int64_t maxsize, uint64_t seed, int64_t* sizes) {
  Rng rng(seed);
  for (int64_t g = 0; g < G; g++) {
    for (;;) {
      int64_t s = 1;
      while (rng.below((uint64_t)mean) != 0) s++;
      if (s <= maxsize) {
        sizes[g] = s;
        break;
      }
    }
  }
}

// members of every group, uniform without replacement over [0, V);
// gptr must hold the prefix sums of the sizes.
void gstd_group_members(int64_t V, int64_t G, const int64_t* gptr, uint64_t seed, int32_t* gmem) {
  Rng rng(seed);
  std::vector<int32_t> tmp;
  for (int64_t g = 0; g < G; g++) {
    const int64_t n = gptr[g + 1] - gptr[g];
    tmp.clear();
    if (n * 4 > V) {
      // dense draw: partial Fisher-Yates over an explicit id array
      std::vector<int32_t> ids(V);
      std::iota(ids.begin(), ids.end(), 0);
      for (int64_t i = 0; i < n; i++) {
        const int64_t j = i + (int64_t)rng.below((uint64_t)(V - i));
        std::swap(ids[i], ids[j]);
        tmp.push_back(ids[i]);
      }
    } else {
      while ((int64_t)tmp.size() < n) {
        const int32_t v = (int32_t)rng.below((uint64_t)V);
        if (std::find(tmp.begin(), tmp.end(), v) == tmp.end()) tmp.push_back(v);
      }
    }
    std::sort(tmp.begin(), tmp.end());
    std::copy(tmp.begin(), tmp.end(), gmem + gptr[g]);
  }
}

// nq queries of k distinct ids drawn uniformly from eligible[0..ne).
void gstd_queries(int64_t nq, int32_t k, const int32_t* eligible, int64_t ne, uint64_t seed, int32_t* out) {
  Rng rng(seed);
  for (int64_t q = 0; q < nq; q++) {
    int32_t* row = out + q * k;
    for (int32_t i = 0; i < k; i++) {
      for (;;) {
        const int32_t c = eligible[rng.below((uint64_t)ne)];
        bool dup = false;
        for (int32_t j = 0; j < i; j++) dup |= (row[j] == c);
        if (!dup || ne < k) {
          row[i] = c;
          break;
        }
      }
    }
  }
}

// uniform integers in [lo, hi] (k values for mixed-k workloads, arrivals ...)
void gstd_uniform_ints(int64_t n, int64_t lo, int64_t hi, uint64_t seed, int64_t* out) {
  Rng rng(seed);
  for (int64_t i = 0; i < n; i++) out[i] = lo + (int64_t)rng.below((uint64_t)(hi - lo + 1));
}

// exponential inter-arrival gaps with the given mean (Poisson process)
void gstd_exponential(int64_t n, double mean, uint64_t seed, double* out) {
  Rng rng(seed);
  for (int64_t i = 0; i < n; i++) out[i] = -mean * std::log(1.0 - rng.unit());
}

}  // extern "C"
