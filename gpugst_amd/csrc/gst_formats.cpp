// gst_formats.cpp -- readers of the reference's text formats, multithreaded so
// that the serving-size graphs (117M edges, 2M groups) load in seconds.
// Host-only C++ with C linkage (prefix gstd_), built into libgstdata.so.
//
//  X.in   : first line "V ..." (V read as a number, code/GPUGST/include/
//           graph_hash_of_mixed_weighted_read_for_GSTP.h:83-86), then one
//           undirected edge per line "v1 v2 ec"; lines with another number of
//           fields are skipped, as there (:91-100).
//  X.g    : line "gN:m1 m2 ..." lists the members of group N-1 (:190-239;
//           code/GPUGST+/include/read_group.hpp:9-48); a trailing blank is
//           ignored.  Groups absent from the file are empty.
//  Xk.csv : one query per line, group ids separated by blanks (:240-286).
//
// Fields are split on any run of blanks (space, tab, CR); an integer field
// must be an optional '-' followed by digits.  Every array a reader returns is
// malloc'd and owned by the caller, released with gstd_free.  Return values:
// 0 on success, -1 when the file cannot be read, -2 on a malformed line (its
// 1-based number in *bad_line), -3 when out of memory.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace {

struct Buf {
  char* p = nullptr;
  size_t n = 0;
  ~Buf() { free(p); }
};

bool slurp(const char* path, Buf* b) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return false;
  }
  const long sz = ftell(f);
  if (sz < 0 || fseek(f, 0, SEEK_SET) != 0) {
    fclose(f);
    return false;
  }
  b->p = (char*)malloc((size_t)sz + 1);
  if (!b->p) {
    fclose(f);
    return false;
  }
  b->n = fread(b->p, 1, (size_t)sz, f);
  fclose(f);
  b->p[b->n] = '\n';  // every line ends with a newline
  return b->n == (size_t)sz;
}

inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Parses one integer field at s (no leading blanks); returns the end or null.
inline const char* parse_int(const char* s, const char* e, int64_t* out) {
  bool neg = false;
  if (s < e && *s == '-') {
    neg = true;
    s++;
  }
  if (s >= e || *s < '0' || *s > '9') return nullptr;
  int64_t v = 0;
  while (s < e && *s >= '0' && *s <= '9') {
    v = v * 10 + (*s - '0');
    if (v > ((int64_t)1 << 40)) return nullptr;
    s++;
  }
  if (s < e && !blank(*s)) return nullptr;
  *out = neg ? -v : v;
  return s;
}

// Integer fields of the line [s, e); false if one is malformed.
bool fields(const char* s, const char* e, std::vector<int64_t>* out) {
  out->clear();
  while (s < e) {
    while (s < e && blank(*s)) s++;
    if (s >= e) break;
    int64_t v = 0;
    s = parse_int(s, e, &v);
    if (!s) return false;
    out->push_back(v);
  }
  return true;
}

// Splits [b, e) into about nt pieces at line starts.
std::vector<const char*> split_lines(const char* b, const char* e, int nt) {
  std::vector<const char*> cut{b};
  const size_t n = (size_t)(e - b);
  for (int t = 1; t < nt; t++) {
    const char* c = b + n * (size_t)t / (size_t)nt;
    if (c <= cut.back()) continue;
    while (c < e && c[-1] != '\n') c++;
    if (c < e && c > cut.back()) cut.push_back(c);
  }
  cut.push_back(e);
  return cut;
}

int n_threads(size_t bytes) {
  if (bytes < (size_t)(8 << 20)) return 1;
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::min<unsigned>(hc ? hc : 1, 16);
}

// Runs body(piece, begin, end) over the line-aligned pieces of [b, e), in
// parallel; returns the pieces' count.
template <class F>
int run_pieces(const char* b, const char* e, F body) {
  const std::vector<const char*> cut = split_lines(b, e, n_threads((size_t)(e - b)));
  const int np = (int)cut.size() - 1;
  std::vector<std::thread> th;
  for (int i = 0; i < np; i++) th.emplace_back([&, i]() { body(i, cut[i], cut[i + 1]); });
  for (auto& t : th) t.join();
  return np;
}

// 1-based line number of position pos
int64_t line_of(const char* b, const char* pos) {
  return 1 + (int64_t)std::count(b, pos, '\n');
}

template <class T>
T* copy_out(const std::vector<T>& v) {
  T* p = (T*)malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (p && !v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

}  // namespace

extern "C" {

void gstd_free(void* p) { free(p); }

int gstd_read_in(const char* path, int64_t* V, int64_t* m, int32_t** eu, int32_t** ev, int32_t** ew,
                 int64_t* bad_line) {
  *bad_line = 0;
  Buf b;
  if (!slurp(path, &b)) return -1;
  const char* e = b.p + b.n;
  const char* nl = (const char*)memchr(b.p, '\n', b.n + 1);
  {
    // V is the first field of the first line, read as a number
    const char* s = b.p;
    while (s < nl && blank(*s)) s++;
    char* end = nullptr;
    const double v = strtod(s, &end);
    if (end == s || v < 1 || v > 2147483646.0) {
      *bad_line = 1;
      return -2;
    }
    *V = (int64_t)v;
  }
  struct Part {
    std::vector<int32_t> u, v, w;
    const char* bad = nullptr;
  };
  std::vector<Part> parts(64);
  const int64_t nv = *V;
  const int np = run_pieces(nl + 1, e, [&](int i, const char* s, const char* pe) {
    Part& P = parts[i];
    std::vector<int64_t> f;
    while (s < pe) {
      const char* le = (const char*)memchr(s, '\n', (size_t)(pe - s));
      if (!le) le = pe;
      if (!fields(s, le, &f)) {
        if (!P.bad) P.bad = s;
      } else if (f.size() == 3) {
        if (f[0] < 0 || f[0] >= nv || f[1] < 0 || f[1] >= nv || f[2] < INT32_MIN || f[2] > INT32_MAX) {
          if (!P.bad) P.bad = s;
        } else {
          P.u.push_back((int32_t)f[0]);
          P.v.push_back((int32_t)f[1]);
          P.w.push_back((int32_t)f[2]);
        }
      }
      s = le + 1;
    }
  });
  size_t tot = 0;
  for (int i = 0; i < np; i++) {
    if (parts[i].bad) {
      *bad_line = line_of(b.p, parts[i].bad);
      return -2;
    }
    tot += parts[i].u.size();
  }
  int32_t* A = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>(tot, 1));
  int32_t* B = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>(tot, 1));
  int32_t* C = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>(tot, 1));
  if (!A || !B || !C) {
    free(A);
    free(B);
    free(C);
    return -3;
  }
  size_t o = 0;
  for (int i = 0; i < np; i++) {
    const Part& P = parts[i];
    if (P.u.empty()) continue;
    memcpy(A + o, P.u.data(), sizeof(int32_t) * P.u.size());
    memcpy(B + o, P.v.data(), sizeof(int32_t) * P.u.size());
    memcpy(C + o, P.w.data(), sizeof(int32_t) * P.u.size());
    o += P.u.size();
  }
  *m = (int64_t)tot;
  *eu = A;
  *ev = B;
  *ew = C;
  return 0;
}

int gstd_read_groups(const char* path, int64_t* n_groups, int64_t** gptr, int32_t** gmem, int64_t* bad_line) {
  *bad_line = 0;
  Buf b;
  if (!slurp(path, &b)) return -1;
  struct Part {
    std::vector<int64_t> g;     // group of each line
    std::vector<int64_t> off;   // its first member in mem
    std::vector<int32_t> mem;
    const char* bad = nullptr;
  };
  std::vector<Part> parts(64);
  const int np = run_pieces(b.p, b.p + b.n, [&](int i, const char* s, const char* pe) {
    Part& P = parts[i];
    std::vector<int64_t> f;
    while (s < pe) {
      const char* le = (const char*)memchr(s, '\n', (size_t)(pe - s));
      if (!le) le = pe;
      const char* t = s;
      while (t < le && blank(*t)) t++;
      if (t < le) {
        const char* colon = (const char*)memchr(t, ':', (size_t)(le - t));
        char* end = nullptr;
        const double gv = (colon && *t == 'g') ? strtod(t + 1, &end) : 0.0;
        if (!colon || *t != 'g' || end != colon || gv < 1 || gv > 2147483647.0 || !fields(colon + 1, le, &f)) {
          if (!P.bad) P.bad = s;
        } else {
          P.g.push_back((int64_t)gv - 1);
          P.off.push_back((int64_t)P.mem.size());
          for (int64_t x : f) {
            if (x < 0 || x > INT32_MAX) {
              if (!P.bad) P.bad = s;
              break;
            }
            P.mem.push_back((int32_t)x);
          }
        }
      }
      s = le + 1;
    }
  });
  int64_t G = 0;
  for (int i = 0; i < np; i++) {
    if (parts[i].bad) {
      *bad_line = line_of(b.p, parts[i].bad);
      return -2;
    }
    for (int64_t g : parts[i].g) G = std::max(G, g + 1);
  }
  std::vector<int64_t> ptr((size_t)G + 1, 0);
  for (int i = 0; i < np; i++) {
    const Part& P = parts[i];
    for (size_t j = 0; j < P.g.size(); j++) {
      const int64_t end = j + 1 < P.g.size() ? P.off[j + 1] : (int64_t)P.mem.size();
      ptr[(size_t)P.g[j] + 1] += end - P.off[j];  // a repeated line adds to its group
    }
  }
  for (int64_t g = 0; g < G; g++) ptr[(size_t)g + 1] += ptr[(size_t)g];
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  std::vector<int32_t> mem((size_t)ptr[(size_t)G]);
  for (int i = 0; i < np; i++) {
    const Part& P = parts[i];
    for (size_t j = 0; j < P.g.size(); j++) {
      const int64_t end = j + 1 < P.g.size() ? P.off[j + 1] : (int64_t)P.mem.size();
      int64_t& at = fill[(size_t)P.g[j]];
      for (int64_t x = P.off[j]; x < end; x++) mem[(size_t)at++] = P.mem[(size_t)x];
    }
  }
  // members sorted within each group (read_Group sorts them, :226)
  for (int64_t g = 0; g < G; g++) std::sort(mem.begin() + ptr[(size_t)g], mem.begin() + ptr[(size_t)g + 1]);
  *gptr = copy_out(ptr);
  *gmem = copy_out(mem);
  if (!*gptr || !*gmem) {
    free(*gptr);
    free(*gmem);
    return -3;
  }
  *n_groups = G;
  return 0;
}

int gstd_read_queries(const char* path, int64_t* nq, int64_t** qptr, int32_t** ids, int64_t* bad_line) {
  *bad_line = 0;
  Buf b;
  if (!slurp(path, &b)) return -1;
  std::vector<int64_t> ptr{0};
  std::vector<int32_t> out;
  std::vector<int64_t> f;
  const char* s = b.p;
  const char* e = b.p + b.n;
  int64_t line = 0;
  while (s < e) {
    const char* le = (const char*)memchr(s, '\n', (size_t)(e - s));
    if (!le) le = e;
    line++;
    if (!fields(s, le, &f)) {
      *bad_line = line;
      return -2;
    }
    if (!f.empty()) {
      for (int64_t x : f) {
        if (x < 0 || x > INT32_MAX) {
          *bad_line = line;
          return -2;
        }
        out.push_back((int32_t)x);
      }
      ptr.push_back((int64_t)out.size());
    }
    s = le + 1;
  }
  *qptr = copy_out(ptr);
  *ids = copy_out(out);
  if (!*qptr || !*ids) {
    free(*qptr);
    free(*ids);
    return -3;
  }
  *nq = (int64_t)ptr.size() - 1;
  return 0;
}

}  // extern "C"
