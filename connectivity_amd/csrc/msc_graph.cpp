// Host-side graph construction and checking for libmsc:
//   * msc_graph_read_edge_list / msc_graph_from_edges: read_edge + create_csr
//     (multistep/scc_main.cpp:94-191), including its pivot rule (:169-183);
//   * synthetic generators (R-MAT, grid digraph, planted SCCs) with counter-based random
//     streams, so a graph depends only on its seed, never on the thread count;
//   * msc_scc_verify / msc_cc_verify (multistep/scc_verify.cpp:47-77, cc_verify.cpp:47-74)
//     and msc_canonicalize.
// None of this is timed by the reference (scc_main.cpp:368-378 times run_scc only).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/msc.h"

namespace {

// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) as a counter-based stream: each
// (seed, stream) pair gives an independent sequence.
struct Rng {
  uint64_t s;
  static uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  Rng(uint64_t seed, uint64_t stream)
      : s(mix(seed + 0x9e3779b97f4a7c15ULL) ^ mix(stream * 0xd1b54a32d192ed03ULL + 0x632be59bd9b4e019ULL)) {}
  uint64_t next() {
    s += 0x9e3779b97f4a7c15ULL;
    return mix(s);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  uint64_t below(uint64_t k) { return (uint64_t)(((unsigned __int128)next() * k) >> 64); }
};

enum : uint64_t {
  kStreamPerm = 0xffffffff00000001ULL,
  kStreamArcs = 0xffffffff00000002ULL,
  kStreamSizes = 0xffffffff00000003ULL,
  kStreamOrder = 0xffffffff00000004ULL,
  kStreamBatch = 0xffffffff00000005ULL,
};

msc_host_graph* alloc_graph(int32_t n, int64_t m) {
  msc_host_graph* g = (msc_host_graph*)std::calloc(1, sizeof(msc_host_graph));
  if (!g) return nullptr;
  g->n = n;
  g->m = m;
  g->out_off = (uint32_t*)std::malloc(sizeof(uint32_t) * ((size_t)n + 1));
  g->in_off = (uint32_t*)std::malloc(sizeof(uint32_t) * ((size_t)n + 1));
  g->out_adj = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  g->in_adj = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  g->max_deg_vert = -1;
  if (!g->out_off || !g->in_off || !g->out_adj || !g->in_adj) {
    msc_graph_free(g);
    return nullptr;
  }
  return g;
}

// create_csr's pivot and average degree (multistep/scc_main.cpp:169-183): first vertex
// with strictly largest (double)out*(double)in; the degree sum is an exact integer.
void pivot_rule(msc_host_graph* g) {
  int32_t best = -1;
  double best_val = 0.0;
  for (int32_t v = 0; v < g->n; ++v) {
    double p = (double)(g->out_off[v + 1] - g->out_off[v]) *
               (double)(g->in_off[v + 1] - g->in_off[v]);
    if (p > best_val) {
      best_val = p;
      best = v;
    }
  }
  g->max_deg_vert = best;
  g->avg_degree = g->n > 0 ? (double)g->m / (double)g->n : 0.0;
}

// Exclusive scan of counts[0..n) into off[0..n]; false if the total reaches 2^32.
bool scan_offsets(const uint32_t* counts, int32_t n, uint32_t* off) {
  uint64_t run = 0;
  for (int32_t v = 0; v < n; ++v) {
    off[v] = (uint32_t)run;
    run += counts[v];
  }
  off[n] = (uint32_t)run;
  return run < (1ULL << 32);
}

// Transposed CSR from a CSR whose lists are sorted: sources within each column end up
// sorted too (columns are filled by a parallel scatter, then each is sorted).
void build_transpose(msc_host_graph* g) {
  int32_t n = g->n;
  std::vector<uint32_t> cnt(n, 0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < g->m; ++e) __atomic_fetch_add(&cnt[g->out_adj[e]], 1u, __ATOMIC_RELAXED);
  scan_offsets(cnt.data(), n, g->in_off);
  std::vector<uint32_t> cur(g->in_off, g->in_off + n);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < n; ++v)
    for (uint32_t j = g->out_off[v]; j < g->out_off[v + 1]; ++j) {
      uint32_t at = __atomic_fetch_add(&cur[g->out_adj[j]], 1u, __ATOMIC_RELAXED);
      g->in_adj[at] = v;
    }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < n; ++v) std::sort(g->in_adj + g->in_off[v], g->in_adj + g->in_off[v + 1]);
}

// CSR + CSC from an edge list with self-loops dropped (optional) and duplicates removed;
// adjacency lists sorted ascending.  Parallel and deterministic.
int build_dedupe(int32_t n, int64_t m, const int32_t* src, const int32_t* dst, bool drop_loops,
                 msc_host_graph** out) {
  std::vector<uint32_t> cnt(n, 0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; ++e)
    if (!(drop_loops && src[e] == dst[e])) __atomic_fetch_add(&cnt[src[e]], 1u, __ATOMIC_RELAXED);
  std::vector<uint32_t> off(n + 1);
  if (!scan_offsets(cnt.data(), n, off.data())) return MSC_E_CAPACITY;
  int32_t* tmp = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(off[n] > 0 ? off[n] : 1));
  if (!tmp) return MSC_E_OOM;
  {
    std::vector<uint32_t> cur(off.begin(), off.begin() + n);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
      if (drop_loops && src[e] == dst[e]) continue;
      uint32_t at = __atomic_fetch_add(&cur[src[e]], 1u, __ATOMIC_RELAXED);
      tmp[at] = dst[e];
    }
  }
  // sort + unique each list; cnt <- kept length
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < n; ++v) {
    int32_t* b = tmp + off[v];
    int32_t* e = tmp + off[v + 1];
    std::sort(b, e);
    cnt[v] = (uint32_t)(std::unique(b, e) - b);
  }
  uint64_t total = 0;
  for (int32_t v = 0; v < n; ++v) total += cnt[v];
  msc_host_graph* g = alloc_graph(n, (int64_t)total);
  if (!g) {
    std::free(tmp);
    return MSC_E_OOM;
  }
  scan_offsets(cnt.data(), n, g->out_off);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < n; ++v)
    std::memcpy(g->out_adj + g->out_off[v], tmp + off[v], sizeof(int32_t) * cnt[v]);
  std::free(tmp);
  build_transpose(g);
  pivot_rule(g);
  *out = g;
  return MSC_OK;
}

// create_csr (multistep/scc_main.cpp:127-167): stable counting sort, lists keep the input
// order, duplicates and self-loops kept.
int build_stable(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                 msc_host_graph** out) {
  msc_host_graph* g = alloc_graph(n, m);
  if (!g) return MSC_E_OOM;
  std::vector<uint32_t> cnt(n, 0);
  for (int64_t e = 0; e < m; ++e) ++cnt[src[e]];
  if (!scan_offsets(cnt.data(), n, g->out_off)) {
    msc_graph_free(g);
    return MSC_E_CAPACITY;
  }
  std::vector<uint32_t> cur(g->out_off, g->out_off + n);
  for (int64_t e = 0; e < m; ++e) g->out_adj[cur[src[e]]++] = dst[e];
  std::fill(cnt.begin(), cnt.end(), 0u);
  for (int64_t e = 0; e < m; ++e) ++cnt[dst[e]];
  scan_offsets(cnt.data(), n, g->in_off);
  cur.assign(g->in_off, g->in_off + n);
  for (int64_t e = 0; e < m; ++e) g->in_adj[cur[dst[e]]++] = src[e];
  pivot_rule(g);
  *out = g;
  return MSC_OK;
}

// Random permutation of [0, n) (Fisher-Yates, one stream).
std::vector<int32_t> permutation(int32_t n, uint64_t seed) {
  std::vector<int32_t> p(n);
  for (int32_t i = 0; i < n; ++i) p[i] = i;
  Rng r(seed, kStreamPerm);
  for (int32_t i = n - 1; i > 0; --i) std::swap(p[i], p[(int32_t)r.below((uint64_t)i + 1)]);
  return p;
}

// R-MAT draws over [0, n) for any n >= 1: quadrant sampling in the 2^ceil(log2 n) id
// space, a draw with an endpoint >= n drawn again from the same stream (after 64 tries
// the endpoints are reduced mod n); ids permuted; self-loops and duplicates removed.
int rmat_any_n(int32_t n, int64_t draws, double a, double b, double c, uint64_t seed, msc_host_graph** out) {
  int32_t scale = 0;
  while ((1LL << scale) < n) ++scale;
  const uint64_t ta = (uint64_t)(a * 4294967296.0), tab = (uint64_t)((a + b) * 4294967296.0),
                 tabc = (uint64_t)((a + b + c) * 4294967296.0);
  std::vector<int32_t> src((size_t)draws), dst((size_t)draws);
  std::vector<int32_t> perm = permutation(n, seed);
  for (int64_t e = 0; e < draws; ++e) {
    Rng r(seed, (uint64_t)e);
    uint32_t u = 0, v = 0;
    for (int tries = 0; tries < 64; ++tries) {
      u = v = 0;
      uint64_t bits = 0;
      for (int32_t k = 0; k < scale; ++k) {
        if ((k & 1) == 0) bits = r.next();
        uint64_t x = (k & 1) ? (bits >> 32) : (bits & 0xffffffffULL);
        uint32_t su = x >= tabc || (x >= tab && x < tabc) ? 1u : 0u;
        uint32_t sv = x >= tabc || (x >= ta && x < tab) ? 1u : 0u;
        u = (u << 1) | su;
        v = (v << 1) | sv;
      }
      if (u < (uint32_t)n && v < (uint32_t)n) break;
    }
    src[e] = perm[u % (uint32_t)n];
    dst[e] = perm[v % (uint32_t)n];
  }
  return build_dedupe(n, draws, src.data(), dst.data(), true, out);
}

bool read_i64(FILE* f, int64_t* x) {
  int c;
  do {
    c = std::fgetc(f);
  } while (c == ' ' || c == '\n' || c == '\r' || c == '\t');
  if (c == EOF) return false;
  bool neg = false;
  if (c == '-') {
    neg = true;
    c = std::fgetc(f);
  }
  if (c < '0' || c > '9') return false;
  int64_t v = 0;
  while (c >= '0' && c <= '9') {
    v = v * 10 + (c - '0');
    c = std::fgetc(f);
  }
  *x = neg ? -v : v;
  return true;
}

void count_components(int32_t n, const int32_t* labels, msc_verify* out) {
  std::vector<int32_t> counts(n > 0 ? n : 1, 0);
  int32_t unassigned = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (labels[v] >= 0 && labels[v] < n)
      ++counts[labels[v]];
    else
      ++unassigned;
  }
  msc_verify r = {0, 0, 0, unassigned};
  for (int32_t v = 0; v < n; ++v) {
    if (!counts[v]) continue;
    ++r.num_components;
    if (counts[v] > 1) ++r.nontrivial;
    if (counts[v] > r.max_size) r.max_size = counts[v];
  }
  *out = r;
}

}  // namespace

extern "C" {

void msc_graph_free(msc_host_graph* g) {
  if (!g) return;
  std::free(g->out_off);
  std::free(g->in_off);
  std::free(g->out_adj);
  std::free(g->in_adj);
  std::free(g);
}

int msc_graph_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                         int32_t flags, msc_host_graph** out) {
  if (!out || n < 0 || m < 0 || (m > 0 && (!src || !dst))) return MSC_E_INVALID;
  if (m >= (1LL << 32)) return MSC_E_CAPACITY;
  for (int64_t e = 0; e < m; ++e)
    if (src[e] < 0 || src[e] >= n || dst[e] < 0 || dst[e] >= n) return MSC_E_INVALID;
  try {
    if (flags & MSC_BUILD_DEDUPE)
      return build_dedupe(n, m, src, dst, (flags & MSC_BUILD_DROP_SELF_LOOPS) != 0, out);
    if (flags & MSC_BUILD_DROP_SELF_LOOPS) {
      std::vector<int32_t> s, d;
      for (int64_t e = 0; e < m; ++e)
        if (src[e] != dst[e]) {
          s.push_back(src[e]);
          d.push_back(dst[e]);
        }
      return build_stable(n, (int64_t)s.size(), s.data(), d.data(), out);
    }
    return build_stable(n, m, src, dst, out);
  } catch (const std::bad_alloc&) {
    return MSC_E_OOM;
  }
}

int msc_graph_read_edge_list(const char* path, msc_host_graph** out) {
  if (!path || !out) return MSC_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return MSC_E_IO;
  std::vector<char> buf(1 << 22);
  setvbuf(f, buf.data(), _IOFBF, buf.size());
  int64_t n = 0, m = 0;
  if (!read_i64(f, &n) || !read_i64(f, &m) || n < 0 || n > INT32_MAX || m < 0) {
    std::fclose(f);
    return MSC_E_IO;
  }
  int rc = MSC_OK;
  try {
    std::vector<int32_t> s((size_t)m), d((size_t)m);
    for (int64_t e = 0; e < m; ++e) {
      int64_t u, v;
      if (!read_i64(f, &u) || !read_i64(f, &v)) {
        rc = MSC_E_IO;
        break;
      }
      if (u < 0 || u >= n || v < 0 || v >= n) {
        rc = MSC_E_INVALID;
        break;
      }
      s[e] = (int32_t)u;
      d[e] = (int32_t)v;
    }
    std::fclose(f);
    if (rc != MSC_OK) return rc;
    return build_stable((int32_t)n, m, s.data(), d.data(), out);
  } catch (const std::bad_alloc&) {
    std::fclose(f);
    return MSC_E_OOM;
  }
}

int msc_write_labels(const char* path, int32_t n, const int32_t* labels) {
  if (!path || n < 0 || (n > 0 && !labels)) return MSC_E_INVALID;
  FILE* f = std::fopen(path, "w");
  if (!f) return MSC_E_IO;
  for (int32_t v = 0; v < n; ++v) std::fprintf(f, "%d\n", labels[v]);
  return std::fclose(f) == 0 ? MSC_OK : MSC_E_IO;
}

int msc_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                 uint64_t seed, int32_t permute, msc_host_graph** out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 0 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return MSC_E_INVALID;
  int32_t n = (int32_t)(1u << scale);
  int64_t draws = (int64_t)edge_factor << scale;
  if (draws >= (1LL << 32)) return MSC_E_CAPACITY;
  const uint64_t ta = (uint64_t)(a * 4294967296.0), tab = (uint64_t)((a + b) * 4294967296.0),
                 tabc = (uint64_t)((a + b + c) * 4294967296.0);
  try {
    std::vector<int32_t> src((size_t)draws), dst((size_t)draws);
    std::vector<int32_t> perm;
    if (permute) perm = permutation(n, seed);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < draws; ++e) {
      Rng r(seed, (uint64_t)e);
      uint32_t u = 0, v = 0;
      uint64_t bits = 0;
      for (int32_t k = 0; k < scale; ++k) {
        if ((k & 1) == 0) bits = r.next();
        uint64_t x = (k & 1) ? (bits >> 32) : (bits & 0xffffffffULL);
        uint32_t su = 0, sv = 0;
        if (x < ta) {
        } else if (x < tab) {
          sv = 1;
        } else if (x < tabc) {
          su = 1;
        } else {
          su = 1;
          sv = 1;
        }
        u = (u << 1) | su;
        v = (v << 1) | sv;
      }
      src[e] = permute ? perm[u] : (int32_t)u;
      dst[e] = permute ? perm[v] : (int32_t)v;
    }
    return build_dedupe(n, draws, src.data(), dst.data(), true, out);
  } catch (const std::bad_alloc&) {
    return MSC_E_OOM;
  }
}

int msc_gen_grid(int32_t rows, int32_t cols, double p_one, double p_both, double arc_frac,
                 uint64_t seed, msc_host_graph** out) {
  if (!out || rows < 1 || cols < 1 || p_one < 0 || p_both < 0 || p_one + p_both > 1.0 ||
      arc_frac < 0)
    return MSC_E_INVALID;
  int64_t n64 = (int64_t)rows * cols;
  if (n64 > INT32_MAX) return MSC_E_CAPACITY;
  int32_t n = (int32_t)n64;
  int64_t arcs = (int64_t)std::llround(arc_frac * (double)n);
  try {
    // two candidate grid edges per vertex (right, down), each yielding up to 2 arcs
    std::vector<int32_t> src((size_t)(4 * n64 + arcs)), dst((size_t)(4 * n64 + arcs));
    std::vector<uint32_t> cnt((size_t)n + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n64; ++v) {
      int32_t r = (int32_t)(v / cols), c = (int32_t)(v % cols);
      for (int32_t dir = 0; dir < 2; ++dir) {
        int64_t w = dir == 0 ? (c + 1 < cols ? v + 1 : -1) : (r + 1 < rows ? v + cols : -1);
        int64_t at = 4 * v + 2 * dir;
        Rng g(seed, (uint64_t)(2 * v + dir));
        double x = g.uniform();
        if (w < 0) {
          src[at] = dst[at] = src[at + 1] = dst[at + 1] = -1;
        } else if (x < p_one) {
          bool fwd = (g.next() >> 63) != 0;
          src[at] = fwd ? (int32_t)v : (int32_t)w;
          dst[at] = fwd ? (int32_t)w : (int32_t)v;
          src[at + 1] = dst[at + 1] = -1;
        } else if (x < p_one + p_both) {
          src[at] = (int32_t)v; dst[at] = (int32_t)w;
          src[at + 1] = (int32_t)w; dst[at + 1] = (int32_t)v;
        } else {
          src[at] = dst[at] = src[at + 1] = dst[at + 1] = -1;
        }
      }
    }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < arcs; ++k) {
      Rng g(seed ^ kStreamArcs, (uint64_t)k);
      src[4 * n64 + k] = (int32_t)g.below((uint64_t)n);
      dst[4 * n64 + k] = (int32_t)g.below((uint64_t)n);
    }
    // drop the unused slots, keep order
    size_t kept = 0;
    for (size_t i = 0; i < src.size(); ++i)
      if (src[i] >= 0) {
        src[kept] = src[i];
        dst[kept] = dst[i];
        ++kept;
      }
    return build_dedupe(n, (int64_t)kept, src.data(), dst.data(), true, out);
  } catch (const std::bad_alloc&) {
    return MSC_E_OOM;
  }
}

int msc_gen_planted(int32_t n, double zipf_s, int32_t max_size, double extra_per_vertex,
                    int64_t dag_arcs, uint64_t seed, msc_host_graph** out, int32_t* comp_out) {
  if (!out || n < 1 || zipf_s <= 1.0 || max_size < 1 || extra_per_vertex < 0 || dag_arcs < 0)
    return MSC_E_INVALID;
  try {
    // component sizes: inverse CDF of a truncated Zipf law
    std::vector<double> cdf(max_size);
    double acc = 0;
    for (int32_t k = 1; k <= max_size; ++k) {
      acc += std::pow((double)k, -zipf_s);
      cdf[k - 1] = acc;
    }
    Rng rs(seed, kStreamSizes);
    std::vector<int32_t> start;  // component c covers [start[c], start[c+1])
    int32_t at = 0;
    while (at < n) {
      double x = rs.uniform() * acc;
      int32_t k = (int32_t)(std::lower_bound(cdf.begin(), cdf.end(), x) - cdf.begin()) + 1;
      k = std::min(k, n - at);
      start.push_back(at);
      at += k;
    }
    int32_t ncomp = (int32_t)start.size();
    start.push_back(n);
    std::vector<int32_t> comp(n);
    for (int32_t c = 0; c < ncomp; ++c)
      for (int32_t v = start[c]; v < start[c + 1]; ++v) comp[v] = c;
    // random topological rank of the components
    std::vector<int32_t> rank(ncomp);
    for (int32_t c = 0; c < ncomp; ++c) rank[c] = c;
    {
      Rng ro(seed, kStreamOrder);
      for (int32_t i = ncomp - 1; i > 0; --i) std::swap(rank[i], rank[(int32_t)ro.below((uint64_t)i + 1)]);
    }
    std::vector<int32_t> src, dst;
    for (int32_t c = 0; c < ncomp; ++c) {
      int32_t s = start[c], k = start[c + 1] - s;
      if (k < 2) continue;
      for (int32_t i = 0; i < k; ++i) {
        src.push_back(s + i);
        dst.push_back(s + (i + 1) % k);
      }
      int64_t extra = (int64_t)std::llround(extra_per_vertex * k);
      Rng g(seed, (uint64_t)c);
      for (int64_t e = 0; e < extra; ++e) {
        src.push_back(s + (int32_t)g.below((uint64_t)k));
        dst.push_back(s + (int32_t)g.below((uint64_t)k));
      }
    }
    Rng gd(seed ^ kStreamArcs, 0);
    for (int64_t e = 0; e < dag_arcs; ++e) {
      int32_t x = (int32_t)gd.below((uint64_t)n), y = (int32_t)gd.below((uint64_t)n);
      if (comp[x] == comp[y]) continue;
      if (rank[comp[x]] > rank[comp[y]]) std::swap(x, y);
      src.push_back(x);
      dst.push_back(y);
    }
    std::vector<int32_t> perm = permutation(n, seed);
    for (size_t e = 0; e < src.size(); ++e) {
      src[e] = perm[src[e]];
      dst[e] = perm[dst[e]];
    }
    if (comp_out)
      for (int32_t v = 0; v < n; ++v) comp_out[perm[v]] = comp[v];
    return build_dedupe(n, (int64_t)src.size(), src.data(), dst.data(), true, out);
  } catch (const std::bad_alloc&) {
    return MSC_E_OOM;
  }
}

int msc_batch_pack(int32_t num_graphs, const msc_host_graph* const* graphs, msc_host_graph** out,
                   int64_t* vert_base) {
  if (num_graphs < 1 || !graphs || !out || !vert_base) return MSC_E_INVALID;
  std::vector<int64_t> ebase((size_t)num_graphs + 1);
  vert_base[0] = 0;
  ebase[0] = 0;
  for (int32_t i = 0; i < num_graphs; ++i) {
    const msc_host_graph* g = graphs[i];
    if (!g || g->n < 0 || g->m < 0) return MSC_E_INVALID;
    vert_base[i + 1] = vert_base[i] + g->n;
    ebase[i + 1] = ebase[i] + g->m;
  }
  const int64_t n = vert_base[num_graphs], m = ebase[num_graphs];
  if (n > INT32_MAX || m >= (1LL << 32)) return MSC_E_CAPACITY;
  msc_host_graph* u = alloc_graph((int32_t)n, m);
  if (!u) return MSC_E_OOM;
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t i = 0; i < num_graphs; ++i) {
    const msc_host_graph* g = graphs[i];
    const int64_t vb = vert_base[i], eb = ebase[i];
    for (int32_t v = 0; v < g->n; ++v) {
      u->out_off[vb + v] = (uint32_t)(eb + g->out_off[v]);
      u->in_off[vb + v] = (uint32_t)(eb + g->in_off[v]);
    }
    for (int64_t e = 0; e < g->m; ++e) {
      u->out_adj[eb + e] = (int32_t)(vb + g->out_adj[e]);
      u->in_adj[eb + e] = (int32_t)(vb + g->in_adj[e]);
    }
  }
  u->out_off[n] = u->in_off[n] = (uint32_t)m;
  pivot_rule(u);
  *out = u;
  return MSC_OK;
}

int msc_gen_batch(int32_t num_graphs, double median_n, double sigma, int32_t min_n, int32_t max_n, uint64_t seed,
                  msc_host_graph** out, int64_t* vert_base, int32_t* kinds) {
  if (num_graphs < 1 || !out || !vert_base || !(median_n >= 1.0) || sigma < 0 || min_n < 1 || max_n < min_n)
    return MSC_E_INVALID;
  try {
    // sizes, kinds and edge factors from one stream (Box-Muller normals)
    std::vector<int32_t> ns(num_graphs), kind(num_graphs), ef(num_graphs);
    Rng rs(seed, kStreamBatch);
    for (int32_t i = 0; i < num_graphs; ++i) {
      double u1 = 1.0 - rs.uniform(), u2 = rs.uniform();
      double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      double x = std::round(median_n * std::exp(sigma * z));
      ns[i] = (int32_t)std::min<double>(std::max<double>(x, (double)min_n), (double)max_n);
      kind[i] = (int32_t)(rs.next() >> 63);
      ef[i] = 8 + (int32_t)rs.below(25);
    }
    std::vector<msc_host_graph*> gs(num_graphs, nullptr);
    int err = MSC_OK;
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < num_graphs; ++i) {
      uint64_t gseed = Rng::mix(seed ^ Rng::mix(0x5851f42d4c957f2dULL + (uint64_t)i));
      int rc;
      try {
        rc = kind[i] == 0 ? rmat_any_n(ns[i], (int64_t)ef[i] * ns[i], 0.57, 0.19, 0.19, gseed, &gs[i])
                          : msc_gen_planted(ns[i], 1.5, ns[i], 1.0, 4LL * ns[i], gseed, &gs[i], nullptr);
      } catch (const std::bad_alloc&) {
        rc = MSC_E_OOM;
      }
      if (rc) {
#pragma omp critical
        err = rc;
      }
    }
    int rc = err;
    if (!rc) rc = msc_batch_pack(num_graphs, gs.data(), out, vert_base);
    for (msc_host_graph* g : gs) msc_graph_free(g);
    if (!rc && kinds) std::copy(kind.begin(), kind.end(), kinds);
    return rc;
  } catch (const std::bad_alloc&) {
    return MSC_E_OOM;
  }
}

int msc_format_edge_list(const msc_host_graph* g, char* out, int64_t cap, int64_t* bytes) {
  if (!g || !bytes || g->n < 0 || g->m < 0) return MSC_E_INVALID;
  auto digits = [](uint64_t x) {
    int d = 1;
    while (x >= 10) {
      x /= 10;
      ++d;
    }
    return d;
  };
  char head[64];
  const int hl = std::snprintf(head, sizeof(head), "%d %lld\n", g->n, (long long)g->m);
  // bytes of each vertex's lines, then their prefix: every vertex writes its own slice
  std::vector<int64_t> at((size_t)g->n + 1, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < g->n; ++v) {
    int64_t b = 0;
    const int dv = digits((uint64_t)v);
    for (uint32_t j = g->out_off[v]; j < g->out_off[v + 1]; ++j) b += dv + 1 + digits((uint64_t)g->out_adj[j]) + 1;
    at[(size_t)v + 1] = b;
  }
  at[0] = hl;
  for (int32_t v = 0; v < g->n; ++v) at[(size_t)v + 1] += at[v];
  *bytes = at[g->n];
  if (!out) return MSC_OK;
  if (cap < at[g->n]) return MSC_E_CAPACITY;
  std::memcpy(out, head, (size_t)hl);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < g->n; ++v) {
    char* p = out + at[v];
    char num[16];
    const int dv = std::snprintf(num, sizeof(num), "%d", v);
    for (uint32_t j = g->out_off[v]; j < g->out_off[v + 1]; ++j) {
      std::memcpy(p, num, (size_t)dv);
      p += dv;
      *p++ = ' ';
      p += std::snprintf(p, 16, "%d", g->out_adj[j]);  // writes its own NUL, overwritten next
      *p++ = '\n';
    }
  }
  return MSC_OK;
}

int msc_scc_verify(int32_t n, const int32_t* labels, msc_verify* out) {
  if (n < 0 || !out || (n > 0 && !labels)) return MSC_E_INVALID;
  try {
    count_components(n, labels, out);
  } catch (const std::bad_alloc&) {
    return MSC_E_OOM;
  }
  return MSC_OK;
}

int msc_cc_verify(int32_t n, const int32_t* labels, msc_verify* out) {
  int rc = msc_scc_verify(n, labels, out);
  if (rc == MSC_OK) out->nontrivial = 0;  // cc_verify prints no nontrivial count
  return rc;
}

int msc_verify_format(const msc_verify* v, int32_t kind, char* buf, int32_t cap) {
  if (!v || !buf || cap <= 0) return MSC_E_INVALID;
  int len;
  if (kind == 0)  // multistep/scc_verify.cpp:73-74
    len = std::snprintf(buf, (size_t)cap, "Num SCCs: %d, Nontrivial: %d, Max SCC: %d, Unassigned %d",
                        v->num_components, v->nontrivial, v->max_size, v->unassigned);
  else  // multistep/cc_verify.cpp:70-71
    len = std::snprintf(buf, (size_t)cap, "Num CCs: %d, Max CC: %d, Unassigned %d",
                        v->num_components, v->max_size, v->unassigned);
  return (len >= 0 && len < cap) ? MSC_OK : MSC_E_CAPACITY;
}

int msc_canonicalize(int32_t n, int32_t* labels) {
  if (n < 0 || (n > 0 && !labels)) return MSC_E_INVALID;
  for (int32_t v = 0; v < n; ++v)
    if (labels[v] >= n) return MSC_E_INVALID;
  try {
    std::vector<int32_t> mins(n > 0 ? n : 1, INT32_MAX);
    for (int32_t v = 0; v < n; ++v)
      if (labels[v] >= 0 && v < mins[labels[v]]) mins[labels[v]] = v;
    for (int32_t v = 0; v < n; ++v)
      if (labels[v] >= 0) labels[v] = mins[labels[v]];
  } catch (const std::bad_alloc&) {
    return MSC_E_OOM;
  }
  return MSC_OK;
}

}  // extern "C"
