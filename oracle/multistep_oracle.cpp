// ORACLE — test infrastructure only (see oracle.h).
//
// CPU restatement of the reference's Multistep SCC and CC/WCC.  Each function cites the
// reference code it follows.  Differences from the reference, all deliberate and none of
// them changing the partition or the deterministic raw labels:
//   * flag claims use atomic exchange, so frontiers carry no duplicate vertices (the
//     reference's `fw[out] = true` is a benign race that can enqueue a vertex twice,
//     multistep/scc_fwbw.cpp:97-100);
//   * colours are raised with an atomic max, so no update is lost and the source vertex
//     need not re-queue itself (multistep/scc_color.cpp:116-131 does, to repair races);
//   * every compaction keeps the input order, so the Tarjan tail's DFS roots are those of
//     a one-thread reference run;
//   * flags of vertices outside the current valid set read as false (the reference leaves
//     them uninitialised, e.g. multistep/scc_fwbw.cpp:59-61);
//   * Tarjan is iterative (the reference's is recursive, multistep/scc_serial.cpp:216-258,
//     and its depth reaches n on deep graphs);
//   * CC's update_valid walks num_valid entries (multistep/cc_bfs.cpp:231 walks g.n and
//     reads past the initialised list).
#include "oracle.h"

#include <omp.h>
#include <sys/time.h>

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

using u8 = uint8_t;

inline u8 ld(const u8* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void st(u8* p, u8 v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
inline u8 xchg(u8* p, u8 v) { return __atomic_exchange_n(p, v, __ATOMIC_RELAXED); }
inline int32_t ldi(const int32_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }

// Raise *p to at least v; returns the value seen before.
inline int32_t atomic_max(int32_t* p, int32_t v) {
  int32_t cur = ldi(p);
  while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED)) {
  }
  return cur;
}

double now() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

// Adjacency seen by a search: one CSR, or two concatenated per vertex (CC's mirror).
struct Lists {
  const uint32_t* o1;
  const int32_t* a1;
  const uint32_t* o2;
  const int32_t* a2;
  // Calls f(w) for each neighbour; stops early and returns true when f returns true.
  template <class F>
  bool any(int32_t v, F&& f) const {
    for (uint32_t j = o1[v], e = o1[v + 1]; j < e; ++j)
      if (f(a1[j])) return true;
    if (o2)
      for (uint32_t j = o2[v], e = o2[v + 1]; j < e; ++j)
        if (f(a2[j])) return true;
    return false;
  }
};

// Per-thread staging buffer appended to a shared array with one atomic add per flush.
struct Stage {
  int32_t* out;
  int64_t* count;
  std::vector<int32_t> buf;
  Stage(int32_t* o, int64_t* c) : out(o), count(c) { buf.reserve(2048); }
  void push(int32_t v) {
    buf.push_back(v);
    if (buf.size() == 2048) flush();
  }
  void flush() {
    if (buf.empty()) return;
    int64_t at = __atomic_fetch_add(count, (int64_t)buf.size(), __ATOMIC_RELAXED);
    std::memcpy(out + at, buf.data(), buf.size() * sizeof(int32_t));
    buf.clear();
  }
};

// Order-preserving parallel filter: keeps in[i] for which keep(in[i]) holds.
template <class P>
std::vector<int32_t> filter_ordered(const int32_t* in, int64_t count, P&& keep) {
  int nt = omp_get_max_threads();
  std::vector<int64_t> part(nt + 1, 0);
  std::vector<int32_t> out;
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    int64_t lo = count * t / nt, hi = count * (t + 1) / nt, c = 0;
    for (int64_t i = lo; i < hi; ++i) c += keep(in[i]) ? 1 : 0;
    part[t + 1] = c;
#pragma omp barrier
#pragma omp single
    {
      for (int k = 0; k < nt; ++k) part[k + 1] += part[k];
      out.resize(part[nt]);
    }
    int64_t at = part[t];
    for (int64_t i = lo; i < hi; ++i)
      if (keep(in[i])) out[at++] = in[i];
  }
  return out;
}

void trace_level(orc_trace* tr, int phase, int level, bool bu, int64_t next) {
  if (!tr) return;
  if (tr->levels && tr->count < tr->cap) {
    orc_level& L = tr->levels[tr->count];
    L.phase = phase;
    L.level = level;
    L.bottom_up = bu ? 1 : 0;
    L.next_size = (int32_t)next;
  }
  tr->count++;
}

// Direction-optimising reachability search from `root`, after multistep/scc_fwbw.cpp:47-227
// (forward), :230-409 (backward) and multistep/cc_bfs.cpp:47-214 (CC).
//   top-down   (scc_fwbw.cpp:87-114): expand `down` lists of the frontier; claim w when
//              allow[w] && !seen[w].
//   bottom-up  (scc_fwbw.cpp:118-148): for each v of `verts` with allow[v] && !seen[v], scan
//              its `up` list and claim v at the first u with seen[u] && allow[u].
//   switch     (scc_fwbw.cpp:164-195): after a top-down level go bottom-up when
//              (n - found)*avg/alpha < next*avg; from bottom-up return to top-down for good
//              when n/beta > next (integer division).  `found` excludes the root.
// Returns the number of vertices discovered (root excluded, like the reference's counter).
int64_t reach(const Lists& down, const Lists& up, const u8* allow, u8* seen,
              const int32_t* verts, int64_t nverts, int32_t root, int32_t n_total, double avg,
              double alpha, int32_t beta, int phase, orc_trace* tr) {
  std::vector<int32_t> q(nverts + 1), qn(nverts + 1);
  int64_t qs = 0;
  q[qs++] = root;
  st(&seen[root], 1);
  bool bottom_up = false, switched_back = false;
  int64_t found = 0;
  int level = 0;
  while (qs > 0) {
    int64_t next = 0;
#pragma omp parallel
    {
      Stage s(qn.data(), &next);
      if (!bottom_up) {
#pragma omp for schedule(guided) nowait
        for (int64_t i = 0; i < qs; ++i) {
          down.any(q[i], [&](int32_t w) {
            if (ld(&allow[w]) && !ld(&seen[w]) && !xchg(&seen[w], 1)) s.push(w);
            return false;
          });
        }
      } else {
#pragma omp for schedule(guided) nowait
        for (int64_t i = 0; i < nverts; ++i) {
          int32_t v = verts[i];
          if (!ld(&allow[v]) || ld(&seen[v])) continue;
          bool hit = up.any(v, [&](int32_t u) { return ld(&seen[u]) && ld(&allow[u]); });
          if (hit) {
            st(&seen[v], 1);
            s.push(v);
          }
        }
      }
      s.flush();
    }
    found += next;
    ++level;
    trace_level(tr, phase, level, bottom_up, next);
    if (!switched_back && !bottom_up) {
      double edges_frontier = (double)next * avg;
      double edges_remainder = (double)(n_total - found) * avg;
      if (edges_remainder / alpha < edges_frontier) bottom_up = true;
    } else if (!switched_back && (n_total / beta) > next) {
      bottom_up = false;
      switched_back = true;
    }
    std::swap(q, qn);
    qs = next;
  }
  return found;
}

// Max-label fixpoint over `lists` restricted to valid vertices, after
// multistep/scc_color.cpp:47-163 and multistep/cc_color.cpp:47-175:
// colors[w] = max id of a valid vertex that reaches w.  Returns the number of sweeps.
int propagate_max(const Lists& lists, const u8* valid, const std::vector<int32_t>& verts,
                  int32_t* colors, int32_t n) {
  std::vector<u8> inq(n, 0), inq_next(n, 0);
  std::vector<int32_t> q(verts), qn(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)verts.size(); ++i) {
    colors[verts[i]] = verts[i];
    inq[verts[i]] = 1;
  }
  int64_t qs = (int64_t)verts.size();
  int sweeps = 0;
  while (qs > 0) {
    int64_t next = 0;
#pragma omp parallel
    {
      Stage s(qn.data(), &next);
#pragma omp for schedule(guided) nowait
      for (int64_t i = 0; i < qs; ++i) {
        int32_t v = q[i];
        st(&inq[v], 0);
        int32_t c = ldi(&colors[v]);
        lists.any(v, [&](int32_t w) {
          if (ld(&valid[w]) && ldi(&colors[w]) < c && atomic_max(&colors[w], c) < c &&
              !xchg(&inq_next[w], 1))
            s.push(w);
          return false;
        });
      }
      s.flush();
    }
    std::swap(q, qn);
    std::swap(inq, inq_next);
    qs = next;
    ++sweeps;
  }
  return sweeps;
}

// Iterative Tarjan over `order`'s vertices restricted to valid ones, after
// multistep/scc_serial.cpp:160-172 (run_valid) and :216-258 (scc_serial_valid).  Visits
// vertices and out-edges in the same order as the recursive version, so each component is
// labelled with the same DFS root.  Returns the number of components.
// Complete trim (multistep/scc_trim.cpp:141-297, TRIM_LEVEL 2): a vertex goes when it has no
// valid out-neighbour other than itself or no valid in-neighbour other than itself, until
// none goes.  Restated as peeling with counters: cnt_out[v] / cnt_in[v] count v's list
// entries that are valid and not v (duplicates counted each time); a vertex whose count
// reaches zero is trimmed and decrements the counters of its valid neighbours.  The
// surviving set is the largest one in which every vertex keeps a valid out- and in-
// neighbour besides itself, so it does not depend on the order of removals (the
// reference's thread-order races only change the order).  Sequential on purpose.
int64_t trim_complete(int32_t n, const uint32_t* oo, const int32_t* oa, const uint32_t* io, const int32_t* ia,
                      u8* valid, int32_t* labels) {
  std::vector<int64_t> cnt_out(n, 0), cnt_in(n, 0);
  for (int32_t v = 0; v < n; ++v) {
    valid[v] = 1;
    for (uint32_t j = oo[v]; j < oo[v + 1]; ++j) cnt_out[v] += oa[j] != v;
    for (uint32_t j = io[v]; j < io[v + 1]; ++j) cnt_in[v] += ia[j] != v;
  }
  std::vector<int32_t> stack;
  for (int32_t v = 0; v < n; ++v)
    if (cnt_out[v] == 0 || cnt_in[v] == 0) stack.push_back(v);
  int64_t trimmed = 0;
  while (!stack.empty()) {
    int32_t v = stack.back();
    stack.pop_back();
    if (!valid[v]) continue;
    valid[v] = 0;
    ++trimmed;
    // v no longer counts as an in-neighbour of its out-neighbours, nor as an
    // out-neighbour of its in-neighbours
    for (uint32_t j = oo[v]; j < oo[v + 1]; ++j) {
      int32_t w = oa[j];
      if (w != v && valid[w] && --cnt_in[w] == 0) stack.push_back(w);
    }
    for (uint32_t j = io[v]; j < io[v + 1]; ++j) {
      int32_t w = ia[j];
      if (w != v && valid[w] && --cnt_out[w] == 0) stack.push_back(w);
    }
  }
  for (int32_t v = 0; v < n; ++v) labels[v] = valid[v] ? -1 : v;
  return trimmed;
}

int tarjan_valid(int32_t n, const uint32_t* oo, const int32_t* oa, const u8* valid,
                 const int32_t* order, int64_t count, int32_t* labels) {
  std::vector<int32_t> index(n, -1), low(n, -1), stack;
  std::vector<u8> on_stack(n, 0);
  std::vector<std::pair<int32_t, uint32_t>> call;  // (vertex, next edge position)
  int32_t next_index = 0;
  int comps = 0;
  for (int64_t i = 0; i < count; ++i) {
    int32_t s = order[i];
    if (index[s] != -1) continue;
    index[s] = low[s] = next_index++;
    stack.push_back(s);
    on_stack[s] = 1;
    call.emplace_back(s, oo[s]);
    while (!call.empty()) {
      int32_t v = call.back().first;
      uint32_t& j = call.back().second;
      bool descended = false;
      while (j < oo[v + 1]) {
        int32_t w = oa[j++];
        if (valid && !valid[w]) continue;
        if (index[w] == -1) {
          index[w] = low[w] = next_index++;
          stack.push_back(w);
          on_stack[w] = 1;
          call.emplace_back(w, oo[w]);
          descended = true;
          break;
        } else if (on_stack[w]) {
          low[v] = std::min(low[v], index[w]);
        }
      }
      if (descended) continue;
      if (low[v] == index[v]) {
        int32_t w;
        do {
          w = stack.back();
          stack.pop_back();
          on_stack[w] = 0;
          labels[w] = v;
        } while (w != v);
        ++comps;
      }
      call.pop_back();
      if (!call.empty()) {
        int32_t parent = call.back().first;
        low[parent] = std::min(low[parent], low[v]);
      }
    }
  }
  return comps;
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }

// multistep/scc_main.cpp:169-183.  The degree sum is an exact integer in a double, so
// avg = m/n whatever the summation order.
int orc_pivot_scc(int32_t n, const uint32_t* oo, const uint32_t* io, int32_t* pivot,
                  double* avg) {
  int32_t best = -1;
  double best_val = 0.0;
  double sum = 0.0;
  for (int32_t v = 0; v < n; ++v) {
    uint32_t od = oo[v + 1] - oo[v], id = io[v + 1] - io[v];
    double prod = (double)od * (double)id;
    sum += (double)od;
    if (prod > best_val) {
      best_val = prod;
      best = v;
    }
  }
  *pivot = best;
  *avg = n > 0 ? sum / (double)n : 0.0;
  return 0;
}

// multistep/cc_main.cpp:148-160 on the mirrored CSR: degree = out + in.
int orc_pivot_cc(int32_t n, const uint32_t* oo, const uint32_t* io, int32_t* pivot,
                 double* avg) {
  int32_t best = -1;
  uint32_t best_deg = 0;
  double sum = 0.0;
  for (int32_t v = 0; v < n; ++v) {
    uint32_t d = (oo[v + 1] - oo[v]) + (io[v + 1] - io[v]);
    sum += (double)d;
    if (d > best_deg) {
      best_deg = d;
      best = v;
    }
  }
  *pivot = best;
  *avg = n > 0 ? sum / (double)n : 0.0;
  return 0;
}

int orc_scc_run(int32_t n, const uint32_t* oo, const int32_t* oa, const uint32_t* io,
                const int32_t* ia, int32_t pivot, double avg, int32_t vert_cutoff,
                double alpha, int32_t beta, int32_t* labels, orc_trace* tr) {
  return orc_scc_run_trim(n, oo, oa, io, ia, pivot, avg, vert_cutoff, alpha, beta, 1, labels, tr);
}

int orc_trim_complete(int32_t n, const uint32_t* oo, const int32_t* oa, const uint32_t* io,
                      const int32_t* ia, int32_t* labels) {
  if (n < 0) return -1;
  std::vector<u8> valid(n > 0 ? n : 1);
  return (int)trim_complete(n, oo, oa, io, ia, valid.data(), labels);
}

int orc_scc_run_trim(int32_t n, const uint32_t* oo, const int32_t* oa, const uint32_t* io,
                     const int32_t* ia, int32_t pivot, double avg, int32_t vert_cutoff,
                     double alpha, int32_t beta, int32_t trim_level, int32_t* labels, orc_trace* tr) {
  if (n < 0 || (pivot >= n) || trim_level < 1 || trim_level > 2) return -1;
  if (tr) {
    tr->count = 0;
    tr->num_trim = tr->num_fwbw = tr->num_color = tr->num_serial = 0;
    tr->color_rounds = tr->color_iters = 0;
  }
  if (n == 0) return 0;
  std::vector<u8> valid(n, 1);

  // Trim (multistep/scc_trim.cpp:89-139): out- or in-degree 0 in the full graph; or the
  // complete trim of TRIM_LEVEL 2 (scc_main.cpp:253-258), after which the valid list is
  // rebuilt in id order.  The pivot stays whatever the caller chose, trimmed or not.
  auto keeps = [&](int32_t v) {
    return trim_level == 2 ? valid[v] != 0 : (oo[v + 1] > oo[v]) && (io[v + 1] > io[v]);
  };
  if (trim_level == 2) {
    trim_complete(n, oo, oa, io, ia, valid.data(), labels);
  } else {
#pragma omp parallel for schedule(static)
    for (int32_t v = 0; v < n; ++v) {
      bool k = keeps(v);
      valid[v] = k ? 1 : 0;
      labels[v] = k ? -1 : v;
    }
  }
  std::vector<int32_t> all(n);
  for (int32_t v = 0; v < n; ++v) all[v] = v;
  std::vector<int32_t> verts = filter_ordered(all.data(), n, [&](int32_t v) { return keeps(v); });
  std::vector<int32_t>().swap(all);
  if (tr) tr->num_trim = n - (int32_t)verts.size();

  // FW-BW from the pivot (multistep/scc_fwbw.cpp:469-513).
  if (pivot >= 0) {
    std::vector<u8> fw(n, 0), scc(n, 0);
    Lists outs{oo, oa, nullptr, nullptr}, ins{io, ia, nullptr, nullptr};
    reach(outs, ins, valid.data(), fw.data(), verts.data(), (int64_t)verts.size(), pivot, n,
          avg, alpha, beta, ORC_PHASE_FW, tr);
    reach(ins, outs, fw.data(), scc.data(), verts.data(), (int64_t)verts.size(), pivot, n, avg,
          alpha, beta, ORC_PHASE_BW, tr);
    // update_valid (multistep/scc_fwbw.cpp:412-466)
    int64_t before = (int64_t)verts.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < before; ++i) {
      int32_t v = verts[i];
      if (scc[v]) {
        valid[v] = 0;
        labels[v] = pivot;
      }
    }
    verts = filter_ordered(verts.data(), before, [&](int32_t v) { return valid[v] != 0; });
    if (tr) tr->num_fwbw = (int32_t)(before - (int64_t)verts.size());
  }

  // Colouring rounds while more than vert_cutoff remain (multistep/scc_color.cpp:340-373).
  Lists outs{oo, oa, nullptr, nullptr};
  std::vector<int32_t> colors(n, -1), roots(n), q2(n);
  while ((int64_t)verts.size() > (int64_t)vert_cutoff) {
    int sweeps = propagate_max(outs, valid.data(), verts, colors.data(), n);
    // roots: colors[v] == v (scc_color.cpp:165-212)
    int64_t nroots = 0;
    for (int32_t v : verts)
      if (colors[v] == v) {
        valid[v] = 0;
        labels[v] = v;
        roots[nroots++] = v;
      }
    // backward search within each colour (scc_color.cpp:214-289)
    std::vector<int32_t> q(roots.begin(), roots.begin() + nroots);
    int64_t qs = nroots;
    while (qs > 0) {
      int64_t next = 0;
#pragma omp parallel
      {
        Stage s(q2.data(), &next);
#pragma omp for schedule(guided) nowait
        for (int64_t i = 0; i < qs; ++i) {
          int32_t v = q[i];
          int32_t c = colors[v];
          for (uint32_t j = io[v]; j < io[v + 1]; ++j) {
            int32_t u = ia[j];
            if (ld(&valid[u]) && colors[u] == c && xchg(&valid[u], 0)) {
              labels[u] = c;
              s.push(u);
            }
          }
        }
        s.flush();
      }
      q.assign(q2.begin(), q2.begin() + next);
      qs = next;
    }
    // update_valid (scc_color.cpp:292-337)
    verts = filter_ordered(verts.data(), (int64_t)verts.size(),
                           [&](int32_t v) { return valid[v] != 0; });
    if (tr) {
      tr->num_color += (int32_t)nroots;
      tr->color_rounds += 1;
      tr->color_iters += sweeps;
    }
  }

  // Serial tail (multistep/scc_main.cpp:305-308, scc_serial.cpp:160-172).
  int comps = tarjan_valid(n, oo, oa, valid.data(), verts.data(), (int64_t)verts.size(), labels);
  if (tr) tr->num_serial = comps;
  return 0;
}

int orc_cc_run(int32_t n, const uint32_t* oo, const int32_t* oa, const uint32_t* io,
               const int32_t* ia, int32_t pivot, double avg, double alpha, int32_t beta,
               int32_t* labels, orc_trace* tr) {
  if (n < 0 || pivot >= n) return -1;
  if (tr) {
    tr->count = 0;
    tr->num_trim = tr->num_fwbw = tr->num_color = tr->num_serial = 0;
    tr->color_rounds = tr->color_iters = 0;
  }
  if (n == 0) return 0;
  Lists mirror{oo, oa, io, ia};
  std::vector<u8> valid(n, 1);
  // cc_trim (multistep/cc_trim.cpp:47-97): mirrored degree 0.
  auto keeps = [&](int32_t v) { return (oo[v + 1] > oo[v]) || (io[v + 1] > io[v]); };
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < n; ++v) {
    bool k = keeps(v);
    valid[v] = k ? 1 : 0;
    labels[v] = k ? -1 : v;
  }
  std::vector<int32_t> all(n);
  for (int32_t v = 0; v < n; ++v) all[v] = v;
  std::vector<int32_t> verts = filter_ordered(all.data(), n, [&](int32_t v) { return keeps(v); });
  std::vector<int32_t>().swap(all);
  if (tr) tr->num_trim = n - (int32_t)verts.size();

  // cc_bfs (multistep/cc_bfs.cpp:273-291): search from the pivot, then update_valid.
  if (pivot >= 0) {
    std::vector<u8> seen(n, 0);
    reach(mirror, mirror, valid.data(), seen.data(), verts.data(), (int64_t)verts.size(), pivot,
          n, avg, alpha, beta, ORC_PHASE_CC, tr);
    int64_t before = (int64_t)verts.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < before; ++i) {
      int32_t v = verts[i];
      if (seen[v]) {
        valid[v] = 0;
        labels[v] = pivot;
      }
    }
    verts = filter_ordered(verts.data(), before, [&](int32_t v) { return valid[v] != 0; });
    if (tr) tr->num_fwbw = (int32_t)(before - (int64_t)verts.size());
  }

  // cc_color (multistep/cc_color.cpp:197-214): max-label fixpoint written into the labels.
  int sweeps = propagate_max(mirror, valid.data(), verts, labels, n);
  if (tr) {
    tr->color_rounds = verts.empty() ? 0 : 1;
    tr->color_iters = sweeps;
    int32_t roots = 0;
    for (int32_t v : verts) roots += (labels[v] == v) ? 1 : 0;
    tr->num_color = roots;
  }
  return 0;
}

void orc_canonicalize(int32_t n, int32_t* labels) {
  std::vector<int32_t> mins(n, INT32_MAX);
  for (int32_t v = 0; v < n; ++v)
    if (labels[v] >= 0 && v < mins[labels[v]]) mins[labels[v]] = v;
  for (int32_t v = 0; v < n; ++v)
    if (labels[v] >= 0) labels[v] = mins[labels[v]];
}

// multistep/scc_verify.cpp:47-77
void orc_verify_counts(int32_t n, const int32_t* labels, int32_t* out) {
  std::vector<int32_t> counts(n > 0 ? n : 1, 0);
  int32_t unassigned = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (labels[v] >= 0)
      ++counts[labels[v]];
    else
      ++unassigned;
  }
  int32_t comps = 0, nontrivial = 0, biggest = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (!counts[v]) continue;
    ++comps;
    if (counts[v] > 1) ++nontrivial;
    biggest = std::max(biggest, counts[v]);
  }
  out[0] = comps;
  out[1] = nontrivial;
  out[2] = biggest;
  out[3] = unassigned;
}

int orc_tarjan(int32_t n, const uint32_t* oo, const int32_t* oa, int32_t* labels) {
  std::vector<int32_t> order(n);
  for (int32_t v = 0; v < n; ++v) order[v] = v;
  int comps = tarjan_valid(n, oo, oa, nullptr, order.data(), n, labels);
  orc_canonicalize(n, labels);
  return comps;
}

int orc_wcc_unionfind(int32_t n, const uint32_t* oo, const int32_t* oa, int32_t* labels) {
  std::vector<int32_t> parent(n);
  for (int32_t v = 0; v < n; ++v) parent[v] = v;
  auto find = [&](int32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (int32_t v = 0; v < n; ++v)
    for (uint32_t j = oo[v]; j < oo[v + 1]; ++j) {
      int32_t a = find(v), b = find(oa[j]);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
  int comps = 0;
  for (int32_t v = 0; v < n; ++v) {
    labels[v] = find(v);
    comps += labels[v] == v;
  }
  orc_canonicalize(n, labels);
  return comps;
}

double orc_scc_time(int32_t n, const uint32_t* oo, const int32_t* oa, const uint32_t* io,
                    const int32_t* ia, int32_t pivot, double avg, int32_t vert_cutoff,
                    int32_t* labels) {
  double t0 = now();
  orc_scc_run(n, oo, oa, io, ia, pivot, avg, vert_cutoff, 15.0, 25, labels, nullptr);
  return now() - t0;
}

}  // extern "C"
