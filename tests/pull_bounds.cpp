// CPU restatement of the index arithmetic of one pull level (pull_group and its grid loop
// in k_level<OP_PUSH> of connectivity_amd/csrc/msc_kernels.hip), with a bounds check on every
// index it forms.  Test infrastructure: compiled and run by tests/test_pull_bounds.py on the
// host; nothing here runs on a GPU.
//
// For every grid position p < xcd_positions(groups) the kernel takes word group
// g = xcd_group(p, groups) (skipped when g >= groups) and, for every 256 set bits of the
// valid words [g*256, g*256 + 256) below nwords, forms: the candidate v (the k-th set bit),
// its list bounds up_off[v], up_off[v + 1] (and up_off2 when mirrored), the exclusive prefix
// of the candidates' list lengths, for every edge e < T the candidate lo found by the
// kernel's binary lifting over that prefix, the list index (up_adj[beg[lo] + i], or
// up_adj2[beg2[lo] + i - len1[lo]]), the neighbour u, its valid word u >> 5 and colour u,
// the LDS slots pref[lo], acc[lo], and the stage bound of the raised vertices.
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

constexpr uint32_t kBlock = 256;

uint32_t xcd_positions(uint32_t groups) { return (groups + 7u) & ~7u; }
uint32_t xcd_group(uint32_t pos, uint32_t groups) {
  const uint32_t per = (groups + 7u) / 8u;
  return (pos % 8u) * per + pos / 8u;
}

struct Err {
  uint64_t count = 0;
  void fail(const char* what, uint64_t a, uint64_t b) {
    if (count < 8) std::fprintf(stderr, "pull_bounds: %s: %llu >= %llu\n", what, (unsigned long long)a,
                                (unsigned long long)b);
    ++count;
  }
  void lt(const char* what, uint64_t a, uint64_t b) {
    if (!(a < b)) fail(what, a, b);
  }
};

}  // namespace

// Returns the number of out-of-bounds indices found (0: every index is in bounds); *checks
// receives how many list entries were walked.
extern "C" long long pull_bounds_check(int32_t n, const uint32_t* up_off, const int32_t* up_adj,
                                       const uint32_t* up_off2, const int32_t* up_adj2, const uint32_t* down_off,
                                       const uint32_t* down_off2, const uint32_t* valid, int64_t nwords,
                                       int64_t* checks) {
  Err err;
  const uint64_t m1 = up_off[n], m2 = up_off2 ? up_off2[n] : 0;
  const uint32_t groups = (uint32_t)((nwords + kBlock - 1) / kBlock), pos = xcd_positions(groups);
  int64_t walked = 0;
  std::vector<uint32_t> vs, beg, len1, beg2, pref(kBlock + 1);
  for (uint32_t p = 0; p < pos; ++p) {
    const uint32_t g = xcd_group(p, groups);
    if (g >= groups) continue;
    const int64_t wb = (int64_t)g * kBlock;
    // candidates: set bits of the group's valid words, in order
    vs.clear();
    for (uint32_t t = 0; t < kBlock; ++t) {
      const int64_t w = wb + t;
      if (w >= nwords) break;
      for (uint32_t b = 0; b < 32; ++b)
        if ((valid[w] >> b) & 1u) vs.push_back((uint32_t)(w * 32 + b));
    }
    for (size_t k0 = 0; k0 < vs.size(); k0 += kBlock) {
      const uint32_t nb = (uint32_t)std::min<size_t>(vs.size() - k0, kBlock);
      beg.assign(kBlock, 0);
      len1.assign(kBlock, 0);
      beg2.assign(kBlock, 0);
      uint32_t T = 0;
      for (uint32_t t = 0; t < kBlock; ++t) {
        uint32_t l = 0;
        if (t < nb) {
          const uint32_t v = vs[k0 + t];
          err.lt("candidate", v, (uint64_t)n);
          err.lt("up_off[v + 1]", (uint64_t)v + 1, (uint64_t)n + 1);
          beg[t] = up_off[v];
          len1[t] = up_off[v + 1] - up_off[v];
          l = len1[t];
          if (up_off2) {
            beg2[t] = up_off2[v];
            l += up_off2[v + 1] - up_off2[v];
          }
          err.lt("down_off[v + 1]", (uint64_t)v + 1, (uint64_t)n + 1);
          (void)down_off;
          (void)down_off2;
        }
        pref[t] = T;
        T += l;
      }
      pref[kBlock] = T;
      for (uint32_t e = 0; e < T; ++e) {
        uint32_t lo = 0;
        for (uint32_t step = kBlock / 2; step >= 1; step >>= 1)
          if (lo + step < nb && pref[lo + step] <= e) lo += step;
        err.lt("lo", lo, nb);
        const uint32_t i = e - pref[lo];
        const uint32_t l = (lo + 1 < nb ? pref[lo + 1] : T) - pref[lo];
        err.lt("edge within the candidate's lists", i, l);
        uint64_t u;
        if (up_off2 && i >= len1[lo]) {
          const uint64_t at = (uint64_t)beg2[lo] + (i - len1[lo]);
          err.lt("up_adj2 index", at, m2);
          u = at < m2 ? (uint64_t)(uint32_t)up_adj2[at] : 0;
        } else {
          const uint64_t at = (uint64_t)beg[lo] + i;
          err.lt("up_adj index", at, m1);
          u = at < m1 ? (uint64_t)(uint32_t)up_adj[at] : 0;
        }
        err.lt("neighbour u (colour index)", u, (uint64_t)n);
        err.lt("valid word of u", u >> 5, (uint64_t)nwords);
        err.lt("LDS acc slot", lo, kBlock);
        ++walked;
      }
      // every candidate may be raised and staged once: at most nb on top of the
      // 2048 - 256 the stage may hold after a flush
      err.lt("stage slots", (uint64_t)(2048 - 256) + nb - 1, 2048);
    }
  }
  *checks = walked;
  return (long long)err.count;
}
