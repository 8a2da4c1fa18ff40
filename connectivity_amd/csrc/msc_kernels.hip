// gfx950 kernels of the Multistep SCC / WCC path.
//
// Design (for a reader with only these sources):
//  * Vertex state lives in bitmaps (valid / claimable / worklist membership): one bit per
//    vertex, 2 MB at 16.8M vertices, so random probes are served from L2 / MALL.  Claims
//    are atomicOr / atomicAnd test-and-set on the containing word, so a frontier never
//    holds a vertex twice.  Sweeps over "all candidates" walk the bitmap itself: each
//    workgroup iteration loads kBW words, prefix-sums their popcounts in LDS and hands the
//    set bits to its threads, so no compacted vertex list has to be maintained.
//  * Same-address atomics serialise in L2 (~10 ns each when every workgroup hits one
//    counter at once), and an agent-scope fence on multi-XCD gfx950 costs about as much
//    again.  So nothing here takes one per workgroup per launch on a shared word:
//    workgroups write partial results to their own slot of a partials array, and a
//    one-workgroup kernel reduces them; levels are closed by such a kernel (the launch
//    boundary provides the visibility a fence would).
//  * A frontier is a queue plus, per slot, the exclusive prefix of the slots' list lengths
//    ("edge offsets").  Claims are staged per workgroup in LDS; when the stage fills (and
//    at the end of the launch) the workgroup scans the staged lengths and reserves queue
//    slots and edge offsets together with ONE 64-bit atomic ((count << 32) | edge sum).
//  * While writing a slot, the workgroup also records, for every kChunk-edge chunk that
//    begins inside that vertex's list, which slot owns the chunk ("starts").  A top-down
//    level gives each workgroup whole chunks of exactly kChunk edges: it maps every edge
//    to its source vertex with an LDS max-scan and reads the adjacency coalesced.  Hubs
//    (the s24 pivot has ~240k out-edges) spread over hundreds of workgroups.
//  * No MFMA: no step is a dense contraction.  Everything is HBM / L2 bound integer work.
#include <hip/hip_runtime.h>

#include "msc_internal.h"

// bottom-up candidates per thread per round (all their loads in flight together)
#ifndef MSC_BU_KR
#define MSC_BU_KR 4
#endif
// list entries a bottom-up candidate probes at a time after its first probe missed
#ifndef MSC_BU_SCAN
#define MSC_BU_SCAN 2
#endif
// 1: a level where both searches of a pivot are bottom-up runs as one fused sweep (bu_pair)
#ifndef MSC_BU_PAIR
#define MSC_BU_PAIR 1
#endif

namespace msc {
namespace {

// ------------------------------------------------------------------------------------
// wave64 and bitmap helpers

__device__ __forceinline__ int lane() { return (int)__lane_id(); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x >> 6); }

__device__ __forceinline__ unsigned long long lanes_below() { return (1ull << lane()) - 1ull; }

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(x, o, 64);
    if (lane() >= o) x += t;
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_incl_max(uint32_t x) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(x, o, 64);
    if (lane() >= o) x = max(x, t);
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor(x, o, 64));
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor((uint32_t)x, o, 64);
    uint32_t hi = __shfl_xor((uint32_t)(x >> 32), o, 64);
    x += ((unsigned long long)hi << 32) | lo;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_max64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor((uint32_t)x, o, 64);
    uint32_t hi = __shfl_xor((uint32_t)(x >> 32), o, 64);
    unsigned long long y = ((unsigned long long)hi << 32) | lo;
    x = y > x ? y : x;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_min64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor((uint32_t)x, o, 64);
    uint32_t hi = __shfl_xor((uint32_t)(x >> 32), o, 64);
    unsigned long long y = ((unsigned long long)hi << 32) | lo;
    x = y < x ? y : x;
  }
  return x;
}

__device__ __forceinline__ bool bit_test(const uint32_t* b, uint32_t v) {
  return (b[v >> 5] >> (v & 31u)) & 1u;
}
// true when this call turned the bit on
__device__ __forceinline__ bool bit_set(uint32_t* b, uint32_t v) {
  uint32_t m = 1u << (v & 31u);
  return (atomicOr(&b[v >> 5], m) & m) == 0u;
}
// true when this call turned the bit off
__device__ __forceinline__ bool bit_clear(uint32_t* b, uint32_t v) {
  uint32_t m = 1u << (v & 31u);
  return (atomicAnd(&b[v >> 5], ~m) & m) != 0u;
}

// Bitmap words of a pull or bottom-up level's word group: a.group_words, at most kBlock (one
// word per thread).
__device__ __forceinline__ int64_t group_words(const LevelArgs& a) {
  return a.group_words >= 1u && a.group_words <= (uint32_t)kBlock ? (int64_t)a.group_words : (int64_t)kBlock;
}

// Words of a (not wide) bottom-up level's group: narrow groups need one word per thread.
__device__ __forceinline__ int64_t bu_words(const LevelArgs& a) { return kWPT == 1 ? group_words(a) : (int64_t)kBW; }

__device__ __forceinline__ uint32_t list_len(const uint32_t* off, const uint32_t* off2, uint32_t v) {
  uint32_t d = off[v + 1] - off[v];
  if (off2) d += off2[v + 1] - off2[v];
  return d;
}

// Block-wide sum of a per-thread value; result valid in thread 0.
__device__ __forceinline__ unsigned long long block_sum64(unsigned long long x, unsigned long long* s_red) {
  x = wave_sum64(x);
  if (lane() == 0) s_red[wave_id()] = x;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) t += s_red[i];
  __syncthreads();
  return t;
}

// ------------------------------------------------------------------------------------
// workgroup-level walk over the set bits of a bitmap

// A group of kBlock * WPT bitmap words: the words and the exclusive prefix of their popcounts.
template <int WPT>
struct BitIterT {
  static constexpr int kWords = kBlock * WPT;
  uint32_t word[kWords];
  uint32_t pref[kWords + 1];
  uint32_t wsum[kBlock / 64];
};
using BitIter = BitIterT<kWPT>;  // the bottom-up search's word groups (kBW words)
static_assert(BitIter::kWords == kBW, "bottom-up group width");

// Loads words [wb, wb + kBlock * WPT) (WPT consecutive words per thread) and the exclusive
// prefix of their popcounts; returns the set bits in the group.  All threads call; ends
// with a barrier.
template <int WPT>
__device__ __forceinline__ uint32_t bits_scan(BitIterT<WPT>& B, const uint32_t* w);

template <int WPT>
__device__ __forceinline__ uint32_t bits_load(BitIterT<WPT>& B, const uint32_t* bits, int64_t nwords, int64_t wb) {
  const int t = threadIdx.x;
  const int64_t w0 = wb + (int64_t)t * WPT;
  uint32_t w[WPT];
#pragma unroll
  for (int j = 0; j < WPT; ++j) w[j] = w0 + j < nwords ? bits[w0 + j] : 0u;
  return bits_scan(B, w);
}

// The group's words (WPT per thread, given) into B with the prefix of their popcounts;
// returns the set bits in the group.  All threads call; ends with a barrier.
template <int WPT>
__device__ __forceinline__ uint32_t bits_scan(BitIterT<WPT>& B, const uint32_t* w) {
  constexpr int kW = BitIterT<WPT>::kWords;
  const int t = threadIdx.x;
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < WPT; ++j) {
    B.word[t * WPT + j] = w[j];
    cnt += (uint32_t)__popc(w[j]);
  }
  const uint32_t incl = wave_incl_sum(cnt);
  if (lane() == 63) B.wsum[wave_id()] = incl;
  __syncthreads();
  uint32_t carry = 0;
  for (int i = 0; i < wave_id(); ++i) carry += B.wsum[i];
  uint32_t run = incl - cnt + carry;  // exclusive prefix of this thread's first word
#pragma unroll
  for (int j = 0; j < WPT; ++j) {
    run += (uint32_t)__popc(w[j]);
    B.pref[t * WPT + j + 1] = run;
  }
  if (t == 0) B.pref[0] = 0u;
  __syncthreads();
  return B.pref[kW];
}

// Vertex id of the k-th set bit (k < total) of the loaded words.
template <int WPT>
__device__ __forceinline__ uint32_t bits_pick(const BitIterT<WPT>& B, int64_t wb, uint32_t k) {
  constexpr int kW = BitIterT<WPT>::kWords;
  int lo = 0, hi = kW;  // pref[lo] <= k < pref[hi]
#pragma unroll
  for (int s = 0; (1 << s) < kW; ++s) {
    int mid = (lo + hi) >> 1;
    if (B.pref[mid] <= k) lo = mid;
    else hi = mid;
  }
  uint32_t w = B.word[lo], j = k - B.pref[lo], pos = 0;
#pragma unroll
  for (int width = 16; width >= 1; width >>= 1) {  // j-th set bit of w
    uint32_t low = w & ((1u << width) - 1u);
    uint32_t c = (uint32_t)__popc(low);
    if (j >= c) {
      j -= c;
      w >>= width;
      pos += (uint32_t)width;
    } else {
      w = low;
    }
  }
  return (uint32_t)((wb + lo) * 32 + pos);
}

// ------------------------------------------------------------------------------------
// per-workgroup staging of claims

struct Stage {
  int32_t v[kStage];
  uint32_t d[kStage];  // list length, then (after the scan) local edge offset
  uint32_t part[kBlock / 64];
  uint32_t cnt;
  uint32_t base_hi, base_lo;
  uint32_t go;  // local rounds: thread 0's decision to run another one
};

// Wave-cooperative: appends the lanes with `take` to the stage (one LDS atomic per wave).
// The caller guarantees room (stage_flush runs before the stage can overflow).
__device__ __forceinline__ void stage_push(Stage& S, bool take, uint32_t w, uint32_t deg) {
  unsigned long long mask = __ballot(take);
  if (mask == 0ull) return;
  int leader = __ffsll((long long)mask) - 1;
  uint32_t at = 0;
  if (lane() == leader) at = atomicAdd(&S.cnt, (uint32_t)__popcll(mask));
  at = __shfl(at, leader, 64);
  if (take) {
    uint32_t slot = at + (uint32_t)__popcll(mask & lanes_below());
    S.v[slot] = (int32_t)w;
    S.d[slot] = deg;
  }
}

// Sum of the list lengths of the n staged claims (every thread gets it; LDS only).
__device__ __forceinline__ uint32_t stage_sum(Stage& S, uint32_t n) {
  uint32_t sum = 0;
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) sum += S.d[i];
  sum = wave_incl_sum(sum);
  if (lane() == 63) S.part[wave_id()] = sum;
  __syncthreads();
  uint32_t total = 0;
  for (int w = 0; w < kBlock / 64; ++w) total += S.part[w];
  __syncthreads();  // S.part is reused by the next scan
  return total;
}

// Turns the list lengths of the n staged claims into exclusive local edge offsets and
// returns their sum (every thread gets it).  No trailing barrier.
__device__ __forceinline__ uint32_t stage_scan(Stage& S, uint32_t n) {
  constexpr int per = (int)(kStage / kBlock);
  const int t = threadIdx.x;
  uint32_t loc[per];
  uint32_t sum = 0;
#pragma unroll
  for (int r = 0; r < per; ++r) {
    uint32_t i = (uint32_t)t * per + r;
    loc[r] = sum;
    sum += i < n ? S.d[i] : 0u;
  }
  uint32_t incl = wave_incl_sum(sum);
  if (lane() == 63) S.part[wave_id()] = incl;
  __syncthreads();
  uint32_t carry = 0, total = 0;
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave_id()) carry += S.part[w];
    total += S.part[w];
  }
  const uint32_t excl = incl - sum + carry;
#pragma unroll
  for (int r = 0; r < per; ++r) {
    uint32_t i = (uint32_t)t * per + r;
    if (i < n) S.d[i] = excl + loc[r];
  }
  return total;
}

// The colour tail's share of a seed flush (TailSeed): the n scanned entries at slots
// [bh, bh + n) whose slot is below ts.maxv record their slots, and every list entry of
// theirs (one thread per entry, its owner by binary search over the scanned offsets) whose
// target is valid and not the vertex itself is appended to ts.src / ts.dst.
__device__ __forceinline__ void tail_collect(const Stage& S, uint32_t n, uint32_t total, uint32_t bh,
                                             const TailSeed& ts) {
  if (bh >= ts.maxv) return;  // uniform
  const uint32_t ne = min(n, ts.maxv - bh);
  for (uint32_t i = threadIdx.x; i < ne; i += kBlock) ts.slots[S.v[i]] = (int32_t)(bh + i);
  if (!ts.src) return;  // uniform
  const uint32_t edges = ne < n ? S.d[ne] : total;
  for (uint32_t e0 = 0; e0 < edges; e0 += kBlock) {
    const uint32_t e = e0 + threadIdx.x;
    bool ok = false;
    uint32_t pos = 0;
    int32_t w = 0;
    if (e < edges) {
      // the last entry whose offset is at most e (offsets rise: listed vertices have lists)
      uint32_t lo = 0, hi = ne;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (S.d[mid] <= e) lo = mid;
        else hi = mid;
      }
      const int32_t v = S.v[lo];
      const uint32_t j = e - S.d[lo];
      const uint32_t beg = ts.off[v], len = ts.off[v + 1] - beg;  // out-list part (CC: then in-list)
      if (j < len) {
        w = ts.adj[beg + j];
        ok = w != v && bit_test(ts.valid, (uint32_t)w);
      }
      pos = bh + lo;
    }
    const unsigned long long mask = __ballot(ok);
    if (mask == 0ull) continue;  // uniform over the wave
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t at = 0;
    if (lane() == leader) at = atomicAdd(ts.count, (uint32_t)__popcll(mask));
    at = __shfl(at, leader, 64) + (uint32_t)__popcll(mask & lanes_below());
    if (ok && at < ts.cap) {
      ts.src[at] = pos;
      ts.dst[at] = w;
    }
  }
}

// Writes the n scanned claims to frontier nx at queue slots [bh, bh + n) and edge offsets
// from bl: ids, offsets, the chunk starts of every kChunk-edge chunk that begins inside a
// claim's list, and the ids to `list` when given (and the colour tail's share, with ts).
// Empties the stage.  Call after a barrier that follows stage_scan; ends with a barrier.
__device__ __forceinline__ void stage_write(Stage& S, uint32_t n, uint32_t total, const Frontier& nx, uint32_t bh,
                                            uint32_t bl, int32_t* list, const TailSeed* ts) {
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
    uint32_t lo = S.d[i];
    uint32_t hi = i + 1 < n ? S.d[i + 1] : total;
    uint32_t pos = bh + i, off = bl + lo;
    nx.q[pos] = S.v[i];
    nx.qo[pos] = off;
    if (list) list[pos] = S.v[i];
    uint64_t end = (uint64_t)off + (hi - lo);
    for (uint64_t ch = ((uint64_t)off + kChunk - 1) / kChunk; ch * kChunk < end; ++ch) nx.starts[ch] = pos;
  }
  if (ts) tail_collect(S, n, total, bh, *ts);
  __syncthreads();
  if (threadIdx.x == 0) S.cnt = 0u;
  __syncthreads();
}

// Workgroup-wide (every thread calls it): when more than `keep_below` claims are staged,
// reserves their queue slots and edge offsets with one 64-bit atomic and writes ids,
// offsets and chunk starts to frontier `nx` (and the ids to `list` at the same slots, when
// given).  Always a barrier.  The frontier comes by value: where the call is not inlined,
// a reference would make the caller keep its (runtime-selected) frontier in scratch memory.
__device__ void stage_flush(Stage& S, Frontier nx, Ctrl* c, uint32_t keep_below, int32_t* list = nullptr,
                            const TailSeed* ts = nullptr) {
  __syncthreads();
  const uint32_t n = S.cnt;
  __syncthreads();  // no wave may push again before every wave has read the count
  if (n <= keep_below || n == 0u) return;  // uniform
  const uint32_t total = stage_scan(S, n);
  if (threadIdx.x == 0) {
    unsigned long long b = atomicAdd(&c->next_packed, ((unsigned long long)n << 32) | (unsigned long long)total);
    S.base_hi = (uint32_t)(b >> 32);
    S.base_lo = (uint32_t)b;
  }
  __syncthreads();
  stage_write(S, n, total, nx, S.base_hi, S.base_lo, list, ts);
}

// ------------------------------------------------------------------------------------
// learned policy: 12 -> 32 -> 32 -> 3 ReLU MLP, evaluated by one wave

// Raw features of the level just closed (same order as the training code; the blob
// carries the standardisation).  Counts enter as logs of their fraction of the graph
// (vertices of n, edges of avg * n), so a policy trained on small graphs sees the same
// values on a large graph in the same state.
__device__ void level_features(const Ctrl* c, const LevelArgs& a, uint32_t next, uint32_t nedges, bool bu,
                               float* f, unsigned long long nvl) {
  float n = a.n_total > 0 ? (float)a.n_total : 1.0f;
  float nv = nvl > 0ull ? (float)nvl : 1.0f;
  float m = fmaxf((float)a.avg * n, 1.0f);
  float found = (float)c->found;
  f[0] = logf(fmaxf((float)next, 1.0f) / n);
  f[1] = logf(fmaxf((float)nedges, 1.0f) / m);
  f[2] = logf(fmaxf(nv - found, 1.0f) / n);
  f[3] = log1pf((float)c->level);
  f[4] = (float)next / n;
  f[5] = (float)nedges / fmaxf((float)a.avg * n, 1.0f);
  f[6] = next > 0 ? log1pf((float)nedges / (float)next) : 0.0f;
  f[7] = log1pf((float)a.avg);
  f[8] = bu ? 1.0f : 0.0f;
  f[9] = (float)c->switched_back;
  f[10] = found / nv;
  f[11] = (float)c->phase;
}

// ------------------------------------------------------------------------------------
// kernels

__global__ void k_fill_u32(uint32_t* p, uint32_t value, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void k_iota(int32_t* p, int64_t count, const uint32_t* gate) {
  if (gate && !*gate) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

// Word-wise set algebra on bitmaps: out = a & ~b (b null: out = a), optionally mirrored
// into out2.
__global__ void k_andnot(uint32_t* __restrict__ out, uint32_t* __restrict__ out2, const uint32_t* __restrict__ a,
                         const uint32_t* __restrict__ b, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t x = a[i] & (b ? ~b[i] : 0xffffffffu);
    out[i] = x;
    if (out2) out2[i] = x;
  }
}

// Sums columns 0..2 and takes the max of column 3 over nparts rows of 4 (one workgroup).
__global__ void __launch_bounds__(kBlock) k_reduce(const unsigned long long* __restrict__ part, int nparts,
                                                   unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_red[kBlock / 64];
  unsigned long long s0 = 0, s1 = 0, s2 = 0, mx = 0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    s0 += part[4 * i];
    s1 += part[4 * i + 1];
    s2 += part[4 * i + 2];
    mx = part[4 * i + 3] > mx ? part[4 * i + 3] : mx;
  }
  s0 = block_sum64(s0, s_red);
  s1 = block_sum64(s1, s_red);
  s2 = block_sum64(s2, s_red);
  mx = wave_max64(mx);
  __shared__ unsigned long long s_mx[kBlock / 64];
  if (lane() == 0) s_mx[wave_id()] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kBlock / 64; ++w) mx = s_mx[w] > mx ? s_mx[w] : mx;
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
    out[3] = mx;
  }
}

// Degree trim (multistep/scc_trim.cpp:89-139; kokkos_scc/scc_trim.cpp:57-140) and the CC
// variant (multistep/cc_trim.cpp:47-97).  Each wave round covers 64 consecutive vertices,
// whose ballot is two words of the valid bitmap; the searches' bitmaps start in the same
// words, when given: claimable cand1 / cand2 = valid, unvisited seen1 / seen2 = all ones.
// part col 0 = survivors of the workgroup.
template <bool MIRROR>
__global__ void __launch_bounds__(kBlock) k_trim(int32_t n, const uint32_t* __restrict__ oo,
                                                 const uint32_t* __restrict__ io, uint32_t* __restrict__ valid,
                                                 uint32_t* __restrict__ cand1, uint32_t* __restrict__ cand2,
                                                 uint32_t* __restrict__ seen1, uint32_t* __restrict__ seen2,
                                                 int32_t* __restrict__ labels, unsigned long long* __restrict__ part) {
  __shared__ unsigned long long s_red[kBlock / 64];
  const int64_t nw = ((int64_t)n + 31) / 32;
  unsigned long long kept = 0;
  // a wave takes kU consecutive 64-vertex slices per iteration, all offset loads first
  constexpr int kU = 4;
  const int64_t stride = (int64_t)gridDim.x * kBlock * kU;
  for (int64_t b0 = ((int64_t)blockIdx.x * kBlock + (threadIdx.x & ~63)) * kU; b0 < n; b0 += stride) {
    uint32_t od[kU], id[kU];
#pragma unroll
    for (int r = 0; r < kU; ++r) {  // past n: reload vertex n - 1 (no branch between the loads)
      const int64_t v0 = b0 + 64 * r + lane();
      const int64_t v = v0 < n ? v0 : (int64_t)n - 1;
      od[r] = oo[v + 1] - oo[v];
      id[r] = io[v + 1] - io[v];
    }
    // the kU ballots are the wave's 2 * kU consecutive bitmap words: lane j < 2 * kU takes
    // word j, so each bitmap gets one contiguous store per wave and iteration
    uint32_t mine = 0u;
#pragma unroll
    for (int r = 0; r < kU; ++r) {
      const int64_t v = b0 + 64 * r + lane();
      bool keep = false;
      if (v < n) {
        keep = MIRROR ? (od[r] + id[r]) > 0u : (od[r] > 0u && id[r] > 0u);
        if (labels) labels[v] = keep ? -1 : (int32_t)v;
      }
      const unsigned long long mask = __ballot(keep);
      if ((lane() >> 1) == r) mine = (uint32_t)(mask >> (32 * (lane() & 1)));
      kept += keep ? 1 : 0;
    }
    const int64_t wi = (b0 >> 5) + lane();
    if (lane() < 2 * kU && wi < nw) {
      valid[wi] = mine;
      if (cand1) cand1[wi] = mine;
      if (cand2) cand2[wi] = mine;
      if (seen1) seen1[wi] = 0xffffffffu;
      if (seen2) seen2[wi] = 0xffffffffu;
    }
  }
  kept = block_sum64(kept, s_red);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = kept;
    part[4 * blockIdx.x + 1] = part[4 * blockIdx.x + 2] = part[4 * blockIdx.x + 3] = 0;
  }
}

// m as normalised by member_src (both pointers set): no branch between the loads, so a
// thread's member words of several slices load together
__device__ __forceinline__ uint32_t member_word(const MemberSrc& m, uint32_t valid_word, int64_t i, uint32_t keep) {
  return valid_word & ~m.ca[i] & ~m.cb[i] & keep;
}

// All ones unless a search named in m was abandoned (then its component is not committed).
__device__ __forceinline__ uint32_t member_keep(const MemberSrc& m) {
  return ((m.g0 && m.g0->abandoned) || (m.g1 && m.g1->abandoned)) ? 0u : 0xffffffffu;
}

// part col 3 = 0xffffffff - smallest member in the workgroup's words (0 when none), so
// that k_reduce's max yields the smallest member overall.
// With `largest`, col 3 = 1 + largest member instead (0 when none).
__global__ void __launch_bounds__(kBlock) k_min_member(MemberSrc m, const uint32_t* __restrict__ valid,
                                                       int64_t nwords, unsigned long long* __restrict__ part,
                                                       const RoundCtl* rc, int32_t largest) {
  __shared__ unsigned long long s_red[kBlock / 64];
  if (rc && !rc->go) return;  // uniform: the searches have not finished
  const uint32_t keep = member_keep(m);
  unsigned long long best = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nwords; i += (int64_t)gridDim.x * kBlock) {
    uint32_t w = member_word(m, valid[i], i, keep);
    if (w) {
      unsigned long long key = largest ? 1ull + (unsigned long long)(i * 32 + (31 - __clz(w)))
                                       : 0xffffffffull - (unsigned long long)(i * 32 + (__ffs(w) - 1));
      best = key > best ? key : best;
    }
  }
  best = wave_max64(best);
  if (lane() == 0) s_red[wave_id()] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kBlock / 64; ++i) best = s_red[i] > best ? s_red[i] : best;
    unsigned long long* p = part + 4 * blockIdx.x;
    p[0] = p[1] = p[2] = 0;
    p[3] = best;
  }
}

// Canonical labels of the colouring phase, once it has ended: every vertex coloured (the
// list of colouring round 0, rc->listed0 long) gets the smallest member of its colour
// root's SCC, tracked in cmin by the per-colour backward search.
__global__ void __launch_bounds__(kBlock) k_canon_list(const int32_t* __restrict__ list, const RoundCtl* rc,
                                                       int32_t* __restrict__ labels, const int32_t* __restrict__ cmin) {
  if (!rc->finished) return;  // uniform: the colouring has not ended yet
  const uint32_t count = rc->listed0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
    const int32_t v = list[i];
    labels[v] = cmin[labels[v]];
  }
}

// CC's canonical labels of the colouring phase, once it has ended (its components carry
// their largest member as label, every listed vertex one of them): pass 1 starts each
// component's minimum at its root, pass 2 lowers it to the smallest listed member (lanes
// of a wave sharing a label combine into one atomic); k_canon_list then applies it.
__global__ void __launch_bounds__(kBlock) k_cc_cmin_init(const int32_t* __restrict__ list, const RoundCtl* rc,
                                                         const int32_t* __restrict__ labels, int32_t* __restrict__ cmin) {
  if (!rc->finished) return;  // uniform
  const uint32_t count = rc->listed0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
    const int32_t v = list[i];
    if (labels[v] == v) cmin[v] = v;
  }
}
__global__ void __launch_bounds__(kBlock) k_cc_cmin(const int32_t* __restrict__ list, const RoundCtl* rc,
                                                    const int32_t* __restrict__ labels, int32_t* __restrict__ cmin) {
  if (!rc->finished) return;  // uniform
  const uint32_t count = rc->listed0;
  for (uint32_t i0 = blockIdx.x * kBlock; i0 < count; i0 += gridDim.x * kBlock) {
    const uint32_t i = i0 + threadIdx.x;
    int32_t v = 0, l = -1;
    if (i < count) {
      v = list[i];
      l = labels[v];
    }
    bool active = l > v;
    for (;;) {
      const unsigned long long mk = __ballot(active);
      if (mk == 0ull) break;
      const int leader = __ffsll((long long)mk) - 1;
      const int32_t key = __shfl(l, leader, 64);
      const bool mine = active && l == key;
      const uint32_t x = wave_min(mine ? (uint32_t)v : 0xffffffffu);
      if (lane() == leader) atomicMin(&cmin[key], (int32_t)x);
      active = active && !mine;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_count_bits(const uint32_t* __restrict__ bits, int64_t nwords,
                                                       unsigned long long* __restrict__ part) {
  __shared__ unsigned long long s_red[kBlock / 64];
  unsigned long long cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nwords; i += (int64_t)gridDim.x * kBlock)
    cnt += (uint32_t)__popc(bits[i]);
  cnt = block_sum64(cnt, s_red);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = cnt;
    part[4 * blockIdx.x + 1] = part[4 * blockIdx.x + 2] = part[4 * blockIdx.x + 3] = 0;
  }
}

// Starts searches p.s[0 .. ns) from `root` (wave s starts search s, phase phase0 + s):
// fresh control block, root claimed and queued alone.  Follows scc_fwbw.cpp:72-73
// (queue = {root}, fw[root] = true, set even if root is invalid).  With rnp > 0 it first
// reduces rnp partial rows into rout as k_reduce does (the trim's survivor count).
constexpr int kInitBlock = 512;  // k_search_init's workgroup: waves 0 and 1 start the searches

__global__ void __launch_bounds__(kInitBlock) k_search_init(ReachArgs p, int32_t ns, int32_t root, int32_t phase0,
                              const unsigned long long* __restrict__ rpart, int32_t rnp,
                              unsigned long long* __restrict__ rout) {
  constexpr int kWaves = kInitBlock / 64;
  __shared__ unsigned long long s_r[4][kWaves];
  if (rnp > 0) {  // uniform
    // a few rows per thread, their loads in flight together (a 2560-row trim took 10 us
    // with 20 dependent rows per thread)
    unsigned long long x[4] = {0ull, 0ull, 0ull, 0ull};
    for (int i0 = threadIdx.x; i0 < rnp; i0 += 4 * kInitBlock) {
      unsigned long long r[4][4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = i0 + j * kInitBlock;
#pragma unroll
        for (int k = 0; k < 4; ++k) r[j][k] = i < rnp ? rpart[4 * i + k] : 0ull;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        for (int k = 0; k < 3; ++k) x[k] += r[j][k];
        x[3] = r[j][3] > x[3] ? r[j][3] : x[3];
      }
    }
    for (int k = 0; k < 3; ++k) x[k] = wave_sum64(x[k]);
    x[3] = wave_max64(x[3]);
    if (lane() == 0)
      for (int k = 0; k < 4; ++k) s_r[k][wave_id()] = x[k];
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long y[4] = {0ull, 0ull, 0ull, 0ull};
      for (int w = 0; w < kWaves; ++w) {
        for (int k = 0; k < 3; ++k) y[k] += s_r[k][w];
        y[3] = s_r[3][w] > y[3] ? s_r[3][w] : y[3];
      }
      for (int k = 0; k < 4; ++k) rout[k] = y[k];
    }
  }
  const int s = wave_id();
  if (s >= ns || s >= 2) return;
  const LevelArgs& a = p.s[s];
  Ctrl* c = a.ctrl;
  uint32_t deg = list_len(a.down_off, a.down_off2, (uint32_t)root);
  if (lane() == 0) {
    c->next_packed = 0ull;
    c->q_size = 1u;
    c->q_edges = deg;
    c->cur = 0u;
    c->bottom_up = 0u;
    c->bu_wide = 0u;
    c->from_bits = 0u;
    c->switched_back = 0u;
    c->done = 0u;
    c->level = 0u;
    c->closes = 0u;
    c->found = 0;
    c->phase = (uint32_t)(phase0 + s);
    c->abandoned = 0u;
    c->within = 0u;
    c->t_start = ~0ull;
    uint32_t m = 1u << ((uint32_t)root & 31u);
    a.cand[(uint32_t)root >> 5] &= ~m;
    a.seen[(uint32_t)root >> 5] &= ~m;
    a.fr0.q[0] = root;
    a.fr0.qo[0] = 0u;
  }
  uint32_t nch = (deg + kChunk - 1) / kChunk;
  for (uint32_t ch = lane(); ch < nch; ch += 64) a.fr0.starts[ch] = 0u;
}

// Thread 0: a fresh control block for a seeding kernel (list_len is kept: the previous
// seed's list length).
__device__ __forceinline__ void ctrl_fresh(Ctrl* c, int32_t phase) {
  c->next_packed = 0ull;
  c->q_size = 0u;
  c->q_edges = 0u;
  c->cur = 0u;
  c->bottom_up = 0u;
  c->bu_wide = 0u;
  c->from_bits = 0u;
  c->switched_back = 0u;
  c->done = 0u;
  c->level = 0u;
  c->closes = 0u;
  c->found = 0;
  c->phase = (uint32_t)phase;
  c->abandoned = 0u;
  c->within = 0u;
  c->t_start = ~0ull;
}

// First kernel of a colouring round (one thread).  Round 0 (flags bit 0) follows the
// searches: it is admitted once their control blocks s0 / s1 (null: none) say done.  With
// flags bit 1 (CC's single fixpoint) it ends the colouring once the push has settled.  A
// later round is admitted once the previous round's push and colour-BW have settled and
// its roots ran; it books that round (multistep/scc_color.cpp:340-373 loops while vertices
// are left) and ends the colouring when none is.  An admitted round gets a fresh push
// control block; a round that cannot be admitted sets rc->stall, which every later step
// honours until the host clears it.  With `prog` (host memory) the outcome is published
// there for the host that enqueues rounds ahead: epoch << 48 | (stall | finished << 1 |
// error << 2) << 32 | rounds admitted.
__device__ void round_decide(RoundCtl* rc, Ctrl* push, const Ctrl* cbw, const Ctrl* s0, const Ctrl* s1,
                             int32_t flags) {
  rc->go = 0u;
  if (rc->stall || rc->finished || rc->error) return;
  if (flags & 1) {
    if ((s0 && !s0->done) || (s1 && !s1->done)) {
      rc->stall = 1u;
      return;
    }
  } else if (flags & 2) {
    // a single max-label fixpoint (CC, multistep/cc_color.cpp:47-175): it ends when it settles
    if (!push->done) {
      rc->stall = 1u;
      return;
    }
    rc->rounds = 1;
    rc->push_levels = push->level;
    rc->listed0 = push->list_len;
    rc->finished = 1u;
    return;
  } else {
    if (!push->done || !cbw->done || rc->roots_in != rc->admitted) {
      rc->stall = 1u;
      return;
    }
    const unsigned long long listed = push->list_len, roots = cbw->list_len;
    const unsigned long long claimed = (unsigned long long)cbw->found;
    rc->rounds += 1;
    rc->push_levels += push->level;
    rc->cbw_levels += cbw->level;
    rc->listed_total += listed;
    rc->roots_total += roots;
    rc->push_closes = push->closes;
    rc->cbw_closes = cbw->closes;
    if (rc->admitted == 1u) rc->listed0 = (uint32_t)listed;
    // a settled push fixpoint over a non-empty list always has a root (its largest vertex)
    if ((roots == 0ull && listed > 0ull) || roots + claimed > listed) {
      rc->error = 1u;
      return;
    }
    rc->left = listed - roots - claimed;
    if (rc->left == 0ull) {
      rc->finished = 1u;
      return;
    }
  }
  ctrl_fresh(push, 3);
  rc->admitted += 1u;
  rc->go = 1u;
}

__global__ void k_round_begin(RoundCtl* rc, Ctrl* push, Ctrl* cbw, const Ctrl* s0, const Ctrl* s1,
                              int32_t flags, unsigned long long* prog, uint32_t epoch) {
  if (threadIdx.x != 0) return;
  if (flags & 4) {  // a fresh colouring first (k_rc_set mode 0)
    *rc = RoundCtl{};
    push->done = 1u;
    cbw->done = 1u;
  }
  round_decide(rc, push, cbw, s0, s1, flags & 3);
  if (prog) {
    const unsigned long long st = (rc->stall ? 1ull : 0ull) | (rc->finished ? 2ull : 0ull) | (rc->error ? 4ull : 0ull);
    __hip_atomic_store(prog, ((unsigned long long)(epoch & 0xffffu) << 48) | (st << 32) | rc->admitted,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Before the colour roots of the admitted round (one thread): the push fixpoint must have
// settled, else the round stalls; then a fresh colour-BW control block.
__global__ void k_roots_begin(RoundCtl* rc, const Ctrl* push, Ctrl* cbw) {
  if (threadIdx.x != 0 || !rc->go) return;
  if (!push->done) {
    rc->go = 0u;
    rc->stall = 1u;
    return;
  }
  ctrl_fresh(cbw, 4);
  rc->roots_in = rc->admitted;
}

// Host-directed round state changes: 0 = a fresh colouring (push / colour-BW blocks marked
// done so that levels enqueued before anything is seeded do nothing), 1 = resume the
// admitted round (its push was finished by the host), 2 = clear a stall whose round the
// host finished.
__global__ void k_rc_set(RoundCtl* rc, Ctrl* push, Ctrl* cbw, int32_t mode) {
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    *rc = RoundCtl{};
    push->done = 1u;
    cbw->done = 1u;
  } else if (mode == 1) {
    rc->stall = 0u;
    rc->go = 1u;
  } else {
    rc->stall = 0u;
    rc->go = 0u;
  }
}

// Resets the control block before a seeding kernel fills frontier buffer 1.
__global__ void k_ctrl_reset(Ctrl* c, int32_t phase) {
  if (threadIdx.x == 0) {
    c->next_packed = 0ull;
    c->q_size = 0u;
    c->q_edges = 0u;
    c->cur = 0u;
    c->bottom_up = 0u;
    c->bu_wide = 0u;
    c->from_bits = 0u;
    c->switched_back = 0u;
    c->done = 0u;
    c->level = 0u;
    c->closes = 0u;
    c->found = 0;
    c->phase = (uint32_t)phase;
    c->abandoned = 0u;
    c->within = 0u;
    c->t_start = ~0ull;
  }
}

// Closes a seeding launch: the queue built in buffer 1 becomes the frontier (and the
// length of the vertex list the seed wrote beside it, if any).  With rc, only in an
// admitted round (rc->go).  With pull_min (colour push seeds), a first level over at
// least pull_min edges is a pull level over the seed's ids.
__global__ void k_close_seed(Ctrl* c, const RoundCtl* rc, uint32_t pull_min) {
  if (rc && !rc->go) return;
  if (threadIdx.x == 0) {
    unsigned long long p = c->next_packed;
    uint32_t next = (uint32_t)(p >> 32);
    c->list_len = next;
    c->q_size = next;
    c->q_edges = (uint32_t)p;
    c->round_edges = (uint32_t)p;
    if (pull_min) c->bottom_up = (uint32_t)p >= pull_min ? 2u : 0u;
    c->cur = 1u;
    c->next_packed = 0ull;
    c->done = next == 0u ? 1u : 0u;
  }
}

// LDS of one top-down chunk.
// The probes, claims and staging of kPer top-down edges per thread (targets w with the
// source's colour col; ok: the lane has that edge): one state probe per edge with all
// loads in flight together, then per edge a claim (reach: cand and seen bits; colour-BW:
// the valid bit, the label and the colour's minimum member) or a raise (push: atomicMax
// of the colour, then the next worklist bit), and every winner staged with its down-list
// length.  Every thread calls (stage_push is wave-cooperative); the caller flushes.
template <int OP, bool MIRROR>
__device__ __forceinline__ void td_claim(const LevelArgs& a, uint32_t* inq_nxt, const uint32_t (&w)[kPer],
                                         const int32_t (&col)[kPer], const bool (&ok)[kPer], Stage& S,
                                         unsigned long long& my_raised) {
  constexpr bool REACH = OP == OP_FW || OP == OP_BW || OP == OP_CC;
  constexpr bool COLOR = OP == OP_PUSH || OP == OP_CBW;
  uint32_t word[kPer], excl[kPer];
  int32_t wcol[kPer];
  // a backward search restricted to the finished forward set: its claimable bitmap still
  // covers the whole valid set, so a probe also reads the forward search's
  const uint32_t* within = (REACH && a.within_cand && a.ctrl->within) ? a.within_cand : nullptr;
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    word[r] = 0;
    excl[r] = 0;
    wcol[r] = 0;
    if (ok[r]) {
      word[r] = (REACH ? a.cand : a.valid)[w[r] >> 5];
      if (within) excl[r] = within[w[r] >> 5];
      if (COLOR) wcol[r] = a.colors[w[r]];
    }
  }
#pragma unroll
  for (int r = 0; r < kPer; ++r) word[r] &= ~excl[r];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    bool in = ok[r] && ((word[r] >> (w[r] & 31u)) & 1u);
    bool got = false;  // claimed (reach, colour-BW) / raised (push) by this lane
    if (REACH) {
      got = in && bit_clear(a.cand, w[r]);
      // and visited for bottom-up probes (no result needed: only this claim clears it)
      if (got) __hip_atomic_fetch_and(&a.seen[w[r] >> 5], ~(1u << (w[r] & 31u)), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    } else if (OP == OP_PUSH) {
      if (in && wcol[r] < col[r] && atomicMax(&a.colors[w[r]], col[r]) < col[r]) {
        ++my_raised;
        got = true;
      }
    } else {  // OP_CBW
      if (in && wcol[r] == col[r] && bit_clear(a.valid, w[r])) {
        a.labels[w[r]] = col[r];
        got = true;
        if (a.cmin && (int32_t)w[r] < a.cmin[col[r]]) atomicMin(&a.cmin[col[r]], (int32_t)w[r]);
      }
    }
    bool take = got;
    if (OP == OP_PUSH && got) take = bit_set(inq_nxt, w[r]);
    uint32_t deg = take ? list_len(a.down_off, MIRROR ? a.down_off2 : nullptr, w[r]) : 0u;
    stage_push(S, take, w[r], deg);
  }
}

static_assert(kChunk < 65536u, "a chunk's owner table holds 16-bit slot indices");
template <int OP, bool MIRROR>
struct ChunkLds {
  static constexpr bool kColor = OP == OP_PUSH || OP == OP_CBW;
  union {
    uint16_t own[kChunk];  // td_chunk: the slot that owns each edge of the chunk
    uint32_t pre[kBlock];  // td_slots: exclusive prefix of the slots' list lengths
  };
  uint32_t base[kChunk + 1];
  int32_t col[kColor ? kChunk + 1 : 1];
  uint32_t split[MIRROR ? kChunk + 1 : 1];
  uint32_t base2[MIRROR ? kChunk + 1 : 1];
  uint32_t scan[kBlock / 64];
};

// Expands chunk `ch` (kChunk edges of the frontier's edge space) with the whole
// workgroup: the slots overlapping the chunk scatter their start positions into LDS, a
// block max-scan turns them into a per-edge owner table (the "Manhattan collapse" of
// kokkos_scc/fwbw_manhattan_global.cpp:162-289 without its per-edge binary search), the
// adjacency is read coalesced, and claims / raises are staged for frontier N.  Ends with
// a workgroup barrier.  Every queued vertex has a non-empty list, so at most len + 1
// slots overlap a chunk.  A push expansion takes the vertex off the current worklist
// bitmap: its bit is cleared atomically, or with PLAIN_CLEAR its whole word is stored as
// zero (only where nothing reads or sets that bitmap until the launch ends).
template <int OP, bool MIRROR, bool LOCAL = false, bool PLAIN_CLEAR = false>
__device__ __forceinline__ void td_chunk(const LevelArgs& a, const Frontier& F, const Frontier& N, uint32_t* inq_cur,
                                         uint32_t* inq_nxt, uint32_t ch, uint32_t nchunks, uint32_t q_size,
                                         uint32_t q_edges, Stage& S, ChunkLds<OP, MIRROR>& L,
                                         unsigned long long& my_raised) {
  constexpr bool COLOR = OP == OP_PUSH || OP == OP_CBW;
  const int t = threadIdx.x;
  const uint32_t e0 = ch * kChunk;
  const uint32_t len = min(kChunk, q_edges - e0);
  const uint32_t k0 = F.starts[ch];
  const uint32_t kend = (ch + 1 < nchunks) ? F.starts[ch + 1] : q_size - 1u;
  // the clamp only keeps LDS writes in bounds
  const uint32_t nk = min(kend - k0 + 1u, kChunk + 1u);
  for (uint32_t i = t; i < len; i += kBlock) L.own[i] = 0u;
  __syncthreads();
  for (uint32_t j = t; j < nk; j += kBlock) {
    uint32_t k = k0 + j;
    uint32_t qo = F.qo[k];
    uint32_t v = (uint32_t)F.q[k];
    uint32_t s = qo > e0 ? qo - e0 : 0u;
    if (s < len) {
      // slot j owns the edges from its start on; of slots starting at one offset (all but
      // the last would have empty lists) the last one writes.  Slot 0 holds edge e0, so no
      // later slot starts at offset 0 and its entry is the cleared one.
      const uint32_t qn = k + 1u < q_size ? F.qo[k + 1u] : q_edges;
      if (j > 0u && qn > qo) L.own[s] = (uint16_t)j;
      if (OP == OP_PUSH) {
        if (PLAIN_CLEAR) inq_cur[v >> 5] = 0u;
        else bit_clear(inq_cur, v);
      }
    }
    uint32_t b1 = a.down_off[v];
    L.base[j] = b1 - qo;
    if (MIRROR) {
      uint32_t split = qo + (a.down_off[v + 1] - b1);
      L.split[j] = split;
      L.base2[j] = a.down_off2[v] - split;
    }
    if (COLOR) {
      if (LOCAL && OP == OP_PUSH) {
        // a workgroup-local expansion: the colour after the worklist bit was cleared, read
        // where atomics are performed (a raise that found the bit set relies on this read)
        __builtin_amdgcn_s_waitcnt(0);
        L.col[j] = __hip_atomic_load(&a.colors[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        L.col[j] = a.colors[v];
      }
    }
  }
  __syncthreads();
  // inclusive max-scan of own[0..len): thread t owns entries [t*kPer, t*kPer+kPer)
  {
    uint32_t vals[kPer];
    uint32_t run = 0;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      uint32_t idx = (uint32_t)t * kPer + r;
      uint32_t x = idx < len ? L.own[idx] : 0u;
      run = max(run, x);
      vals[r] = run;
    }
    uint32_t incl = wave_incl_max(run);
    uint32_t excl = __shfl_up(incl, 1, 64);
    if (lane() == 0) excl = 0u;
    if (lane() == 63) L.scan[wave_id()] = incl;
    __syncthreads();
    uint32_t carry = 0;
    for (int w = 0; w < wave_id(); ++w) carry = max(carry, L.scan[w]);
    excl = max(excl, carry);
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      uint32_t idx = (uint32_t)t * kPer + r;
      if (idx < len) L.own[idx] = (uint16_t)max(vals[r], excl);
    }
  }
  __syncthreads();
  // batched per thread so the loads of all kPer edges are in flight together:
  // adjacency (coalesced), then one state probe per edge, then the claims
  uint32_t w[kPer];
  int32_t col[kPer];
  bool ok[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    uint32_t i = (uint32_t)t + (uint32_t)r * kBlock;
    ok[r] = i < len;
    w[r] = 0;
    col[r] = 0;
    if (ok[r]) {
      uint32_t j = L.own[i];
      uint32_t e = e0 + i;
      if (MIRROR)
        w[r] = (uint32_t)(e < L.split[j] ? a.down_adj[L.base[j] + e] : a.down_adj2[L.base2[j] + e]);
      else
        w[r] = (uint32_t)a.down_adj[L.base[j] + e];
      if (COLOR) col[r] = L.col[j];
    }
  }
  td_claim<OP, MIRROR>(a, inq_nxt, w, col, ok, S, my_raised);
  stage_flush(S, N, a.ctrl, kStage - kChunk);  // also the end-of-chunk barrier
}

// Top-down load balance of a grid-wide level (msc_tuning.balance): how the frontier's edges
// are dealt to threads, as the Kokkos variants selected by kokkos_scc/scc_fwbw.cpp:337-364
// deal them.  Every variant probes and claims the same edges, so labels never depend on it.
enum Balance : int32_t {
  kBalanceVertex = 0,  // a thread per frontier vertex walks its whole list (fwbw_baseline.cpp)
  kBalanceLocal = 1,   // a workgroup deals the edges of its 256 frontier vertices evenly
                       // (fwbw_manhattan_local.cpp)
  kBalanceGlobal = 2,  // kChunk-edge chunks of the whole frontier's edge space, td_chunk
                       // (fwbw_manhattan_global.cpp)
  kBalanceAuto = 3,    // per level: kBalanceLocal while the frontier's lists average at most
                       // MSC_AUTO_DEG entries, else kBalanceGlobal
};
#ifndef MSC_AUTO_DEG
#define MSC_AUTO_DEG 4
#endif

// The variant a level of frontier (q_size vertices, q_edges list entries) runs.  With short
// lists a chunk's owner table (an LDS scatter and max-scan over kChunk entries) costs more
// than the per-slot prefix of a workgroup's own 256 vertices (grid 4096^2 SCC 14.4 ->
// 13.6 ms).  Mirrored levels (WCC: in- and out-list per vertex) keep the chunks: there the
// slot form was slower (grid WCC 5.9 -> 6.2 ms, batch 21.5 -> 22.1 ms).
__device__ __forceinline__ int32_t td_variant(const LevelArgs& a, uint32_t q_size, uint32_t q_edges) {
  if (a.balance != kBalanceAuto) return a.balance;
  if (a.down_off2) return kBalanceGlobal;
  return (unsigned long long)q_edges <= (unsigned long long)MSC_AUTO_DEG * q_size ? kBalanceLocal : kBalanceGlobal;
}

// Work units of a queue-based top-down level: chunks of the frontier's edges, or blocks of
// kBlock frontier vertices.
__device__ __forceinline__ uint32_t td_units(const LevelArgs& a, uint32_t q_size, uint32_t q_edges) {
  return td_variant(a, q_size, q_edges) == kBalanceGlobal ? (q_edges + kChunk - 1) / kChunk
                                                          : (q_size + kBlock - 1) / kBlock;
}

// Expands frontier slots [blk * kBlock, blk * kBlock + kBlock) with the whole workgroup, as
// the balance variants VAR (kBalanceVertex or kBalanceLocal) deal them; the probes, claims
// and staging are td_chunk's (td_claim), and a push expansion clears each expanded vertex's
// worklist word with a plain store, as a grid-wide td_chunk does.  kBalanceVertex: thread t
// walks slot t's lists kPer entries at a time, the workgroup for as many steps as its
// longest list needs.  kBalanceLocal: the slots' list lengths are prefix-summed in LDS and
// the block's edges taken kChunk at a time, an edge's slot found by binary search over the
// prefix (the reference's highest_less_than over its team's trip counts).  Ends with a
// workgroup barrier.
template <int OP, bool MIRROR, int VAR>
__device__ __forceinline__ void td_slots(const LevelArgs& a, const Frontier& F, const Frontier& N, uint32_t* inq_cur,
                                         uint32_t* inq_nxt, uint32_t blk, uint32_t q_size, Stage& S,
                                         ChunkLds<OP, MIRROR>& L, unsigned long long& my_raised) {
  constexpr bool COLOR = OP == OP_PUSH || OP == OP_CBW;
  const int t = threadIdx.x;
  const uint32_t k0 = blk * kBlock;
  const uint32_t nb = min(q_size - k0, (uint32_t)kBlock);
  const bool has = (uint32_t)t < nb;
  uint32_t b1 = 0, l1 = 0, b2 = 0, l2 = 0;
  int32_t own = 0;
  if (has) {
    const uint32_t v = (uint32_t)F.q[k0 + (uint32_t)t];
    if (OP == OP_PUSH) inq_cur[v >> 5] = 0u;
    b1 = a.down_off[v];
    l1 = a.down_off[v + 1] - b1;
    if (MIRROR) {
      b2 = a.down_off2[v];
      l2 = a.down_off2[v + 1] - b2;
    }
    if (COLOR) own = a.colors[v];
  }
  const uint32_t len = l1 + l2;
  if (VAR == kBalanceVertex) {
    const uint32_t wmax = wave_incl_max(len);
    if (lane() == 63) L.scan[wave_id()] = wmax;
    __syncthreads();
    uint32_t steps = 0;
    for (int w = 0; w < kBlock / 64; ++w) steps = max(steps, L.scan[w]);
    for (uint32_t i0 = 0; i0 < steps; i0 += kPer) {
      uint32_t w[kPer];
      int32_t col[kPer];
      bool ok[kPer];
#pragma unroll
      for (int r = 0; r < kPer; ++r) {
        const uint32_t i = i0 + (uint32_t)r;
        ok[r] = i < len;
        w[r] = ok[r] ? (uint32_t)((MIRROR && i >= l1) ? a.down_adj2[b2 + (i - l1)] : a.down_adj[b1 + i]) : 0u;
        col[r] = own;
      }
      td_claim<OP, MIRROR>(a, inq_nxt, w, col, ok, S, my_raised);
      stage_flush(S, N, a.ctrl, kStage - kChunk);  // also a barrier
    }
  } else {
    const uint32_t incl = wave_incl_sum(len);
    if (lane() == 63) L.scan[wave_id()] = incl;
    L.base[t] = b1;
    if (MIRROR) {
      L.split[t] = l1;
      L.base2[t] = b2;
    }
    if (COLOR) L.col[t] = own;
    __syncthreads();
    uint32_t carry = 0, T = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      if (w < wave_id()) carry += L.scan[w];
      T += L.scan[w];
    }
    L.pre[t] = incl - len + carry;  // exclusive prefix; slots past nb have len 0
    __syncthreads();
    for (uint32_t e0 = 0; e0 < T; e0 += kChunk) {
      uint32_t w[kPer];
      int32_t col[kPer];
      bool ok[kPer];
#pragma unroll
      for (int r = 0; r < kPer; ++r) {
        const uint32_t e = e0 + (uint32_t)t + (uint32_t)r * kBlock;
        ok[r] = e < T;
        w[r] = 0u;
        col[r] = 0;
        if (ok[r]) {
          // the last slot whose list starts at or before e (pre[0] = 0 <= e < T)
          uint32_t lo = 0;
#pragma unroll
          for (uint32_t step = kBlock / 2; step >= 1; step >>= 1)
            if (lo + step < nb && L.pre[lo + step] <= e) lo += step;
          const uint32_t i = e - L.pre[lo];
          w[r] = (uint32_t)((MIRROR && i >= L.split[lo]) ? a.down_adj2[L.base2[lo] + (i - L.split[lo])]
                                                         : a.down_adj[L.base[lo] + i]);
          if (COLOR) col[r] = L.col[lo];
        }
      }
      td_claim<OP, MIRROR>(a, inq_nxt, w, col, ok, S, my_raised);
      stage_flush(S, N, a.ctrl, kStage - kChunk);  // also a barrier
    }
  }
  __syncthreads();  // the slots' LDS is reused by the next block
}

// The rest [j, end) of a bottom-up candidate's up-list after its first probe missed:
// MSC_BU_SCAN entries at a time, their adjacency loads and then their seen probes in flight
// together, instead of one dependent load pair per entry (a wave's round lasts as long as
// its longest scan).  `inspected` counts the entries up to the first visited one, as the
// reference's scan with its break does (multistep/scc_fwbw.cpp:124-133).
__device__ __forceinline__ bool bu_scan(const int32_t* __restrict__ adj, const uint32_t* seen, uint32_t j,
                                        uint32_t end, unsigned long long& inspected) {
  constexpr uint32_t K = MSC_BU_SCAN;
  for (; j < end; j += K) {
    uint32_t x[K], w[K];
#pragma unroll
    for (uint32_t i = 0; i < K; ++i) x[i] = j + i < end ? (uint32_t)adj[j + i] : 0u;
#pragma unroll
    for (uint32_t i = 0; i < K; ++i) w[i] = j + i < end ? seen[x[i] >> 5] : 0xffffffffu;
#pragma unroll
    for (uint32_t i = 0; i < K; ++i) {
      if (j + i < end) {
        ++inspected;
        if (!((w[i] >> (x[i] & 31u)) & 1u)) return true;
      }
    }
  }
  return false;
}

// One bottom-up word group of a reach search (multistep/scc_fwbw.cpp:118-148): every set bit
// of cand words [wb, wb + kBW) is a candidate, claimed when some vertex of its up-lists is
// visited (allowed and no longer claimable, the reference's `fw[in] && valid[in]`).  The
// workgroup takes the candidates in rounds of kR per thread; each candidate's first up-list
// entry is checked with all loads of the thread's candidates in flight together, and most
// candidates are settled there (on R-MAT the first in-neighbour is already visited ~90% of
// the time); the rest scan on, MSC_BU_SCAN entries at a time.  Claims collect in LDS word
// masks and are written back with one plain store per word: no other workgroup clears bits
// of these words during a bottom-up level.  Claims are not queued (queueing them cost R-MAT
// s24 1.5%, also when only levels with few claims left did): a following bottom-up level
// needs no frontier, and a following top-down level expands the claims from the group's
// words of a.fresh, which every group writes in full (zero words included).  The thread's claims and their
// down-list lengths are counted into my_next / my_nedges.  Every thread calls.
template <bool MIRROR, int WPT>
__device__ __forceinline__ void bu_group(const LevelArgs& a, int64_t wb, BitIterT<WPT>& B, uint32_t* s_claim,
                                         unsigned long long& my_examined, unsigned long long& my_inspected,
                                         unsigned long long& my_next, unsigned long long& my_nedges) {
  constexpr int kR = MSC_BU_KR;
  const int t = threadIdx.x;
  // the group: kBlock * WPT words, or group_words(a) of them for a narrow (WPT == 1) group
  const int64_t gw = WPT == 1 ? group_words(a) : (int64_t)kBlock * WPT;
  static_assert(WPT != kWPT || kWPT > 1 || kBW == kBlock, "a one-word-per-thread group is at most kBlock words");
  const int64_t wend = wb + gw < a.nwords ? wb + gw : a.nwords;
#pragma unroll
  for (int j = 0; j < WPT; ++j) s_claim[t * WPT + j] = 0u;
  // candidates: the claimable bits (restricted to the forward set for a backward search
  // once ctrl->within is set); this thread's claimable and seen words are kept for the
  // write-back
  uint32_t sw[WPT], cw[WPT], w[WPT];
  const uint32_t* within = (a.within_cand && a.ctrl->within) ? a.within_cand : nullptr;
  {
    const int64_t w0 = wb + (int64_t)t * WPT;
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      const bool in = w0 + j < wend;
      cw[j] = in ? a.cand[w0 + j] : 0u;
      sw[j] = in ? a.seen[w0 + j] : 0u;
      w[j] = cw[j] & ~((in && within) ? within[w0 + j] : 0u);
    }
  }
  const uint32_t total = bits_scan(B, w);  // barrier: s_claim cleared
  for (uint32_t k0 = 0; k0 < total; k0 += kBlock * kR) {
    uint32_t v[kR], beg[kR], end[kR], beg2[kR], end2[kR], u[kR], wc[kR];
    bool has[kR], hit[kR], first[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      uint32_t k = k0 + (uint32_t)(r * kBlock + t);
      has[r] = k < total;
      v[r] = has[r] ? bits_pick(B, wb, k) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      beg[r] = end[r] = beg2[r] = end2[r] = 0u;
      if (has[r]) {
        beg[r] = a.up_off[v[r]];
        end[r] = a.up_off[v[r] + 1];
        if (MIRROR) {
          beg2[r] = a.up_off2[v[r]];
          end2[r] = a.up_off2[v[r] + 1];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      u[r] = 0u;
      first[r] = false;
      if (has[r]) {
        if (beg[r] < end[r]) {
          u[r] = (uint32_t)a.up_adj[beg[r]++];
          first[r] = true;
        } else if (MIRROR && beg2[r] < end2[r]) {
          u[r] = (uint32_t)a.up_adj2[beg2[r]++];
          first[r] = true;
        }
      }
    }
    // a parent is a visited vertex (the reference's `fw[in] && valid[in]`): its seen bit
    // is clear, which only a claim (of an allowed vertex, or the root) does
#pragma unroll
    for (int r = 0; r < kR; ++r) wc[r] = first[r] ? a.seen[u[r] >> 5] : 0xffffffffu;
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      hit[r] = false;
      if (!has[r]) continue;
      ++my_examined;
      if (first[r]) {
        ++my_inspected;
        hit[r] = !(wc[r] & (1u << (u[r] & 31u)));
      }
      if (!hit[r]) hit[r] = bu_scan(a.up_adj, a.seen, beg[r], end[r], my_inspected);
      if (MIRROR && !hit[r]) hit[r] = bu_scan(a.up_adj2, a.seen, beg2[r], end2[r], my_inspected);
      if (hit[r]) atomicOr(&s_claim[(uint32_t)(v[r] - (uint32_t)(wb * 32)) >> 5], 1u << (v[r] & 31u));
    }
    // the claims, counted per wave from a ballot (a point where the wave's lanes reconverge
    // after their divergent list walks: without one, -O3 builds of the mirrored search
    // were seen to lose claims), and their down-list lengths: the next frontier's edges
    // (the policy's features, and the size of a top-down level that expands these claims)
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const unsigned long long mask = __ballot(hit[r]);
      if (lane() == 0) my_next += (unsigned long long)__popcll(mask);
      if (hit[r]) my_nedges += list_len(a.down_off, MIRROR ? a.down_off2 : nullptr, v[r]);
    }
    __syncthreads();  // every claim of the round is in s_claim before the next round's picks
  }
  // every claim of this group is in s_claim; each thread reads only its own words and
  // masks, so the next group's loads need no barrier first.  Words past a narrow group are
  // the next group's.
#pragma unroll
  for (int j = 0; j < WPT; ++j) {
    const int i = t * WPT + j;
    if ((int64_t)i >= gw) continue;
    const uint32_t c = s_claim[i];
    a.fresh[wb + i] = c;
    if (c) {
      a.cand[wb + i] = cw[j] & ~c;
      a.seen[wb + i] = sw[j] & ~c;
    }
  }
}

// LDS of the fused bottom-up group of two searches (bu_pair): the union of their candidate
// words (walked as one), each search's candidate words and its claim masks.
struct BuPairLds {
  BitIter B;
  uint32_t cf[kBW];
  uint32_t cb[kBW];
  uint32_t claim_f[kBW];
  uint32_t claim_b[kBW];
};

// Whether the two searches of a k_reach<2> launch run this level as one fused bottom-up
// sweep: both bottom-up over kBW-word groups of one bitmap, the forward search's up-lists
// being the backward search's down-lists and vice versa (a pivot's FW and BW search).
__device__ __forceinline__ bool bu_pairable(const LevelArgs& f, const LevelArgs& b) {
  const Ctrl* cf = f.ctrl;
  const Ctrl* cb = b.ctrl;
  return !cf->done && !cb->done && cf->bottom_up && cb->bottom_up && !cf->bu_wide && !cb->bu_wide &&
         f.nwords == b.nwords && f.up_off == b.down_off && f.down_off == b.up_off && f.up_off2 == nullptr &&
         bu_words(f) == (int64_t)kBW && bu_words(b) == (int64_t)kBW;
}

// One bottom-up word group [wb, wb + kBW) of a forward search f and a backward search b
// together (each as bu_group does it alone, multistep/scc_fwbw.cpp:118-148 and 301-331).
// The candidates of either search are walked as one set, so a vertex's in- and out-list
// bounds are loaded once for both: f probes its in-list head (f.up), b its out-list head
// (b.up = f.down), and a claim's down-list length, counted for the next frontier's edges,
// is the other pair of bounds.  On R-MAT nearly every unvisited vertex is a candidate of
// both searches at the giant level, and the two offset arrays are each read once instead
// of twice.  Counts, claims and write-backs per search are those of two bu_group calls.
// Every thread calls.
__device__ __forceinline__ void bu_pair(const LevelArgs& f, const LevelArgs& b, int64_t wb, BuPairLds& L,
                                        unsigned long long* cf4, unsigned long long* cb4) {
  constexpr int kR = MSC_BU_KR;
  const int t = threadIdx.x;
  uint32_t wf[kWPT], wbk[kWPT], sf[kWPT], sb[kWPT], un[kWPT];
  const int64_t w0 = wb + (int64_t)t * kWPT;
#pragma unroll
  for (int j = 0; j < kWPT; ++j) {
    const bool in = w0 + j < f.nwords;
    wf[j] = in ? f.cand[w0 + j] : 0u;
    wbk[j] = in ? b.cand[w0 + j] : 0u;
    sf[j] = in ? f.seen[w0 + j] : 0u;
    sb[j] = in ? b.seen[w0 + j] : 0u;
    un[j] = wf[j] | wbk[j];
    L.cf[t * kWPT + j] = wf[j];
    L.cb[t * kWPT + j] = wbk[j];
    L.claim_f[t * kWPT + j] = 0u;
    L.claim_b[t * kWPT + j] = 0u;
  }
  const uint32_t total = bits_scan(L.B, un);  // barrier: words and cleared masks visible
  for (uint32_t k0 = 0; k0 < total; k0 += kBlock * kR) {
    uint32_t v[kR], ib[kR], ie[kR], ob[kR], oe[kR], uf[kR], ub[kR], pf[kR], pb[kR];
    bool has[kR], isf[kR], isb[kR], hf[kR], hb[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const uint32_t k = k0 + (uint32_t)(r * kBlock + t);
      has[r] = k < total;
      v[r] = has[r] ? bits_pick(L.B, wb, k) : 0u;
      const uint32_t i = (v[r] - (uint32_t)(wb * 32)) >> 5, bit = 1u << (v[r] & 31u);
      isf[r] = has[r] && (L.cf[i] & bit);
      isb[r] = has[r] && (L.cb[i] & bit);
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      ib[r] = ie[r] = ob[r] = oe[r] = 0u;
      if (has[r]) {
        ib[r] = f.up_off[v[r]];
        ie[r] = f.up_off[v[r] + 1];
        ob[r] = f.down_off[v[r]];
        oe[r] = f.down_off[v[r] + 1];
      }
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      pf[r] = isf[r] && ib[r] < ie[r];
      pb[r] = isb[r] && ob[r] < oe[r];
      uf[r] = pf[r] ? (uint32_t)f.up_adj[ib[r]] : 0u;
      ub[r] = pb[r] ? (uint32_t)b.up_adj[ob[r]] : 0u;
    }
    uint32_t xf[kR], xb[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      xf[r] = pf[r] ? f.seen[uf[r] >> 5] : 0xffffffffu;
      xb[r] = pb[r] ? b.seen[ub[r] >> 5] : 0xffffffffu;
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      hf[r] = hb[r] = false;
      if (isf[r]) {
        ++cf4[0];
        if (pf[r]) {
          ++cf4[1];
          hf[r] = !(xf[r] & (1u << (uf[r] & 31u)));
          if (!hf[r]) hf[r] = bu_scan(f.up_adj, f.seen, ib[r] + 1u, ie[r], cf4[1]);
        }
      }
      if (isb[r]) {
        ++cb4[0];
        if (pb[r]) {
          ++cb4[1];
          hb[r] = !(xb[r] & (1u << (ub[r] & 31u)));
          if (!hb[r]) hb[r] = bu_scan(b.up_adj, b.seen, ob[r] + 1u, oe[r], cb4[1]);
        }
      }
      const uint32_t i = (v[r] - (uint32_t)(wb * 32)) >> 5, bit = 1u << (v[r] & 31u);
      if (hf[r]) atomicOr(&L.claim_f[i], bit);
      if (hb[r]) atomicOr(&L.claim_b[i], bit);
    }
    // per-search claim counts from ballots (the reconvergence point after the scans, as in
    // bu_group) and the claims' down-list lengths: the other search's up-list bounds
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const unsigned long long mf = __ballot(hf[r]), mb = __ballot(hb[r]);
      if (lane() == 0) {
        cf4[2] += (unsigned long long)__popcll(mf);
        cb4[2] += (unsigned long long)__popcll(mb);
      }
      if (hf[r]) cf4[3] += oe[r] - ob[r];
      if (hb[r]) cb4[3] += ie[r] - ib[r];
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kWPT; ++j) {
    const int i = t * kWPT + j;
    const uint32_t c1 = L.claim_f[i], c2 = L.claim_b[i];
    f.fresh[wb + i] = c1;
    b.fresh[wb + i] = c2;
    if (c1) {
      f.cand[wb + i] = wf[j] & ~c1;
      f.seen[wb + i] = sf[j] & ~c1;
    }
    if (c2) {
      b.cand[wb + i] = wbk[j] & ~c2;
      b.seen[wb + i] = sb[j] & ~c2;
    }
  }
}

// LDS of a top-down level that expands the claims of the bottom-up level before it from the
// fresh bitmap: the word group, and 256 of its vertices at a time with the exclusive prefix
// of their list lengths.
struct TdBitsLds {
  BitIterT<kWPTWide> B;
  uint32_t base[kBlock];
  uint32_t beg[kBlock];
  uint32_t len1[kBlock];
  uint32_t beg2[kBlock];
  uint32_t scan[kBlock / 64];
};

// A top-down level right after a bottom-up one (multistep/scc_fwbw.cpp:87-114 on the
// vertices that level found), for one word group [wb, wb + kBWide) of a.fresh: the group's
// vertices are taken 256 at a time, their list lengths prefix-summed in LDS, and their
// edges expanded kBlock * kPer at a time (owner by binary search over the prefix), with the
// same probes, claims and staging as td_chunk.  The edges of a group stay in its workgroup.
// Every thread calls; claims are staged for frontier N.
template <bool MIRROR>
__device__ __forceinline__ void td_bits_group(const LevelArgs& a, const Frontier& N, int64_t wb, Stage& S,
                                              TdBitsLds& T) {
  const int t = threadIdx.x;
  const uint32_t total = bits_load(T.B, a.fresh, a.nwords, wb);  // ends with a barrier
  for (uint32_t k0 = 0; k0 < total; k0 += kBlock) {
    const uint32_t k = k0 + (uint32_t)t;
    uint32_t b1 = 0, d1 = 0, b2 = 0, d2 = 0;
    if (k < total) {
      const uint32_t v = bits_pick(T.B, wb, k);
      b1 = a.down_off[v];
      d1 = a.down_off[v + 1] - b1;
      if (MIRROR) {
        b2 = a.down_off2[v];
        d2 = a.down_off2[v + 1] - b2;
      }
    }
    const uint32_t d = d1 + d2;
    const uint32_t incl = wave_incl_sum(d);
    if (lane() == 63) T.scan[wave_id()] = incl;
    __syncthreads();
    uint32_t carry = 0, tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      if (w < wave_id()) carry += T.scan[w];
      tot += T.scan[w];
    }
    T.base[t] = incl - d + carry;
    T.beg[t] = b1;
    T.len1[t] = d1;
    T.beg2[t] = b2;
    __syncthreads();
    const int nv = (int)min((uint32_t)kBlock, total - k0);
    for (uint32_t e0 = 0; e0 < tot; e0 += kBlock * kPer) {
      uint32_t w[kPer], word[kPer];
      bool ok[kPer];
#pragma unroll
      for (int r = 0; r < kPer; ++r) {
        const uint32_t i = e0 + (uint32_t)(r * kBlock + t);
        ok[r] = i < tot;
        w[r] = 0u;
        if (ok[r]) {
          int lo = 0, hi = nv;  // the last vertex whose edges start at or before i
#pragma unroll
          for (int st = 0; st < 8; ++st) {  // nv <= 256
            const int mid = (lo + hi) >> 1;
            if (hi - lo > 1) {
              if (T.base[mid] <= i) lo = mid;
              else hi = mid;
            }
          }
          const uint32_t idx = i - T.base[lo];
          if (MIRROR && idx >= T.len1[lo]) w[r] = (uint32_t)a.down_adj2[T.beg2[lo] + (idx - T.len1[lo])];
          else w[r] = (uint32_t)a.down_adj[T.beg[lo] + idx];
        }
      }
#pragma unroll
      for (int r = 0; r < kPer; ++r) word[r] = ok[r] ? a.cand[w[r] >> 5] : 0u;
#pragma unroll
      for (int r = 0; r < kPer; ++r) {
        const bool got = ok[r] && ((word[r] >> (w[r] & 31u)) & 1u) && bit_clear(a.cand, w[r]);
        if (got)
          __hip_atomic_fetch_and(&a.seen[w[r] >> 5], ~(1u << (w[r] & 31u)), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t deg = got ? list_len(a.down_off, MIRROR ? a.down_off2 : nullptr, w[r]) : 0u;
        stage_push(S, got, w[r], deg);
      }
      stage_flush(S, N, a.ctrl, kStage - kBlock * kPer);  // also the barrier before the next step
    }
    __syncthreads();  // T.base and T.scan are rewritten by the next 256 vertices
  }
}

// Local continuation of a colour level (push or per-colour backward): the claims / raises
// this workgroup staged are expanded by the workgroup itself, from a private frontier, for
// up to a.local_rounds more levels while they have at most kLocalEdges edges and the
// launch has run for less than a.local_ticks of the wall clock; whatever is staged when it
// stops goes to the global next frontier.  A path of workgroup-local hops thus costs one
// launch instead of one per hop, and the time budget keeps a long chain in a few
// workgroups from outlasting the next launch, which spreads the work over the whole grid.
// Push stays exact: a local expansion
// clears the vertex's worklist bit before it reads the colour where atomics are performed,
// so a raise that found the bit set is carried by that expansion (max-label propagation
// reaches the same fixpoint in any order).  Colour-BW claims are final (each vertex's
// valid bit is cleared once) and colours no longer change, so their expansion is the
// global one moved earlier.  Every thread calls; (vertices, edges) expanded here are added
// to loc_n / loc_e by thread 0.
// A local frontier of at most kLocalLds claims and one chunk of edges is kept in LDS (a
// chain of single-vertex hops then reads no frontier from memory); larger ones go to the
// workgroup's part of a.loc.
constexpr uint32_t kLocalLds = 256;
struct LocalLds {
  int32_t q[kLocalLds];
  uint32_t qo[kLocalLds];
  uint32_t starts[4];
};

// A tiny local frontier (at most 64 vertices and 64 edges, a chain's usual hop) is walked
// by wave 0 alone, hop by hop in registers, with no LDS staging and no workgroup barrier
// per hop.  The staged vertices come off the worklist as in any local expansion (bit
// cleared, then the colour read where atomics are performed).  A vertex the walk raises
// (push) or claims (colour-BW) is expanded at the next hop with the colour it was given,
// without touching its worklist bit: a larger raise by anyone else later finds the bit
// clear and stages an expansion of its own, and max-label propagation reaches the same
// fixpoint in any order.  The walk stops when a hop has more than 64 edges, after
// kWalkHops hops or once the launch's time budget is spent; the frontier it stops at is
// staged again (push: each vertex whose worklist bit this call turns on, as for any
// raise).  (vertices, edges) expanded are added to loc_n / loc_e by thread 0.  Every
// thread calls; ends with a workgroup barrier.
constexpr int kWalkHops = 1024;
template <int OP>
__device__ void wave_walk(const LevelArgs& a, uint32_t* inq, Stage& S, uint32_t n, unsigned long long& raised,
                          unsigned long long& loc_n, unsigned long long& loc_e, unsigned long long t_begin) {
  if (wave_id() == 0) {
    const uint32_t l = (uint32_t)lane();
    uint32_t v = 0, beg = 0, len = 0;
    int32_t col = 0;
    // colour-BW: the valid word a frontier vertex's claim returned (its bit set: the claim
    // won); the staged vertices were claimed already
    uint32_t cold = 0xffffffffu;
    if (l < n) {
      v = (uint32_t)S.v[l];
      if (OP == OP_PUSH) {
        bit_clear(inq, v);
        __builtin_amdgcn_s_waitcnt(0);
        col = __hip_atomic_load(&a.colors[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        col = a.colors[v];
      }
      beg = a.down_off[v];
      len = a.down_off[v + 1] - beg;
    }
    if (l == 0) S.cnt = 0u;  // the staged vertices are all in registers now
    uint32_t fn = n;
    for (int hop = 0;; ++hop) {
      const uint32_t incl = wave_incl_sum(len);
      const uint32_t total = __shfl(incl, 63, 64);
      if (total == 0u) {
        const uint32_t won = OP == OP_CBW ? (uint32_t)__popcll(__ballot(l < fn && ((cold >> (v & 31u)) & 1u))) : fn;
        if (l == 0) loc_n += won;
        break;
      }
      const bool stop = __builtin_amdgcn_readfirstlane(
                            (total > 64u || hop >= kWalkHops || (wall_clock64() - t_begin) >= a.local_ticks) ? 1
                                                                                                           : 0) != 0;
      if (stop) {
        bool take = l < fn;
        // the last hop's raises are performed before their worklist bits are set: a local
        // expansion that clears a bit and then reads the colour must see the raise
        if (OP == OP_PUSH) __builtin_amdgcn_s_waitcnt(0);
        if (OP == OP_PUSH && take) take = bit_set(inq, v);
        if (OP == OP_CBW) take = take && ((cold >> (v & 31u)) & 1u);  // only the claims it won
        stage_push(S, take, v, len);
        break;
      }
      if (l == 0) loc_e += total;
      if (OP == OP_PUSH && l == 0) loc_n += fn;
      // edge l of this hop belongs to the first lane whose inclusive prefix exceeds l
      uint32_t j = 0;
      for (uint32_t step = 32; step >= 1; step >>= 1)
        if (__shfl(incl, (int)(j + step - 1), 64) <= l) j += step;
      const uint32_t jb = __shfl(beg, (int)j, 64);
      const uint32_t jx = __shfl(incl - len, (int)j, 64);
      const int32_t jc = __shfl(col, (int)j, 64);
      uint32_t w = 0, wb = 0, we = 0, ow = 0u;
      bool got = false;
      if (l < total) w = (uint32_t)a.down_adj[jb + (l - jx)];
      if (OP == OP_CBW) {
        // the claims of the last hop, issued then and read only now (while this hop's
        // adjacency load is in flight): a claimed vertex counts once, for the walk that won it
        const unsigned long long won = __ballot(l < fn && ((cold >> (v & 31u)) & 1u));
        if (l == 0) loc_n += (uint32_t)__popcll(won);
      }
      if (l < total) {
        const uint32_t word = a.valid[w >> 5];
        const int32_t wc = a.colors[w];
        // the target's list bounds with its probes: the next hop needs no load to start
        wb = a.down_off[w];
        we = a.down_off[w + 1];
        const bool in = (word >> (w & 31u)) & 1u;
        if (OP == OP_PUSH) {
          // the raise is not waited for: w is walked on with jc at the next hop whether or
          // not a larger colour reached w first (then that walk is redundant, never wrong),
          // which takes one of a hop's three dependent round trips off the chain
          got = in && wc < jc;
          if (got) {
            atomicMax(&a.colors[w], jc);
            ++raised;
          }
        } else {
          // the claim is not waited for either: w is walked on at once, and if another walk
          // of the same colour won w, expanding it again claims nothing more (each claim is
          // an atomic); the label and the colour's minimum are the same whoever writes them
          got = in && wc == jc;
          if (got) {
            ow = atomicAnd(&a.valid[w >> 5], ~(1u << (w & 31u)));
            a.labels[w] = jc;
            if (a.cmin && (int32_t)w < a.cmin[jc]) atomicMin(&a.cmin[jc], (int32_t)w);
          }
        }
      }
      // the next frontier: lane k takes the (k+1)-th lane that raised / claimed
      const unsigned long long m = __ballot(got);
      fn = (uint32_t)__popcll(m);
      uint32_t src = 0;
      for (uint32_t step = 32; step >= 1; step >>= 1) {
        const uint32_t c = src + step - 1;  // at most 62
        if ((uint32_t)__popcll(m & ((2ull << c) - 1ull)) <= l) src += step;
      }
      const uint32_t nv = __shfl(w, (int)src, 64);
      const int32_t nc = __shfl(jc, (int)src, 64);
      const uint32_t nb = __shfl(wb, (int)src, 64);
      const uint32_t ne = __shfl(we, (int)src, 64);
      const uint32_t no = OP == OP_CBW ? __shfl(ow, (int)src, 64) : 0u;
      v = 0;
      beg = 0;
      len = 0;
      col = 0;
      cold = 0u;
      if (l < fn) {
        v = nv;
        col = nc;
        beg = nb;
        len = ne - nb;
        cold = no;
      }
    }
  }
  __syncthreads();
}

template <int OP, bool MIRROR>
__device__ void local_rounds(const LevelArgs& a, const Frontier& N, uint32_t* inq, Stage& S,
                             ChunkLds<OP, MIRROR>& L, LocalLds& Q, unsigned long long& raised,
                             unsigned long long& loc_n, unsigned long long& loc_e, unsigned long long t_begin) {
  int32_t* mine = a.loc + (size_t)blockIdx.x * kLocalStride;
  for (int lr = 0; lr < a.local_rounds; ++lr) {
    __syncthreads();
    if (threadIdx.x == 0) S.go = (wall_clock64() - t_begin) < a.local_ticks ? 1u : 0u;
    __syncthreads();
    const uint32_t n = S.cnt;
    const bool go = S.go != 0u;
    __syncthreads();
    if (n == 0u || !go) break;  // uniform
    const uint32_t total = stage_sum(S, n);
    if (!MIRROR && n <= 64u && total <= 64u) {  // uniform
      wave_walk<OP>(a, inq, S, n, raised, loc_n, loc_e, t_begin);
      continue;
    }
    if (total > kLocalEdges) break;  // uniform: too many for one workgroup; the grid takes them
    int32_t* pb = mine + (lr & 1) * kLocalHalf;
    const bool in_lds = n <= kLocalLds && total <= kChunk;  // uniform
    const Frontier P = in_lds ? Frontier{Q.q, Q.qo, Q.starts}
                              : Frontier{pb, (uint32_t*)(pb + kStage), (uint32_t*)(pb + 2 * kStage)};
    stage_scan(S, n);
    __syncthreads();
    stage_write(S, n, total, P, 0u, 0u, nullptr, nullptr);
    if (threadIdx.x == 0) {
      loc_n += n;
      loc_e += total;
    }
    const uint32_t nch = (total + kChunk - 1) / kChunk;
    for (uint32_t ch = 0; ch < nch; ++ch)
      td_chunk<OP, MIRROR, true>(a, P, N, inq, inq, ch, nch, n, total, S, L, raised);
  }
}

// Pull levels hand out their bitmap word groups XCD-aware.  Workgroups are dispatched to
// the 8 XCDs round robin (workgroup b to XCD b % 8) and each XCD has its own L2, so with
// the groups dealt out in order every XCD touches every region of the graph at once.  Instead the
// positions 0, 1, 2, ... a grid-stride loop visits are padded to a multiple of 8 and
// position p stands for group (p % 8) * ceil(groups / 8) + p / 8: with a grid that is a
// multiple of 8, XCD x walks the x-th contiguous eighth of the groups, and neighbouring
// vertices (one graph of a batch, one band of a mesh) share an L2.  The map is a
// bijection onto [0, 8 * ceil(groups / 8)); positions past the last group do nothing.
__device__ __forceinline__ uint32_t xcd_positions(uint32_t groups) { return (groups + 7u) & ~7u; }
__device__ __forceinline__ uint32_t xcd_group(uint32_t pos, uint32_t groups) {
  return (pos & 7u) * ((groups + 7u) >> 3) + (pos >> 3);
}

// Edges a pull level's workgroup walks per step (kPullPer per thread, all loads in flight
// together).
constexpr int kPullPer = 8;
constexpr uint32_t kPullEdges = (uint32_t)kBlock * kPullPer;


// LDS of a pull level: one bitmap word group and kBlock of its candidates' in-lists.
template <bool MIRROR>
struct PullLds {
  BitIterT<1> B;
  uint32_t pref[kBlock + 1];  // exclusive prefix of the candidates' list lengths
  uint32_t beg[kBlock];
  uint32_t beg2[MIRROR ? kBlock : 1];
  uint32_t len1[MIRROR ? kBlock : 1];
  int32_t acc[kBlock];
  uint32_t scan[kBlock / 64];
};

// One word group [wb, wb + group_words(a)) of a pull level: a dense push level of max-label
// propagation (multistep/scc_color.cpp:67-163) turned around.  Every valid vertex of the
// group takes the largest colour among its valid in-neighbours (in- and out-neighbours
// when mirrored) and, when that exceeds its own, stores it and is queued for frontier N.
// Only the workgroup that owns a vertex's word writes its colour, so the level needs no
// atomics beyond the queue reservations; a colour read while its owner raises it is
// either value, and both are valid lower bounds of the fixpoint, so the level ends at the
// same fixpoint as a push level.  With `ids` the colours are still the seed's
// (colors[u] == u for every valid u) and a neighbour's id stands for its colour.  A
// candidate's list is walked kPullEdges edges at a time by the whole workgroup, the owner of
// an edge found by binary search over the candidates' prefix in LDS, so one long list
// does not stall a thread.  The group's words of the level's worklist bitmap are cleared:
// no pull level reads it, and the level after needs both worklist bitmaps clear.  Every
// thread calls; raised / candidates / edges / colour reads are per-thread counts.
template <bool MIRROR>
__device__ __forceinline__ void pull_group(const LevelArgs& a, const Frontier& N, uint32_t* inq_cur, int64_t wb,
                                           bool ids, Stage& S, PullLds<MIRROR>& P, unsigned long long& raised,
                                           unsigned long long& cands, unsigned long long& edges,
                                           unsigned long long& reads) {
  const int t = threadIdx.x;
  const int64_t wend = wb + group_words(a) < a.nwords ? wb + group_words(a) : a.nwords;
  if (wb + t < wend) inq_cur[wb + t] = 0u;
  const uint32_t total = bits_load(P.B, a.valid, wend, wb);  // ends with a barrier
  if (t == 0) cands += total;
  for (uint32_t k0 = 0; k0 < total; k0 += kBlock) {
    const uint32_t nb = min(total - k0, (uint32_t)kBlock);
    const bool has = (uint32_t)t < nb;
    uint32_t v = 0, b1 = 0, l1 = 0, b2 = 0, l2 = 0;
    int32_t own = 0;
    if (has) {
      v = bits_pick(P.B, wb, k0 + (uint32_t)t);
      b1 = a.up_off[v];
      l1 = a.up_off[v + 1] - b1;
      if (MIRROR) {
        b2 = a.up_off2[v];
        l2 = a.up_off2[v + 1] - b2;
      }
      own = a.colors[v];
    }
    const uint32_t len = l1 + l2;
    const uint32_t incl = wave_incl_sum(len);
    if (lane() == 63) P.scan[wave_id()] = incl;
    P.beg[t] = b1;
    if (MIRROR) {
      P.beg2[t] = b2;
      P.len1[t] = l1;
    }
    P.acc[t] = own;
    __syncthreads();
    uint32_t carry = 0, T = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      if (w < wave_id()) carry += P.scan[w];
      T += P.scan[w];
    }
    P.pref[t] = incl - len + carry;  // threads past nb have len 0: pref = T
    if (t == 0) {
      P.pref[kBlock] = T;
      edges += T;
    }
    __syncthreads();
    for (uint32_t e0 = 0; e0 < T; e0 += kPullEdges) {
      uint32_t u[kPullPer], j[kPullPer], word[kPullPer];
      int32_t cu[kPullPer];
      bool ok[kPullPer];
#pragma unroll
      for (int r = 0; r < kPullPer; ++r) {
        const uint32_t e = e0 + (uint32_t)t + (uint32_t)r * kBlock;
        ok[r] = e < T;
        u[r] = 0;
        j[r] = 0;
        if (ok[r]) {
          // the last candidate whose list starts at or before e (pref[0] = 0 <= e < T)
          uint32_t lo = 0;
#pragma unroll
          for (uint32_t step = kBlock / 2; step >= 1; step >>= 1)
            if (lo + step < nb && P.pref[lo + step] <= e) lo += step;
          const uint32_t i = e - P.pref[lo];
          j[r] = lo;
          if (MIRROR && i >= P.len1[lo]) u[r] = (uint32_t)a.up_adj2[P.beg2[lo] + (i - P.len1[lo])];
          else u[r] = (uint32_t)a.up_adj[P.beg[lo] + i];
        }
      }
      // valid words first, then the colours of valid neighbours only (reading every
      // neighbour's colour alongside its word was slower: after FW-BW most neighbours are
      // no longer valid and their colour lines are cold)
#pragma unroll
      for (int r = 0; r < kPullPer; ++r) word[r] = ok[r] ? a.valid[u[r] >> 5] : 0u;
#pragma unroll
      for (int r = 0; r < kPullPer; ++r) {
        const bool in = ok[r] && ((word[r] >> (u[r] & 31u)) & 1u);
        cu[r] = (in && !ids) ? a.colors[u[r]] : (int32_t)u[r];
      }
#pragma unroll
      for (int r = 0; r < kPullPer; ++r) {
        const bool in = ok[r] && ((word[r] >> (u[r] & 31u)) & 1u);
        if (in && !ids) ++reads;
        if (in && cu[r] > P.acc[j[r]]) atomicMax(&P.acc[j[r]], cu[r]);
      }
    }
    __syncthreads();
    bool take = false;
    uint32_t deg = 0;
    if (has) {
      const int32_t best = P.acc[t];
      if (best > own) {
        a.colors[v] = best;
        take = true;
        ++raised;
        deg = list_len(a.down_off, MIRROR ? a.down_off2 : nullptr, v);
      }
    }
    stage_push(S, take, v, deg);
    stage_flush(S, N, a.ctrl, kStage - kBlock);  // also the barrier before P is refilled
  }
  __syncthreads();  // B is reloaded next
}

// One colour sweep level (push or per-colour backward; work only, k_close_color ends it):
// edge-balanced expansion of the current frontier, one chunk per workgroup at a time;
// workgroups without a chunk exit at once (the closer sums only the rows of the ones that
// ran); then, with a.loc, the workgroup continues on its own claims (local_rounds).
// part row = (vertices and edges expanded locally, raised, 0).
template <int OP, bool MIRROR, bool SLOTS = false>
__global__ void __launch_bounds__(kBlock) k_level(LevelArgs a) {
  static_assert(OP == OP_PUSH || OP == OP_CBW, "colour ops only; searches run in k_reach");
  __shared__ Stage S;
  // a push level either expands its frontier (ChunkLds) or pulls (PullLds)
  __shared__ union {
    ChunkLds<OP, MIRROR> L;
    PullLds<MIRROR> P;
  } lds;
  ChunkLds<OP, MIRROR>& L = lds.L;
  __shared__ LocalLds Q;  // small local frontiers
  __shared__ unsigned long long s_red[kBlock / 64];

  Ctrl* c = a.ctrl;
  if (c->done) return;  // uniform: the sweep already ended
  if (OP == OP_PUSH && c->bottom_up != 0u) {
    // pull level: every word group of the bitmap; part row = (candidates, edges, raised,
    // colour reads)
    const int64_t pw = group_words(a);
    const uint32_t groups = (uint32_t)((a.nwords + pw - 1) / pw), pos = xcd_positions(groups);
    if (blockIdx.x >= pos) return;  // uniform
    const Frontier N = c->cur ? a.fr0 : a.fr1;
    uint32_t* inq_cur = c->cur ? a.inq1 : a.inq0;
    const bool ids = c->bottom_up == 2u;
    if (threadIdx.x == 0) {
      S.cnt = 0u;
      if (a.recs) atomicMin(&c->t_start, wall_clock64());
    }
    __syncthreads();
    unsigned long long ra = 0, cn = 0, ed = 0, rd = 0;
    for (uint32_t p = blockIdx.x; p < pos; p += gridDim.x) {
      const uint32_t g = xcd_group(p, groups);
      if (g < groups) pull_group<MIRROR>(a, N, inq_cur, (int64_t)g * pw, ids, S, lds.P, ra, cn, ed, rd);
    }
    stage_flush(S, N, c, 0u);
    ra = block_sum64(ra, s_red);
    rd = block_sum64(rd, s_red);
    if (threadIdx.x == 0) {
      unsigned long long* p = a.part + 4 * blockIdx.x;
      p[0] = cn;
      p[1] = ed;
      p[2] = ra;
      p[3] = rd;
    }
    return;
  }
  const uint32_t cur = c->cur;
  const uint32_t q_size = c->q_size, q_edges = c->q_edges;
  const uint32_t units = td_units(a, q_size, q_edges);
  if (blockIdx.x >= units) return;  // uniform: no unit for this workgroup
  const Frontier F = cur ? a.fr1 : a.fr0;
  const Frontier N = cur ? a.fr0 : a.fr1;
  uint32_t* inq_cur = cur ? a.inq1 : a.inq0;
  uint32_t* inq_nxt = cur ? a.inq0 : a.inq1;
  const int t = threadIdx.x;
  unsigned long long my_raised = 0;
  const unsigned long long t_begin = wall_clock64();
  if (t == 0) {
    S.cnt = 0u;
    if (a.recs) atomicMin(&c->t_start, t_begin);
  }
  __syncthreads();
  // nothing reads the current worklist bitmap during a grid-wide level (local rounds use
  // the next one) and the level must leave it clear: each expanded vertex's whole word is
  // cleared with a plain store instead of an atomic on its bit
  // (chunks dealt out XCD-aware, as pull groups are, made no difference on the grid or
  // the batch)
  if (!SLOTS) {
    for (uint32_t ch = blockIdx.x; ch < units; ch += gridDim.x)
      td_chunk<OP, MIRROR, false, true>(a, F, N, inq_cur, inq_nxt, ch, units, q_size, q_edges, S, L, my_raised);
  } else if (td_variant(a, q_size, q_edges) == kBalanceGlobal) {
    for (uint32_t ch = blockIdx.x; ch < units; ch += gridDim.x)
      td_chunk<OP, MIRROR, false, true>(a, F, N, inq_cur, inq_nxt, ch, units, q_size, q_edges, S, L, my_raised);
  } else if (td_variant(a, q_size, q_edges) == kBalanceLocal) {
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x)
      td_slots<OP, MIRROR, kBalanceLocal>(a, F, N, inq_cur, inq_nxt, u, q_size, S, L, my_raised);
  } else {
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x)
      td_slots<OP, MIRROR, kBalanceVertex>(a, F, N, inq_cur, inq_nxt, u, q_size, S, L, my_raised);
  }
  unsigned long long loc_n = 0, loc_e = 0;
  if (a.loc && a.local_rounds > 0)
    local_rounds<OP, MIRROR>(a, N, inq_nxt, S, L, Q, my_raised, loc_n, loc_e, t_begin);
  stage_flush(S, N, c, 0u);
  unsigned long long ra = block_sum64(my_raised, s_red);
  if (t == 0) {
    unsigned long long* p = a.part + 4 * blockIdx.x;
    p[0] = loc_n;
    p[1] = loc_e;
    p[2] = ra;
    p[3] = 0;
  }
}

// Work units of a reach search's coming level: bitmap word groups (bottom-up) or
// kChunk-edge chunks of the frontier (top-down); none once the search is done.
__device__ __forceinline__ uint32_t reach_units(const Ctrl* c, const LevelArgs& a) {
  if (c->done) return 0u;
  if (c->bottom_up || c->from_bits) {
    const int64_t gw = (c->bu_wide || !c->bottom_up) ? kBWide : bu_words(a);
    return (uint32_t)((a.nwords + gw - 1) / gw);
  }
  return td_units(a, c->q_size, c->q_edges);
}

// The units [first, units) of one reach search, stepping by the grid, then a flush of the
// claims staged for it.  Every thread calls (uniform arguments).
template <bool MIRROR, bool SLOTS>
__device__ __forceinline__ void reach_search(const LevelArgs& a, uint32_t first, uint32_t units, Stage& S,
                                             ChunkLds<OP_FW, MIRROR>& L, BitIter& B, uint32_t* s_claim,
                                             BitIterT<kWPTWide>& BW, uint32_t* s_claimw, TdBitsLds& T,
                                             unsigned long long& ex, unsigned long long& in,
                                             unsigned long long& nx, unsigned long long& ne) {
  Ctrl* c = a.ctrl;
  const uint32_t cur = c->cur, q_size = c->q_size, q_edges = c->q_edges;
  const bool bu = c->bottom_up != 0u;
  const Frontier F = cur ? a.fr1 : a.fr0;
  const Frontier N = cur ? a.fr0 : a.fr1;
  unsigned long long unused = 0ull;
  if (bu) {
    // groups in plain order: dealt out XCD-aware (as pull levels are) they cost R-MAT s24
    // 2.5% and gave a batch nothing.  A level with few candidates left walks wider
    // groups: its time is the per-group chain of barriers and dependent loads, not the
    // candidates, and the grid then covers the bitmap in one round of groups
    if (c->bu_wide) {
      for (uint32_t lu = first; lu < units; lu += gridDim.x)
        bu_group<MIRROR, kWPTWide>(a, (int64_t)lu * kBWide, BW, s_claimw, ex, in, nx, ne);
    } else {
      for (uint32_t lu = first; lu < units; lu += gridDim.x)
        bu_group<MIRROR, kWPT>(a, (int64_t)lu * bu_words(a), B, s_claim, ex, in, nx, ne);
    }
    __syncthreads();  // nothing staged; the workgroup may move on to the other search's LDS
    return;
  }
  if (c->from_bits) {
    for (uint32_t lu = first; lu < units; lu += gridDim.x) td_bits_group<MIRROR>(a, N, (int64_t)lu * kBWide, S, T);
  } else if (!SLOTS) {
    for (uint32_t lu = first; lu < units; lu += gridDim.x)
      td_chunk<OP_FW, MIRROR>(a, F, N, nullptr, nullptr, lu, units, q_size, q_edges, S, L, unused);
  } else if (td_variant(a, q_size, q_edges) == kBalanceGlobal) {
    for (uint32_t lu = first; lu < units; lu += gridDim.x)
      td_chunk<OP_FW, MIRROR>(a, F, N, nullptr, nullptr, lu, units, q_size, q_edges, S, L, unused);
  } else if (td_variant(a, q_size, q_edges) == kBalanceLocal) {
    for (uint32_t lu = first; lu < units; lu += gridDim.x)
      td_slots<OP_FW, MIRROR, kBalanceLocal>(a, F, N, nullptr, nullptr, lu, q_size, S, L, unused);
  } else {
    for (uint32_t lu = first; lu < units; lu += gridDim.x)
      td_slots<OP_FW, MIRROR, kBalanceVertex>(a, F, N, nullptr, nullptr, lu, q_size, S, L, unused);
  }
  stage_flush(S, N, c, 0u);
}

// One level of NS reach searches (FW / BW / CC) in a single launch.  The searches' work
// units are laid end to end (search 0's, then search 1's) and the grid strides over the
// union, so a forward and a backward search of one pivot share every launch: half the
// launches and closers of running them one after the other, and no closer waits for a
// compute unit behind the other search's level.  Each search keeps its own direction,
// control block, frontiers and bitmaps; a workgroup flushes the claims staged for one
// search before it moves on to the other's units.  Workgroups without a unit exit at once;
// every other one writes one partial row per search, (examined, inspected, claims, their
// down-list lengths) -- the last two counted by bottom-up levels, which queue nothing.
template <int NS, bool MIRROR, bool SLOTS = false>
__global__ void __launch_bounds__(kBlock) k_reach(ReachArgs p) {
  __shared__ Stage S;
  // a search's level is either top-down (ChunkLds, or TdBitsLds right after a bottom-up
  // level) or bottom-up (BitIter + claim masks), and a workgroup moves between searches only
  // across a barrier: they share LDS
  __shared__ union {
    ChunkLds<OP_FW, MIRROR> L;
    struct {
      BitIter B;
      uint32_t claim[kBW];
    } bu;
    struct {
      BitIterT<kWPTWide> B;
      uint32_t claim[kBWide];
    } buw;
    TdBitsLds T;
    BuPairLds pair;
  } U;
  __shared__ unsigned long long s_red[kBlock / 64];
  ChunkLds<OP_FW, MIRROR>& L = U.L;
  BitIter& B = U.bu.B;
  BitIterT<kWPTWide>& BW = U.buw.B;
  uint32_t* s_claim = U.bu.claim;
  uint32_t* s_claimw = U.buw.claim;

  const uint32_t u0 = reach_units(p.s[0].ctrl, p.s[0]);
  const uint32_t u1 = NS == 2 ? reach_units(p.s[1].ctrl, p.s[1]) : 0u;
  // both searches bottom-up over the same word groups: one fused sweep (bu_pair), u0 units
  const bool pair = MSC_BU_PAIR && NS == 2 && !MIRROR && bu_pairable(p.s[0], p.s[NS - 1]);
  const uint32_t b = blockIdx.x, G = gridDim.x;
  if (b >= (pair ? u0 : u0 + u1)) {  // uniform: no unit; zero rows, so that the closer can sum every row
    if (threadIdx.x < 4 * NS) p.s[threadIdx.x >> 2].part[4 * b + (threadIdx.x & 3)] = 0ull;
    return;
  }
  const int t = threadIdx.x;
  if (t == 0) {
    S.cnt = 0u;
    if (u0 && p.s[0].recs) atomicMin(&p.s[0].ctrl->t_start, (unsigned long long)wall_clock64());
    if (NS == 2 && u1 && p.s[1].recs) atomicMin(&p.s[1].ctrl->t_start, (unsigned long long)wall_clock64());
  }
  __syncthreads();
  unsigned long long c0[4] = {0ull, 0ull, 0ull, 0ull}, c1[4] = {0ull, 0ull, 0ull, 0ull};
  if (pair) {
    for (uint32_t lu = b; lu < u0; lu += G) bu_pair(p.s[0], p.s[NS - 1], (int64_t)lu * kBW, U.pair, c0, c1);
    __syncthreads();  // the LDS is reused by the partial sums below
  } else if (b < u0) {
    reach_search<MIRROR, SLOTS>(p.s[0], b, u0, S, L, B, s_claim, BW, s_claimw, U.T, c0[0], c0[1], c0[2], c0[3]);
  }
  if (NS == 2 && !pair) {
    // this workgroup's first unit of the union at or past u0, in search 1's numbering
    const uint32_t first = b >= u0 ? b - u0 : b + (u0 - b + G - 1) / G * G - u0;
    if (first < u1)
      reach_search<MIRROR, SLOTS>(p.s[1], first, u1, S, L, B, s_claim, BW, s_claimw, U.T, c1[0], c1[1], c1[2], c1[3]);
  }
  for (int s = 0; s < NS; ++s) {
    unsigned long long* cs = s == 0 ? c0 : c1;
    unsigned long long r4[4];
    for (int k = 0; k < 4; ++k) r4[k] = block_sum64(cs[k], s_red);
    if (t == 0) {
      unsigned long long* r = p.s[s].part + 4 * b;
      for (int k = 0; k < 4; ++k) r[k] = r4[k];
    }
  }
}

// Rows of the partials array written by the colour level that just ran: those of the
// workgroups that had a unit.
__device__ __forceinline__ int active_rows(int nparts, const LevelArgs& a, uint32_t q_size, uint32_t q_edges) {
  return (int)min((uint32_t)nparts, td_units(a, q_size, q_edges));
}

// Direction of the coming colour level: a push level whose frontier has at least
// a.pull_min list entries and 1/a.pull_div of the round's runs as a pull level (1); else 0.
__device__ __forceinline__ int colour_action(const Ctrl* c, const LevelArgs& a, int op, uint32_t nedges) {
  if (op != OP_PUSH || a.pull_div == 0u || nedges < a.pull_min) return 0;
  return (unsigned long long)nedges * a.pull_div >= (unsigned long long)c->round_edges ? 1 : 0;
}

// Thread 0 only: books the level that just closed (next frontier from next_packed,
// algorithmic bytes, level record) and turns the control block to the next level with
// direction `act` (0 top-down, 1 bottom-up, 2 abandon the search).
__device__ void finish_level(Ctrl* c, const LevelArgs& a, int op, bool bu, int act, uint32_t q_size,
                             uint32_t q_edges, unsigned long long ex, unsigned long long in, unsigned long long ra,
                             uint32_t next, uint32_t nedges, const float* feat, unsigned long long rd = 0ull,
                             unsigned long long nvl = 0ull) {
  // algorithmic bytes of this level
  unsigned long long bytes;
  // colour levels: ex / in are the vertices / edges a workgroup expanded locally; a pull
  // level (bu) sweeps the valid and worklist bitmaps, reads each candidate's list bounds
  // and colour (ex), every list entry and its valid bit (in), rd neighbour colours, and
  // writes ra colours and the queue
  if (op == OP_PUSH && bu)
    bytes = 8ull * (unsigned long long)a.nwords + 12ull * ex + 5ull * in + 4ull * rd + 4ull * ra + 8ull * next;
  else if (op == OP_PUSH)
    bytes = 12ull * (q_size + ex) + 9ull * (q_edges + in) + 4ull * ra + 4ull * next;
  else if (op == OP_CBW)
    bytes = 12ull * (q_size + ex) + 9ull * (q_edges + in) + 5ull * next;
  else if (bu)  // candidate and fresh words, list bounds, probes, claims' list bounds (+ queue)
    bytes = 8ull * (unsigned long long)a.nwords + 8ull * ex + 5ull * in + 8ull * next;
  else if (c->from_bits)  // fresh words instead of queue ids
    bytes = 4ull * (unsigned long long)a.nwords + 8ull * q_size + 5ull * q_edges + 5ull * next;
  else
    bytes = 12ull * q_size + 5ull * q_edges + 5ull * next;
  c->bytes_op[op] += bytes;
  if (bu) c->bu_levels += 1;
  if (a.recs && c->nrec < a.rec_cap) {
    msc_level_rec& r = a.recs[c->nrec];
    r.phase = (int32_t)c->phase;
    r.level = (int32_t)c->level;
    r.bottom_up = bu ? 1 : 0;
    r.next_size = (int32_t)next;
    r.q_size = q_size;
    r.q_edges = q_edges;
    r.next_edges = nedges;
    r.examined = (uint32_t)ex;
    r.inspected = in;
    r.found = c->found;
    r.n_total = a.n_total;
    r.num_valid = (int32_t)nvl;
    r.action = act;
    r.pad = 0;
    for (int i = 0; i < kMlpIn; ++i) r.feat[i] = feat ? feat[i] : 0.0f;
    unsigned long long t0 = c->t_start, t1 = wall_clock64();
    r.level_us = (t0 != ~0ull && t1 > t0 && a.clock_mhz > 0.0f) ? (float)(t1 - t0) / a.clock_mhz : 0.0f;
    r.pad2 = 0;
  }
  c->t_start = ~0ull;
  c->nrec += 1;
  c->q_size = next;
  c->q_edges = nedges;
  c->cur ^= 1u;
  c->from_bits = bu ? 1u : 0u;  // a top-down level next expands this level's claims from a.fresh
  c->bottom_up = act == 1 ? 1u : 0u;
  c->next_packed = 0ull;
  c->done = (next == 0u || act == 2) ? 1u : 0u;
  if (act == 2) c->abandoned = 1u;  // the host discards this search's result
}

// One lane's share of the policy MLP, held in registers: lane l < kMlpH owns hidden unit l
// of both layers, lane l < kMlpOut output l, lane l < kMlpIn the standardisation of
// feature l.
struct MlpLane {
  float w1[kMlpIn], w2[kMlpH], w3[kMlpH];
  float b1, b2, b3, mean, sd;
};

__device__ __forceinline__ void mlp_lane_load(const float* blob, MlpLane& m) {
  const float* W1 = blob;
  const float* B1 = W1 + kMlpIn * kMlpH;
  const float* W2 = B1 + kMlpH;
  const float* B2 = W2 + kMlpH * kMlpH;
  const float* W3 = B2 + kMlpH;
  const float* B3 = W3 + kMlpH * kMlpOut;
  const float* mean = blob + kMlpParams;
  const float* sd = mean + kMlpIn;
  const int l = lane();
  const int h = l < kMlpH ? l : 0, o = l < kMlpOut ? l : 0, f = l < kMlpIn ? l : 0;
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) m.w1[i] = W1[h * kMlpIn + i];
#pragma unroll
  for (int i = 0; i < kMlpH; ++i) m.w2[i] = W2[h * kMlpH + i];
#pragma unroll
  for (int i = 0; i < kMlpH; ++i) m.w3[i] = W3[o * kMlpH + i];
  m.b1 = B1[h];
  m.b2 = B2[h];
  m.b3 = B3[o];
  m.mean = mean[f];
  m.sd = sd[f];
}

// The MLP's action for the raw features in s_feat (whole wave; same arithmetic, term
// order and argmax as the training code's forward pass).  s_z / s_h: wave-private LDS.
__device__ __forceinline__ int mlp_eval(const MlpLane& m, const float* s_feat, float* s_z, float* s_h) {
  const int l = lane();
  if (l < kMlpIn) s_z[l] = (s_feat[l] - m.mean) / m.sd;
  __builtin_amdgcn_wave_barrier();
  float h1 = 0.0f;
  if (l < kMlpH) {
    float acc = m.b1;
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) acc += m.w1[i] * s_z[i];
    h1 = fmaxf(acc, 0.0f);
    s_h[l] = h1;
  }
  __builtin_amdgcn_wave_barrier();
  float h2 = 0.0f;
  if (l < kMlpH) {
    float acc = m.b2;
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) acc += m.w2[i] * s_h[i];
    h2 = fmaxf(acc, 0.0f);
  }
  __builtin_amdgcn_wave_barrier();
  if (l < kMlpH) s_h[l] = h2;
  __builtin_amdgcn_wave_barrier();
  float zo = -3.0e38f;
  if (l < kMlpOut) {
    float acc = m.b3;
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) acc += m.w3[i] * s_h[i];
    zo = acc;
  }
  float z0 = __shfl(zo, 0, 64), z1 = __shfl(zo, 1, 64), z2 = __shfl(zo, 2, 64);
  int best = 0;
  float bv = z0;
  if (z1 > bv) {
    bv = z1;
    best = 1;
  }
  if (z2 > bv) best = 2;
  return best;
}

// Closes one reach search's level; one whole wave calls it.  Lane 0 books the level and
// applies the direction rule -- the reference's (multistep/scc_fwbw.cpp:168-195,
// cc_bfs.cpp:167-194), the MLP (when `learned`), or a forced schedule.
__device__ void close_search(const LevelArgs& a, Ctrl* c, unsigned long long nvalid, unsigned long long ex,
                             unsigned long long in, unsigned long long nx, unsigned long long ne, bool learned,
                             const MlpLane& m, float* s_feat, float* s_z, float* s_h) {
  int act = 0;
  uint32_t q_size = 0, q_edges = 0, next = 0, nedges = 0;
  bool bu = false;
  if (lane() == 0) {
    q_size = c->q_size;
    q_edges = c->q_edges;
    bu = c->bottom_up != 0u;
    // a bottom-up level counted its claims in the partial rows (it queues nothing); a
    // top-down level's queue counter holds them
    const unsigned long long np = c->next_packed;
    next = bu ? (uint32_t)nx : (uint32_t)(np >> 32);
    nedges = bu ? (uint32_t)ne : (uint32_t)np;
    c->found += next;
    c->level += 1;
    c->closes += 1;
    act = bu ? 1 : 0;
    if (!c->switched_back && !bu) {
      double edges_frontier = (double)next * a.avg;
      double edges_remainder = (double)((int64_t)a.n_total - c->found) * a.avg;
      if (edges_remainder / a.alpha < edges_frontier) act = 1;
    } else if (!c->switched_back && (a.n_total / a.beta) > (int32_t)next) {
      act = 0;
      c->switched_back = 1;
    }
    level_features(c, a, next, nedges, bu, s_feat, nvalid);
  }
  __builtin_amdgcn_wave_barrier();
  if (learned) {
    int mact = mlp_eval(m, s_feat, s_z, s_h);
    if (lane() == 0) act = mact;
  }
  if (lane() != 0) return;
  if (a.sched && c->level <= (uint32_t)MSC_SCHED_LEVELS) {
    int f = a.sched[c->phase * MSC_SCHED_LEVELS + c->level - 1];
    if (f >= 0 && f <= 2) act = f;
  }
  finish_level(c, a, a.op, bu, act, q_size, q_edges, ex, in, 0ull, next, nedges, s_feat, 0ull, nvalid);
  // a bottom-up level over fewer than one unvisited vertex in 8 walks wide word groups
  // (unvisited: what the trim left, less what the search found; an estimate when the
  // allowed set is smaller)
  const unsigned long long nvl = a.nvalid ? nvalid : (unsigned long long)a.n_total;
  const long long left = (long long)nvl - (long long)c->found;
  c->bu_wide = c->bottom_up && left * 8ll < (long long)a.nwords * 32ll ? 1u : 0u;
}

// Closes a level of NS reach searches (one workgroup).  Everything it reads comes in one
// round trip: both control blocks (copied into LDS, worked on there and copied back once),
// the valid counts, every partial row (a level kernel's idle workgroups write zero rows) and
// the MLP weights (into registers).  Then all threads sum the rows and wave s closes search
// s, the searches in parallel.
constexpr int kCloseBlock = 512;  // k_close_reach's workgroup

template <int NS>
__global__ void __launch_bounds__(kCloseBlock) k_close_reach(ReachArgs p, int nparts) {
  constexpr int kCW = (int)(sizeof(Ctrl) / 8);
  static_assert(sizeof(Ctrl) % 8 == 0 && NS * kCW <= kCloseBlock, "control blocks copied as 8-byte words");
  __shared__ unsigned long long s_ctl[2][kCW];
  __shared__ unsigned long long s_nv[2];
  __shared__ unsigned long long s_sum[2][4][kCloseBlock / 64];
  __shared__ float s_feat[2][kMlpIn];
  __shared__ float s_z[2][kMlpIn];
  __shared__ float s_h[2][kMlpH];
  const int t = threadIdx.x, w = wave_id();
  if (t < NS * kCW) s_ctl[t / kCW][t % kCW] = reinterpret_cast<const unsigned long long*>(p.s[t / kCW].ctrl)[t % kCW];
  if (t >= kCloseBlock - NS) {
    const LevelArgs& a = p.s[kCloseBlock - 1 - t];
    s_nv[kCloseBlock - 1 - t] = a.nvalid ? *a.nvalid : 0ull;
  }
  const bool learned0 = p.s[0].policy == MSC_POLICY_LEARNED && p.s[0].mlp;
  const bool learned1 = NS == 2 && p.s[1].policy == MSC_POLICY_LEARNED && p.s[1].mlp;
  MlpLane m;
  if (w == 0 && learned0) mlp_lane_load(p.s[0].mlp, m);
  if (NS == 2 && w == 1 && learned1) mlp_lane_load(p.s[1].mlp, m);
  // up to four rows per thread in flight together: with the level grid's 1280 rows one
  // round trip instead of five dependent ones
  unsigned long long acc[2][4] = {{0ull, 0ull, 0ull, 0ull}, {0ull, 0ull, 0ull, 0ull}};
  for (int r0 = t; r0 < nparts; r0 += 4 * kCloseBlock) {
    unsigned long long x[4][2][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + j * kCloseBlock;
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) x[j][s][k] = r < nparts ? p.s[s].part[4 * r + k] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[s][k] += x[j][s][k];
  }
  for (int s = 0; s < NS; ++s)
    for (int k = 0; k < 4; ++k) {
      const unsigned long long x = wave_sum64(acc[s][k]);
      if (lane() == 0) s_sum[s][k][w] = x;
    }
  __syncthreads();
  Ctrl* c0 = reinterpret_cast<Ctrl*>(s_ctl[0]);
  Ctrl* c1 = reinterpret_cast<Ctrl*>(s_ctl[1]);
  const uint32_t u0 = reach_units(c0, p.s[0]);
  const uint32_t u1 = NS == 2 ? reach_units(c1, p.s[1]) : 0u;
  if (u0 + u1 == 0u) return;  // uniform: every search is done
  if (w < NS && (w == 0 ? u0 : u1)) {
    unsigned long long tot[4] = {0ull, 0ull, 0ull, 0ull};
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < kCloseBlock / 64; ++j) tot[k] += s_sum[w][k][j];
    if (w == 0)
      close_search(p.s[0], c0, s_nv[0], tot[0], tot[1], tot[2], tot[3], learned0, m, s_feat[0], s_z[0], s_h[0]);
    else
      close_search(p.s[1], c1, s_nv[1], tot[0], tot[1], tot[2], tot[3], learned1, m, s_feat[1], s_z[1], s_h[1]);
  }
  __syncthreads();
  // once the forward search has finished (and was not abandoned), a backward search beside
  // it claims only vertices of the forward set: a path from an SCC member to the pivot never
  // leaves the SCC, so the SCC (forward set & backward set) is unchanged, and the backward
  // search no longer explores what reaches the pivot from outside the forward set
  if (NS == 2 && t == 0 && p.s[1].within_cand && c0->done && !c0->abandoned && !c1->done) c1->within = 1u;
  __syncthreads();
  if (t < NS * kCW && (t < kCW ? u0 : u1))
    reinterpret_cast<unsigned long long*>(p.s[t / kCW].ctrl)[t % kCW] = s_ctl[t / kCW][t % kCW];
}

// Closes a colour sweep (push or per-colour backward level), then keeps going in this one
// workgroup while the frontier has at most a.small_edges edges: each such level is the
// same top-down chunk expansion as in k_level.  All of a level's claims, raises and queue
// writes come from this workgroup (one CU, one L1), so a workgroup barrier is all the next
// level needs to see them; no agent-scope fence (an L2 write-back on multi-XCD gfx950) is
// taken.  Small frontiers thus cost no launch and no grid-wide dispatch; the loop hands
// back to the grid-wide kernel when the frontier grows past the threshold, and stops
// after kSmallLevels levels so a launch stays bounded.
constexpr int kSmallLevels = 4096;

template <int OP, bool MIRROR>
__global__ void __launch_bounds__(kBlock) k_close_color(LevelArgs a, int nparts) {
  __shared__ Stage S;
  __shared__ ChunkLds<OP, MIRROR> L;
  __shared__ LocalLds Q;
  __shared__ unsigned long long s_red[kBlock / 64];
  __shared__ uint32_t s_ctl[4];  // done, cur, q_size, q_edges
  Ctrl* c = a.ctrl;
  if (c->done) return;
  const int t = threadIdx.x;
  {
    const uint32_t q_size = c->q_size, q_edges = c->q_edges;
    const bool pulled = OP == OP_PUSH && c->bottom_up != 0u;
    const int rows = pulled ? (int)min((uint32_t)nparts, xcd_positions((uint32_t)((a.nwords + kBlock - 1) / kBlock)))
                            : active_rows(nparts, a, q_size, q_edges);
    unsigned long long ln = 0, le = 0, ra = 0, rd = 0;
    for (int i = t; i < rows; i += kBlock) {
      ln += a.part[4 * i];
      le += a.part[4 * i + 1];
      ra += a.part[4 * i + 2];
      if (pulled) rd += a.part[4 * i + 3];
    }
    ln = block_sum64(ln, s_red);
    le = block_sum64(le, s_red);
    ra = block_sum64(ra, s_red);
    rd = block_sum64(rd, s_red);
    if (t == 0) {
      const unsigned long long p = c->next_packed;
      const uint32_t next = (uint32_t)(p >> 32), nedges = (uint32_t)p;
      // claims expanded locally never reach the frontier; a pull level's row is its sweep
      c->found += next + (pulled ? 0 : (int64_t)ln);
      c->level += 1;
      c->closes += 1;
      finish_level(c, a, OP, pulled, colour_action(c, a, OP, nedges), q_size, q_edges, ln, le, ra, next, nedges,
                   nullptr, rd);
      s_ctl[0] = c->done;
      s_ctl[1] = c->cur;
      s_ctl[2] = c->q_size;
      s_ctl[3] = c->q_edges;
      S.cnt = 0u;
    }
    __syncthreads();
  }
  for (int it = 0; it < kSmallLevels; ++it) {
    const uint32_t done = s_ctl[0], cur = s_ctl[1], q_size = s_ctl[2], q_edges = s_ctl[3];
    if (done || q_edges > a.small_edges) break;  // uniform
    if (t == 0 && a.recs) c->t_start = wall_clock64();
    const Frontier F = cur ? a.fr1 : a.fr0;
    const Frontier N = cur ? a.fr0 : a.fr1;
    uint32_t* inq_cur = cur ? a.inq1 : a.inq0;
    uint32_t* inq_nxt = cur ? a.inq0 : a.inq1;
    const uint32_t nchunks = (q_edges + kChunk - 1) / kChunk;
    unsigned long long raised = 0, loc_n = 0, loc_e = 0;
    const unsigned long long t_begin = wall_clock64();
    for (uint32_t ch = 0; ch < nchunks; ++ch)
      td_chunk<OP, MIRROR>(a, F, N, inq_cur, inq_nxt, ch, nchunks, q_size, q_edges, S, L, raised);
    // then, as a level launch's workgroup does, on from its own claims (nothing else runs)
    if (a.loc && a.local_rounds > 0)
      local_rounds<OP, MIRROR>(a, N, inq_nxt, S, L, Q, raised, loc_n, loc_e, t_begin);
    stage_flush(S, N, c, 0u);
    raised = block_sum64(raised, s_red);  // ends with a workgroup barrier
    if (t == 0) {
      const unsigned long long p = c->next_packed;
      const uint32_t next = (uint32_t)(p >> 32), nedges = (uint32_t)p;
      c->found += next + (int64_t)loc_n;
      c->level += 1;
      finish_level(c, a, OP, false, colour_action(c, a, OP, nedges), q_size, q_edges, loc_n, loc_e, raised, next,
                   nedges, nullptr);
      s_ctl[0] = c->done;
      s_ctl[1] = c->cur;
      s_ctl[2] = c->q_size;
      s_ctl[3] = c->q_edges;
    }
    __syncthreads();
  }
}

// Vertices per thread per step of the seeding and root sweeps: up to kListR, fewer while
// the grid has a step of one per thread for each.
constexpr int kListR = 4;
static_assert(kListR * kBlock <= (int)kStage, "a step's claims fit the stage");
__device__ __forceinline__ uint32_t list_steps(uint32_t n) {
  const uint32_t per = (n + gridDim.x * kBlock - 1) / (gridDim.x * kBlock);
  return per < 1u ? 1u : per > (uint32_t)kListR ? (uint32_t)kListR : per;
}

// Colour initialisation of a later colouring round (multistep/scc_color.cpp:56-63): the
// previous round's vertex list (length a.ctrl->list_len, which k_ctrl_reset keeps) filtered
// to the vertices still valid; each gets colors[v] = v and is queued, and the queue is also
// written to `out` as this round's list.  Round 0 is seeded by k_commit.  Needs
// k_ctrl_reset first, k_close_seed after.
__global__ void __launch_bounds__(kBlock) k_seed_list(LevelArgs a, const int32_t* __restrict__ in,
                                                      int32_t* __restrict__ out) {
  __shared__ Stage S;
  if (a.rc && !a.rc->go) return;  // uniform
  const uint32_t in_n = a.ctrl->list_len;
  if (threadIdx.x == 0) S.cnt = 0u;
  __syncthreads();
  // up to kListR entries per thread per step, each step's loads in flight together: one entry
  // at a time made a long sweep a chain of dependent round trips (grid 4096^2, round 1:
  // 107 us); a short list keeps one per thread, spread over more workgroups
  const uint32_t rr = list_steps(in_n);
  for (uint32_t i0 = blockIdx.x * kBlock * rr; i0 < in_n; i0 += gridDim.x * kBlock * rr) {
    uint32_t v[kListR], w[kListR], b1[kListR], e1[kListR], b2[kListR], e2[kListR];
    bool take[kListR];
#pragma unroll
    for (int r = 0; r < kListR; ++r) {
      const uint32_t i = i0 + (uint32_t)r * kBlock + threadIdx.x;
      v[r] = (r < (int)rr && i < in_n) ? (uint32_t)in[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kListR; ++r) {
      const uint32_t i = i0 + (uint32_t)r * kBlock + threadIdx.x;
      w[r] = (r < (int)rr && i < in_n) ? a.valid[v[r] >> 5] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kListR; ++r) {
      take[r] = (w[r] >> (v[r] & 31u)) & 1u;
      b1[r] = e1[r] = b2[r] = e2[r] = 0u;
      if (take[r]) {
        a.colors[v[r]] = (int32_t)v[r];
        b1[r] = a.down_off[v[r]];
        e1[r] = a.down_off[v[r] + 1];
        if (a.down_off2) {
          b2[r] = a.down_off2[v[r]];
          e2[r] = a.down_off2[v[r] + 1];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kListR; ++r) stage_push(S, take[r], v[r], (e1[r] - b1[r]) + (e2[r] - b2[r]));
    stage_flush(S, a.fr1, a.ctrl, kStage - kListR * kBlock, out);  // also the barrier before the next step
  }
  stage_flush(S, a.fr1, a.ctrl, 0u, out);
}

// Colouring round 0's seed (multistep/scc_color.cpp:56-63, cc_color.cpp:62-69), from the
// valid bitmap the commit left: every valid vertex gets colors[v] = v and is queued, and
// the queue is also written to `list`.  A workgroup takes 1024 words (4 per thread) at a
// time, so after FW-BW, when few vertices are left, a sweep costs one load per thread.
// Needs a fresh control block (k_round_begin), k_close_seed after; with a.rc only in an
// admitted round.
__global__ void __launch_bounds__(kBlock) k_seed_bits(LevelArgs a, int32_t* __restrict__ list) {
  __shared__ Stage S;
  __shared__ BitIterT<4> B;
  if (a.rc && !a.rc->go) return;  // uniform
  if (threadIdx.x == 0) S.cnt = 0u;
  __syncthreads();
  constexpr int kW = BitIterT<4>::kWords;
  for (int64_t wb = (int64_t)blockIdx.x * kW; wb < a.nwords; wb += (int64_t)gridDim.x * kW) {
    const uint32_t total = bits_load(B, a.valid, a.nwords, wb);
    // kListR vertices per thread per step, their loads in flight together
    for (uint32_t k0 = 0; k0 < total; k0 += kBlock * kListR) {
      uint32_t v[kListR], b1[kListR], e1[kListR], b2[kListR], e2[kListR];
      bool take[kListR];
#pragma unroll
      for (int r = 0; r < kListR; ++r) {
        const uint32_t k = k0 + (uint32_t)r * kBlock + threadIdx.x;
        take[r] = k < total;
        v[r] = take[r] ? bits_pick(B, wb, k) : 0u;
      }
#pragma unroll
      for (int r = 0; r < kListR; ++r) {
        b1[r] = e1[r] = b2[r] = e2[r] = 0u;
        if (take[r]) {
          a.colors[v[r]] = (int32_t)v[r];
          b1[r] = a.down_off[v[r]];
          e1[r] = a.down_off[v[r] + 1];
          if (a.down_off2) {
            b2[r] = a.down_off2[v[r]];
            e2[r] = a.down_off2[v[r] + 1];
          }
        }
      }
#pragma unroll
      for (int r = 0; r < kListR; ++r) stage_push(S, take[r], v[r], (e1[r] - b1[r]) + (e2[r] - b2[r]));
      stage_flush(S, a.fr1, a.ctrl, kStage - kListR * kBlock, list, a.tail.slots ? &a.tail : nullptr);  // a barrier
    }
    __syncthreads();  // B is reloaded next
  }
  stage_flush(S, a.fr1, a.ctrl, 0u, list, a.tail.slots ? &a.tail : nullptr);
}

// Colour roots (multistep/scc_color.cpp:165-212) among this round's vertex list (length
// gate->list_len): colors[v] == v labels v, removes it and queues it for the per-colour
// backward search of control block a.ctrl.  Runs only once the push fixpoint of control
// block `gate` has settled (gate->done); otherwise it does nothing and the host repeats it
// after finishing the push levels.  Needs k_ctrl_reset first, k_close_seed after.
__global__ void __launch_bounds__(kBlock) k_roots_list(LevelArgs a, const int32_t* __restrict__ list,
                                                       const Ctrl* __restrict__ gate) {
  __shared__ Stage S;
  if (!gate->done || (a.rc && !a.rc->go)) return;  // uniform
  const uint32_t nl = gate->list_len;
  if (threadIdx.x == 0) S.cnt = 0u;
  __syncthreads();
  // up to kListR entries per thread per step (as k_seed_list), their loads and claims in
  // flight together
  const uint32_t rr = list_steps(nl);
  for (uint32_t i0 = blockIdx.x * kBlock * rr; i0 < nl; i0 += gridDim.x * kBlock * rr) {
    uint32_t v[kListR], b[kListR], e[kListR];
    int32_t col[kListR];
    bool take[kListR];
#pragma unroll
    for (int r = 0; r < kListR; ++r) {
      const uint32_t i = i0 + (uint32_t)r * kBlock + threadIdx.x;
      v[r] = (r < (int)rr && i < nl) ? (uint32_t)list[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kListR; ++r) {
      const uint32_t i = i0 + (uint32_t)r * kBlock + threadIdx.x;
      col[r] = (r < (int)rr && i < nl) ? a.colors[v[r]] : -1;
    }
    // every list entry is still valid here (the seed listed valid vertices, and nothing but
    // the roots and the colour-BW search after them removes any), so a root's removal needs
    // no result: the atomic is not waited for
#pragma unroll
    for (int r = 0; r < kListR; ++r) {
      take[r] = col[r] == (int32_t)v[r];
      if (take[r]) __hip_atomic_fetch_and(&a.valid[v[r] >> 5], ~(1u << (v[r] & 31u)), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int r = 0; r < kListR; ++r) {
      b[r] = e[r] = 0u;
      if (take[r]) {
        a.labels[v[r]] = (int32_t)v[r];
        if (a.cmin) a.cmin[v[r]] = (int32_t)v[r];
        b[r] = a.down_off[v[r]];
        e[r] = a.down_off[v[r] + 1];
      }
    }
#pragma unroll
    for (int r = 0; r < kListR; ++r) stage_push(S, take[r], v[r], e[r] - b[r]);
    stage_flush(S, a.fr1, a.ctrl, kStage - kListR * kBlock);  // also the barrier before the next step
  }
  stage_flush(S, a.fr1, a.ctrl, 0u);
}

// ------------------------------------------------------------------------------------
// colour tail: when FW-BW leaves few vertices (R-MAT s24: 12 390, joined by 61 edges), one
// workgroup settles all of them, instead of the ~30 launches of the colouring rounds
// (multistep/scc_color.cpp:340-373 loops rounds; scc_main.cpp:257-262 hands the last
// vert_cutoff vertices to a serial tail).  A listed vertex that no edge joins to another
// one is its own component and already carries its id as label (the commit wrote it);
// only the endpoints of the joining edges can share a component, and they are coloured in
// LDS.  The seed of colouring round 0 collects those edges (tail_collect); k_tail_solve
// does the rest.

struct TailLds {
  uint32_t mark[kTailV / 32];  // list slots that are endpoints of a joining edge
  uint32_t pre[kTailV / 32];   // exclusive prefix of mark's popcounts: node of a slot
  uint32_t edge[kTailE];       // node(u) << 16 | node(w)
  int32_t gid[kTailN];         // vertex id of every node
  uint32_t col[kTailN];        // colour (a node); once settled, the node's final label
  int32_t rl[kTailN];          // per root: its component's label (largest or smallest id)
  uint8_t st[kTailN];          // bit 0: still valid, bit 1: in the current root's component
  uint32_t part[kTailBlock / 64];
  uint32_t flag[3];            // sweep i sets flag[i % 3] on a change and clears flag[(i + 1) % 3]
  uint32_t count;
};

// One sweep's end: every thread learns whether the sweep changed anything.  flag[i % 3]
// is cleared two sweeps after it was read, so one barrier per sweep suffices.
__device__ __forceinline__ bool tail_sweep_end(TailLds& T, uint32_t i) {
  __syncthreads();
  return T.flag[i % 3u] != 0u;
}

// Colour tail (one workgroup of kTailBlock threads).  First the two steps the tail path
// does not launch: the seed's hand-over (k_close_seed: the queue becomes the push frontier,
// which a declined tail leaves to the colouring rounds) and, when cnp > 0, the commit's
// count (k_reduce of its cnp rows into cred).  Then the components among the endpoints of
// the joining edges the seed collected (src slot, dst vertex: slots[dst] is its slot).  SCC: rounds of max-colour push to a
// fixpoint, roots (colour == self), and the per-colour backward search from the roots
// (multistep/scc_color.cpp:47-289), each a sweep over the edges in LDS; colours are node
// indices, and each component is labelled with its largest vertex id (raw: what colouring
// gives) or its smallest (canonical).  CC: max-colour propagation over the edges both ways
// (multistep/cc_color.cpp:47-175).  Labels are written only once everything is settled, so
// a declined tail (too many edges or endpoints, or more than kTailSweeps sweeps) leaves
// every buffer as the seed left it: rc->tail = 2, and the host runs the colouring rounds.
// A settled tail ends the colouring: rc->tail = 1, rc->finished.
template <bool CC>
__global__ void __launch_bounds__(kTailBlock) k_tail_solve(RoundCtl* rc, Ctrl* push, const int32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ psrc,
                                                           const int32_t* __restrict__ pdst,
                                                           const int32_t* __restrict__ slots, uint32_t cap,
                                                           uint32_t maxv, int32_t* __restrict__ labels,
                                                           int32_t canonical, uint32_t pull_min,
                                                           const unsigned long long* __restrict__ cpart, int32_t cnp,
                                                           unsigned long long* __restrict__ cred) {
  __shared__ TailLds T;
  __shared__ unsigned long long s_red[4][kTailBlock / 64];
  __shared__ uint32_t s_seed[2];
  if (!rc->go) return;  // uniform: round 0 was not admitted
  const int t = threadIdx.x;
  {
    unsigned long long x[4] = {0ull, 0ull, 0ull, 0ull};
    for (int r = t; r < cnp; r += kTailBlock) {
      x[0] += cpart[4 * r];
      x[1] += cpart[4 * r + 1];
      x[2] += cpart[4 * r + 2];
      x[3] = cpart[4 * r + 3] > x[3] ? cpart[4 * r + 3] : x[3];
    }
    for (int k = 0; k < 3; ++k) x[k] = wave_sum64(x[k]);
    x[3] = wave_max64(x[3]);
    if (lane() == 0)
      for (int k = 0; k < 4; ++k) s_red[k][wave_id()] = x[k];
    if (t == 0) {
      const unsigned long long p = push->next_packed;
      const uint32_t next = (uint32_t)(p >> 32);
      push->list_len = next;
      push->q_size = next;
      push->q_edges = (uint32_t)p;
      push->round_edges = (uint32_t)p;
      if (pull_min) push->bottom_up = (uint32_t)p >= pull_min ? 2u : 0u;
      push->cur = 1u;
      push->next_packed = 0ull;
      push->done = next == 0u ? 1u : 0u;
      s_seed[0] = next;
      s_seed[1] = (uint32_t)p;
    }
    __syncthreads();
    if (t == 0 && cnp > 0) {
      unsigned long long o[4] = {0ull, 0ull, 0ull, 0ull};
      for (int w = 0; w < kTailBlock / 64; ++w) {
        for (int k = 0; k < 3; ++k) o[k] += s_red[k][w];
        o[3] = s_red[3][w] > o[3] ? s_red[3][w] : o[3];
      }
      for (int k = 0; k < 4; ++k) cred[k] = o[k];
    }
  }
  const uint32_t nv = s_seed[0], ne = rc->tail_edges;
  if (nv > maxv || ne > kTailE || ne > cap) {  // uniform
    if (t == 0) rc->tail = 2u;
    return;
  }
  const uint32_t nwd = (nv + 31) / 32;
  for (uint32_t i = t; i < nwd; i += kTailBlock) T.mark[i] = 0u;
  if (t < 3) T.flag[t] = 0u;
  if (t == 0) T.count = 0u;
  __syncthreads();
  for (uint32_t e = t; e < ne; e += kTailBlock) {
    const uint32_t u = psrc[e], w = (uint32_t)slots[pdst[e]];
    if (u >= nv || w >= nv) {  // cannot happen: every listed vertex's slot is below nv <= maxv
      T.count = 1u;
      continue;
    }
    T.edge[e] = (u << 16) | w;
    atomicOr(&T.mark[u >> 5], 1u << (u & 31));
    atomicOr(&T.mark[w >> 5], 1u << (w & 31));
  }
  __syncthreads();
  if (T.count) {  // uniform
    if (t == 0) rc->tail = 2u;
    return;
  }
  {
    // nodes: the marked slots in slot order (kTailV / 32 words, two per thread)
    const uint32_t i0 = 2u * (uint32_t)t, i1 = i0 + 1u;
    const uint32_t c0 = i0 < nwd ? (uint32_t)__popc(T.mark[i0]) : 0u;
    const uint32_t c1 = i1 < nwd ? (uint32_t)__popc(T.mark[i1]) : 0u;
    const uint32_t incl = wave_incl_sum(c0 + c1);
    if (lane() == 63) T.part[wave_id()] = incl;
    __syncthreads();
    uint32_t carry = 0, total = 0;
    for (int w = 0; w < kTailBlock / 64; ++w) {
      if (w < wave_id()) carry += T.part[w];
      total += T.part[w];
    }
    const uint32_t excl = incl - (c0 + c1) + carry;
    if (i0 < nwd) T.pre[i0] = excl;
    if (i1 < nwd) T.pre[i1] = excl + c0;
    if (total > kTailN) {  // uniform
      if (t == 0) rc->tail = 2u;
      return;
    }
    if (t == 0) T.count = total;
  }
  __syncthreads();
  const uint32_t nn = T.count;
  for (uint32_t e = t; e < ne; e += kTailBlock) {
    const uint32_t p = T.edge[e], u = p >> 16, w = p & 0xffffu;
    const uint32_t nu = T.pre[u >> 5] + (uint32_t)__popc(T.mark[u >> 5] & ((1u << (u & 31)) - 1u));
    const uint32_t nw = T.pre[w >> 5] + (uint32_t)__popc(T.mark[w >> 5] & ((1u << (w & 31)) - 1u));
    T.edge[e] = (nu << 16) | nw;
    T.gid[nu] = list[u];
    T.gid[nw] = list[w];
  }
  for (uint32_t x = t; x < nn; x += kTailBlock) {
    T.st[x] = 1u;
    T.col[x] = x;
  }
  __syncthreads();
  uint32_t sweeps = 0, push_sweeps = 0, bw_sweeps = 0, rounds = 0;
  bool settled = true;
  if (CC) {
    for (;;) {
      const uint32_t i = sweeps++;
      if (t == 0) T.flag[(i + 1) % 3u] = 0u;
      for (uint32_t e = t; e < ne; e += kTailBlock) {
        const uint32_t p = T.edge[e], u = p >> 16, w = p & 0xffffu;
        const uint32_t cu = T.col[u], cw = T.col[w];
        if (cu > cw) {
          atomicMax(&T.col[w], cu);
          T.flag[i % 3u] = 1u;
        } else if (cw > cu) {
          atomicMax(&T.col[u], cw);
          T.flag[i % 3u] = 1u;
        }
      }
      ++push_sweeps;
      if (!tail_sweep_end(T, i)) break;
      if (sweeps >= kTailSweeps) {
        settled = false;
        break;
      }
    }
    if (settled) {
      rounds = 1;
      for (uint32_t x = t; x < nn; x += kTailBlock) T.rl[x] = T.gid[x];
      __syncthreads();
      for (uint32_t x = t; x < nn; x += kTailBlock) {
        if (canonical) atomicMin(&T.rl[T.col[x]], T.gid[x]);
        else atomicMax(&T.rl[T.col[x]], T.gid[x]);
      }
      __syncthreads();
      for (uint32_t x = t; x < nn; x += kTailBlock) T.col[x] = (uint32_t)T.rl[T.col[x]];
      __syncthreads();
    }
  } else {
    for (;;) {
      ++rounds;
      // push: the largest colour along u -> w among valid nodes, to the fixpoint
      for (;;) {
        const uint32_t i = sweeps++;
        if (t == 0) T.flag[(i + 1) % 3u] = 0u;
        for (uint32_t e = t; e < ne; e += kTailBlock) {
          const uint32_t p = T.edge[e], u = p >> 16, w = p & 0xffffu;
          if (T.st[u] & T.st[w] & 1u) {
            const uint32_t cu = T.col[u];
            if (cu > T.col[w]) {
              atomicMax(&T.col[w], cu);
              T.flag[i % 3u] = 1u;
            }
          }
        }
        ++push_sweeps;
        if (!tail_sweep_end(T, i)) break;
        if (sweeps >= kTailSweeps) {
          settled = false;
          break;
        }
      }
      if (!settled) break;
      // roots, then their components: valid u with an edge u -> w into the component and
      // w's colour
      for (uint32_t x = t; x < nn; x += kTailBlock) {
        if (T.st[x] == 1u && T.col[x] == x) {
          T.st[x] = 3u;
          T.rl[x] = T.gid[x];
        }
      }
      __syncthreads();
      for (;;) {
        const uint32_t i = sweeps++;
        if (t == 0) T.flag[(i + 1) % 3u] = 0u;
        for (uint32_t e = t; e < ne; e += kTailBlock) {
          const uint32_t p = T.edge[e], u = p >> 16, w = p & 0xffffu;
          if (T.st[w] == 3u && T.st[u] == 1u && T.col[u] == T.col[w]) {
            T.st[u] = 3u;
            T.flag[i % 3u] = 1u;
          }
        }
        ++bw_sweeps;
        if (!tail_sweep_end(T, i)) break;
        if (sweeps >= kTailSweeps) {
          settled = false;
          break;
        }
      }
      if (!settled) break;
      for (uint32_t x = t; x < nn; x += kTailBlock) {
        if (T.st[x] == 3u) {
          if (canonical) atomicMin(&T.rl[T.col[x]], T.gid[x]);
          else atomicMax(&T.rl[T.col[x]], T.gid[x]);
        }
      }
      if (t == 0) T.count = 0u;
      __syncthreads();
      // the component found settles with its label; the rest start the next round
      bool left = false;
      for (uint32_t x = t; x < nn; x += kTailBlock) {
        const uint32_t s = T.st[x];
        if (s == 3u) {
          T.col[x] = (uint32_t)T.rl[T.col[x]];
          T.st[x] = 0u;
        } else if (s == 1u) {
          T.col[x] = x;
          left = true;
        }
      }
      if (left) T.count = 1u;
      __syncthreads();
      const bool more = T.count != 0u;
      __syncthreads();  // T.count is cleared again before the next round's read
      if (!more) break;
    }
  }
  if (!settled) {  // uniform: nothing was written outside LDS
    if (t == 0) rc->tail = 2u;
    return;
  }
  for (uint32_t x = t; x < nn; x += kTailBlock) {
    const int32_t lab = (int32_t)T.col[x];
    if (lab != T.gid[x]) labels[T.gid[x]] = lab;
  }
  if (t == 0) {
    rc->tail = 1u;
    rc->tail_nodes = nn;
    rc->rounds = rounds;
    rc->push_levels = push_sweeps;
    rc->cbw_levels = bw_sweeps;
    rc->listed0 = nv;
    rc->left = 0ull;
    rc->finished = 1u;
  }
}

// Pivot of create_csr (multistep/scc_main.cpp:169-183): largest (double)out*(double)in,
// first vertex on ties.  Pass 1: per-workgroup max of the product's bit pattern
// (non-negative doubles order like their bits), reduced by k_reduce; pass 2: smallest
// vertex attaining it.
__global__ void __launch_bounds__(kBlock) k_pivot_max(int32_t n, const uint32_t* __restrict__ oo,
                                                      const uint32_t* __restrict__ io, unsigned long long* part,
                                                      const uint32_t* __restrict__ mask) {
  __shared__ unsigned long long s_red[kBlock / 64];
  unsigned long long best = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
    if (mask && !bit_test(mask, (uint32_t)v)) continue;
    double p = (double)(oo[v + 1] - oo[v]) * (double)(io[v + 1] - io[v]);
    unsigned long long b = (unsigned long long)__double_as_longlong(p);
    best = b > best ? b : best;
  }
  best = wave_max64(best);
  if (lane() == 0) s_red[wave_id()] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kBlock / 64; ++i) best = s_red[i] > best ? s_red[i] : best;
    unsigned long long* p = part + 4 * blockIdx.x;
    p[0] = p[1] = p[2] = 0;
    p[3] = best;
  }
}
// red[3] holds the max; red[1] receives the first vertex attaining it (~0 when none).  With
// a mask, only the vertices it holds are candidates (both passes).
__global__ void k_pivot_first(int32_t n, const uint32_t* __restrict__ oo, const uint32_t* __restrict__ io,
                              unsigned long long* red, const uint32_t* __restrict__ mask) {
  const unsigned long long target = red[3];
  if (target == 0ull) return;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    if (mask && !bit_test(mask, (uint32_t)v)) continue;
    double p = (double)(oo[v + 1] - oo[v]) * (double)(io[v + 1] - io[v]);
    if ((unsigned long long)__double_as_longlong(p) == target) atomicMin(&red[1], (unsigned long long)v);
  }
}
// CC pivot (multistep/cc_main.cpp:148-160): largest out+in degree, first vertex on ties,
// as the max of (degree << 32 | ~v).
__global__ void __launch_bounds__(kBlock) k_pivot_cc(int32_t n, const uint32_t* __restrict__ oo,
                                                     const uint32_t* __restrict__ io, unsigned long long* part) {
  __shared__ unsigned long long s_red[kBlock / 64];
  unsigned long long best = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
    uint32_t d = (oo[v + 1] - oo[v]) + (io[v + 1] - io[v]);
    if (d == 0u) continue;
    unsigned long long key = ((unsigned long long)d << 32) | (unsigned long long)(0xffffffffu - (uint32_t)v);
    best = key > best ? key : best;
  }
  best = wave_max64(best);
  if (lane() == 0) s_red[wave_id()] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kBlock / 64; ++i) best = s_red[i] > best ? s_red[i] : best;
    unsigned long long* p = part + 4 * blockIdx.x;
    p[0] = p[1] = p[2] = 0;
    p[3] = best;
  }
}

// Input check: offsets start at 0, never decrease, end at m; ids in [0, n).
__global__ void k_validate(int32_t n, int64_t m, const uint32_t* __restrict__ off, const int32_t* __restrict__ adj,
                           uint32_t* err) {
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  bool bad = false;
  if (tid == 0) bad = off[0] != 0u || (int64_t)off[n] != m;
  for (int64_t v = tid; v < n; v += stride) bad |= off[v + 1] < off[v];
  for (int64_t e = tid; e < m; e += stride) bad |= (uint32_t)adj[e] >= (uint32_t)n;
  if (bad) atomicOr(err, 1u);
}

// ------------------------------------------------------------------------------------
// complete trim (multistep/scc_trim.cpp:141-297, TRIM_LEVEL 2)

// All vertices valid (bits past n clear), labels -1, fresh control block.
__global__ void k_trim2_init(LevelArgs a) {
  const int64_t nw = a.nwords, n = a.n_total;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = i * 32;
    a.valid[i] = lo + 32 <= n ? 0xffffffffu : (lo < n ? (1u << (uint32_t)(n - lo)) - 1u : 0u);
    a.inq0[i] = 0u;
    a.inq1[i] = 0u;
  }
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    a.labels[v] = -1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ctrl* c = a.ctrl;
    c->next_packed = 0ull;
    c->q_size = 0u;
    c->q_edges = 0u;
    c->cur = 0u;
    c->bottom_up = 0u;
    c->bu_wide = 0u;
    c->from_bits = 0u;
    c->switched_back = 0u;
    c->done = 0u;
    c->level = 0u;
    c->closes = 0u;
    c->found = 0;
    c->phase = 5u;
    c->abandoned = 0u;
    c->t_start = ~0ull;
  }
}

// One round: a vertex with no valid out-neighbour or no valid in-neighbour other than
// itself is labelled with its own id and leaves the valid set; one trimmed for lack of
// out-neighbours re-queues its in-neighbours (they lost an out-neighbour), one trimmed
// for lack of in-neighbours its out-neighbours (scc_trim.cpp:203-256).  Round 1 examines
// every vertex, later rounds the queue.  Removals only shrink neighbour sets, and whoever
// loses a needed neighbour is re-queued by the vertex that leaves, so the rounds end at the
// largest set in which every vertex keeps both kinds of neighbour, whatever the order.
// Re-queues are deduplicated by the worklist bitmaps and reserved with one atomic per wave.
__global__ void __launch_bounds__(kBlock) k_trim2(LevelArgs a) {
  Ctrl* c = a.ctrl;
  if (c->done) return;
  const bool first = c->level == 0u;
  const uint32_t cur = c->cur;
  const int64_t items = first ? (int64_t)a.n_total : (int64_t)c->q_size;
  const Frontier F = cur ? a.fr1 : a.fr0;
  const Frontier N = cur ? a.fr0 : a.fr1;
  uint32_t* inq_cur = cur ? a.inq1 : a.inq0;
  uint32_t* inq_nxt = cur ? a.inq0 : a.inq1;
  for (int64_t wb = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~63); wb < items;
       wb += (int64_t)gridDim.x * kBlock) {
    const int64_t i = wb + lane();
    uint32_t v = 0;
    bool act = false;
    if (i < items) {
      v = first ? (uint32_t)i : (uint32_t)F.q[i];
      if (!first) bit_clear(inq_cur, v);
      act = bit_test(a.valid, v);
    }
    bool has_out = false, has_in = false;
    if (act) {
      for (uint32_t j = a.down_off[v], e = a.down_off[v + 1]; j < e && !has_out; ++j) {
        uint32_t w = (uint32_t)a.down_adj[j];
        has_out = w != v && bit_test(a.valid, w);
      }
      for (uint32_t j = a.up_off[v], e = a.up_off[v + 1]; j < e && !has_in; ++j) {
        uint32_t w = (uint32_t)a.up_adj[j];
        has_in = w != v && bit_test(a.valid, w);
      }
    }
    const bool trim = act && !(has_out && has_in);
    if (trim) {
      a.labels[v] = (int32_t)v;
      bit_clear(a.valid, v);
    }
    uint32_t rb = 0, re = 0;
    const int32_t* adj = a.down_adj;
    if (trim && has_out != has_in) {
      if (!has_out) {
        rb = a.up_off[v];
        re = a.up_off[v + 1];
        adj = a.up_adj;
      } else {
        rb = a.down_off[v];
        re = a.down_off[v + 1];
      }
    }
    for (;;) {  // wave-uniform walk over the lists being re-queued
      const bool more = rb < re;
      if (__ballot(more) == 0ull) break;
      bool push = false;
      uint32_t w = 0;
      if (more) {
        w = (uint32_t)adj[rb++];
        push = w != v && bit_test(a.valid, w) && bit_set(inq_nxt, w);
      }
      const unsigned long long m = __ballot(push);
      if (m) {
        const int leader = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if (lane() == leader)
          base = (uint32_t)(atomicAdd(&c->next_packed, (unsigned long long)__popcll(m) << 32) >> 32);
        base = __shfl(base, leader, 64);
        if (push) N.q[base + (uint32_t)__popcll(m & lanes_below())] = (int32_t)w;
      }
    }
  }
}

// Closes a trim round: the re-queued vertices become the next round's worklist.
__global__ void k_trim2_close(Ctrl* c) {
  if (threadIdx.x != 0 || c->done) return;
  const uint32_t next = (uint32_t)(c->next_packed >> 32);
  c->level += 1;
  c->closes += 1;
  c->q_size = next;
  c->q_edges = 0u;
  c->cur ^= 1u;
  c->next_packed = 0ull;
  c->done = next == 0u ? 1u : 0u;
}

// ------------------------------------------------------------------------------------
// batches: many independent graphs packed into one CSR
//
// Graph g owns vertices [base[g], base[g+1]) (base[0] = 0, base[ng] = n, non-decreasing)
// and every edge stays inside its graph.  The whole path runs on the union: trim,
// colouring and canonicalisation are per-vertex or follow edges, so they never mix
// graphs; FW-BW and the CC search start from every graph's pivot at once, and because the
// graphs are disjoint, the union's reach sets are the per-graph reach sets side by side.

// Last g in [lo, hi) with base[g] <= v, given base[lo] <= v < base[hi] (an empty graph
// shares its base with the next one, and the later graph is the one holding v).
__device__ __forceinline__ uint32_t graph_of(const int64_t* __restrict__ base, uint32_t lo, uint32_t hi, int64_t v) {
  while (hi - lo > 1u) {
    uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Graph of vertex wbase + lane (clamped to n - 1) for a wave walking 64 consecutive ids:
// two searches per wave, and a per-lane one only when the wave straddles graphs.
__device__ __forceinline__ uint32_t wave_graph_of(const int64_t* __restrict__ base, uint32_t ng, int64_t wbase,
                                                  int64_t n) {
  const int64_t last = wbase + 63 < n ? wbase + 63 : n - 1;
  const uint32_t glo = graph_of(base, 0u, ng, wbase);
  const uint32_t ghi = graph_of(base, glo, ng, last);
  int64_t v = wbase + lane();
  if (v > last) v = last;
  return glo == ghi ? glo : graph_of(base, glo, ghi + 1u, v);
}

// The per-graph pivots walk the union the way k_localise does: each wave takes a contiguous
// run of 64-vertex chunks, finds the graph of its first vertex with one binary search over
// `base`, and every lane steps forward over the graph boundaries its vertices cross.  The
// wave folds its lanes' values per graph (graphs appear in increasing order across the lanes
// and the run) and issues one atomic per graph it leaves, not one per wave-sized slice: a
// search per 64 vertices and an atomic per slice made the CC pivots of the 8192-graph batch
// take 2.3 ms.  WaveRun holds the pending (graph, value) of the wave.
struct WaveRun {
  uint32_t g = 0xffffffffu;
  unsigned long long x = 0ull;
};
// Folds value x of every `active` lane (lane graph g) into the run; MAX: keep the largest,
// else the smallest; flush(graph, value) is called by lane 0 for every graph the run leaves.
template <bool MAX, typename Flush>
__device__ __forceinline__ void run_fold(WaveRun& run, bool active, uint32_t g, unsigned long long x, Flush flush) {
  for (;;) {
    const unsigned long long mk = __ballot(active);
    if (mk == 0ull) break;
    const uint32_t key = __shfl(g, __ffsll((long long)mk) - 1, 64);
    const bool mine = active && g == key;
    const unsigned long long y = MAX ? wave_max64(mine ? x : 0ull) : wave_min64(mine ? x : ~0ull);
    if (key != run.g) {
      if (run.g != 0xffffffffu && lane() == 0) flush(run.g, run.x);
      run.g = key;
      run.x = y;
    } else {
      run.x = MAX ? (y > run.x ? y : run.x) : (y < run.x ? y : run.x);
    }
    active = active && !mine;
  }
}

// The wave's run of chunks [c0, c1) and the graph of its first vertex; false: no run.
__device__ __forceinline__ bool wave_run(int32_t n, const int64_t* __restrict__ base, uint32_t ng, int64_t& c0,
                                         int64_t& c1, uint32_t& g) {
  const int64_t chunks = ((int64_t)n + 63) / 64;
  const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
  const int64_t per = (chunks + waves - 1) / waves;
  const int64_t wv = (int64_t)blockIdx.x * (kBlock / 64) + wave_id();
  c0 = wv * per;
  c1 = c0 + per < chunks ? c0 + per : chunks;
  if (c0 >= c1) return false;
  g = graph_of(base, 0u, ng, c0 * 64);
  return true;
}

// Per-graph pivot, pass 1: keys[g] = max over the graph of the product's bit pattern
// (create_csr's rule, multistep/scc_main.cpp:169-183, applied to each graph).
__global__ void __launch_bounds__(kBlock) k_bpivot_max(int32_t n, const uint32_t* __restrict__ oo,
                                                       const uint32_t* __restrict__ io,
                                                       const int64_t* __restrict__ base, uint32_t ng,
                                                       unsigned long long* __restrict__ keys) {
  int64_t c0, c1;
  uint32_t g;
  if (!wave_run(n, base, ng, c0, c1, g)) return;
  WaveRun run;
  auto flush = [&](uint32_t gg, unsigned long long x) { atomicMax(&keys[gg], x); };
  for (int64_t c = c0; c < c1; ++c) {
    const int64_t v = c * 64 + lane();
    const bool here = v < n;  // all lanes stay for the wave-wide fold
    unsigned long long p = 0ull;
    if (here) {
      while (base[g + 1] <= v) ++g;  // base[ng] = n > v ends it
      p = (unsigned long long)__double_as_longlong((double)(oo[v + 1] - oo[v]) * (double)(io[v + 1] - io[v]));
    }
    run_fold<true>(run, p != 0ull, g, p, flush);
  }
  if (run.g != 0xffffffffu && lane() == 0) flush(run.g, run.x);
}
// Pass 2: piv[g] (initially 0xffffffff, i.e. -1) = first vertex of g attaining keys[g] > 0.
__global__ void __launch_bounds__(kBlock) k_bpivot_first(int32_t n, const uint32_t* __restrict__ oo,
                                                         const uint32_t* __restrict__ io,
                                                         const int64_t* __restrict__ base, uint32_t ng,
                                                         const unsigned long long* __restrict__ keys,
                                                         uint32_t* __restrict__ piv) {
  int64_t c0, c1;
  uint32_t g;
  if (!wave_run(n, base, ng, c0, c1, g)) return;
  WaveRun run;
  auto flush = [&](uint32_t gg, unsigned long long x) { atomicMin(&piv[gg], (uint32_t)x); };
  unsigned long long kg = keys[g];
  for (int64_t c = c0; c < c1; ++c) {
    const int64_t v = c * 64 + lane();
    const bool here = v < n;
    bool hit = false;
    if (here) {
      bool moved = false;
      while (base[g + 1] <= v) {
        ++g;
        moved = true;
      }
      if (moved) kg = keys[g];
      const unsigned long long p =
          (unsigned long long)__double_as_longlong((double)(oo[v + 1] - oo[v]) * (double)(io[v + 1] - io[v]));
      hit = p != 0ull && p == kg;
    }
    run_fold<false>(run, hit, g, (unsigned long long)v, flush);
  }
  if (run.g != 0xffffffffu && lane() == 0) flush(run.g, run.x);
}
// CC pivot per graph (multistep/cc_main.cpp:148-160): max of (out+in degree << 32 | ~v).
__global__ void __launch_bounds__(kBlock) k_bpivot_cc(int32_t n, const uint32_t* __restrict__ oo,
                                                      const uint32_t* __restrict__ io,
                                                      const int64_t* __restrict__ base, uint32_t ng,
                                                      unsigned long long* __restrict__ keys) {
  int64_t c0, c1;
  uint32_t g;
  if (!wave_run(n, base, ng, c0, c1, g)) return;
  WaveRun run;
  auto flush = [&](uint32_t gg, unsigned long long x) { atomicMax(&keys[gg], x); };
  for (int64_t c = c0; c < c1; ++c) {
    const int64_t v = c * 64 + lane();
    const bool here = v < n;
    unsigned long long key = 0ull;
    if (here) {
      while (base[g + 1] <= v) ++g;
      const uint32_t d = (oo[v + 1] - oo[v]) + (io[v + 1] - io[v]);
      if (d) key = ((unsigned long long)d << 32) | (unsigned long long)(0xffffffffu - (uint32_t)v);
    }
    run_fold<true>(run, key != 0ull, g, key, flush);
  }
  if (run.g != 0xffffffffu && lane() == 0) flush(run.g, run.x);
}
__global__ void k_bpivot_cc_final(uint32_t ng, const unsigned long long* __restrict__ keys, int32_t* __restrict__ piv) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    unsigned long long k = keys[g];
    piv[g] = k ? (int32_t)(0xffffffffu - (uint32_t)(k & 0xffffffffull)) : -1;
  }
}

// Starts a search from every graph's pivot (piv[g] < 0: the graph has none).  Each root
// is claimed and queued, like the single root of k_search_init.  Needs k_ctrl_reset
// first, k_close_seed after.
template <bool MIRROR>
__global__ void __launch_bounds__(kBlock) k_seed_pivots(LevelArgs a, const int32_t* __restrict__ piv, uint32_t ng) {
  __shared__ Stage S;
  if (threadIdx.x == 0) S.cnt = 0u;
  __syncthreads();
  for (uint32_t g0 = blockIdx.x * kBlock; g0 < ng; g0 += gridDim.x * kBlock) {
    const uint32_t g = g0 + threadIdx.x;
    const int32_t p = g < ng ? piv[g] : -1;
    uint32_t deg = 0;
    if (p >= 0) {
      atomicAnd(&a.cand[(uint32_t)p >> 5], ~(1u << ((uint32_t)p & 31u)));
      atomicAnd(&a.seen[(uint32_t)p >> 5], ~(1u << ((uint32_t)p & 31u)));
      deg = list_len(a.down_off, MIRROR ? a.down_off2 : nullptr, (uint32_t)p);
    }
    stage_push(S, deg > 0u, (uint32_t)p, deg);  // a root without a list reaches nothing more
    stage_flush(S, a.fr1, a.ctrl, kStage - kBlock);
  }
  stage_flush(S, a.fr1, a.ctrl, 0u);
}

// Commit of the pivot's component (multistep/scc_fwbw.cpp:412-466, cc_bfs.cpp:217-271):
// every member gets the label and leaves the valid set, every other vertex gets its own id
// (a trimmed vertex's final label; colouring overwrites the ones still valid).  part col 0
// = members of the workgroup.  One vertex per thread, so the label stores of a wave are
// one contiguous 256-byte run and every label is written in full lines; the first lane of
// each word writes the word's new valid bits (no other writer).  The label is `label`, or
// the smallest member when min_red is given (0xffffffff - min_red[3], canonical labels), or
// in a batch the pivot of the vertex's own graph.  With a.rc, only in an admitted round.
template <bool VEC = false>
__global__ void __launch_bounds__(kBlock) k_commit(MemberSrc m, uint32_t* __restrict__ valid, int32_t n,
                                                   int32_t* __restrict__ labels, int32_t label,
                                                   const unsigned long long* __restrict__ min_red,
                                                   unsigned long long* __restrict__ part,
                                                   const RoundCtl* rc, const unsigned long long* __restrict__ max_red,
                                                   int32_t members_only, const unsigned long long* __restrict__ min_part,
                                                   int32_t min_np) {
  __shared__ unsigned long long s_red[kBlock / 64];
  if (rc && !rc->go) return;  // uniform: the searches have not finished
  const uint32_t keep = member_keep(m);
  if (min_red) label = (int32_t)(0xffffffffu - (uint32_t)min_red[3]);
  if (min_np > 0) {
    // k_min_member's rows, reduced by every workgroup (col 3: 0xffffffff - smallest member)
    unsigned long long mx = 0;
    for (int r = threadIdx.x; r < min_np; r += kBlock) mx = min_part[4 * r + 3] > mx ? min_part[4 * r + 3] : mx;
    mx = wave_max64(mx);
    if (lane() == 0) s_red[wave_id()] = mx;
    __syncthreads();
    for (int w = 0; w < kBlock / 64; ++w) mx = s_red[w] > mx ? s_red[w] : mx;
    __syncthreads();  // s_red is reused below
    label = (int32_t)(0xffffffffu - (uint32_t)mx);
  }
  if (max_red) label = (int32_t)(max_red[3] - 1ull);
  unsigned long long cnt = 0;
  // kU slices of kBlock consecutive vertices per iteration, with all their loads issued
  // before any store (one memory latency per iteration, not per slice)
  constexpr int kU = 4;
  if (VEC) {
    // four consecutive vertices per thread: one 16-byte label store (labels 16-byte aligned)
    constexpr int64_t kStep = (int64_t)kBlock * 4 * kU;
    for (int64_t wg = (int64_t)blockIdx.x * kStep; wg < n; wg += (int64_t)gridDim.x * kStep) {
      uint32_t vw[kU], w[kU];
#pragma unroll
      for (int r = 0; r < kU; ++r) {
        const int64_t wi = (wg + 4 * (r * kBlock + threadIdx.x)) >> 5;  // slack words past n
        vw[r] = valid[wi];
        w[r] = member_word(m, vw[r], wi, keep);
      }
#pragma unroll
      for (int r = 0; r < kU; ++r) {
        const int64_t v0 = wg + 4 * (r * kBlock + threadIdx.x);
        const uint32_t sh = (uint32_t)(v0 & 31), bits = (w[r] >> sh) & 0xfu;
        if (v0 + 3 < n) {
          int4 o;
          o.x = (bits & 1u) ? label : (int32_t)v0;
          o.y = (bits & 2u) ? label : (int32_t)(v0 + 1);
          o.z = (bits & 4u) ? label : (int32_t)(v0 + 2);
          o.w = (bits & 8u) ? label : (int32_t)(v0 + 3);
          *reinterpret_cast<int4*>(labels + v0) = o;
        } else {
          for (int k = 0; k < 4; ++k)
            if (v0 + k < n) labels[v0 + k] = ((bits >> k) & 1u) ? label : (int32_t)(v0 + k);
        }
        if (sh == 0 && v0 < n && w[r]) {  // the word's first thread writes its valid bits
          valid[v0 >> 5] = vw[r] & ~w[r];
          cnt += (uint32_t)__popc(w[r]);
        }
      }
    }
  }
  for (int64_t wg = (int64_t)blockIdx.x * kBlock * kU; !VEC && wg < n; wg += (int64_t)gridDim.x * kBlock * kU) {
    uint32_t vw[kU], w[kU];
#pragma unroll
    for (int r = 0; r < kU; ++r) {
      const int64_t wi = (wg + r * kBlock + threadIdx.x) >> 5;  // bitmaps carry slack words past n
      vw[r] = valid[wi];
      w[r] = member_word(m, vw[r], wi, keep);
    }
#pragma unroll
    for (int r = 0; r < kU; ++r) {
      const int64_t v = wg + r * kBlock + threadIdx.x;
      const int64_t wi = v >> 5;
      const bool in = v < n && ((w[r] >> (v & 31)) & 1u);
      if (v < n && (in || !members_only)) labels[v] = in ? label : (int32_t)v;
      if ((lane() & 31) == 0 && v < n && w[r]) {
        valid[wi] = vw[r] & ~w[r];
        cnt += (uint32_t)__popc(w[r]);
      }
    }
  }
  cnt = block_sum64(cnt, s_red);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = cnt;
    part[4 * blockIdx.x + 1] = part[4 * blockIdx.x + 2] = part[4 * blockIdx.x + 3] = 0;
  }
}

// Commit of a batch's pivot components (k_commit for every graph at once): each member of
// graph g (valid and in both of its searches' sets) gets g's pivot as its label and leaves
// the valid set; every other vertex gets its own id.  Each wave walks a contiguous run of
// 64-vertex chunks, four at a time with their bitmap loads in flight together, and finds
// the graph of its first vertex with one binary search over `base`; from there every lane
// steps forward over the graph boundaries its vertices cross, as k_localise does (a binary
// search per chunk made this pass 0.68 ms on the 8192-graph batch).  With cmin (canonical
// labels; cmin[v] = v beforehand) each graph's smallest member is folded into cmin[pivot]:
// the wave keeps a running minimum for the graph its lanes are in and issues one atomicMin
// when it moves on to the next graph or ends.  Part row col 0: the vertices committed by
// this workgroup.
__global__ void __launch_bounds__(kBlock) k_commit_batch(MemberSrc m, uint32_t* __restrict__ valid, int32_t n,
                                                         int32_t* __restrict__ labels,
                                                         const int32_t* __restrict__ piv,
                                                         const int64_t* __restrict__ base, uint32_t ng,
                                                         unsigned long long* __restrict__ part, const RoundCtl* rc,
                                                         int32_t* __restrict__ cmin) {
  __shared__ unsigned long long s_red[kBlock / 64];
  if (rc && !rc->go) return;  // uniform: the searches have not finished
  const uint32_t keep = member_keep(m);
  constexpr int kU = 4;
  const int64_t chunks = ((int64_t)n + 63) / 64;
  const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
  const int64_t per = (chunks + waves - 1) / waves;
  const int64_t wv = (int64_t)blockIdx.x * (kBlock / 64) + wave_id();
  const int64_t c0 = wv * per, c1 = c0 + per < chunks ? c0 + per : chunks;
  unsigned long long cnt = 0;
  if (c0 < c1) {  // per wave
    uint32_t g = graph_of(base, 0u, ng, c0 * 64);
    uint32_t run_g = 0xffffffffu, run_min = 0xffffffffu;  // the wave's pending minimum
    for (int64_t c = c0; c < c1; c += kU) {
      uint32_t vw[kU], w[kU];
#pragma unroll
      for (int r = 0; r < kU; ++r) {
        const int64_t wi = ((c + r) * 64 + lane()) >> 5;  // bitmaps carry slack words past n
        vw[r] = c + r < c1 ? valid[wi] : 0u;
        w[r] = c + r < c1 ? member_word(m, vw[r], wi, keep) : 0u;
      }
#pragma unroll
      for (int r = 0; r < kU; ++r) {
        if (c + r >= c1) break;  // uniform
        const int64_t v = (c + r) * 64 + lane();
        const bool here = v < n;  // all lanes stay for the wave-wide steps below
        if (here)
          while (base[g + 1] <= v) ++g;  // base[ng] = n > v ends it
        const bool in = here && ((w[r] >> (v & 31)) & 1u);
        if (here) labels[v] = in ? piv[g] : (int32_t)v;
        if (here && (lane() & 31) == 0 && w[r]) {
          valid[v >> 5] = vw[r] & ~w[r];
          cnt += (uint32_t)__popc(w[r]);
        }
        if (cmin) {
          // the members' graphs, in increasing order across the lanes and the run
          bool active = in;
          for (;;) {
            const unsigned long long mk = __ballot(active);
            if (mk == 0ull) break;
            const uint32_t key = __shfl(g, __ffsll((long long)mk) - 1, 64);
            const bool mine = active && g == key;
            const uint32_t mn = wave_min(mine ? (uint32_t)v : 0xffffffffu);
            if (key != run_g) {
              if (run_g != 0xffffffffu && lane() == 0) atomicMin(&cmin[piv[run_g]], (int32_t)run_min);
              run_g = key;
              run_min = mn;
            } else {
              run_min = min(run_min, mn);
            }
            active = active && !mine;
          }
        }
      }
    }
    if (cmin && run_g != 0xffffffffu && lane() == 0) atomicMin(&cmin[piv[run_g]], (int32_t)run_min);
  }
  cnt = block_sum64(cnt, s_red);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = cnt;
    part[4 * blockIdx.x + 1] = part[4 * blockIdx.x + 2] = part[4 * blockIdx.x + 3] = 0;
  }
}

// Final labels of a batch, once *gate != 0: labels[v] = mins[labels[v]] (canonical; mins
// null: raw) minus the base of v's graph (its own ids).  Each wave takes a contiguous run of
// 64-vertex chunks and finds the graph of its first vertex with one binary search over
// `base`; from there every lane steps forward over the graph boundaries its vertices cross
// (graphs are contiguous id ranges in order), so no vertex needs a search of its own.
__global__ void __launch_bounds__(kBlock) k_localise(int32_t n, int32_t* __restrict__ labels,
                                                     const int32_t* __restrict__ mins,
                                                     const int64_t* __restrict__ base, uint32_t ng,
                                                     const uint32_t* gate) {
  if (gate && !*gate) return;
  const int64_t chunks = ((int64_t)n + 63) / 64;
  const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
  const int64_t per = (chunks + waves - 1) / waves;
  const int64_t w = (int64_t)blockIdx.x * (kBlock / 64) + wave_id();
  const int64_t c0 = w * per, c1 = c0 + per < chunks ? c0 + per : chunks;
  if (c0 >= c1) return;
  uint32_t g = graph_of(base, 0u, ng, c0 * 64);
  for (int64_t c = c0; c < c1; ++c) {
    const int64_t v = c * 64 + lane();
    if (v >= n) break;
    while (base[g + 1] <= v) ++g;  // base[ng] = n > v ends it
    const int32_t l = labels[v];
    if (l >= 0) labels[v] = (int32_t)((int64_t)(mins ? mins[l] : l) - base[g]);
  }
}

// Batch input check: every edge of graph g ends inside [base[g], base[g+1]).
__global__ void __launch_bounds__(kBlock) k_validate_batch(int32_t n, const uint32_t* __restrict__ off,
                                                           const int32_t* __restrict__ adj,
                                                           const int64_t* __restrict__ base, uint32_t ng,
                                                           uint32_t* err) {
  bool bad = false;
  for (int64_t wb = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~63); wb < n; wb += (int64_t)gridDim.x * kBlock) {
    const int64_t v = wb + lane();
    const uint32_t g = wave_graph_of(base, ng, wb, n);
    if (v < n) {
      const int64_t lo = base[g], hi = base[g + 1];
      for (uint32_t j = off[v]; j < off[v + 1]; ++j) {
        int64_t w = adj[j];
        bad |= w < lo || w >= hi;
      }
    }
  }
  if (bad) atomicOr(err, 1u);
}

// Every launch goes through here: while the runtime replays a HIP graph captured from the
// same launches (msc_opts.replay), the launchers only rebuild the host-side state.
#define MSC_LAUNCH(L, ...)                            \
  do {                                                \
    if (!(L).dry) hipLaunchKernelGGL(__VA_ARGS__); \
  } while (0)

inline int grid_for(int64_t items, int64_t per_block, int cap) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

}  // namespace

// ------------------------------------------------------------------------------------
// launchers

void launch_fill_u32(const Launch& L, uint32_t* p, uint32_t value, int64_t count) {
  if (count <= 0) return;
  MSC_LAUNCH(L, k_fill_u32, dim3(grid_for(count, kBlock, L.grid * 4)), dim3(kBlock), 0, L.stream, p, value,
                     count);
}

int launch_trim(const Launch& L, bool mirror, int32_t n, const uint32_t* oo, const uint32_t* io, uint32_t* valid,
                uint32_t* cand1, uint32_t* cand2, uint32_t* seen1, uint32_t* seen2, int32_t* labels,
                unsigned long long* part) {
  int g = grid_for(n, kBlock * 8, L.grid * 2);
  if (mirror)
    MSC_LAUNCH(L, k_trim<true>, dim3(g), dim3(kBlock), 0, L.stream, n, oo, io, valid, cand1, cand2, seen1,
                       seen2, labels, part);
  else
    MSC_LAUNCH(L, k_trim<false>, dim3(g), dim3(kBlock), 0, L.stream, n, oo, io, valid, cand1, cand2, seen1,
                       seen2, labels, part);
  return g;
}

void launch_reduce(const Launch& L, const unsigned long long* part, int nparts, unsigned long long* out) {
  MSC_LAUNCH(L, k_reduce, dim3(1), dim3(kBlock), 0, L.stream, part, nparts, out);
}

void launch_trim2_init(const Launch& L, const LevelArgs& a) {
  MSC_LAUNCH(L, k_trim2_init, dim3(grid_for(a.n_total > a.nwords ? a.n_total : a.nwords, kBlock, L.grid * 4)),
                     dim3(kBlock), 0, L.stream, a);
}

void launch_search_init(const Launch& L, const LevelArgs& a, int32_t root, int32_t phase,
                        const unsigned long long* rpart, int rnp, unsigned long long* rout) {
  ReachArgs p{};
  p.s[0] = a;
  MSC_LAUNCH(L, k_search_init, dim3(1), dim3(kInitBlock), 0, L.stream, p, 1, root, phase, rpart, rnp, rout);
}

void launch_search_init_pair(const Launch& L, const LevelArgs& a0, const LevelArgs& a1, int32_t root,
                             const unsigned long long* rpart, int rnp, unsigned long long* rout) {
  ReachArgs p{};
  p.s[0] = a0;
  p.s[1] = a1;
  MSC_LAUNCH(L, k_search_init, dim3(1), dim3(kInitBlock), 0, L.stream, p, 2, root, 0, rpart, rnp, rout);
}

void launch_level(const Launch& L, int op, bool mirror, const LevelArgs& a) {
  dim3 g((unsigned)L.grid), b(kBlock);
  // the instantiations with the other balance variants (td_slots) only where a level may
  // use one: mirrored levels under kBalanceAuto always take the chunks (td_variant)
  const bool slots = a.balance != kBalanceGlobal && !(a.balance == kBalanceAuto && mirror);
  switch (op) {
    case OP_PUSH:
      if (slots) {
        if (mirror) MSC_LAUNCH(L, (k_level<OP_PUSH, true, true>), g, b, 0, L.stream, a);
        else MSC_LAUNCH(L, (k_level<OP_PUSH, false, true>), g, b, 0, L.stream, a);
      } else {
        if (mirror) MSC_LAUNCH(L, (k_level<OP_PUSH, true>), g, b, 0, L.stream, a);
        else MSC_LAUNCH(L, (k_level<OP_PUSH, false>), g, b, 0, L.stream, a);
      }
      break;
    case OP_CBW:
      if (slots) MSC_LAUNCH(L, (k_level<OP_CBW, false, true>), g, b, 0, L.stream, a);
      else MSC_LAUNCH(L, (k_level<OP_CBW, false>), g, b, 0, L.stream, a);
      break;
    case OP_TRIM2: MSC_LAUNCH(L, k_trim2, g, b, 0, L.stream, a); break;
    default: break;
  }
}

void launch_close(const Launch& L, int op, bool mirror, const LevelArgs& a) {
  dim3 g(1), b(kBlock);
  switch (op) {
    case OP_PUSH:
      if (mirror) MSC_LAUNCH(L, (k_close_color<OP_PUSH, true>), g, b, 0, L.stream, a, L.grid);
      else MSC_LAUNCH(L, (k_close_color<OP_PUSH, false>), g, b, 0, L.stream, a, L.grid);
      break;
    case OP_CBW: MSC_LAUNCH(L, (k_close_color<OP_CBW, false>), g, b, 0, L.stream, a, L.grid); break;
    case OP_TRIM2: MSC_LAUNCH(L, k_trim2_close, g, dim3(64), 0, L.stream, a.ctrl); break;
    default: break;
  }
}

void launch_reach(const Launch& L, bool mirror, const ReachArgs& p, int ns) {
  dim3 g((unsigned)L.grid), b(kBlock);
  // the other balance variants (td_slots), as in launch_level
  if (p.s[0].balance != kBalanceGlobal && !(p.s[0].balance == kBalanceAuto && mirror)) {
    if (ns == 2) MSC_LAUNCH(L, (k_reach<2, false, true>), g, b, 0, L.stream, p);
    else if (mirror) MSC_LAUNCH(L, (k_reach<1, true, true>), g, b, 0, L.stream, p);
    else MSC_LAUNCH(L, (k_reach<1, false, true>), g, b, 0, L.stream, p);
    return;
  }
  if (ns == 2) MSC_LAUNCH(L, (k_reach<2, false>), g, b, 0, L.stream, p);
  else if (mirror) MSC_LAUNCH(L, (k_reach<1, true>), g, b, 0, L.stream, p);
  else MSC_LAUNCH(L, (k_reach<1, false>), g, b, 0, L.stream, p);
}

void launch_close_reach(const Launch& L, const ReachArgs& p, int ns) {
  if (ns == 2) MSC_LAUNCH(L, k_close_reach<2>, dim3(1), dim3(kCloseBlock), 0, L.stream, p, L.grid);
  else MSC_LAUNCH(L, k_close_reach<1>, dim3(1), dim3(kCloseBlock), 0, L.stream, p, L.grid);
}

// Both pointers set: no member (ca = valid), one bitmap (cb = ca) or two.
MemberSrc member_src(const MemberSrc& m, const uint32_t* valid) {
  MemberSrc r = m;
  if (!r.ca) r.ca = valid;
  if (!r.cb) r.cb = r.ca;
  return r;
}

int launch_commit(const Launch& L, const MemberSrc& m, uint32_t* valid, int32_t n, int32_t* labels, int32_t label,
                  const unsigned long long* min_red, unsigned long long* part, const LevelArgs* seed,
                  int32_t* list, const unsigned long long* max_red, bool members_only,
                  const unsigned long long* min_part, int min_np) {
  LevelArgs a = seed ? *seed : LevelArgs{};
  // the seed runs as its own sparse sweep (k_seed_bits): fused into this pass it made
  // every iteration wait on the stage (s24: 51 us fused, 22 + 12 us apart)
  int g;
  if (((uintptr_t)labels & 15u) == 0 && !members_only) {
    g = grid_for(n, kBlock * 64, L.grid * 2);
    MSC_LAUNCH(L, k_commit<true>, dim3(g), dim3(kBlock), 0, L.stream, member_src(m, valid), valid, n, labels, label,
               min_red, part, a.rc, max_red, 0, min_part, min_np);
  } else {
    g = grid_for(n, kBlock * 16, L.grid * 2);
    MSC_LAUNCH(L, k_commit<false>, dim3(g), dim3(kBlock), 0, L.stream, member_src(m, valid), valid, n, labels,
               label, min_red, part, a.rc, max_red, members_only ? 1 : 0, min_part, min_np);
  }
  if (seed)
    MSC_LAUNCH(L, k_seed_bits, dim3(grid_for((n + 31) / 32, BitIterT<4>::kWords, L.grid)), dim3(kBlock), 0,
                       L.stream, a, list);
  return g;
}

void launch_seed_bits(const Launch& L, const LevelArgs& a, int32_t n, int32_t* list) {
  MSC_LAUNCH(L, k_seed_bits, dim3(grid_for((n + 31) / 32, BitIterT<4>::kWords, L.grid)), dim3(kBlock), 0,
                     L.stream, a, list);
}

int launch_commit_batch(const Launch& L, const MemberSrc& m, uint32_t* valid, int32_t n, int32_t* labels,
                        const int32_t* piv, const int64_t* base, uint32_t ng, unsigned long long* part,
                        const LevelArgs* seed, int32_t* list, int32_t* cmin) {
  int g = grid_for(n, kBlock * 16, L.grid * 2);
  LevelArgs a = seed ? *seed : LevelArgs{};
  if (cmin)  // cmin = identity, when the commit itself runs (the round's go)
    MSC_LAUNCH(L, k_iota, dim3(grid_for(n, kBlock, L.grid * 4)), dim3(kBlock), 0, L.stream, cmin, (int64_t)n,
               a.rc ? &a.rc->go : nullptr);
  MSC_LAUNCH(L, k_commit_batch, dim3(g), dim3(kBlock), 0, L.stream, member_src(m, valid), valid, n, labels, piv,
             base, ng, part, a.rc, cmin);
  if (seed)
    MSC_LAUNCH(L, k_seed_bits, dim3(grid_for((n + 31) / 32, BitIterT<4>::kWords, L.grid)), dim3(kBlock), 0,
                       L.stream, a, list);
  return g;
}

int launch_min_member(const Launch& L, const MemberSrc& m, const uint32_t* valid, int64_t nwords,
                      unsigned long long* part, const RoundCtl* rc, bool largest) {
  int g = grid_for(nwords, kBlock * 4, L.grid * 2);
  MSC_LAUNCH(L, k_min_member, dim3(g), dim3(kBlock), 0, L.stream, member_src(m, valid), valid, nwords, part,
                     rc, largest ? 1 : 0);
  return g;
}

void launch_cc_cmin(const Launch& L, const int32_t* list, const RoundCtl* rc, int64_t bound, const int32_t* labels,
                    int32_t* cmin) {
  if (bound <= 0) return;
  const dim3 g(grid_for(bound, kBlock, L.grid * 8));
  MSC_LAUNCH(L, k_cc_cmin_init, g, dim3(kBlock), 0, L.stream, list, rc, labels, cmin);
  MSC_LAUNCH(L, k_cc_cmin, g, dim3(kBlock), 0, L.stream, list, rc, labels, cmin);
}

void launch_canon_list(const Launch& L, const int32_t* list, const RoundCtl* rc, int64_t bound, int32_t* labels,
                       const int32_t* cmin) {
  if (bound <= 0) return;
  MSC_LAUNCH(L, k_canon_list, dim3(grid_for(bound, kBlock, L.grid * 8)), dim3(kBlock), 0, L.stream, list, rc,
                     labels, cmin);
}

void launch_round_begin(const Launch& L, RoundCtl* rc, Ctrl* push, Ctrl* cbw, const Ctrl* s0, const Ctrl* s1,
                        bool first, bool single, unsigned long long* prog, uint32_t epoch, bool fresh) {
  MSC_LAUNCH(L, k_round_begin, dim3(1), dim3(64), 0, L.stream, rc, push, cbw, s0, s1,
             (first ? 1 : 0) | (single ? 2 : 0) | (fresh ? 4 : 0), prog, epoch);
}

void launch_rc_set(const Launch& L, RoundCtl* rc, Ctrl* push, Ctrl* cbw, int mode) {
  MSC_LAUNCH(L, k_rc_set, dim3(1), dim3(64), 0, L.stream, rc, push, cbw, mode);
}

int launch_count_bits(const Launch& L, const uint32_t* bits, int64_t nwords, unsigned long long* part) {
  int g = grid_for(nwords, kBlock * 4, L.grid * 2);
  MSC_LAUNCH(L, k_count_bits, dim3(g), dim3(kBlock), 0, L.stream, bits, nwords, part);
  return g;
}

void launch_andnot(const Launch& L, uint32_t* out, uint32_t* out2, const uint32_t* a, const uint32_t* b,
                   int64_t words) {
  if (words <= 0) return;
  MSC_LAUNCH(L, k_andnot, dim3(grid_for(words, kBlock, L.grid * 4)), dim3(kBlock), 0, L.stream, out, out2, a,
                     b, words);
}

void launch_seed_end(const Launch& L, Ctrl* c, const RoundCtl* rc, uint32_t pull_min) {
  MSC_LAUNCH(L, k_close_seed, dim3(1), dim3(64), 0, L.stream, c, rc, pull_min);
}

void launch_seed_list(const Launch& L, const LevelArgs& a, const int32_t* in, int32_t* out, int64_t bound) {
  MSC_LAUNCH(L, k_seed_list, dim3(grid_for(bound, kBlock, L.grid)), dim3(kBlock), 0, L.stream, a, in, out);
  MSC_LAUNCH(L, k_close_seed, dim3(1), dim3(64), 0, L.stream, a.ctrl, a.rc, a.pull_div ? a.pull_min : 0u);
}

void launch_roots_list(const Launch& L, const LevelArgs& a, const int32_t* list, const Ctrl* gate, int64_t bound) {
  MSC_LAUNCH(L, k_roots_begin, dim3(1), dim3(64), 0, L.stream, const_cast<RoundCtl*>(a.rc), gate, a.ctrl);
  MSC_LAUNCH(L, k_roots_list, dim3(grid_for(bound, kBlock, L.grid)), dim3(kBlock), 0, L.stream, a, list, gate);
  MSC_LAUNCH(L, k_close_seed, dim3(1), dim3(64), 0, L.stream, a.ctrl, a.rc, 0u);
}

void launch_tail(const Launch& L, bool cc, const LevelArgs& a, RoundCtl* rc, const int32_t* list, int32_t* labels,
                 bool canonical, const unsigned long long* commit_part, int commit_np, unsigned long long* commit_red) {
  const TailSeed& ts = a.tail;
  const uint32_t pull_min = a.pull_div ? a.pull_min : 0u;
  if (cc)
    MSC_LAUNCH(L, k_tail_solve<true>, dim3(1), dim3(kTailBlock), 0, L.stream, rc, a.ctrl, list, ts.src, ts.dst,
               ts.slots, ts.cap, ts.maxv, labels, canonical ? 1 : 0, pull_min, commit_part, commit_np, commit_red);
  else
    MSC_LAUNCH(L, k_tail_solve<false>, dim3(1), dim3(kTailBlock), 0, L.stream, rc, a.ctrl, list, ts.src, ts.dst,
               ts.slots, ts.cap, ts.maxv, labels, canonical ? 1 : 0, pull_min, commit_part, commit_np, commit_red);
}

int launch_pivot_scc(const Launch& L, int32_t n, const uint32_t* oo, const uint32_t* io, unsigned long long* part,
                     const uint32_t* mask) {
  int g = grid_for(n, kBlock * 8, L.grid * 2);
  MSC_LAUNCH(L, k_pivot_max, dim3(g), dim3(kBlock), 0, L.stream, n, oo, io, part, mask);
  return g;
}

void launch_pivot_first(const Launch& L, int32_t n, const uint32_t* oo, const uint32_t* io, unsigned long long* red,
                        const uint32_t* mask) {
  MSC_LAUNCH(L, k_pivot_first, dim3(grid_for(n, kBlock, L.grid * 4)), dim3(kBlock), 0, L.stream, n, oo, io,
                     red, mask);
}

int launch_pivot_cc(const Launch& L, int32_t n, const uint32_t* oo, const uint32_t* io, unsigned long long* part) {
  int g = grid_for(n, kBlock * 8, L.grid * 2);
  MSC_LAUNCH(L, k_pivot_cc, dim3(g), dim3(kBlock), 0, L.stream, n, oo, io, part);
  return g;
}

void launch_batch_pivots(const Launch& L, bool cc, int32_t n, const uint32_t* oo, const uint32_t* io,
                         const int64_t* base, uint32_t ng, unsigned long long* keys, int32_t* piv) {
  launch_fill_u32(L, (uint32_t*)keys, 0u, 2 * (int64_t)ng);
  dim3 g(grid_for(n, kBlock * 8, L.grid * 2)), b(kBlock);
  if (cc) {
    MSC_LAUNCH(L, k_bpivot_cc, g, b, 0, L.stream, n, oo, io, base, ng, keys);
    MSC_LAUNCH(L, k_bpivot_cc_final, dim3(grid_for(ng, kBlock, L.grid)), b, 0, L.stream, ng, keys, piv);
  } else {
    launch_fill_u32(L, (uint32_t*)piv, 0xffffffffu, ng);
    MSC_LAUNCH(L, k_bpivot_max, g, b, 0, L.stream, n, oo, io, base, ng, keys);
    MSC_LAUNCH(L, k_bpivot_first, g, b, 0, L.stream, n, oo, io, base, ng, keys, (uint32_t*)piv);
  }
}

void launch_seed_pivots(const Launch& L, bool mirror, const LevelArgs& a, const int32_t* piv, uint32_t ng,
                        int32_t phase) {
  MSC_LAUNCH(L, k_ctrl_reset, dim3(1), dim3(64), 0, L.stream, a.ctrl, phase);
  dim3 g((unsigned)grid_for(ng, kBlock, L.grid));
  if (mirror) MSC_LAUNCH(L, k_seed_pivots<true>, g, dim3(kBlock), 0, L.stream, a, piv, ng);
  else MSC_LAUNCH(L, k_seed_pivots<false>, g, dim3(kBlock), 0, L.stream, a, piv, ng);
  MSC_LAUNCH(L, k_close_seed, dim3(1), dim3(64), 0, L.stream, a.ctrl, nullptr, 0u);
}

void launch_localise(const Launch& L, int32_t n, int32_t* labels, const int32_t* mins, const int64_t* base,
                     uint32_t ng, const uint32_t* gate) {
  if (n <= 0) return;
  MSC_LAUNCH(L, k_localise, dim3(grid_for(n, kBlock * 16, L.grid * 2)), dim3(kBlock), 0, L.stream, n, labels,
                     mins, base, ng, gate);
}

void launch_validate_batch(const Launch& L, int32_t n, const uint32_t* off, const int32_t* adj, const int64_t* base,
                           uint32_t ng, uint32_t* err) {
  if (n <= 0) return;
  MSC_LAUNCH(L, k_validate_batch, dim3(grid_for(n, kBlock, L.grid * 4)), dim3(kBlock), 0, L.stream, n, off,
                     adj, base, ng, err);
}

void launch_validate(const Launch& L, int32_t n, int64_t m, const uint32_t* off, const int32_t* adj, uint32_t* err) {
  MSC_LAUNCH(L, k_validate, dim3(grid_for(m > n ? m : n, kBlock, L.grid * 4)), dim3(kBlock), 0, L.stream, n, m,
                     off, adj, err);
}

}  // namespace msc
