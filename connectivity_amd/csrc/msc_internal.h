// Internal declarations shared by the gfx950 kernels (msc_kernels.hip, msc_mlp.hip) and
// the host runtime (msc_runtime.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/msc.h"

namespace msc {

constexpr int kBlock = 256;                    // threads per workgroup: 4 wave64s
constexpr uint32_t kChunk = 1024;              // edges per edge-balanced work item
constexpr int kPer = (int)(kChunk / kBlock);   // edges per thread per work item
constexpr uint32_t kStage = 2048;              // claims staged in LDS per workgroup
#ifndef MSC_BU_WPT
#define MSC_BU_WPT 1
#endif
constexpr int kWPT = MSC_BU_WPT;               // bitmap words per thread of a bottom-up word group
constexpr int kBW = kBlock * kWPT;             // bitmap words per bottom-up word group
#ifndef MSC_BU_WPT_WIDE
#define MSC_BU_WPT_WIDE 4
#endif
constexpr int kWPTWide = MSC_BU_WPT_WIDE;      // ... and of a sparse bottom-up level (few candidates)
constexpr int kBWide = kBlock * kWPTWide;
constexpr int kMaxParts = 8192;                // per-workgroup partial results
// workgroup-local continuation of colour levels: private frontiers of at most kLocalEdges
// edges, two per workgroup (ids, offsets, chunk starts)
#ifndef MSC_LOCAL_EDGES
#define MSC_LOCAL_EDGES 4096
#endif
constexpr uint32_t kLocalEdges = MSC_LOCAL_EDGES;
constexpr int kLocalHalf = (int)(2 * kStage + kLocalEdges / kChunk + 2);
constexpr int kLocalStride = 2 * kLocalHalf;
// colour tail (the seed of colouring round 0, then k_tail_solve): one workgroup settles
// colouring round 0's list when it has at most kTailV vertices (slots fit 16 bits), at most
// kTailE edges join two of them, those edges have at most kTailN endpoints, and the
// colouring of those takes at most kTailSweeps sweeps
constexpr uint32_t kTailV = 65536;
constexpr uint32_t kTailE = 8192;
constexpr uint32_t kTailN = 8192;
constexpr uint32_t kTailSweeps = 4096;
constexpr int kTailBlock = 1024;
constexpr int kMlpIn = 12, kMlpH = 32, kMlpOut = 3;
constexpr int kMlpParams = kMlpIn * kMlpH + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH * kMlpOut + kMlpOut;
// feature standardisation (mean, std) follows the weights
constexpr int kMlpBlob = kMlpParams + 2 * kMlpIn;

// What one level kernel does with each edge it walks.  Reach ops keep two bitmaps per
// search: cand = allow & ~visited ("claimable": a top-down edge costs a single probe, a
// claim is one atomicAnd that clears the bit, and the set bits of cand are the bottom-up
// candidates) and seen = ~visited (all ones at the start; a bottom-up probe of a
// neighbour is a single load of it, where allow & ~cand would take two).  A claim clears
// both: an extra atomic without result in a top-down level, an extra plain word store in
// a bottom-up level.
enum Op : int {
  OP_FW = 0,   // forward reach:  allow = valid, top-down on out-lists, bottom-up on in-lists
  OP_BW = 1,   // backward reach: allow = fw set, top-down on in-lists, bottom-up on out-lists
  OP_CC = 2,   // undirected reach over out+in lists
  OP_PUSH = 3, // max-label push (colour propagation) along down lists
  OP_CBW = 4,  // per-colour backward search: claim valid in-neighbours of equal colour
  OP_TRIM2 = 5,  // one round of the complete trim (down = out-lists, up = in-lists)
};

// Double-buffered frontier: queue of vertices, exclusive prefix of their list lengths,
// and the first queue slot of every kChunk-edge chunk of that prefix space.
struct Frontier {
  int32_t* q;
  uint32_t* qo;
  uint32_t* starts;
};

// Device-resident loop state.  Level kernels read it at entry; a one-workgroup closing
// kernel after each level reads the level's counters and prepares the next level, so the
// host never looks at it between levels.
struct Ctrl {
  unsigned long long next_packed;  // (vertices enqueued << 32) | their list lengths
  uint32_t q_size;                 // frontier to expand
  uint32_t q_edges;                // its list lengths summed
  uint32_t cur;                    // Frontier buffer holding it
  uint32_t bottom_up;              // direction of the coming level (colour push: 0 push,
                                   // 1 pull, 2 pull while colours are still the seed's ids)
  uint32_t switched_back;          // the one-way return to top-down happened
  uint32_t done;
  uint32_t level;                  // levels closed so far
  uint32_t nrec;                   // level records written
  int64_t found;                   // vertices discovered, root excluded
  uint32_t phase;                  // msc_level_rec.phase of this search
  uint32_t abandoned;              // the policy stopped the search; its result is void
  unsigned long long reduced[4];   // results of k_reduce
  unsigned long long bytes_op[6];  // algorithmic bytes of closed levels, per Op
  unsigned long long bu_levels;
  unsigned long long t_start;      // earliest workgroup start of the running launch
  uint32_t closes;                 // closing-kernel launches that closed a level
  uint32_t list_len;               // vertex list written by the last seeding launch (kept by resets)
  uint32_t round_edges;            // colour push: list lengths of the round's seeded frontier
  uint32_t bu_wide;                // reach: the coming bottom-up level walks kBWide-word groups
  uint32_t from_bits;              // reach: the coming top-down level expands the vertices the
                                   // bottom-up level before it claimed (LevelArgs::fresh), not a queue
  uint32_t within;                  // reach: claims only vertices the other search of the pivot
                                   // visited (LevelArgs::within_cand), set once that one finished
  // counts the host reads with the next control-block copy instead of a round trip each
  unsigned long long red_trim[4];    // col 0: trim survivors
  unsigned long long red_commit[4];  // col 0: vertices committed by FW-BW / the CC search
  unsigned long long red_left[4];    // col 0: valid vertices after the last colour round
  unsigned long long red_min[4];     // col 3: 0xffffffff - smallest member of the committed set
};

// Device state of the colouring rounds, which the host enqueues ahead of time: each
// round's first kernel (k_round_begin) admits the round only if the steps before it have
// finished, and every later step of the round checks `go`.  A step that finds its
// predecessor unfinished sets `stall`; nothing after it runs until the host has finished
// that step and cleared it.
struct RoundCtl {
  uint32_t go;        // the round admitted last may run its steps
  uint32_t stall;     // a step found its predecessor unfinished
  uint32_t finished;  // colouring complete: no valid vertex left
  uint32_t admitted;  // rounds admitted so far (the current one is admitted - 1)
  uint32_t roots_in;  // 1 + index of the last round whose roots ran (0: none)
  uint32_t error;     // a round removed nothing, or its push settled without a root
  uint32_t listed0;   // length of round 0's vertex list (the canonical pass walks it)
  uint32_t tail;      // colour tail: 0 not run, 1 settled the colouring, 2 declined it
  unsigned long long rounds;        // rounds completed
  unsigned long long push_levels;   // push levels of the completed rounds
  unsigned long long cbw_levels;    // colour-BW levels of the completed rounds
  unsigned long long listed_total;  // vertices listed by the completed rounds
  unsigned long long roots_total;   // roots of the completed rounds
  unsigned long long push_closes;   // closer launches of the last completed round's push
  unsigned long long cbw_closes;    // ... and of its colour-BW
  unsigned long long left;          // valid vertices after the last completed round
  uint32_t tail_edges;              // colour tail: edges joining two listed vertices
  uint32_t tail_nodes;              // ... and their endpoints
};

// What colouring round 0's seed collects for the colour tail, for the vertices in the first
// maxv slots of the round's list: slots[v] = v's slot, and every out-list entry u -> w whose
// target is still valid (listed as well) and not u itself, as src = slot(u), dst = w (cap
// entries; *count counts them all).
struct TailSeed {
  int32_t* slots;
  uint32_t* src;
  int32_t* dst;
  uint32_t* count;
  uint32_t cap, maxv;
  const uint32_t* off;
  const int32_t* adj;
  const uint32_t* valid;
};

struct LevelArgs {
  const uint32_t* down_off;   // top-down lists (first half for OP_CC / mirrored push)
  const int32_t* down_adj;
  const uint32_t* down_off2;  // second list for mirrored ops, else null
  const int32_t* down_adj2;
  const uint32_t* up_off;     // bottom-up lists
  const int32_t* up_adj;
  const uint32_t* up_off2;
  const int32_t* up_adj2;
  uint32_t* allow;            // bitmap: vertices the search may visit
  uint32_t* cand;             // bitmap: allowed and not yet visited (cleared on claim)
  uint32_t* seen;             // bitmap: not yet visited (cleared on claim; allowed or not)
  uint32_t* fresh;            // reach: bitmap of the vertices the last bottom-up level claimed
  uint32_t* valid;            // bitmap: vertices without a label yet
  uint32_t* inq0;             // colour worklist membership, buffer 0 / 1
  uint32_t* inq1;
  int32_t* colors;
  int32_t* labels;
  Frontier fr0, fr1;
  Ctrl* ctrl;
  unsigned long long* part;   // kMaxParts x 4 per-workgroup partials
  msc_level_rec* recs;
  uint32_t rec_cap;
  int32_t n_total;
  int64_t nwords;             // bitmap words
  const unsigned long long* nvalid;  // device: valid vertices when the search started (null: 0)
  int32_t beta;
  int32_t policy;
  double avg;
  double alpha;
  const float* mlp;           // kMlpBlob floats (learned policy)
  const int8_t* sched;        // forced actions, 3 x MSC_SCHED_LEVELS (null: none)
  float clock_mhz;            // wall_clock64() rate, for per-level durations
  uint32_t small_edges;       // colour levels over at most this many edges run inside the closer
  int32_t* cmin;              // canonical labels in flight: min member per colour root (null: off)
  int32_t op;                 // Op of this search or sweep (books its algorithmic bytes)
  int32_t* loc;               // colour levels: kLocalStride ints per workgroup for local rounds (null: none)
  int32_t local_rounds;       // colour levels: local continuation rounds per launch
  uint64_t local_ticks;       // colour levels: no local round starts this long (wall clock) after a launch began
  const RoundCtl* rc;         // colour seeds and roots: run only when rc->go (null: always)
  uint32_t pull_div;          // colour push: a level over at least 1/pull_div of the round's
                              // edges runs as a pull level (0: never)
  uint32_t pull_min;          // ... and over at least this many edges
  uint32_t group_words;       // bitmap words of a pull or (narrow) bottom-up level's word group (1..kBlock)
  TailSeed tail;              // colouring round 0's seed: what the colour tail needs (tail.slots null: off)
  int32_t balance;            // top-down load balance of grid-wide levels (msc_tuning.balance)
  const uint32_t* within_cand;  // backward search beside the forward one: the forward search's
                                // claimable bitmap; once ctrl->within is set, only vertices it no
                                // longer holds (the forward set) may be claimed (null: never)
};

// The reach searches one k_reach launch advances by a level: one (FW, BW or CC alone) or
// two (the forward and backward searches of one pivot, side by side).
struct ReachArgs {
  LevelArgs s[2];
};

struct Launch {
  hipStream_t stream;
  int grid;          // workgroups for level kernels
  bool dry = false;  // the launches are already in a captured graph that replays them: skip
};

// ---- context accessors for the policy code (msc_runtime.cpp) ----
int ctx_stream(msc_ctx* ctx, hipStream_t* stream, int* device);
int ctx_set_policy(msc_ctx* ctx, const float* blob, int nets);  // nets x kMlpBlob floats, validated by the caller

// ---- kernel launchers (msc_kernels.hip) ----
// Launchers that leave per-workgroup partials return how many workgroups wrote them.
void launch_fill_u32(const Launch& L, uint32_t* p, uint32_t value, int64_t count);
// valid bitmap (and, when given, the searches' cand1 / cand2 = valid, seen1 / seen2 = all
// ones), trimmed labels; part col 0 = survivors
int launch_trim(const Launch& L, bool mirror, int32_t n, const uint32_t* out_off, const uint32_t* in_off,
                uint32_t* valid, uint32_t* cand1, uint32_t* cand2, uint32_t* seen1, uint32_t* seen2,
                int32_t* labels, unsigned long long* part);
// out[i] = sums (cols 0..2) and max (col 3) of part[0..nparts)[i]
void launch_reduce(const Launch& L, const unsigned long long* part, int nparts, unsigned long long* out);
// (with rnp > 0, first reduces rnp rows of rpart into rout, as launch_reduce)
void launch_search_init(const Launch& L, const LevelArgs& a, int32_t root, int32_t phase,
                        const unsigned long long* rpart = nullptr, int rnp = 0, unsigned long long* rout = nullptr);
// both searches of a pivot at once: a0 gets phase 0 (FW), a1 phase 1 (BW)
void launch_search_init_pair(const Launch& L, const LevelArgs& a0, const LevelArgs& a1, int32_t root,
                             const unsigned long long* rpart = nullptr, int rnp = 0,
                             unsigned long long* rout = nullptr);
// complete trim: every vertex valid and unlabelled, control block reset (round 1 examines all)
void launch_trim2_init(const Launch& L, const LevelArgs& a);
// colour sweeps (OP_PUSH, OP_CBW) and complete-trim rounds (OP_TRIM2)
void launch_level(const Launch& L, int op, bool mirror, const LevelArgs& a);  // work only
void launch_close(const Launch& L, int op, bool mirror, const LevelArgs& a);  // one workgroup
// reach searches (OP_FW, OP_BW, OP_CC): p.s[0 .. ns) advance by one level together
void launch_reach(const Launch& L, bool mirror, const ReachArgs& p, int ns);  // work only
void launch_close_reach(const Launch& L, const ReachArgs& p, int ns);        // one workgroup
// Members of the pivot's component: valid & ~ca & ~cb, the valid vertices both searches
// visited (ca, cb: their claimable bitmaps; cb null: one search).  ca null: no member; none
// either when a search whose control block is named in g0 / g1 was abandoned.
struct MemberSrc {
  const uint32_t* ca;
  const uint32_t* cb;
  const Ctrl* g0;
  const Ctrl* g1;
};
// labels[v] = label and valid &= ~member for every member, labels[v] = v for every other
// vertex (with members_only: other labels untouched); part col 0 = count.  With min_red,
// the label is the smallest member found by launch_min_member (0xffffffff - min_red[3]);
// with max_red the largest (max_red[3] - 1).  With `seed` (the colouring's arguments),
// every vertex still valid afterwards is then seeded for colouring round 0 (colors[v] = v,
// queued in frontier 1 of seed->ctrl, written to `list`), and launch_seed_end closes the
// seed.  With seed->rc, both run only in an admitted round (after launch_round_begin).
// With min_part (min_np rows of launch_min_member, which must not be `part`), every
// workgroup reduces them itself instead of reading min_red.
int launch_commit(const Launch& L, const MemberSrc& m, uint32_t* valid, int32_t n, int32_t* labels, int32_t label,
                  const unsigned long long* min_red, unsigned long long* part, const LevelArgs* seed,
                  int32_t* list, const unsigned long long* max_red = nullptr, bool members_only = false,
                  const unsigned long long* min_part = nullptr, int min_np = 0);
// colouring round 0's seed alone (k_seed_bits), after commits made without one
void launch_seed_bits(const Launch& L, const LevelArgs& a, int32_t n, int32_t* list);
// part col 3 = 0xffffffff - smallest member in each workgroup's words (0: none), or with
// `largest` 1 + largest member; with rc, only in an admitted round
int launch_min_member(const Launch& L, const MemberSrc& m, const uint32_t* valid, int64_t nwords,
                      unsigned long long* part, const RoundCtl* rc, bool largest = false);
// CC: the smallest member of every colouring component into cmin (list: colouring round 0's)
void launch_cc_cmin(const Launch& L, const int32_t* list, const RoundCtl* rc, int64_t bound, const int32_t* labels,
                    int32_t* cmin);
// once rc->finished: labels[v] = cmin[labels[v]] for the rc->listed0 vertices of `list`
// (bound >= that count sizes the grid)
void launch_canon_list(const Launch& L, const int32_t* list, const RoundCtl* rc, int64_t bound, int32_t* labels,
                       const int32_t* cmin);
// the first kernel of a colouring round (k_round_begin); s0 / s1: the searches' control
// blocks that round 0 (`first`) waits for (null: none); `single`: CC's one fixpoint, whose
// closing check ends the colouring; with `prog` (pinned host memory) the outcome is also
// published there, tagged with `epoch`; `fresh`: the round state is reset first (as
// launch_rc_set mode 0)
void launch_round_begin(const Launch& L, RoundCtl* rc, Ctrl* push, Ctrl* cbw, const Ctrl* s0, const Ctrl* s1,
                        bool first, bool single, unsigned long long* prog = nullptr, uint32_t epoch = 0,
                        bool fresh = false);
// host-directed round state: 0 fresh, 1 resume the admitted round, 2 clear a stall
void launch_rc_set(const Launch& L, RoundCtl* rc, Ctrl* push, Ctrl* cbw, int mode);
int launch_count_bits(const Launch& L, const uint32_t* bits, int64_t nwords, unsigned long long* part);
// frontier hand-over after a seeding kernel (with rc: only in an admitted round)
void launch_seed_end(const Launch& L, Ctrl* c, const RoundCtl* rc, uint32_t pull_min);
// colouring round > 0, after launch_round_begin: seed from the previous round's list `in`
// (length a.ctrl->list_len) into `out`, hand-over; bound >= that length sizes the grid
void launch_seed_list(const Launch& L, const LevelArgs& a, const int32_t* in, int32_t* out, int64_t bound);
// colour roots of this round's list (length gate->list_len) into a.ctrl's frontier, once
// the push fixpoint of `gate` is done (else the round stalls); reset and hand-over included
void launch_roots_list(const Launch& L, const LevelArgs& a, const int32_t* list, const Ctrl* gate, int64_t bound);
// out[w] = a[w] & ~b[w] (b may be null: copy), and when out2 is given, out2[w] = out[w]
void launch_andnot(const Launch& L, uint32_t* out, uint32_t* out2, const uint32_t* a, const uint32_t* b,
                   int64_t words);
// Colour tail, after colouring round 0's seed (which collected a.tail) with no seed
// hand-over and no reduce of the commit's count: k_tail_solve does both (the count from
// commit_np rows of commit_part into commit_red), then colours the endpoints of the joining
// edges in one workgroup and writes the labels of every component with more than one vertex
// (largest id; smallest with `canonical`), or declines (rc->tail = 2)
void launch_tail(const Launch& L, bool cc, const LevelArgs& a, RoundCtl* rc, const int32_t* list, int32_t* labels,
                 bool canonical, const unsigned long long* commit_part, int commit_np, unsigned long long* commit_red);
// pivot: part col 3 = max key; results via launch_reduce then launch_pivot_first; with a
// mask, among the vertices it holds only
int launch_pivot_scc(const Launch& L, int32_t n, const uint32_t* out_off, const uint32_t* in_off,
                     unsigned long long* part, const uint32_t* mask = nullptr);
void launch_pivot_first(const Launch& L, int32_t n, const uint32_t* out_off, const uint32_t* in_off,
                        unsigned long long* red, const uint32_t* mask = nullptr);
int launch_pivot_cc(const Launch& L, int32_t n, const uint32_t* out_off, const uint32_t* in_off,
                    unsigned long long* part);
void launch_validate(const Launch& L, int32_t n, int64_t m, const uint32_t* off, const int32_t* adj, uint32_t* err);
// ---- device ingest (msc_ingest.hip) ----
// read_edge + create_csr on the device from the edge-list text in host memory; fills *out
// with device arrays (hipMalloc) owned by the caller.
int ingest_text(hipStream_t stream, const char* text, int64_t bytes, msc_csr* out);

// ---- batches: graph g owns vertices [base[g], base[g+1]) of a packed union CSR ----
// piv[g] = per-graph pivot (union id, -1 = none); keys: ng u64 scratch
void launch_batch_pivots(const Launch& L, bool cc, int32_t n, const uint32_t* oo, const uint32_t* io,
                         const int64_t* base, uint32_t ng, unsigned long long* keys, int32_t* piv);
// fresh control block, every pivot claimed and queued (frontier buffer 1)
void launch_seed_pivots(const Launch& L, bool mirror, const LevelArgs& a, const int32_t* piv, uint32_t ng,
                        int32_t phase);
// launch_commit with labels[v] = piv[graph of v]
int launch_commit_batch(const Launch& L, const MemberSrc& m, uint32_t* valid, int32_t n, int32_t* labels,
                        const int32_t* piv, const int64_t* base, uint32_t ng, unsigned long long* part,
                        const LevelArgs* seed, int32_t* list, int32_t* cmin);
// labels[v] = (mins ? mins[labels[v]] : labels[v]) - base of v's graph, once *gate != 0
void launch_localise(const Launch& L, int32_t n, int32_t* labels, const int32_t* mins, const int64_t* base,
                     uint32_t ng, const uint32_t* gate);
// run only after launch_validate passed (reads adjacency through the offsets)
void launch_validate_batch(const Launch& L, int32_t n, const uint32_t* off, const int32_t* adj, const int64_t* base,
                           uint32_t ng, uint32_t* err);

}  // namespace msc
