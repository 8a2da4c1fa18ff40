/*
 * msc — Multistep strongly / weakly connected components on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libmsc.so.  Plain pointers and sizes only; every entry point returns an int
 * status (MSC_OK or a negative MSC_E_* code) and never throws across the boundary.
 *
 * The reference this replaces has no FFI: its path sits behind plain C++ functions that
 * main() calls directly.  Each declaration below names the reference function it stands
 * in for (file:line in HPCGraphAnalysis/Connectivity).
 *
 * Graph layout (same as the reference's `graph` struct, multistep/scc_main.cpp:75-86):
 * n vertices with int32 ids, CSR offsets uint32[n+1] (`unsigned* out_degree_list`),
 * adjacency int32[m] (`int* out_array`), and the transposed CSR (`in_degree_list`,
 * `in_array`).  m must be below 2^32 (2^31 for msc_cc_run, which walks out+in lists).
 */
#ifndef MSC_H
#define MSC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define MSC_OK 0
#define MSC_E_INVALID (-1)  /* bad argument or malformed graph */
#define MSC_E_OOM (-2)      /* host or device allocation failed */
#define MSC_E_DEVICE (-3)   /* HIP runtime error or no usable gfx950 device */
#define MSC_E_CAPACITY (-4) /* a fixed-capacity buffer would overflow */
#define MSC_E_IO (-5)       /* file could not be read or written */

/* ---- enums ----------------------------------------------------------------------- */
enum msc_where { MSC_HOST = 0, MSC_DEVICE = 1 };
enum msc_policy {
  MSC_POLICY_PLAIN = 0,   /* the reference's ALPHA/BETA switch rule */
  MSC_POLICY_LEARNED = 1  /* the MLP loaded with msc_policy_set_weights */
};
enum msc_label_mode {
  MSC_LABELS_CANONICAL = 0, /* each component labelled by its minimum vertex id */
  MSC_LABELS_RAW = 1        /* the reference's per-phase labels (see msc_scc_run) */
};

/* ---- graph ----------------------------------------------------------------------- */
/* Reference: `graph` (multistep/scc_main.cpp:75-86); Kokkos views n/out_degree_list/
 * out_array/in_degree_list/in_array (kokkos_scc/scc_run.cpp:48-52). */
typedef struct msc_csr {
  int32_t n;
  int64_t m;
  const uint32_t* out_off; /* n+1 */
  const int32_t* out_adj;  /* m   */
  const uint32_t* in_off;  /* n+1 */
  const int32_t* in_adj;   /* m   */
  int32_t where;           /* MSC_HOST or MSC_DEVICE: where all four arrays live */
} msc_csr;

/* Host graph owned by the library (built by the generators and the edge-list reader). */
typedef struct msc_host_graph {
  int32_t n;
  int64_t m;
  uint32_t* out_off;
  int32_t* out_adj;
  uint32_t* in_off;
  int32_t* in_adj;
  int32_t max_deg_vert; /* create_csr's pivot (multistep/scc_main.cpp:169-183), -1 if none */
  double avg_degree;    /* m / n */
} msc_host_graph;

/* ---- options and statistics ------------------------------------------------------ */
typedef struct msc_ctx msc_ctx; /* per-device, per-stream workspace; see msc_ctx_create */

typedef struct msc_opts {
  double alpha;         /* ALPHA = 15.0 (multistep/scc_main.cpp:64) */
  int32_t beta;         /* BETA  = 25   (multistep/scc_main.cpp:65) */
  int32_t trim_level;   /* TRIM_LEVEL = 1 (multistep/scc_main.cpp:61): scc_trim; 2: the complete
                           trim scc_trim_complete (scc_trim.cpp:141-297), SCC only */
  int32_t policy;       /* enum msc_policy */
  int32_t label_mode;   /* enum msc_label_mode */
  int32_t labels_where; /* MSC_HOST or MSC_DEVICE: where the `labels` output lives */
  int32_t device;       /* HIP device ordinal (used when ctx is NULL) */
  int32_t validate;     /* nonzero: check offsets and ids on device before running */
  int32_t timing;       /* nonzero: per-kernel HIP-event timing into msc_stats.k[] */
  int32_t record;       /* nonzero: keep per-level records (msc_ctx_levels) */
  msc_ctx* ctx;         /* NULL: the library's default context for `device` */
  /* Forced decisions (policy training): schedule[phase * MSC_SCHED_LEVELS + (L - 1)] is the
   * action taken after level L of search `phase` (0 FW, 1 BW, 2 CC): 0 top-down, 1
   * bottom-up, 2 abandon the search, -1 the policy's own choice.  NULL: none. */
  const int8_t* schedule;
  int32_t schedule_len;
  /* FW-BW rounds before colouring (no counterpart in the reference, which runs one):
   * while the last round's SCC held at least 1/16 of the vertices left before it, up to
   * this many rounds each take a new pivot (create_csr's rule among the vertices left) and
   * commit its SCC labelled with its largest member (its smallest with canonical labels),
   * the label colouring would give it, so every label is unchanged.  0 or 1: one round.
   * SCC of one graph with concurrent_fwbw only; ignored elsewhere. */
  int32_t fwbw_rounds;
  /* Nonzero: the context may capture the call's launches in a HIP graph and replay that graph
   * on later calls with the same arguments, tuning and workspace (no reference counterpart).
   * Applies to one graph in device memory with a given pivot (max_deg_vert >= 0), labels in
   * device memory, trim_level 1, one FW-BW round, searches side by side (tuning
   * concurrent_fwbw = 1, SCC) and no timing, record, schedule or validate; other calls run
   * as usual.  Labels and statistics are those of an ordinary call. */
  int32_t replay;
} msc_opts;
#define MSC_SCHED_LEVELS 64

/* Kernel classes timed when msc_opts.timing is set. */
enum msc_kernel {
  MSC_K_TRIM = 0,   /* degree trim + survivor compaction */
  MSC_K_LEVEL,      /* one BFS level (top-down or bottom-up), FW/BW/CC */
  MSC_K_COMMIT,     /* FW-BW / CC-BFS label commit + compaction */
  MSC_K_COLOR_PROP, /* one max-label push sweep */
  MSC_K_COLOR_BW,   /* one level of the per-colour backward search */
  MSC_K_ROOTS,      /* colour roots + compaction */
  MSC_K_OTHER,      /* init, pivot, canonicalisation, small fills */
  MSC_K_COUNT
};

typedef struct msc_kstat {
  int64_t launches;
  double ms;    /* summed HIP-event durations */
  double bytes; /* summed algorithmic bytes: each kernel class books the bytes its work units
                   must read and write (offsets, list entries, probes, claims), not cache lines */
} msc_kstat;

typedef struct msc_stats {
  double wall_ms;      /* host wall time of the call, stream-synchronised */
  int32_t pivot;       /* pivot actually used (-1: FW-BW skipped) */
  double avg_degree;   /* average degree actually used */
  int32_t num_trim;    /* vertices removed by trim */
  int32_t num_fwbw;    /* vertices in the pivot's component */
  int32_t num_color;   /* components found by colouring */
  int32_t color_rounds;
  int32_t color_iters; /* push sweeps summed over rounds */
  int32_t fw_levels, bw_levels, bu_levels, cbw_levels;
  int32_t host_syncs;  /* host round trips taken by the level loops */
  double bytes;        /* algorithmic bytes over all kernels */
  msc_kstat k[MSC_K_COUNT];
  int32_t color_tail;  /* 1: one workgroup settled the colouring (msc_tuning.tail_vertices);
                          color_rounds / color_iters / cbw_levels then count its sweeps */
} msc_stats;

/* Defaults equal the reference's compile-time settings. */
void msc_opts_default(msc_opts* o);

/* ---- context --------------------------------------------------------------------- */
/* A context owns a HIP stream on `device` (or borrows `stream`, a hipStream_t, when not
 * NULL), the device workspace (grown on demand, reused across calls) and the policy
 * weights.  Calls on one context are serialised by the caller; distinct contexts may be
 * used from distinct threads. */
int msc_ctx_create(int32_t device, void* stream, msc_ctx** out);
int msc_ctx_destroy(msc_ctx* ctx);
/* Per-level records of the last run with opts.record set (see msc_level_rec). */
typedef struct msc_level_rec {
  int32_t phase;      /* 0 FW, 1 BW, 2 CC */
  int32_t level;      /* 1-based */
  int32_t bottom_up;  /* direction this level ran in */
  int32_t next_size;  /* vertices discovered */
  uint32_t q_size;    /* frontier expanded (top-down) */
  uint32_t q_edges;   /* its list lengths summed */
  uint32_t next_edges;/* list lengths of the discovered vertices */
  uint32_t examined;  /* bottom-up candidates examined */
  uint64_t inspected; /* bottom-up edges inspected */
  int64_t found;      /* discovered so far, root excluded */
  int32_t n_total;
  int32_t num_valid;
  int32_t action;     /* direction chosen for the next level (0 TD, 1 BU, 2 stop) */
  int32_t pad;
  float feat[12];     /* the learned policy's raw input features at this decision */
  float level_us;     /* duration of this level's launch (opts.timing), else 0 */
  int32_t pad2;
} msc_level_rec;
int msc_ctx_levels(msc_ctx* ctx, msc_level_rec* out, int32_t cap, int32_t* count);

/* ---- learned direction / phase policy ------------------------------------------- */
/* No reference counterpart (the reference's switch is the fixed ALPHA/BETA rule).  A
 * 12 -> 32 -> 32 -> 3 ReLU MLP in FP32 picks, after each BFS level, top-down, bottom-up or
 * abandoning FW-BW for colouring; it changes only the schedule, never the labels.
 * Blob layout (MSC_MLP_BLOB floats): W1[32][12] b1[32] W2[32][32] b2[32] W3[3][32] b3[3],
 * then the input standardisation: mean[12], std[12]. */
#define MSC_MLP_IN 12
#define MSC_MLP_HIDDEN 32
#define MSC_MLP_OUT 3
#define MSC_MLP_PARAMS 1571
#define MSC_MLP_BLOB 1595
/* count = MSC_MLP_BLOB: one network for every search; count = 2 * MSC_MLP_BLOB: the first
 * blob decides the SCC searches (FW, BW) and the second the CC search, whose mirrored
 * sweeps trade off differently. */
int msc_policy_set_weights(msc_ctx* ctx, const float* blob, int32_t count);
/* Batched inference on device: logits_out[count][3] from raw features x[count][12]. */
int msc_mlp_forward(msc_ctx* ctx, const float* blob, const float* x, int64_t count, float* logits_out);
/* Gradient of sum_b w[b] * CE(softmax(f(x_b)), y_b) w.r.t. the MSC_MLP_PARAMS weights, on
 * device (the training kernels' own forward and backward); loss_out receives the sum. */
int msc_mlp_grad(msc_ctx* ctx, const float* blob, const float* x, const int32_t* y, const float* w,
                 int64_t count, float* loss_out, float* grad_out);
typedef struct msc_train_opts {
  int32_t epochs;
  int32_t batch;
  float lr;
  float beta1, beta2, eps; /* Adam */
  float weight_decay;
  uint64_t seed;           /* mini-batch shuffling */
} msc_train_opts;
/* Trains the blob's weights in place on device with Adam on the weighted cross-entropy
 * (mean over each mini-batch); mean/std in the blob are inputs, not trained.
 * loss_curve_out (may be NULL) receives the mean training loss of each epoch. */
int msc_mlp_train(msc_ctx* ctx, float* blob, const float* x, const int32_t* y, const float* w,
                  int64_t count, const msc_train_opts* opts, float* loss_curve_out);

/* Launch and schedule tunables of a context (no reference counterpart; picked by the
 * autotuner, connectivity_amd/autotune.py).  Defaults are the values the library ships. */
typedef struct msc_tuning {
  int32_t blocks_per_cu;    /* level-kernel workgroups per compute unit (1..8) */
  int32_t small_push_edges; /* colour push sweeps over at most this many frontier edges run
                               inside the one-workgroup closing kernel, several per launch */
  int32_t small_cbw_edges;  /* same for the per-colour backward search (0: only empty ones) */
  int32_t first_batch;      /* levels enqueued before the first host check of a loop */
  int32_t max_batch;        /* levels enqueued between later host checks */
  int32_t concurrent_fwbw;  /* 1: the backward search advances beside the forward one, both in
                               one launch per level, over the valid set (SCC = FW set & BW set);
                               0: after it, restricted to the FW set, the reference's order;
                               2 (the default): as 1 for one graph, but once the forward search
                               has finished the backward one claims only vertices of the FW set;
                               as 0 for a batch */
  int32_t search_batch;     /* levels of a FW-BW / CC search enqueued before its first host
                               check (colour loops start with first_batch) */
  int32_t local_rounds;     /* levels a workgroup of a colour launch continues on its own claims
                               before handing them to the next launch (0..256; 0: off) */
  int32_t local_us;         /* ... while the launch has run for less than this many us (1..100000) */
  int32_t pull_div;         /* a colour push level whose frontier holds at least 1/pull_div of
                               the round's list entries runs as a pull level: every valid vertex
                               takes the largest colour of its valid in-neighbours, no atomics
                               (0..1024; 0: push only) */
  int32_t tail_vertices;    /* colouring round 0 with at most this many vertices is settled by
                               one workgroup (joining edges coloured in LDS) instead of rounds of
                               level launches, when its joining edges fit (0..65536; 0: never) */
  int32_t balance;          /* top-down load balance of the grid-wide search and colour levels,
                               the Kokkos driver's variants (kokkos_scc/scc_fwbw.cpp:337-364,
                               scc_color.cpp:579-655): 0 a thread per frontier vertex (fwbw_baseline),
                               1 a workgroup deals the edges of its 256 frontier vertices
                               (fwbw_manhattan_local), 2 fixed-size chunks of the whole
                               frontier's edges (fwbw_manhattan_global), 3 per level: 1 while
                               the frontier's lists average at most 4 entries, else 2; WCC
                               levels, which walk two lists per vertex, 2 (the default).
                               Labels do not depend on it. */
  int32_t group_words;      /* bitmap words (32 vertices each) a workgroup of a dense bitmap sweep
                               (a pull level, a bottom-up search level) takes at a time (1..256),
                               or 0 (the default): 64 for a batch, whose probes stay inside each
                               graph, so narrower groups keep an XCD's working set in its L2,
                               and 256 for one graph */
} msc_tuning;
int msc_ctx_get_tuning(msc_ctx* ctx, msc_tuning* out);
int msc_ctx_set_tuning(msc_ctx* ctx, const msc_tuning* t);

/* ---- the hot path ---------------------------------------------------------------- */
/* Strongly connected components.
 * Replaces run_scc(graph&, int*& scc_maps, int max_deg_vert, double avg_degree,
 * int vert_cutoff) (multistep/scc_main.cpp:210) and Kokkos do_run<ExecSpace>(...)
 * (kokkos_scc/scc_run.cpp:48-59).  Phases: trim (scc_trim.cpp:89-139) -> FW-BW from the
 * pivot (scc_fwbw.cpp:469-513) -> colouring until nothing is left (the Kokkos loop,
 * kokkos_scc/scc_color.cpp:553; the reference's CPU serial tail is not used).
 *   max_deg_vert: pivot; -1 computes create_csr's rule on device.
 *   avg_degree:   < 0 computes m/n.
 *   labels:       n int32 written by the call (caller-owned; host or device per
 *                 opts.labels_where).  MSC_LABELS_RAW gives the reference's raw labels
 *                 except for its order-dependent Tarjan tail: trimmed vertex -> itself,
 *                 pivot's SCC -> pivot, every other SCC -> its maximum vertex id.
 *   stats:        may be NULL.
 * Graph arrays may be on host (copied in) or device (used in place). */
int msc_scc_run(const msc_csr* g, int32_t max_deg_vert, double avg_degree, int32_t* labels,
                const msc_opts* opts, msc_stats* stats);

/* msc_scc_run without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, labels and statuses as msc_scc_run;
 * every call uses one static device array of the library (136 MiB per device, held until
 * the process ends) as its workspace instead of the context's, so it calls no hipMalloc,
 * hipFree or other allocating routine.  Requirements (else MSC_E_INVALID):
 * opts->ctx is a context made beforehand with msc_ctx_create (its creation is the set-up
 * that allocates), the graph and labels are device-resident (g->where and
 * opts->labels_where MSC_DEVICE), opts->timing = opts->record = 0.  Limits (else
 * MSC_E_CAPACITY): n <= 1,048,576 vertices and m <= 16,777,216 edges.  All such calls on
 * one device share that array, which keeps the last call's contents; each call starts by
 * clearing it, and calls that overlap on one device race. */
int msc_scc_run_noalloc(const msc_csr* g, int32_t max_deg_vert, double avg_degree, int32_t* labels,
                        const msc_opts* opts, msc_stats* stats);

/* Weakly connected components of a digraph (connected components of its mirror).
 * Replaces run_cc(graph&, int*& cc_maps, int max_deg_vert, double avg_degree)
 * (multistep/cc_main.cpp:184) over the mirrored edge list read_edge builds
 * (cc_main.cpp:102-118); out and in lists are walked together instead of being mirrored.
 *   max_deg_vert: -1 computes cc_main.cpp:148-160's rule (first max out+in degree).
 *   avg_degree:   < 0 computes 2m/n (the mirrored graph's average).
 * MSC_LABELS_RAW: isolated vertex -> itself, pivot's component -> pivot, others -> max id. */
int msc_cc_run(const msc_csr* g, int32_t max_deg_vert, double avg_degree, int32_t* labels,
               const msc_opts* opts, msc_stats* stats);

/* Batches: many independent graphs decomposed in one pass.  Replaces a loop of
 * run_scc / run_cc calls (multistep/scc_main.cpp:210, cc_main.cpp:184), one per graph.
 *   g:           the graphs packed into one CSR: graph i owns the vertices
 *                [vert_base[i], vert_base[i+1]) and its edges end inside that range
 *                (msc_batch_pack builds this from separate graphs).
 *   vert_base:   host, num_graphs + 1 entries, vert_base[0] = 0, non-decreasing,
 *                vert_base[num_graphs] = g->n (empty graphs allowed).
 *   pivots:      host, num_graphs pivots in each graph's own ids (-1: that graph has
 *                none), or NULL to apply create_csr's rule to every graph on device.
 *   labels:      g->n int32; each graph's slice holds exactly what msc_scc_run returns
 *                for that graph alone (ids local to the graph, raw or canonical).
 * All graphs share every launch: trim and colouring are per vertex, and FW-BW starts from
 * all pivots at once (the graphs are disjoint, so the reach sets do not mix).
 * stats->pivot is -1; the other counters cover the whole batch. */
int msc_scc_run_batch(const msc_csr* g, int32_t num_graphs, const int64_t* vert_base, const int32_t* pivots,
                      int32_t* labels, const msc_opts* opts, msc_stats* stats);
/* Weakly connected components of every graph of a batch (per graph as msc_cc_run). */
int msc_cc_run_batch(const msc_csr* g, int32_t num_graphs, const int64_t* vert_base, const int32_t* pivots,
                     int32_t* labels, const msc_opts* opts, msc_stats* stats);

/* Pivot and average degree on device.  Replaces the loop inside create_csr
 * (multistep/scc_main.cpp:169-183; kokkos_scc/scc_main.cpp:228-247). */
int msc_scc_pivot(const msc_csr* g, const msc_opts* opts, int32_t* max_deg_vert,
                  double* avg_degree);
/* The same for CC/WCC: the first vertex of largest degree in the mirrored graph (out + in)
 * and the mirrored graph's average degree 2m/n.  Replaces the loop inside cc_main's
 * create_csr (multistep/cc_main.cpp:148-160), whose results run_cc takes as arguments. */
int msc_cc_pivot(const msc_csr* g, const msc_opts* opts, int32_t* max_deg_vert, double* avg_degree);

/* ---- verification ---------------------------------------------------------------- */
typedef struct msc_verify {
  int32_t num_components;
  int32_t nontrivial; /* components with more than one vertex */
  int32_t max_size;
  int32_t unassigned; /* labels < 0 */
} msc_verify;

/* Replaces scc_verify(graph&, int*) (multistep/scc_verify.cpp:47-77) and
 * scc_verify(int, int*) (kokkos_scc/scc_verify.cpp:47-76): counts only, host labels.
 * The reference prints "Num SCCs: %d, Nontrivial: %d, Max SCC: %d, Unassigned %d";
 * msc_verify_format writes that line. */
int msc_scc_verify(int32_t n, const int32_t* labels, msc_verify* out);
/* Replaces cc_verify (multistep/cc_verify.cpp:47-74). */
int msc_cc_verify(int32_t n, const int32_t* labels, msc_verify* out);
/* kind 0: the SCC line, 1: the CC line ("Num CCs: %d, Max CC: %d, Unassigned %d"). */
int msc_verify_format(const msc_verify* v, int32_t kind, char* buf, int32_t cap);

/* Relabel every component to its minimum vertex id, host labels (no reference
 * counterpart: the reference's raw labels differ by phase, so parity is stated on this
 * canonical form). */
int msc_canonicalize(int32_t n, int32_t* labels);

/* ---- graph construction (host) --------------------------------------------------- */
/* read_edge + create_csr (multistep/scc_main.cpp:94-191): "n m" then m lines "u v".
 * Keeps duplicates and self-loops, adjacency in input order, like the reference. */
int msc_graph_read_edge_list(const char* path, msc_host_graph** out);
/* create_csr from arrays (multistep/scc_main.cpp:127-191).  flags: MSC_BUILD_* */
#define MSC_BUILD_DROP_SELF_LOOPS 1
#define MSC_BUILD_DEDUPE 2 /* also sorts every adjacency list */
int msc_graph_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                         int32_t flags, msc_host_graph** out);
/* output_scc (multistep/scc_main.cpp:193-202): one label per line. */
int msc_write_labels(const char* path, int32_t n, const int32_t* labels);
void msc_graph_free(msc_host_graph* g);

/* Synthetic inputs (no reference counterpart; deterministic for a seed, independent of
 * the thread count).  Self-loops and duplicate edges are removed, lists sorted. */
/* R-MAT: 2^scale vertices, edge_factor*2^scale draws with quadrant probabilities
 * a, b, c (d = 1-a-b-c), vertex ids permuted when permute != 0. */
int msc_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                 uint64_t seed, int32_t permute, msc_host_graph** out);
/* rows x cols grid: each grid edge kept in one random direction with p_one, in both with
 * p_both; plus round(arc_frac * n) uniform random arcs. */
int msc_gen_grid(int32_t rows, int32_t cols, double p_one, double p_both, double arc_frac,
                 uint64_t seed, msc_host_graph** out);
/* Planted SCCs: component sizes drawn from a Zipf(s) law (s > 1) truncated at max_size,
 * each component a directed cycle plus extra_per_vertex random internal arcs, joined by
 * dag_arcs random arcs that respect a random component order; ids permuted.  The SCC
 * partition is the planted one; comp_out (may be NULL, n entries) receives it. */
int msc_gen_planted(int32_t n, double zipf_s, int32_t max_size, double extra_per_vertex,
                    int64_t dag_arcs, uint64_t seed, msc_host_graph** out, int32_t* comp_out);

/* Packs graphs into one CSR for msc_*_run_batch: graph i's vertex v becomes
 * vert_base[i] + v.  vert_base receives num_graphs + 1 entries.  Fails with
 * MSC_E_CAPACITY when the union reaches 2^31 vertices or 2^32 edges. */
int msc_batch_pack(int32_t num_graphs, const msc_host_graph* const* graphs, msc_host_graph** out,
                   int64_t* vert_base);
/* A synthetic batch, packed: vertex counts log-normal (median median_n, log-space
 * sigma), clipped to [min_n, max_n]; each graph is, with equal odds, an R-MAT digraph
 * (a, b, c = .57, .19, .19, edge factor uniform in [8, 32], 2^ceil(log2 n) id space with
 * draws outside [0, n) drawn again) or planted SCCs (Zipf s = 1.5 sizes up to n, one
 * extra internal arc per vertex, 4n DAG arcs).  kinds (may be NULL) receives 0 for R-MAT,
 * 1 for planted, per graph. */
int msc_gen_batch(int32_t num_graphs, double median_n, double sigma, int32_t min_n, int32_t max_n, uint64_t seed,
                  msc_host_graph** out, int64_t* vert_base, int32_t* kinds);

/* Device-side read_edge + create_csr (multistep/scc_main.cpp:94-191): `text` (host memory,
 * `bytes` long) is the edge-list file ("n m" line, then m lines "u v"), parsed on the device
 * with read_edge's rules (getline(' ') + atoi for the source, getline + atoi for the
 * destination); the CSR and transposed CSR are built in device memory in create_csr's order
 * (adjacency in input order, duplicates and self-loops kept), and the pivot / average degree
 * follow create_csr's rule (either pointer may be NULL).  *dev receives device arrays; free
 * them with msc_csr_device_free.  Malformed text, ids outside [0, n) or fewer than m lines:
 * MSC_E_INVALID. */
int msc_graph_ingest_text(msc_ctx* ctx, const char* text, int64_t bytes, msc_csr* dev, int32_t* max_deg_vert,
                          double* avg_degree);
/* The same from a file (read into pinned host memory, then ingested). */
int msc_graph_read_edge_list_device(msc_ctx* ctx, const char* path, msc_csr* dev, int32_t* max_deg_vert,
                                    double* avg_degree);
/* The reference's edge-list text of a graph ("n m\n" then "u v\n" per edge in CSR order),
 * formatted in parallel.  out == NULL: *bytes receives the size needed. */
int msc_format_edge_list(const msc_host_graph* g, char* out, int64_t cap, int64_t* bytes);

/* ---- device graphs --------------------------------------------------------------- */
/* Copy a host graph into device memory of the context's device; fills *dev (where =
 * MSC_DEVICE).  Free with msc_csr_device_free. */
int msc_csr_to_device(msc_ctx* ctx, const msc_csr* host, msc_csr* dev);
int msc_csr_device_free(msc_ctx* ctx, msc_csr* dev);
/* Device buffers for labels kept in HBM (opts.labels_where = MSC_DEVICE) and the copies
 * in and out of them, ordered on the context's stream. */
int msc_device_alloc(msc_ctx* ctx, int64_t bytes, void** out);
int msc_device_free(msc_ctx* ctx, void* p);
int msc_copy_to_host(msc_ctx* ctx, void* dst, const void* src, int64_t bytes);
int msc_copy_to_device(msc_ctx* ctx, void* dst, const void* src, int64_t bytes);
/* Device-to-device copy (also the known-size reference for HBM traffic counters). */
int msc_copy_device(msc_ctx* ctx, void* dst, const void* src, int64_t bytes);
/* Waits for all work queued on the context's stream. */
int msc_ctx_sync(msc_ctx* ctx);

/* ---- version / build info -------------------------------------------------------- */
const char* msc_version(void);
/* Number of visible gfx950 devices (0 when none). */
int msc_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* MSC_H */
