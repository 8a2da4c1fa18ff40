/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement (C++/OpenMP) of the reference's Multistep SCC and CC/WCC algorithms
 * (multistep/scc_*.cpp, multistep/cc_*.cpp).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker.  The product library
 * (connectivity_amd/libmsc.so) never links or calls it.
 *
 * Labels written by orc_scc_run / orc_cc_run are the reference's *raw* labels:
 *   trimmed vertex -> itself, FW-BW (or CC-BFS) component -> the pivot,
 *   coloured component -> its maximum vertex id, Tarjan tail -> the DFS root.
 * Every compaction here preserves order, so all of them (the Tarjan tail included) are
 * deterministic and equal the reference run with one OpenMP thread.
 *
 * Graph: int32 vertex ids, uint32 CSR offsets (multistep/scc_main.cpp:75-82 uses
 * `unsigned` offsets and `int` adjacency).  m is out_off[n].
 */
#ifndef MSC_ORACLE_H
#define MSC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PHASE_FW = 0, ORC_PHASE_BW = 1, ORC_PHASE_CC = 2 };

typedef struct orc_level {
  int32_t phase;      /* ORC_PHASE_* */
  int32_t level;      /* 1-based level within its search */
  int32_t bottom_up;  /* direction this level ran in */
  int32_t next_size;  /* vertices this level discovered */
} orc_level;

typedef struct orc_trace {
  orc_level* levels;  /* caller-owned, may be NULL */
  int32_t cap;
  int32_t count;
  int32_t num_trim;     /* vertices removed by trim */
  int32_t num_fwbw;     /* vertices in the pivot's component */
  int32_t num_color;    /* components found by colouring (roots) */
  int32_t num_serial;   /* components found by the Tarjan tail */
  int32_t color_rounds; /* outer colouring rounds */
  int32_t color_iters;  /* propagation sweeps summed over rounds */
} orc_trace;

/* Pivot and average degree of create_csr (multistep/scc_main.cpp:169-183).  pivot = first
 * vertex with strictly largest (double)out*(double)in; -1 when every product is 0. */
int orc_pivot_scc(int32_t n, const uint32_t* out_off, const uint32_t* in_off,
                  int32_t* pivot, double* avg_degree);
/* CC rule (multistep/cc_main.cpp:148-160) on the mirrored graph: first strictly largest
 * out+in degree; avg = 2m/n. */
int orc_pivot_cc(int32_t n, const uint32_t* out_off, const uint32_t* in_off,
                 int32_t* pivot, double* avg_degree);

/* run_scc (multistep/scc_main.cpp:210-317).  pivot < 0 skips FW-BW. */
int orc_scc_run(int32_t n, const uint32_t* out_off, const int32_t* out_adj,
                const uint32_t* in_off, const int32_t* in_adj, int32_t pivot, double avg_degree,
                int32_t vert_cutoff, double alpha, int32_t beta, int32_t* labels,
                orc_trace* trace);

/* run_cc (multistep/cc_main.cpp:184-245) on the digraph's mirror (out + in lists). */
/* Same with the trim level of multistep/scc_main.cpp:241-258: 1 (scc_trim) or 2 (the
 * complete trim scc_trim_complete, scc_trim.cpp:141-297). */
int orc_scc_run_trim(int32_t n, const uint32_t* out_off, const int32_t* out_adj,
                     const uint32_t* in_off, const int32_t* in_adj, int32_t pivot, double avg,
                     int32_t vert_cutoff, double alpha, int32_t beta, int32_t trim_level,
                     int32_t* labels, orc_trace* trace);
/* The complete trim alone: labels[v] = v for trimmed vertices, -1 for survivors; returns
 * the number trimmed. */
int orc_trim_complete(int32_t n, const uint32_t* out_off, const int32_t* out_adj,
                      const uint32_t* in_off, const int32_t* in_adj, int32_t* labels);

int orc_cc_run(int32_t n, const uint32_t* out_off, const int32_t* out_adj,
               const uint32_t* in_off, const int32_t* in_adj, int32_t pivot, double avg_degree,
               double alpha, int32_t beta, int32_t* labels, orc_trace* trace);

/* Relabel every component to its minimum vertex id (labels must be in [0,n) or -1). */
void orc_canonicalize(int32_t n, int32_t* labels);

/* scc_verify counts (multistep/scc_verify.cpp:47-77):
 * out = {num_components, nontrivial, max_size, unassigned}. */
void orc_verify_counts(int32_t n, const int32_t* labels, int32_t* out);

/* Independent verifier: iterative Tarjan over the whole digraph; canonical labels. */
int orc_tarjan(int32_t n, const uint32_t* out_off, const int32_t* out_adj, int32_t* labels);

/* Independent verifier: union-find weak components; canonical labels. */
int orc_wcc_unionfind(int32_t n, const uint32_t* out_off, const int32_t* out_adj,
                      int32_t* labels);

/* Wall time of orc_scc_run in seconds (for the CPU baseline leg). */
double orc_scc_time(int32_t n, const uint32_t* out_off, const int32_t* out_adj,
                    const uint32_t* in_off, const int32_t* in_adj, int32_t pivot,
                    double avg_degree, int32_t vert_cutoff, int32_t* labels);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
