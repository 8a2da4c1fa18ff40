// Test-infrastructure only: compiles the reference's own multistep CC translation unit
// (multistep/cc_main.cpp, which #includes cc_trim/cc_bfs/cc_color/cc_verify) unchanged, with
// its main() renamed, and exposes run_cc behind a C entry point.  The caller passes the
// directed edge list; it is mirrored here the way multistep/cc_main.cpp:102-118 (read_edge)
// mirrors every edge, and handed to the reference's create_csr (cc_main.cpp:123-165).
//
// Array allocations made by this library are zero-filled, as in ref_driver_scc.cpp (the
// reference reads cc[] of vertices it never initialised, multistep/cc_bfs.cpp:59-61).
// Known hazard left as is: update_valid walks g.n entries of a list holding num_valid
// (cc_bfs.cpp:231); past num_valid it sees vertex 0 repeatedly, which overflows the new
// list when vertex 0 is valid but outside the pivot's component.  Callers avoid that case.
#include <cstdlib>
#include <new>
void* operator new[](std::size_t bytes) {
  void* p = std::calloc(1, bytes ? bytes : 1);
  if (!p) throw std::bad_alloc();
  return p;
}
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

#define main msc_reference_cc_main
#include "cc_main.cpp"
#undef main

#define REF_API extern "C" __attribute__((visibility("default")))

static void mirror_and_build(int n, unsigned m, const int* srcs, const int* dsts, graph& g,
                             int& piv, double& avg)
{
  unsigned mm = 2u * m;
  int* s2 = new int[mm];
  int* d2 = new int[mm];
  for (unsigned i = 0; i < m; ++i) {
    s2[2 * i] = srcs[i];     d2[2 * i] = dsts[i];
    s2[2 * i + 1] = dsts[i]; d2[2 * i + 1] = srcs[i];
  }
  int* oa = nullptr;
  unsigned* oo = nullptr;
  piv = -1;
  avg = 0.0;
  create_csr(n, mm, s2, d2, oa, oo, piv, avg);
  delete[] s2;
  delete[] d2;
  g.n = n; g.m = mm; g.out_array = oa; g.out_degree_list = oo;
}

// run_cc(graph&, int*&, int, double): multistep/cc_main.cpp:184
REF_API int ref_cc_run(int n, unsigned m, const int* srcs, const int* dsts, int* labels_out,
                       int* pivot_out)
{
  graph g;
  int piv;
  double avg;
  mirror_and_build(n, m, srcs, dsts, g, piv, avg);
  int* maps = nullptr;
  run_cc(g, maps, piv, avg);
  for (int i = 0; i < n; ++i) labels_out[i] = maps[i];
  if (pivot_out) *pivot_out = piv;
  delete[] maps;
  delete[] g.out_array;
  delete[] g.out_degree_list;
  return 0;
}

REF_API double ref_cc_time(int n, unsigned m, const int* srcs, const int* dsts, int* labels_out)
{
  graph g;
  int piv;
  double avg;
  mirror_and_build(n, m, srcs, dsts, g, piv, avg);
  int* maps = nullptr;
  double t0 = timer();
  run_cc(g, maps, piv, avg);
  double t1 = timer();
  if (labels_out)
    for (int i = 0; i < n; ++i) labels_out[i] = maps[i];
  delete[] maps;
  delete[] g.out_array;
  delete[] g.out_degree_list;
  return t1 - t0;
}
