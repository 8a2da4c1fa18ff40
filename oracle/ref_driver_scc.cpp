// Test-infrastructure only: compiles the reference's own multistep SCC translation unit
// (multistep/scc_main.cpp, which #includes scc_trim/scc_fwbw/scc_color/scc_serial/scc_verify)
// unchanged, with its main() renamed, and exposes its driver and CSR builder behind C entry
// points so the tests can run the reference on in-memory graphs with any vert_cutoff.
// The sources are read where they lie (include path set by oracle/Makefile); nothing from
// them is copied here.  Output goes to oracle/_ref/ only.
//
// The reference reads flags it never initialised: scc_fwbw_fw clears fw[] only for valid
// vertices (multistep/scc_fwbw.cpp:59-61) and scc_fwbw_bw then tests fw[in] for every
// in-neighbour, trimmed ones included (:269).  In a fresh process those arrays come from
// zeroed pages and read false; inside a long-lived process the heap is dirty and trimmed
// vertices can pull their in-neighbours into the pivot's SCC.  To run the reference the
// way its authors did, array allocations made by the code in this library are zero-filled
// (the operators below bind only inside this .so: it is built with hidden visibility).
#include <cstdlib>
#include <new>
void* operator new[](std::size_t bytes) {
  void* p = std::calloc(1, bytes ? bytes : 1);
  if (!p) throw std::bad_alloc();
  return p;
}
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

#define main msc_reference_scc_main
#include "scc_main.cpp"
#undef main

#define REF_API extern "C" __attribute__((visibility("default")))

// run_scc(graph&, int*&, int, double, int): multistep/scc_main.cpp:210
REF_API int ref_scc_run(int n, unsigned m, const unsigned* out_off, const int* out_adj,
                        const unsigned* in_off, const int* in_adj, int max_deg_vert,
                        double avg_degree, int vert_cutoff, int* labels_out)
{
  graph g = {n, m, const_cast<int*>(out_adj), const_cast<int*>(in_adj),
             const_cast<unsigned*>(out_off), const_cast<unsigned*>(in_off)};
  int* maps = nullptr;
  run_scc(g, maps, max_deg_vert, avg_degree, vert_cutoff);
  for (int i = 0; i < n; ++i) labels_out[i] = maps[i];
  delete[] maps;
  return 0;
}

// Same call, timed exactly like the reference's own timed region (scc_main.cpp:368-378).
REF_API double ref_scc_time(int n, unsigned m, const unsigned* out_off, const int* out_adj,
                            const unsigned* in_off, const int* in_adj, int max_deg_vert,
                            double avg_degree, int vert_cutoff, int* labels_out)
{
  graph g = {n, m, const_cast<int*>(out_adj), const_cast<int*>(in_adj),
             const_cast<unsigned*>(out_off), const_cast<unsigned*>(in_off)};
  int* maps = nullptr;
  double t0 = timer();
  run_scc(g, maps, max_deg_vert, avg_degree, vert_cutoff);
  double t1 = timer();
  if (labels_out)
    for (int i = 0; i < n; ++i) labels_out[i] = maps[i];
  delete[] maps;
  return t1 - t0;
}

// scc_trim_complete(...): multistep/scc_trim.cpp:141-297 (the TRIM_LEVEL 2 trim) on a fresh
// valid set, as run_scc sets it up (scc_main.cpp:225-238); labels_out[v] = v for trimmed
// vertices and -1 for survivors.
REF_API int ref_trim_complete(int n, unsigned m, const unsigned* out_off, const int* out_adj,
                              const unsigned* in_off, const int* in_adj, int* labels_out)
{
  graph g = {n, m, const_cast<int*>(out_adj), const_cast<int*>(in_adj),
             const_cast<unsigned*>(out_off), const_cast<unsigned*>(in_off)};
  int* maps = new int[n > 0 ? n : 1];
  bool* valid = new bool[n > 0 ? n : 1];
  int* valid_verts = new int[n > 0 ? n : 1];
  int num_valid = n;
  for (int i = 0; i < n; ++i) { maps[i] = -1; valid[i] = true; valid_verts[i] = i; }
  int trimmed = scc_trim_complete(g, valid, valid_verts, num_valid, maps);
  for (int i = 0; i < n; ++i) labels_out[i] = maps[i];
  delete[] maps; delete[] valid; delete[] valid_verts;
  return trimmed;
}

// create_csr(...): multistep/scc_main.cpp:127-191 (CSR + transposed CSR, pivot, avg degree)
REF_API int ref_create_csr(int n, unsigned m, const int* srcs, const int* dsts,
                           unsigned* out_off, int* out_adj, unsigned* in_off, int* in_adj,
                           int* max_deg_vert, double* avg_degree)
{
  int *oa = nullptr, *ia = nullptr;
  unsigned *oo = nullptr, *io = nullptr;
  int piv = -1;
  double avg = 0.0;
  create_csr(n, m, const_cast<int*>(srcs), const_cast<int*>(dsts), oa, ia, oo, io, piv, avg);
  for (unsigned i = 0; i < m; ++i) { out_adj[i] = oa[i]; in_adj[i] = ia[i]; }
  for (int i = 0; i <= n; ++i) { out_off[i] = oo[i]; in_off[i] = io[i]; }
  *max_deg_vert = piv;
  *avg_degree = avg;
  delete[] oa; delete[] ia; delete[] oo; delete[] io;
  return 0;
}
