// kscc <graph> <alg to run> [<graph name> <timing iterations>]
//
// Command-line counterpart of the reference's Kokkos driver (kokkos_scc/scc_main.cpp:254-400,
// kokkos_scc/scc_run.cpp:60-184): reads the edge list, builds CSR + transposed CSR, runs
// the decomposition with the graph resident on the device and prints the scc_verify line.
// With a graph name and an iteration count it reproduces the reference's TIMING build: all
// three variants are run `iterations` times each and "<name>, Baseline|ManhattanLocal|
// ManhattanGlobal, <average s>" is printed per variant (scc_run.cpp:171-180).
//
// The reference's variants differ only in how their top-down kernels spread work and
// compute the same labels.  The selector picks msc's matching load balance
// (msc_tuning.balance) for the grid-wide search and colour levels: 0 a thread per frontier
// vertex (fwbw_baseline / color_baseline), 1 a workgroup deals the edges of its 256 frontier
// vertices (fwbw_manhattan_local / color_manhattan_local), 2 fixed-size chunks of the whole
// frontier's edges (fwbw_manhattan_global / color_manhattan_global, msc's default).  An alg
// outside 0..2 prints the usage line and exits, as the reference does.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/msc.h"

static const char* const kNames[3] = {"Baseline", "ManhattanLocal", "ManhattanGlobal"};

static void usage(const char* argv0) {
  std::printf("Usage: %s <graph> <alg to run> [<graph name> <timing iterations>]\n", argv0);
  std::exit(0);
}

int main(int argc, char** argv) {
  setbuf(stdout, nullptr);
  if (argc < 3) usage(argv[0]);
  const int alg = std::atoi(argv[2]);
  if (alg < 0 || alg > 2) usage(argv[0]);
  const bool timing = argc >= 5;
  const char* name = timing ? argv[3] : nullptr;
  const int iters = timing ? std::atoi(argv[4]) : 1;
  if (timing && iters < 1) usage(argv[0]);

  msc_host_graph* h = nullptr;
  int rc = msc_graph_read_edge_list(argv[1], &h);
  if (rc) {
    std::fprintf(stderr, "kscc: cannot read %s (error %d)\n", argv[1], rc);
    return 1;
  }
  msc_opts o;
  msc_opts_default(&o);
  o.label_mode = MSC_LABELS_RAW;
  msc_ctx* ctx = nullptr;
  if ((rc = msc_ctx_create(0, nullptr, &ctx))) {
    std::fprintf(stderr, "kscc: no usable gfx950 device (error %d)\n", rc);
    return 1;
  }
  o.ctx = ctx;
  msc_csr host = {h->n, h->m, h->out_off, h->out_adj, h->in_off, h->in_adj, MSC_HOST};
  msc_csr dev;
  if ((rc = msc_csr_to_device(ctx, &host, &dev))) {
    std::fprintf(stderr, "kscc: device upload failed (error %d)\n", rc);
    return 1;
  }
  int32_t* labels_dev = nullptr;
  if ((rc = msc_device_alloc(ctx, (int64_t)(h->n > 0 ? h->n : 1) * 4, (void**)&labels_dev))) {
    std::fprintf(stderr, "kscc: device allocation failed (error %d)\n", rc);
    return 1;
  }
  o.labels_where = MSC_DEVICE;

  // every variant in the TIMING build, only the chosen one otherwise
  const int first = timing ? 0 : alg, last = timing ? 2 : alg;
  msc_tuning tune;
  if ((rc = msc_ctx_get_tuning(ctx, &tune))) {
    std::fprintf(stderr, "kscc: cannot read the tuning (error %d)\n", rc);
    return 1;
  }
  for (int a = first; a <= last; ++a) {
    tune.balance = a;
    if ((rc = msc_ctx_set_tuning(ctx, &tune))) {
      std::fprintf(stderr, "kscc: cannot select variant %d (error %d)\n", a, rc);
      return 1;
    }
    double total = 0.0;
    for (int t = 0; t < iters; ++t) {
      auto t0 = std::chrono::steady_clock::now();
      rc = msc_scc_run(&dev, h->max_deg_vert, h->avg_degree, labels_dev, &o, nullptr);
      auto t1 = std::chrono::steady_clock::now();
      if (rc) {
        std::fprintf(stderr, "kscc: run failed (error %d)\n", rc);
        return 1;
      }
      total += std::chrono::duration<double>(t1 - t0).count();
    }
    if (timing) std::printf("%s, %s, %2.4lf\n", name, kNames[a], total / iters);
  }

  std::vector<int32_t> labels((size_t)h->n);
  if (h->n > 0 && (rc = msc_copy_to_host(ctx, labels.data(), labels_dev, (int64_t)h->n * 4))) {
    std::fprintf(stderr, "kscc: copy failed (error %d)\n", rc);
    return 1;
  }
  msc_verify v;
  char line[256];
  msc_scc_verify(h->n, labels.data(), &v);
  msc_verify_format(&v, 0, line, sizeof(line));
  std::printf("%s\n", line);
  msc_device_free(ctx, labels_dev);
  msc_csr_device_free(ctx, &dev);
  msc_ctx_destroy(ctx);
  msc_graph_free(h);
  return 0;
}
