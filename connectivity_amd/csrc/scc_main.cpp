// scc [graph] [optional: output file] [--canonical] [--learned] [--device N]
//
// Command-line counterpart of the reference's `scc` (multistep/scc_main.cpp:319-395):
// reads the "n m" + m lines "u v" edge list, builds CSR + transposed CSR with create_csr's
// pivot rule, times the decomposition alone (scc_main.cpp:368-378; here the device graph
// is uploaded before the clock starts, labels come back inside it), prints
// "Multistep SCC time: %9.6lf" and the scc_verify line, and writes one label per line.
// Labels are the reference's raw ones unless --canonical is given.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/msc.h"

static void usage(const char* argv0) {
  std::printf("Usage: %s [graph] [optional: output file]\n", argv0);
  std::exit(0);
}

int main(int argc, char** argv) {
  setbuf(stdout, nullptr);
  const char* graph = nullptr;
  const char* out = nullptr;
  msc_opts o;
  msc_opts_default(&o);
  o.label_mode = MSC_LABELS_RAW;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--canonical")) o.label_mode = MSC_LABELS_CANONICAL;
    else if (!std::strcmp(argv[i], "--learned")) o.policy = MSC_POLICY_LEARNED;
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) o.device = std::atoi(argv[++i]);
    else if (!graph) graph = argv[i];
    else if (!out) out = argv[i];
  }
  if (!graph) usage(argv[0]);
  msc_host_graph* h = nullptr;
  int rc = msc_graph_read_edge_list(graph, &h);
  if (rc) {
    std::fprintf(stderr, "scc: cannot read %s (error %d)\n", graph, rc);
    return 1;
  }
  msc_ctx* ctx = nullptr;
  if ((rc = msc_ctx_create(o.device, nullptr, &ctx))) {
    std::fprintf(stderr, "scc: no usable gfx950 device (error %d)\n", rc);
    return 1;
  }
  o.ctx = ctx;
  msc_csr host = {h->n, h->m, h->out_off, h->out_adj, h->in_off, h->in_adj, MSC_HOST};
  msc_csr dev;
  if ((rc = msc_csr_to_device(ctx, &host, &dev))) {
    std::fprintf(stderr, "scc: device upload failed (error %d)\n", rc);
    return 1;
  }
  std::vector<int32_t> labels((size_t)h->n);
  auto t0 = std::chrono::steady_clock::now();
  rc = msc_scc_run(&dev, h->max_deg_vert, h->avg_degree, labels.data(), &o, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  if (rc) {
    std::fprintf(stderr, "scc: run failed (error %d)\n", rc);
    return 1;
  }
  std::printf("Multistep SCC time: %9.6lf\n", std::chrono::duration<double>(t1 - t0).count());
  msc_verify v;
  char line[256];
  msc_scc_verify(h->n, labels.data(), &v);
  msc_verify_format(&v, 0, line, sizeof(line));
  std::printf("%s\n", line);
  if (out && (rc = msc_write_labels(out, h->n, labels.data()))) {
    std::fprintf(stderr, "scc: cannot write %s\n", out);
    return 1;
  }
  msc_csr_device_free(ctx, &dev);
  msc_ctx_destroy(ctx);
  msc_graph_free(h);
  return 0;
}
