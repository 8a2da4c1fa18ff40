// cc [graph] [optional: output file] [--canonical] [--device N]
//
// Command-line counterpart of the reference's `cc` (multistep/cc_main.cpp:247-319): the
// edge list is read as a digraph and its weak components are computed (the reference
// mirrors every edge in read_edge, cc_main.cpp:102-118; here out- and in-lists are walked
// together).  Prints "Multistep CC time: %9.6lf" and the cc_verify line, writes one label
// per line (raw labels unless --canonical).
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
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) o.device = std::atoi(argv[++i]);
    else if (!graph) graph = argv[i];
    else if (!out) out = argv[i];
  }
  if (!graph) usage(argv[0]);
  msc_host_graph* h = nullptr;
  int rc = msc_graph_read_edge_list(graph, &h);
  if (rc) {
    std::fprintf(stderr, "cc: cannot read %s (error %d)\n", graph, rc);
    return 1;
  }
  msc_ctx* ctx = nullptr;
  if ((rc = msc_ctx_create(o.device, nullptr, &ctx))) {
    std::fprintf(stderr, "cc: no usable gfx950 device (error %d)\n", rc);
    return 1;
  }
  o.ctx = ctx;
  msc_csr host = {h->n, h->m, h->out_off, h->out_adj, h->in_off, h->in_adj, MSC_HOST};
  msc_csr dev;
  if ((rc = msc_csr_to_device(ctx, &host, &dev))) {
    std::fprintf(stderr, "cc: device upload failed (error %d)\n", rc);
    return 1;
  }
  std::vector<int32_t> labels((size_t)h->n);
  auto t0 = std::chrono::steady_clock::now();
  rc = msc_cc_run(&dev, -1, -1.0, labels.data(), &o, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  if (rc) {
    std::fprintf(stderr, "cc: run failed (error %d)\n", rc);
    return 1;
  }
  std::printf("Multistep CC time: %9.6lf\n", std::chrono::duration<double>(t1 - t0).count());
  msc_verify v;
  char line[256];
  msc_cc_verify(h->n, labels.data(), &v);
  msc_verify_format(&v, 1, line, sizeof(line));
  std::printf("%s\n", line);
  if (out && (rc = msc_write_labels(out, h->n, labels.data()))) {
    std::fprintf(stderr, "cc: cannot write %s\n", out);
    return 1;
  }
  msc_csr_device_free(ctx, &dev);
  msc_ctx_destroy(ctx);
  msc_graph_free(h);
  return 0;
}
