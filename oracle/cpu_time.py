"""ORACLE -- test infrastructure only.

CPU baseline timer, run in a process of its own so that the OpenMP settings it is timed
under (OMP_NUM_THREADS, OMP_PROC_BIND, OMP_PLACES: read by libgomp when it loads) apply to
it alone.  It times the reference's own run_scc (``_ref/libref_scc.so``, compiled from
multistep's sources; the restatement in ``_build/liboracle.so`` where that was not built)
exactly around the call, as the reference times it (multistep/scc_main.cpp:368-378), on a
graph the parent wrote as .npy files, and prints one JSON line.

    python -m oracle.cpu_time --dir DIR --pivot P --avg A --runs R
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
from types import SimpleNamespace

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", required=True, help="out_off.npy, out_adj.npy, in_off.npy, in_adj.npy")
    ap.add_argument("--pivot", type=int, required=True)
    ap.add_argument("--avg", type=float, required=True)
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--vert-cutoff", type=int, default=10000)
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import oracle

    arr = {k: np.load(os.path.join(args.dir, f"{k}.npy"), mmap_mode="r") for k in
           ("out_off", "out_adj", "in_off", "in_adj")}
    g = SimpleNamespace(n=int(arr["out_off"].shape[0] - 1), m=int(arr["out_adj"].shape[0]),
                        **{k: np.ascontiguousarray(v) for k, v in arr.items()})
    use_ref = oracle.have_ref()
    times = []
    for _ in range(args.runs):
        if use_ref:
            times.append(oracle.ref_time_scc(g, args.pivot, args.avg, args.vert_cutoff))
        else:
            times.append(oracle.scc_time(g, args.pivot, args.avg, args.vert_cutoff))
    print(json.dumps({"kind": "reference" if use_ref else "port", "times_s": times,
                      "median_s": statistics.median(times)}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
