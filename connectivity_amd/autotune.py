"""Autotuner for a context's launch parameters (msc_tuning).

The level loops have thirteen tunables that no rule of the reference fixes and that trade
against each other differently on shallow power-law graphs and on deep meshes:

  blocks_per_cu     workgroups per compute unit of every level kernel (grid = CUs x this);
  small_push_edges  colour push sweeps whose frontier has at most this many edges run
                    inside the one-workgroup closing kernel, many levels per launch;
  small_cbw_edges   the same for the per-colour backward searches;
  first_batch       levels enqueued before a loop first reads its `done` word back;
  max_batch         levels enqueued between later reads (batches double up to this);
  concurrent_fwbw   backward search beside the forward one in the same launches (1), after it (0),
                    or beside it and restricted to the forward set once that is done (2);
  search_batch      levels of a FW-BW / CC search enqueued before its first host check;
  local_rounds      levels a workgroup of a colour launch continues on its own claims;
  local_us          ... while that launch has run for less than this many microseconds;
  pull_div          a colour level over at least 1/pull_div of its round's edges pulls (0: never);
  tail_vertices     a colouring of at most this many vertices is settled by one workgroup when
                    the edges joining them fit its LDS (0: always colour in rounds);
  balance           top-down load balance of grid-wide levels: 1 a workgroup deals its 256
                    frontier vertices' edges, 2 chunks of the whole frontier's edges, 3 one or
                    the other per level by the frontier's mean list length (0, a thread per
                    frontier vertex, is left out: it walks a hub's list on one thread);
  group_words       bitmap words a pull or bottom-up level's workgroup takes at a time (0: 64
                    for a batch, 256 for one graph).

None of them changes a label: every candidate's canonical labels are compared with the
default tuning's and a candidate that differs is reported as an error, not skipped.
Search: coordinate descent from the library defaults, each field swept over its candidate
values with the others held, best (median wall time of `reps` runs) kept, two passes.

CLI:  python -m connectivity_amd.autotune --workload rmat:24:16 --alg scc --out tuning.json
Workload specs: rmat:SCALE:EDGE_FACTOR, grid:ROWS:COLS, planted:N:MAX_SIZE.
"""
from __future__ import annotations

import argparse
import json
import statistics
import sys

import numpy as np

from . import (LABELS_CANONICAL, POLICY_LEARNED, POLICY_PLAIN, Context, DeviceGraph, DeviceLabels, MscError,
               gen_grid, gen_planted, gen_rmat, load_policy, run_cc, run_scc)

CANDIDATES = {
    "blocks_per_cu": [2, 4, 5, 6, 8],
    "small_push_edges": [0, 1024, 2048, 4096, 8192, 16384],
    "small_cbw_edges": [0, 1024, 2048, 4096, 8192, 16384],
    "first_batch": [2, 4, 8],
    "max_batch": [16, 32, 64],
    "concurrent_fwbw": [0, 1, 2],
    "search_batch": [2, 4, 6, 8],
    "local_rounds": [0, 4, 16, 32, 64],
    "local_us": [40, 80, 160, 240, 320],
    "pull_div": [0, 2, 4, 8, 16],
    "tail_vertices": [0, 4096, 65536],
    "balance": [1, 2, 3],
    "group_words": [0, 32, 64, 128, 256],
}


def parse_workload(spec: str):
    """rmat:S:EF | grid:R:C | planted:N:MAX -> host Graph (seed 1, the bench generators)."""
    kind, *nums = spec.split(":")
    vals = [int(x) for x in nums]
    if kind == "rmat":
        return gen_rmat(vals[0], vals[1] if len(vals) > 1 else 16, seed=1)
    if kind == "grid":
        return gen_grid(vals[0], vals[1] if len(vals) > 1 else vals[0], seed=1)
    if kind == "planted":
        return gen_planted(vals[0], **({"max_size": vals[1]} if len(vals) > 1 else {}), seed=1)[0]
    raise ValueError(f"unknown workload {spec!r}")


class Bench:
    """Times one algorithm on one device-resident graph under a given tuning."""

    def __init__(self, ctx: Context, g, alg: str, policy: int):
        self.ctx, self.g, self.alg, self.policy = ctx, g, alg, policy
        self.dg = DeviceGraph(ctx, g)
        self.out = DeviceLabels(ctx, g.n)
        self.fn = run_scc if alg == "scc" else run_cc

    def once(self) -> float:
        piv = self.g.max_deg_vert if self.alg == "scc" else None
        avg = self.g.avg_degree if self.alg == "scc" else None
        _, st = self.fn(self.dg, piv, avg, ctx=self.ctx, out=self.out, policy=self.policy, stats=True,
                        labels=LABELS_CANONICAL)
        return st.wall_ms

    def measure(self, tuning: dict, reps: int) -> float:
        self.ctx.set_tuning(**tuning)
        self.once()  # warm-up under this tuning
        return statistics.median(self.once() for _ in range(reps))

    def labels(self) -> np.ndarray:
        return self.out.to_host()

    def free(self) -> None:
        self.out.free()
        self.dg.free()


def tune(ctx: Context, g, alg: str = "scc", policy: int = POLICY_PLAIN, reps: int = 5, passes: int = 2,
         log=print) -> dict:
    """Returns {"tuning": best, "ms": its median, "default_ms": the defaults' median, "trials": [...]}."""
    b = Bench(ctx, g, alg, policy)
    try:
        default = ctx.tuning()
        ref_ms = b.measure(default, reps)
        ref = b.labels()
        best, best_ms = dict(default), ref_ms
        trials = [(dict(default), ref_ms)]
        log(f"default {default}: {ref_ms:.3f} ms")
        for _ in range(passes):
            changed = False
            for field, values in CANDIDATES.items():
                for v in values:
                    if v == best[field]:
                        continue
                    cand = dict(best, **{field: v})
                    if cand["max_batch"] < cand["first_batch"]:
                        continue
                    ms = b.measure(cand, reps)
                    if not np.array_equal(b.labels(), ref):
                        raise MscError(-1, f"tuning {cand} changed the labels")
                    trials.append((cand, ms))
                    log(f"  {field}={v}: {ms:.3f} ms")
                    if ms < best_ms * 0.98:
                        best, best_ms, changed = cand, ms, True
            if not changed:
                break
        ctx.set_tuning(**best)
        return {"tuning": best, "ms": best_ms, "default_ms": ref_ms, "trials": trials}
    finally:
        b.free()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m connectivity_amd.autotune")
    ap.add_argument("--workload", action="append", required=True, help="rmat:S:EF, grid:R:C, planted:N:MAX")
    ap.add_argument("--alg", choices=["scc", "cc"], default="scc")
    ap.add_argument("--policy", choices=["plain", "learned"], default="plain")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=None, help="JSON file to merge results into")
    args = ap.parse_args(argv)
    ctx = Context(0)
    defaults = ctx.tuning()
    policy = POLICY_PLAIN
    if args.policy == "learned":
        load_policy(ctx)
        policy = POLICY_LEARNED
    results = {}
    if args.out:
        try:
            with open(args.out) as f:
                results = json.load(f)
        except FileNotFoundError:
            pass
    for spec in args.workload:
        print(f"== {spec} {args.alg} {args.policy}", flush=True)
        g = parse_workload(spec)
        r = tune(ctx, g, args.alg, policy, reps=args.reps, log=lambda m: print(m, flush=True))
        print(f"best {r['tuning']}: {r['ms']:.3f} ms (defaults {r['default_ms']:.3f} ms)", flush=True)
        results[f"{spec}/{args.alg}/{args.policy}"] = {"tuning": r["tuning"], "ms": r["ms"],
                                                       "default_ms": r["default_ms"]}
        ctx.set_tuning(**defaults)  # every workload starts from the library defaults
    if args.out:
        with open(args.out, "w") as f:
            json.dump(results, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
