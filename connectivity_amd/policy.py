"""Learned direction / phase policy: data, training and evaluation.

The policy is an FP32 MLP 12 -> 32 -> 32 -> 3 evaluated on device after every BFS level of
the FW, BW and CC searches (msc_kernels.hip: close_level).  Its actions are: next level
top-down, next level bottom-up, or abandon the search and leave its vertices to colouring.
It only changes the schedule; labels are identical whatever it picks.

Training data come from the library's own runs:
  1. for each seeded training graph, the plain run (the reference's ALPHA/BETA rule) is
     recorded level by level (features the device computes, action taken, level time);
  2. for each search (FW, BW, CC) a family of forced schedules is timed on device
     (switch to bottom-up after level k, back after level j; all top-down; abandon after
     level s), with the other searches on the plain rule;
  3. the fastest schedule's levels give (features, action) samples; the plain run's levels
     up to the first level where it departs from that schedule are added with the same
     labels (once it departs, its states are not ones the fast schedule visits).  Samples of
     one (graph, search) are weighted to sum to the relative time the best schedule saves.
The network is trained on the GPU by msc_mlp_train (hand-written forward / backward / Adam
kernels, msc_mlp.hip).

CLI:  python -m connectivity_amd.policy all --graphs 210 --reps 2 [--max-batch-graphs 1024] --out ...
      python -m connectivity_amd.policy train --data <collected.npz> --seed 2 --epochs 600 --out ...
      python -m connectivity_amd.policy combine --scc <a.npz> --cc <b.npz> --out ...  (a network per search kind)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import (ACT_STOP, DeviceGraph, DeviceLabels, LABELS_CANONICAL, MLP_BLOB, MLP_H, MLP_IN, MLP_OUT,
               MLP_PARAMS, POLICY_LEARNED, POLICY_PLAIN, POLICY_WEIGHTS, SCHED_LEVELS, Context, batch_pivots,
               gen_batch, gen_grid, gen_planted, gen_rmat, mlp_train, run_cc, run_cc_batch, run_scc, run_scc_batch,
               schedule_array, set_policy)

FEATURES = ["log(next/n)", "log(next_edges/(avg*n))", "log(unvisited/n)", "log1p(level)", "next/n", "next_edges/(avg*n)",
            "log1p(next_edges/next)", "log1p(avg_degree)", "bottom_up", "switched_back", "found/num_valid",
            "phase"]
PHASES = {0: "fw", 1: "bw", 2: "cc"}


def log(msg: str) -> None:
    print(f"[policy] {msg}", flush=True)


def init_blob(mean: np.ndarray, std: np.ndarray, seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, and the given input standardisation."""
    rng = np.random.default_rng(seed)

    def he(fan_in, shape):
        lim = math.sqrt(6.0 / fan_in)
        return rng.uniform(-lim, lim, size=shape)

    parts = [he(MLP_IN, MLP_H * MLP_IN), np.zeros(MLP_H), he(MLP_H, MLP_H * MLP_H), np.zeros(MLP_H),
             he(MLP_H, MLP_OUT * MLP_H), np.zeros(MLP_OUT), mean, std]
    blob = np.concatenate(parts).astype(np.float32)
    assert blob.shape == (MLP_BLOB,)
    return blob


# ---- training graphs ---------------------------------------------------------------------
def training_graphs(count: int, seed: int = 7, max_batch_graphs: int = 8192):
    """`count` seeded specs: R-MAT (scale 14-20, edge factor 8-32), grid digraphs (side
    128-1024), planted-SCC graphs and packed batches (128-8192 R-MAT / planted graphs, whose
    merged searches meet states no single graph shows: most unvisited vertices lie in graphs
    whose search has ended; the largest make the bottom-up sweeps of the CC search pay where
    the small ones do not), in a fixed interleaved order.  ``max_batch_graphs`` caps the
    batches at a power of two (the shipped SCC network's data: 1024)."""
    top = int(math.log2(max_batch_graphs)) + 1
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        kind = ("rmat", "rmat", "grid", "rmat", "grid", "planted", "batch")[i % 7]
        s = int(rng.integers(1, 1 << 30))
        if kind == "batch":
            specs.append(("batch", {"graphs": int(2 ** rng.integers(7, top)), "median_n": int(2 ** rng.integers(10, 13)),
                                    "seed": s}))
        elif kind == "rmat":
            specs.append(("rmat", {"scale": int(rng.integers(14, 21)), "edge_factor": int(rng.choice([8, 12, 16, 24, 32])),
                                   "seed": s}))
        elif kind == "grid":
            side = int(round(2 ** (rng.integers(14, 21) / 2)))
            specs.append(("grid", {"rows": side, "cols": side, "p_one": float(rng.uniform(0.4, 0.6)),
                                   "p_both": float(rng.uniform(0.05, 0.15)),
                                   "arc_frac": float(rng.uniform(0.002, 0.01)), "seed": s}))
        else:
            specs.append(("planted", {"n": int(2 ** rng.integers(14, 19)), "max_size": int(2 ** rng.integers(8, 13)),
                                      "seed": s}))
    return specs


def build(spec):
    """The spec's graph; a batch spec gives the packed Graph with `.batch` = (vert_base, pivots)."""
    kind, kw = spec
    if kind == "batch":
        g, base, _ = gen_batch(kw["graphs"], median_n=kw["median_n"], sigma=1.4, min_n=256, max_n=200_000,
                               seed=kw["seed"])
        g.batch = (base, batch_pivots(g, base))
        return g
    if kind == "rmat":
        return gen_rmat(kw["scale"], kw["edge_factor"], seed=kw["seed"])
    if kind == "grid":
        return gen_grid(kw["rows"], kw["cols"], kw["p_one"], kw["p_both"], kw["arc_frac"], seed=kw["seed"])
    return gen_planted(kw["n"], max_size=kw["max_size"], seed=kw["seed"])[0]


# ---- schedule search ---------------------------------------------------------------------
class Runner:
    """Times decompositions of one device-resident graph."""

    def __init__(self, ctx: Context, g, kind: str):
        self.ctx, self.g, self.kind = ctx, g, kind
        self.dg = DeviceGraph(ctx, g)
        self.out = DeviceLabels(ctx, g.n)
        self.batch = getattr(g, "batch", None)
        if self.batch is not None:
            self.fn = run_scc_batch if kind == "scc" else run_cc_batch
        else:
            self.fn = run_scc if kind == "scc" else run_cc

    def run(self, schedule=None, policy=POLICY_PLAIN, record=False):
        kw = dict(ctx=self.ctx, out=self.out, schedule=schedule, policy=policy, record=record, stats=True,
                  labels=LABELS_CANONICAL)
        if self.batch is not None:
            # SCC takes the batch's create_csr pivots; CC applies its own rule per graph
            base, pivots = self.batch
            _, st = self.fn(self.dg, base, pivots if self.kind == "scc" else None, **kw)
        else:
            piv = self.g.max_deg_vert if self.kind == "scc" else None
            avg = self.g.avg_degree if self.kind == "scc" else None
            _, st = self.fn(self.dg, piv, avg, **kw)
        recs = self.ctx.levels() if record else None
        return st.wall_ms, recs

    def time(self, schedule=None, policy=POLICY_PLAIN, reps=3):
        return min(self.run(schedule, policy)[0] for _ in range(reps))

    def free(self):
        self.out.free()
        self.dg.free()


def phase_schedules(nlev: int):
    """Candidate forced action lists for one search whose plain run took nlev levels."""
    L = SCHED_LEVELS
    cands = [("all_td", [0] * L)]
    kmax = max(1, min(nlev, 6))
    for k in range(1, kmax + 1):
        for j in list(range(k + 1, min(nlev + 2, k + 5) + 1)) + [None]:
            acts = [0 if lv < k or (j is not None and lv >= j) else 1 for lv in range(1, L + 1)]
            cands.append((f"bu{k}-{j}", acts))
    for s in (1, 2, 3):
        if s <= max(nlev, 1):
            cands.append((f"stop{s}", [-1] * (s - 1) + [ACT_STOP]))
    return cands


MIN_GAIN = 0.03  # a forced schedule must beat the plain one by this fraction to be a label


def search_phase(runner: Runner, phase: int, nlev: int, plain_ms: float, reps: int):
    best = ("plain", None, plain_ms)
    for name, acts in phase_schedules(nlev):
        kw = {PHASES[phase]: acts}
        sched = schedule_array(**kw)
        ms = runner.time(sched, reps=reps)
        if ms < best[2]:
            best = (name, sched, ms)
    if best[2] > plain_ms * (1.0 - MIN_GAIN):  # within timing noise of the plain schedule
        return ("plain", None, plain_ms)
    return best


def samples_from(plain_recs, best_recs, phase):
    """(features, labels) of one search: the best run's levels, plus the plain run's levels
    up to the first departure from the best run's actions."""
    br = [r for r in best_recs if r.phase == phase]
    pr = [r for r in plain_recs if r.phase == phase]
    X = [r.feat for r in br]
    Y = [r.action for r in br]
    for i, r in enumerate(pr):
        if i >= len(br):
            break
        if r.action != br[i].action:
            X.append(r.feat)
            Y.append(br[i].action)
            break
    return X, Y


def collect(ctx: Context, specs, reps: int = 3, kinds=("scc", "cc")):
    """Runs the schedule search over the specs; returns X, Y, W, per-group metadata."""
    X, Y, W, meta = [], [], [], []
    t_all = time.time()
    for gi, spec in enumerate(specs):
        t0 = time.time()
        g = build(spec)
        for kind in kinds:
            runner = Runner(ctx, g, kind)
            runner.run()  # warm the workspace
            plain_ms = runner.time(reps=reps)
            _, plain_recs = runner.run(record=True)
            phases = (0, 1) if kind == "scc" else (2,)
            for phase in phases:
                nlev = sum(1 for r in plain_recs if r.phase == phase)
                if nlev == 0:
                    continue
                share = sum(r.level_us for r in plain_recs if r.phase == phase) * 1e-3 / plain_ms
                if share < 0.03:
                    # the search is noise in this run: keep the plain schedule, no search
                    name, sched, best_ms = "plain(skip)", None, plain_ms
                else:
                    name, sched, best_ms = search_phase(runner, phase, nlev, plain_ms, reps)
                _, best_recs = runner.run(schedule=sched, record=True) if sched is not None else (None, plain_recs)
                xs, ys = samples_from(plain_recs, best_recs, phase)
                # each (graph, search) weighs its share of a run plus the time the best
                # schedule saves; its samples split that weight
                gain = max(plain_ms - best_ms, 0.0) / plain_ms
                wgt = (0.3 * min(share, 1.0) + 0.05 + gain) / max(len(xs), 1)
                X += xs
                Y += ys
                W += [wgt] * len(xs)
                meta.append({"graph": gi, "spec": spec, "kind": kind, "phase": phase, "levels": nlev, "best": name,
                             "plain_ms": plain_ms, "best_ms": best_ms, "samples": len(xs)})
            runner.free()
        log(f"graph {gi + 1}/{len(specs)} {spec[0]} n={g.n} m={g.m}: "
            + ", ".join(f"{PHASES[m['phase']]} best {m['best']} {m['plain_ms']:.2f}->{m['best_ms']:.2f} ms"
                        for m in meta if m["graph"] == gi)
            + f" ({time.time() - t0:.1f}s, total {time.time() - t_all:.0f}s)")
        del g
    return (np.asarray(X, np.float32).reshape(-1, MLP_IN), np.asarray(Y, np.int32), np.asarray(W, np.float32),
            meta)


def fit(ctx: Context, X, Y, W, epochs=300, batch=256, lr=3e-3, seed=1):
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std[std < 1e-6] = 1.0
    blob = init_blob(mean, std, seed)
    blob, curve = mlp_train(ctx, blob, X, Y, W, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return blob, curve


def evaluate(ctx: Context, blob, specs, reps: int = 3, kinds=("scc", "cc"), blob_cc=None):
    """Plain vs learned wall time (and identical labels) on each spec."""
    set_policy(ctx, blob, blob_cc)
    rows = []
    for spec in specs:
        g = build(spec)
        for kind in kinds:
            runner = Runner(ctx, g, kind)
            runner.run()
            plain = runner.time(reps=reps)
            lab_plain = runner.out.to_host()
            runner.run(policy=POLICY_LEARNED)
            learned = runner.time(policy=POLICY_LEARNED, reps=reps)
            lab_learned = runner.out.to_host()
            same = bool(np.array_equal(lab_plain, lab_learned))
            rows.append({"spec": spec, "kind": kind, "n": g.n, "m": g.m, "plain_ms": plain, "learned_ms": learned,
                         "speedup": plain / learned, "identical_labels": same})
            log(f"eval {spec[0]} {kind} n={g.n}: plain {plain:.3f} ms, learned {learned:.3f} ms "
                f"({plain / learned:.2f}x), labels identical: {same}")
            runner.free()
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m connectivity_amd.policy")
    ap.add_argument("cmd", choices=["collect", "train", "eval", "all", "combine"])
    ap.add_argument("--scc", help="combine: weights whose network decides the SCC searches")
    ap.add_argument("--cc", help="combine: weights whose network decides the CC search")
    ap.add_argument("--graphs", type=int, default=120)
    ap.add_argument("--max-batch-graphs", type=int, default=8192, help="largest training batch (a power of two)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--data", default="policy_data.npz")
    ap.add_argument("--out", default=POLICY_WEIGHTS)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--seed", type=int, default=1, help="initialisation and minibatch order of the training")
    ap.add_argument("--eval-specs", default="heldout")
    args = ap.parse_args(argv)
    ctx = Context(0) if args.cmd != "combine" else None  # combining two files needs no GPU
    if args.cmd in ("collect", "all"):
        specs = training_graphs(args.graphs, max_batch_graphs=args.max_batch_graphs)
        X, Y, W, meta = collect(ctx, specs, reps=args.reps)
        os.makedirs(os.path.dirname(os.path.abspath(args.data)), exist_ok=True)
        np.savez(args.data, X=X, Y=Y, W=W)
        with open(os.path.splitext(args.data)[0] + ".json", "w") as f:
            json.dump(meta, f, indent=1)
        log(f"collected {len(Y)} samples; label counts {np.bincount(Y, minlength=3).tolist()}")
    if args.cmd in ("train", "all"):
        with np.load(args.data, allow_pickle=False) as z:
            X, Y, W = z["X"], z["Y"], z["W"]
        t0 = time.time()
        blob, curve = fit(ctx, X, Y, W, epochs=args.epochs, seed=args.seed)
        acc = float(np.mean(np.argmax(_logits(ctx, blob, X), axis=1) == Y))
        log(f"trained on {len(Y)} samples in {time.time() - t0:.1f}s: loss {curve[0]:.4f} -> {curve[-1]:.4f}, "
            f"train accuracy {acc:.3f}")
        np.savez(args.out, blob=blob, curve=curve, features=np.array(FEATURES))
        log(f"wrote {args.out}")
    if args.cmd == "combine":
        # one file, two networks: msc_policy_set_weights then takes the first for the SCC
        # searches and the second for the CC search
        with np.load(args.scc, allow_pickle=False) as a, np.load(args.cc, allow_pickle=False) as b:
            np.savez(args.out, blob=a["blob"], curve=a["curve"], blob_cc=b["blob"], curve_cc=b["curve"],
                     features=np.array(FEATURES))
        log(f"wrote {args.out}: SCC network from {args.scc}, CC network from {args.cc}")
    if args.cmd in ("eval", "all"):
        with np.load(args.out, allow_pickle=False) as z:
            blob = z["blob"]
            blob_cc = z["blob_cc"] if "blob_cc" in z.files else None
        specs = heldout_specs()
        rows = evaluate(ctx, blob, specs, reps=args.reps, blob_cc=blob_cc)
        with open(os.path.splitext(args.data)[0] + "_eval.json", "w") as f:
            json.dump(rows, f, indent=1)
    if ctx is not None:
        ctx.close()
    return 0


def _logits(ctx, blob, X):
    from . import mlp_forward

    return mlp_forward(ctx, blob, X)


def heldout_specs():
    """Graphs outside the training seeds: R-MAT s18-s22, grids 1024^2 / 2048^2, a planted
    1 M graph and a batch of 2048 graphs."""
    return [("rmat", {"scale": 18, "edge_factor": 16, "seed": 424242}),
            ("rmat", {"scale": 20, "edge_factor": 16, "seed": 424243}),
            ("rmat", {"scale": 22, "edge_factor": 16, "seed": 424244}),
            ("grid", {"rows": 1024, "cols": 1024, "p_one": 0.5, "p_both": 0.1, "arc_frac": 0.005, "seed": 424245}),
            ("grid", {"rows": 2048, "cols": 2048, "p_one": 0.5, "p_both": 0.1, "arc_frac": 0.005, "seed": 424246}),
            ("planted", {"n": 1 << 20, "max_size": 4096, "seed": 424247}),
            ("batch", {"graphs": 2048, "median_n": 4096, "seed": 424248})]


if __name__ == "__main__":
    sys.exit(main())
