"""Benchmark: Multistep SCC on a directed R-MAT scale-24 graph on one MI355X.

One step = one msc_scc_run over the whole graph (trim -> FW-BW -> colouring to empty,
canonical labels), with the CSR / transposed CSR already resident in HBM and the labels
left in HBM.  `value` is GTEPS = m / t_step (m = directed edges after dedupe) over K timed
steps bracketed by stream synchronisation, with the learned direction policy when its
weights are shipped (connectivity_amd/policy_weights.npz), else the plain rule; both are
timed in the same run and reported with their labels checked identical.

The line also carries:
  roofline      -- dominant kernel class: algorithmic bytes / summed HIP-event durations of
                   its launches (events recorded on the library's stream around every
                   launch, in a second pass of the same K steps, so that `value` carries no
                   event overhead), against 8 TB/s; `traffic` comes from the committed
                   rocprofv3 FETCH_SIZE / WRITE_SIZE summary when present;
  cpu_baseline  -- the reference's own multistep run_scc (oracle/_ref, compiled from its
                   sources) or, where that was not built, the oracle's restatement, timed on
                   this host's cores on the same graph;
  wcc_ms        -- msc_cc_run on the same graph with the pivot and average degree cc_main's
                   create_csr gives (not part of `value`); wcc_parity: its canonical labels
                   against the oracle's.

Workloads: rmat (scale 24, the metric's), grid (4096 x 4096 grid digraph), batch (8192
packed graphs, msc_scc_run_batch).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmat|grid|batch] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SCC wall time + GTEPS, R-MAT s24, 1 MI355X; learned-policy vs plain speedup"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_graph(args, msc):
    if args.workload == "batch":
        g, base, kinds = msc.gen_batch(args.graphs, seed=args.seed)
        desc = (f"synthetic batch of {args.graphs} digraphs, n log-normal (median 4096, sigma 1.4, clipped to "
                f"[256, 500000]), half R-MAT (edge factor 8-32), half planted Zipf(1.5) SCCs joined by 4n DAG "
                f"arcs, packed into one CSR, seed {args.seed}")
        cfg = {"workload": f"batch{args.graphs}_scc", "graphs": args.graphs, "rmat_graphs": int((kinds == 0).sum())}
        return g, desc, cfg, base
    if args.workload == "grid":
        g = msc.gen_grid(args.side, args.side, 0.5, 0.1, 0.005, seed=args.seed)
        desc = (f"synthetic {args.side}x{args.side} grid digraph, edge kept one-way p=0.5 / both p=0.1, "
                f"+0.5% n random arcs, seed {args.seed}")
        cfg = {"workload": f"grid{args.side}_scc", "rows": args.side, "cols": args.side}
    else:
        g = msc.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
        desc = (f"synthetic directed R-MAT (a,b,c,d)=(.57,.19,.19,.05), self-loops/duplicates removed, "
                f"ids permuted, seed {args.seed}")
        cfg = {"workload": f"rmat{args.scale}_scc", "scale": args.scale, "edge_factor": args.edge_factor}
    return g, desc, cfg, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["rmat", "grid", "batch"], default="rmat")
    ap.add_argument("--graphs", type=int, default=8192, help="graphs in the batch workload")
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--policy", choices=["auto", "plain", "learned"], default="auto")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-check", action="store_true", help="skip the oracle parity check")
    ap.add_argument("--check-graphs", type=int, default=64, help="batch graphs checked against the oracle")
    ap.add_argument("--no-compare", action="store_true",
                    help="run the reported policy only (no plain/learned comparison; for profiling)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU work budget of the batch baseline")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run one device-to-device copy of this size first (PMC traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the decomposition runs on a single device (a batch's graphs "
                                   "would be dealt to devices whole; one graph would need cross-device frontiers, "
                                   "which are not built)", "n_gpus": args.gpus}))
        return 2

    import numpy as np

    import connectivity_amd as msc

    if msc.device_count() < 1:
        log("no gfx950 device visible")
        return 1
    t0 = time.time()
    g, desc, cfg, base = build_graph(args, msc)
    # every graph's pivot comes with the batch, as max_deg_vert comes with a single graph
    # (create_csr computes it, untimed)
    pivots = msc.batch_pivots(g, base) if base is not None else None
    log(f"{cfg['workload']}: n={g.n} m={g.m} pivot={g.max_deg_vert} built in {time.time() - t0:.1f}s")
    ctx = msc.Context(0)
    if args.calib_copy_mb:
        nbytes = args.calib_copy_mb << 20
        a, b = msc.DeviceLabels(ctx, nbytes // 4), msc.DeviceLabels(ctx, nbytes // 4)
        msc._lib().msc_copy_device(ctx.handle, b.ptr, a.ptr, nbytes)
        log(f"calibration copy of {nbytes} bytes done")
        a.free()
        b.free()
    dg = msc.DeviceGraph(ctx, g)
    out = msc.DeviceLabels(ctx, g.n)
    have_weights = os.path.exists(msc.POLICY_WEIGHTS)
    if have_weights:
        msc.load_policy(ctx)
    policy_name = args.policy if args.policy != "auto" else ("learned" if have_weights else "plain")
    if policy_name == "learned" and not have_weights:
        log("no policy weights shipped")
        return 1
    policy = msc.POLICY_LEARNED if policy_name == "learned" else msc.POLICY_PLAIN

    def step(pol, timing=True):
        if base is not None:
            _, st = msc.run_scc_batch(dg, base, pivots, ctx=ctx, out=out, policy=pol, timing=timing, stats=True)
        else:
            _, st = msc.run_scc(dg, g.max_deg_vert, g.avg_degree, ctx=ctx, out=out, policy=pol, timing=timing,
                                stats=True)
        return st

    def prepared(pol):
        # the arguments marshalled once: a step is one C ABI call, no statistics
        if base is not None:
            return msc.Prepared("scc", dg, policy=pol, ctx=ctx, out=out, batch=(base, pivots))
        return msc.Prepared("scc", dg, g.max_deg_vert, g.avg_degree, policy=pol, ctx=ctx, out=out)

    def timed(pol, steps, timing):
        """ms per step over `steps` steps between device synchronisations; with `timing`,
        HIP events around every launch and the per-step statistics (a second pass)."""
        msc._lib().msc_ctx_sync(ctx.handle)
        stats = []
        call = None if timing else prepared(pol)
        t_start = time.perf_counter()
        for _ in range(steps):
            if call is None:
                stats.append(step(pol, timing))
            else:
                call()
        msc._lib().msc_ctx_sync(ctx.handle)
        return (time.perf_counter() - t_start) * 1e3 / steps, stats

    for pol in ({policy} if args.no_compare else {policy, msc.POLICY_PLAIN}):
        for _ in range(args.warmup):
            step(pol)
    log(f"warmup {args.warmup} done")
    # the timed region: K steps of the reported policy; then the same K steps again with HIP
    # events around every launch, for the per-kernel durations of the roofline
    ms_per_step, _ = timed(policy, args.steps, False)
    labels = out.to_host()
    ms_events, stats = timed(policy, args.steps, True)
    ms_noevents = ms_per_step
    log(f"timed {args.steps} steps ({policy_name}): {ms_per_step:.3f} ms/step; {ms_events:.3f} with kernel events")
    plain_ms = ms_per_step
    learned = None
    if have_weights and not args.no_compare:
        other = msc.POLICY_PLAIN if policy == msc.POLICY_LEARNED else msc.POLICY_LEARNED
        other_ms, _ = timed(other, args.steps, False)
        other_labels = out.to_host()
        same = bool(np.array_equal(other_labels, labels))
        p_ms = other_ms if policy == msc.POLICY_LEARNED else ms_noevents
        l_ms = ms_noevents if policy == msc.POLICY_LEARNED else other_ms
        plain_ms = p_ms
        learned = {"plain_ms": round(p_ms, 4), "learned_ms": round(l_ms, 4), "speedup": round(p_ms / l_ms, 3),
                   "identical_labels": same}
        log(f"plain {p_ms:.3f} ms, learned {l_ms:.3f} ms: {p_ms / l_ms:.2f}x, labels identical: {same}")
        if not same:
            return 1

    # the plain rule with the searches in the reference's order (FW, then BW restricted to the
    # FW set: scc_fwbw.cpp:469-513), which the plain schedule of the reference follows level by
    # level; the default runs BW beside FW over the whole valid set
    ref_order = None
    if not args.no_compare:
        defaults = ctx.tuning()
        try:
            ctx.set_tuning(concurrent_fwbw=0)
            step(msc.POLICY_PLAIN, timing=False)
            ro_ms, _ = timed(msc.POLICY_PLAIN, args.steps, False)
            ro_same = bool(np.array_equal(out.to_host(), labels))
        finally:
            ctx.set_tuning(**defaults)
        ref_order = {"plain_ms": round(ro_ms, 4), "value_speedup": round(ro_ms / ms_per_step, 3),
                     "identical_labels": ro_same}
        log(f"plain rule, searches in the reference's order: {ro_ms:.3f} ms ({ro_ms / ms_per_step:.2f}x the "
            f"reported run), labels identical: {ro_same}")
        if not ro_same:
            return 1

    # per-kernel-class event time and algorithmic bytes summed over the timed steps
    agg = {}
    for st in stats:
        for name, k in st.kernels.items():
            a = agg.setdefault(name, {"launches": 0, "ms": 0.0, "bytes": 0.0})
            a["launches"] += k["launches"]
            a["ms"] += k["ms"]
            a["bytes"] += k["bytes"]
    dom = max(agg, key=lambda k: agg[k]["ms"])
    d = agg[dom]
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0
    s0 = stats[-1]
    log("kernels: " + ", ".join(f"{k}: {v['launches']} launches {v['ms']:.2f} ms {v['bytes'] / 1e9:.3f} GB"
                                for k, v in agg.items() if v["launches"]))
    log(f"phases: trim {s0.num_trim}, fwbw {s0.num_fwbw}, colour rounds {s0.color_rounds}, fw levels "
        f"{s0.fw_levels}, bw levels {s0.bw_levels}, bu levels {s0.bu_levels}, host syncs {s0.host_syncs}")

    check = "skipped"
    if not args.no_check:
        import oracle

        t1 = time.time()
        if base is not None:
            check = check_batch(msc, oracle, g, base, labels, args.check_graphs)
        else:
            want = oracle.canonical(oracle.scc(g, g.max_deg_vert, g.avg_degree, vert_cutoff=0))
            check = "identical" if np.array_equal(labels, want) else "MISMATCH"
        log(f"parity vs oracle (canonical labels): {check} ({time.time() - t1:.1f}s)")
        if not check.startswith("identical"):
            return 1
    counts = msc.scc_verify(labels) if base is None else None

    # WCC with run_cc's call shape: the pivot and average degree come with the graph
    # (cc_main.cpp's create_csr computes them, untimed), as they do for SCC
    cc_piv = msc.cc_pivot(dg, ctx=ctx) if base is None else (None, None)
    cc_ms = []
    for _ in range(max(2, args.steps // 2)):
        t1 = time.perf_counter()
        if base is not None:
            msc.run_cc_batch(dg, base, ctx=ctx, out=out, policy=policy)
        else:
            msc.run_cc(dg, cc_piv[0], cc_piv[1], ctx=ctx, out=out, policy=policy)
        cc_ms.append((time.perf_counter() - t1) * 1e3)
    cc_check = "skipped"
    if not args.no_check and base is None:
        import oracle

        cc_check = "identical" if np.array_equal(out.to_host(), oracle.canonical(oracle.cc(g))) else "MISMATCH"
        log(f"WCC parity vs oracle (canonical labels): {cc_check}")
        if cc_check != "identical":
            return 1

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(g, base, args.cpu_seconds) if base is not None else cpu_baseline(g, args.cpu_runs)

    gteps = g.m / (ms_per_step * 1e-3) / 1e9
    traffic = read_traffic(cfg["workload"], policy_name, dom)
    cfg.update({"n": g.n, "m": g.m, "pivot": g.max_deg_vert if base is None else "per graph",
                "policy": policy_name, "phases": "trim -> fw-bw -> colour to empty"})
    line = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": desc,
        "config": cfg,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launches": d["launches"], "avg_launch_us": round(d["ms"] * 1e3 / max(d["launches"], 1), 3),
                     "alg_bytes_per_launch": round(d["bytes"] / max(d["launches"], 1), 1)},
        "cpu_baseline": cpu,
        "ms_per_step_with_kernel_events": round(ms_events, 4),
        "plain_ms": round(plain_ms, 4),
        "learned": learned,
        "plain_reference_order": ref_order,
        "scc_counts": ({"num_sccs": counts[0], "nontrivial": counts[1], "max_scc": counts[2]}
                       if counts else None),
        "parity": check,
        "wcc_ms": round(statistics.median(cc_ms), 4),
        "wcc_parity": cc_check,
    }
    print(json.dumps(line), flush=True)
    return 0


def unpack(msc, u, base, i):
    """Graph i of a packed batch, in its own ids."""
    import numpy as np

    b0, b1 = int(base[i]), int(base[i + 1])
    e0, e1 = int(u.out_off[b0]), int(u.out_off[b1])
    f0, f1 = int(u.in_off[b0]), int(u.in_off[b1])
    return msc.Graph(n=b1 - b0, out_off=(u.out_off[b0:b1 + 1] - u.out_off[b0]).astype(np.uint32),
                     out_adj=(u.out_adj[e0:e1] - b0).astype(np.int32),
                     in_off=(u.in_off[b0:b1 + 1] - u.in_off[b0]).astype(np.uint32),
                     in_adj=(u.in_adj[f0:f1] - b0).astype(np.int32))


def check_batch(msc, oracle, u, base, labels, count):
    """Every graph's slice against the oracle run on that graph alone, for `count` graphs
    spread over the batch (the smallest and the largest graph always among them)."""
    import numpy as np

    k = len(base) - 1
    sizes = np.diff(base)
    pick = set(np.linspace(0, k - 1, min(count, k)).astype(int).tolist())
    pick |= {int(np.argmin(sizes)), int(np.argmax(sizes))}
    for i in sorted(pick):
        g = unpack(msc, u, base, i)
        want = oracle.canonical(oracle.scc(g, vert_cutoff=0))
        if not np.array_equal(labels[int(base[i]):int(base[i + 1])], want):
            log(f"graph {i} differs from the oracle")
            return "MISMATCH"
    return f"identical ({len(pick)} of {k} graphs checked)"


def cpu_baseline_batch(u, base, seconds: float):
    """The reference's run_scc over a bounded sample of the batch (graphs in batch order
    until about `seconds` of wall time), one single-threaded call per graph with the graphs
    spread over this host's cores; reported as edges per second of wall time.  Single
    threads per graph: with several threads the reference's racy enqueue (check-then-set of
    its visited flags) can queue a vertex twice and overrun queues sized num_valid on
    graphs whose pivot reaches nearly every vertex (it aborts on some planted graphs)."""
    import concurrent.futures as cf

    import oracle

    import connectivity_amd as msc

    cores = host_cores()
    use_ref = oracle.have_ref()

    def one(i):
        oracle.set_threads(1)
        g = unpack(msc, u, base, i)
        piv, avg = oracle.pivot_scc(g)
        if piv < 0:  # the reference's run_scc needs a pivot; such graphs stay out of the sample
            return 0, 0.0
        t = oracle.ref_time_scc(g, piv, avg) if use_ref else oracle.scc_time(g, piv, avg)
        return g.m, t

    edges, graphs, i, k = 0, 0, 0, len(base) - 1
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        while i < k and time.perf_counter() - t0 < seconds:
            chunk = list(range(i, min(k, i + 4 * cores)))
            for m, _ in ex.map(one, chunk):
                edges += m
                graphs += 1 if m else 0
            i += len(chunk)
    wall = time.perf_counter() - t0
    log(f"cpu baseline ({'reference' if use_ref else 'port'}, {cores} threads x 1): {graphs} graphs, "
        f"{edges} edges in {wall:.2f} s")
    return {"value": round(edges / wall / 1e9, 5), "unit": "GTEPS", "cores": cores,
            "kind": "reference" if use_ref else "port", "ms": round(wall * 1e3, 2), "cpu_model": cpu_model(),
            "sample": f"run_scc (multistep, vert_cutoff 10000) on the first {i} graphs of the batch ({graphs} with "
                      f"a pivot), one single-threaded call per graph, {cores} graphs at a time; edges / wall time"}


def read_traffic(workload: str, policy: str, kernel_class: str):
    """HBM bytes per launch of the dominant kernel from the committed PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{policy}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get("per_launch_bytes", {}).get(kernel_class)
    except (OSError, ValueError):
        return None


def cpu_model() -> str:
    """The host CPU's model name, with its family / model numbers (a virtual machine may
    report only a generic name) and the logical CPUs the machine has."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" in line:
                    k, v = (x.strip() for x in line.split(":", 1))
                    if k in ("model name", "cpu family", "model", "stepping") and k not in fields:
                        fields[k] = v
    except OSError:
        return "unknown"
    name = fields.get("model name", "unknown")
    ids = ", ".join(f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields)
    return f"{name} ({ids}; {os.cpu_count()} logical CPUs)" if ids else name


def host_cores() -> int:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return min(cores, 16)  # this host's share for one GPU


def cpu_time_child(g, threads: int, runs: int) -> dict:
    """oracle/cpu_time.py in a child process with OMP_NUM_THREADS = threads,
    OMP_PROC_BIND = close, OMP_PLACES = cores (libgomp reads them when it loads, so they
    cannot be changed inside this process, whose libgomp is already up); the graph goes to
    it as .npy files in memory-backed temporary storage."""
    import shutil
    import subprocess
    import tempfile

    import numpy as np

    tmp = tempfile.mkdtemp(prefix="msc_cpu_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        for k in ("out_off", "out_adj", "in_off", "in_adj"):
            np.save(os.path.join(tmp, f"{k}.npy"), np.asarray(getattr(g, k)))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_time", "--dir", tmp, "--pivot", str(g.max_deg_vert),
                            "--avg", repr(g.avg_degree), "--runs", str(runs)], cwd=ROOT, env=env,
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"cpu baseline child failed: {r.stderr[-400:]}")
        return json.loads(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(g, runs: int):
    """The reference's run_scc (oracle/_ref; the restatement if absent) timed on this
    host's cores with threads bound close to their cores, and once more on one thread."""
    cores = host_cores()
    full = cpu_time_child(g, cores, runs)
    one = cpu_time_child(g, 1, max(1, min(3, runs)))
    med, med1 = full["median_s"], one["median_s"]
    log(f"cpu baseline ({full['kind']}, {cores} threads, {cpu_model()}): median {med * 1e3:.1f} ms over {runs} "
        f"runs; 1 thread {med1 * 1e3:.1f} ms")
    return {"value": round(g.m / med / 1e9, 5), "unit": "GTEPS", "cores": cores, "kind": full["kind"],
            "ms": round(med * 1e3, 2), "cpu_model": cpu_model(),
            "omp": {"OMP_NUM_THREADS": cores, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"},
            "one_thread": {"value": round(g.m / med1 / 1e9, 5), "ms": round(med1 * 1e3, 2),
                           "runs": len(one["times_s"])},
            "sample": f"{runs} full run_scc calls (multistep, OpenMP, vert_cutoff 10000) on the same graph, median; "
                      f"each timed around the call as scc_main.cpp:368-378 does"}


if __name__ == "__main__":
    sys.exit(main())
