"""GPU parity at the benchmark sizes: directed R-MAT scale 24 (16.8 M vertices, 263 M
edges), the 4096 x 4096 grid digraph and the batch of 8192 graphs (85 M vertices, 935 M
edges), SCC and WCC, plain and learned policies, against the oracle.

Raw labels must equal the oracle's run_scc with vert_cutoff 0 (multistep/scc_main.cpp:
210-317: trimmed vertex -> itself, pivot's SCC -> pivot, other SCCs -> max id) and run_cc
(multistep/cc_main.cpp:184-245); canonical labels must equal the oracle's partition with
each component labelled by its minimum vertex.  Every graph of the batch is checked against
the oracle run on that graph alone, in its own ids.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _say(msg):
    print(msg, flush=True)


def _first_bad(got, want):
    bad = np.flatnonzero(got != want)
    return f"{bad.size} labels differ, first at {bad[:5]}: {got[bad[:5]]} vs {want[bad[:5]]}"


class Single:
    """One graph in HBM with labels kept in HBM, and the oracle's raw labels."""

    def __init__(self, msc, orc, ctx, name, g):
        self.msc, self.ctx, self.name, self.g = msc, ctx, name, g
        _say(f"full size: {name} n={g.n} m={g.m} pivot={g.max_deg_vert}")
        self.scc_raw = orc.scc(g, g.max_deg_vert, g.avg_degree, vert_cutoff=0)
        self.scc_canon = orc.canonical(self.scc_raw)
        self.cc_raw = orc.cc(g)
        self.cc_canon = orc.canonical(self.cc_raw)
        self.dg = msc.DeviceGraph(ctx, g)
        self.out = msc.DeviceLabels(ctx, g.n)

    def scc(self, policy, labels):
        self.msc.run_scc(self.dg, self.g.max_deg_vert, self.g.avg_degree, ctx=self.ctx, out=self.out,
                         policy=policy, labels=labels)
        return self.out.to_host()

    def cc(self, policy, labels):
        self.msc.run_cc(self.dg, ctx=self.ctx, out=self.out, policy=policy, labels=labels)
        return self.out.to_host()

    def close(self):
        self.out.free()
        self.dg.free()


@pytest.fixture(scope="module", params=["rmat24", "grid4096"])
def single(request, msc, orc, gpu_ctx):
    if request.param == "rmat24":
        g = msc.gen_rmat(24, 16, seed=1)
    else:
        g = msc.gen_grid(4096, 4096, 0.5, 0.1, 0.005, seed=1)
    msc.load_policy(gpu_ctx)
    w = Single(msc, orc, gpu_ctx, request.param, g)
    yield w
    w.close()


@pytest.mark.parametrize("policy", ["plain", "learned"])
def test_full_size_scc(msc, single, policy):
    pol = msc.POLICY_LEARNED if policy == "learned" else msc.POLICY_PLAIN
    _say(f"case: {single.name} SCC {policy} raw")
    got = single.scc(pol, msc.LABELS_RAW)
    assert np.array_equal(got, single.scc_raw), _first_bad(got, single.scc_raw)
    _say(f"case: {single.name} SCC {policy} canonical")
    got = single.scc(pol, msc.LABELS_CANONICAL)
    assert np.array_equal(got, single.scc_canon), _first_bad(got, single.scc_canon)


def test_full_size_scc_searches_in_turn(msc, gpu_ctx, single):
    """FW then BW restricted to the FW set (the reference's order) gives the same labels."""
    defaults = gpu_ctx.tuning()
    try:
        gpu_ctx.set_tuning(concurrent_fwbw=0)
        _say(f"case: {single.name} SCC plain raw, searches in turn")
        got = single.scc(msc.POLICY_PLAIN, msc.LABELS_RAW)
    finally:
        gpu_ctx.set_tuning(**defaults)
    assert np.array_equal(got, single.scc_raw), _first_bad(got, single.scc_raw)


@pytest.mark.parametrize("policy", ["plain", "learned"])
def test_full_size_wcc(msc, single, policy):
    pol = msc.POLICY_LEARNED if policy == "learned" else msc.POLICY_PLAIN
    _say(f"case: {single.name} WCC {policy} raw")
    got = single.cc(pol, msc.LABELS_RAW)
    assert np.array_equal(got, single.cc_raw), _first_bad(got, single.cc_raw)
    _say(f"case: {single.name} WCC {policy} canonical")
    got = single.cc(pol, msc.LABELS_CANONICAL)
    assert np.array_equal(got, single.cc_canon), _first_bad(got, single.cc_canon)


def _unpack(msc, u, base, i):
    """Graph i of a packed batch, in its own ids."""
    b0, b1 = int(base[i]), int(base[i + 1])
    e0, e1 = int(u.out_off[b0]), int(u.out_off[b1])
    f0, f1 = int(u.in_off[b0]), int(u.in_off[b1])
    return msc.Graph(n=b1 - b0, out_off=(u.out_off[b0:b1 + 1] - u.out_off[b0]).astype(np.uint32),
                     out_adj=(u.out_adj[e0:e1] - b0).astype(np.int32),
                     in_off=(u.in_off[b0:b1 + 1] - u.in_off[b0]).astype(np.uint32),
                     in_adj=(u.in_adj[f0:f1] - b0).astype(np.int32))


class Batch:
    """The 8192-graph batch in HBM and the oracle's labels of every graph, run alone."""

    def __init__(self, msc, orc, ctx):
        self.msc, self.ctx = msc, ctx
        self.g, self.base, _ = msc.gen_batch(8192, seed=1)
        self.pivots = msc.batch_pivots(self.g, self.base)
        g, base = self.g, self.base
        _say(f"full size: batch of {len(base) - 1} graphs, n={g.n} m={g.m}")
        n = g.n
        self.scc_raw = np.empty(n, np.int32)
        self.scc_canon = np.empty(n, np.int32)
        self.cc_raw = np.empty(n, np.int32)
        self.cc_canon = np.empty(n, np.int32)

        def one(i):
            orc.set_threads(1)
            h = _unpack(msc, g, base, i)
            b0, b1 = int(base[i]), int(base[i + 1])
            r = orc.scc(h, vert_cutoff=0)
            self.scc_raw[b0:b1] = r
            self.scc_canon[b0:b1] = orc.canonical(r)
            c = orc.cc(h)
            self.cc_raw[b0:b1] = c
            self.cc_canon[b0:b1] = orc.canonical(c)

        workers = min(16, os.cpu_count() or 4)
        with cf.ThreadPoolExecutor(workers) as ex:
            list(ex.map(one, range(len(base) - 1)))
        _say("full size: batch oracle done")
        self.dg = msc.DeviceGraph(ctx, g)
        self.out = msc.DeviceLabels(ctx, n)

    def check(self, got, want):
        if np.array_equal(got, want):
            return None
        bad = np.flatnonzero(got != want)
        gi = int(np.searchsorted(self.base, bad[0], side="right") - 1)
        return f"{bad.size} labels differ; first in graph {gi} (n={int(self.base[gi + 1] - self.base[gi])})"

    def close(self):
        self.out.free()
        self.dg.free()


@pytest.fixture(scope="module")
def batch(msc, orc, gpu_ctx):
    msc.load_policy(gpu_ctx)
    b = Batch(msc, orc, gpu_ctx)
    yield b
    b.close()


@pytest.mark.parametrize("policy", ["plain", "learned"])
def test_batch8192_scc_every_graph(msc, gpu_ctx, batch, policy):
    pol = msc.POLICY_LEARNED if policy == "learned" else msc.POLICY_PLAIN
    for labels, want in ((msc.LABELS_RAW, batch.scc_raw), (msc.LABELS_CANONICAL, batch.scc_canon)):
        _say(f"case: batch8192 SCC {policy} labels={labels}")
        msc.run_scc_batch(batch.dg, batch.base, batch.pivots, ctx=gpu_ctx, out=batch.out, policy=pol,
                          labels=labels)
        err = batch.check(batch.out.to_host(), want)
        assert err is None, err


def test_batch8192_wcc_every_graph(msc, gpu_ctx, batch):
    for labels, want in ((msc.LABELS_RAW, batch.cc_raw), (msc.LABELS_CANONICAL, batch.cc_canon)):
        _say(f"case: batch8192 WCC labels={labels}")
        msc.run_cc_batch(batch.dg, batch.base, ctx=gpu_ctx, out=batch.out, labels=labels)
        err = batch.check(batch.out.to_host(), want)
        assert err is None, err
