"""Top-down load balance variants (msc_tuning.balance), the Kokkos driver's `alg` selector
(kokkos_scc/scc_fwbw.cpp:337-364, kokkos_scc/scc_color.cpp:579-655): 0 a thread per
frontier vertex (fwbw_baseline.cpp), 1 a workgroup deals the edges of its 256 frontier
vertices (fwbw_manhattan_local.cpp), 2 fixed-size chunks of the whole frontier's edges
(fwbw_manhattan_global.cpp), 3 (the default) 1 or 2 per level by the frontier's mean list
length.  The reference's three variants compute the same labels; here every variant's raw
and canonical SCC labels and CC labels must equal the oracle's, with the searches side by
side and in turn, on graphs whose frontiers hold hub lists longer than a workgroup's 256
slots and than one 1024-edge chunk."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _hubs(msc, n=6000, seed=5):
    """Cycles through two hubs with 3000-entry lists plus random arcs: a top-down level then
    expands single lists that span many chunks and every slot of a block."""
    rng = np.random.default_rng(seed)
    src = [np.arange(n), np.zeros(3000, np.int64), rng.integers(0, n, 3000), rng.integers(0, n, 2 * n)]
    dst = [(np.arange(n) + 1) % n, rng.integers(0, n, 3000), np.ones(3000, np.int64), rng.integers(0, n, 2 * n)]
    return msc.create_csr(n, np.concatenate(src), np.concatenate(dst))


def _graphs(msc):
    return {
        "rmat14": msc.gen_rmat(14, 16, seed=7),
        "grid128": msc.gen_grid(128, 128, seed=3),
        "planted": msc.gen_planted(1 << 14, max_size=512, seed=4)[0],
        "hubs": _hubs(msc),
    }


@pytest.fixture(scope="module")
def graphs(msc):
    return _graphs(msc)


@pytest.mark.parametrize("concurrent", [1, 0])
@pytest.mark.parametrize("balance", [0, 1, 2, 3])
def test_balance_keeps_labels(msc, orc, gpu_ctx, graphs, balance, concurrent):
    defaults = gpu_ctx.tuning()
    try:
        got_t = gpu_ctx.set_tuning(balance=balance, concurrent_fwbw=concurrent)
        assert got_t["balance"] == balance
        for name, g in graphs.items():
            _say(f"case: balance {balance} concurrent {concurrent} {name} n={g.n} m={g.m}")
            piv, avg = orc.pivot_scc(g)
            want = orc.scc(g, piv, avg, vert_cutoff=0)
            got = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, f"{name}: raw SCC labels differ at {bad.size} vertices, first {bad[:5]}"
            assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx), orc.canonical(want)), name
            cc = orc.cc(g)
            assert np.array_equal(msc.run_cc(g, labels=msc.LABELS_RAW, ctx=gpu_ctx), cc), name
    finally:
        gpu_ctx.set_tuning(**defaults)


@pytest.mark.parametrize("balance", [0, 1, 2])
def test_balance_batch(msc, orc, gpu_ctx, balance):
    """A packed batch of R-MAT and planted graphs: every graph's slice equals the oracle run
    on that graph alone."""
    g, base, _ = msc.gen_batch(48, seed=9)
    defaults = gpu_ctx.tuning()
    try:
        gpu_ctx.set_tuning(balance=balance)
        _say(f"case: balance {balance} batch of 48 n={g.n} m={g.m}")
        got = msc.run_scc_batch(g, base, labels=msc.LABELS_RAW, ctx=gpu_ctx)
        gotc = msc.run_cc_batch(g, base, labels=msc.LABELS_RAW, ctx=gpu_ctx)
    finally:
        gpu_ctx.set_tuning(**defaults)
    for i in range(len(base) - 1):
        b0, b1 = int(base[i]), int(base[i + 1])
        e0, e1 = int(g.out_off[b0]), int(g.out_off[b1])
        f0, f1 = int(g.in_off[b0]), int(g.in_off[b1])
        h = msc.Graph(n=b1 - b0, out_off=(g.out_off[b0:b1 + 1] - g.out_off[b0]).astype(np.uint32),
                      out_adj=(g.out_adj[e0:e1] - b0).astype(np.int32),
                      in_off=(g.in_off[b0:b1 + 1] - g.in_off[b0]).astype(np.uint32),
                      in_adj=(g.in_adj[f0:f1] - b0).astype(np.int32))
        assert np.array_equal(got[b0:b1], orc.scc(h, vert_cutoff=0)), f"graph {i}"
        assert np.array_equal(gotc[b0:b1], orc.cc(h)), f"graph {i}"


def test_balance_learned_policy(msc, orc, gpu_ctx, graphs):
    """The learned schedule (bottom-up levels between top-down ones) with the slot variants."""
    msc.load_policy(gpu_ctx)
    g = graphs["rmat14"]
    want = orc.scc(g, vert_cutoff=0)
    defaults = gpu_ctx.tuning()
    try:
        for balance in (0, 1, 2):
            gpu_ctx.set_tuning(balance=balance)
            _say(f"case: balance {balance} learned rmat14")
            got = msc.run_scc(g, labels=msc.LABELS_RAW, ctx=gpu_ctx, policy=msc.POLICY_LEARNED)
            assert np.array_equal(got, want)
    finally:
        gpu_ctx.set_tuning(**defaults)
