"""Batches: many graphs packed into one CSR and decomposed together.

CPU tests pin the packer and the batch generator (structure, determinism, edges that stay
inside their graph, the union's components equal the per-graph ones).  GPU tests check that
every graph's slice of msc_scc_run_batch / msc_cc_run_batch equals what the oracle gives for
that graph alone: raw labels bit for bit (vert_cutoff = 0), canonical labels too.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components


def _say(msg):
    print(msg, flush=True)


def unpack(msc, u, base, i):
    """Graph i of a packed batch, in its own ids."""
    b0, b1 = int(base[i]), int(base[i + 1])
    e0, e1 = int(u.out_off[b0]), int(u.out_off[b1])
    f0, f1 = int(u.in_off[b0]), int(u.in_off[b1])
    return msc.Graph(n=b1 - b0,
                     out_off=(u.out_off[b0:b1 + 1] - u.out_off[b0]).astype(np.uint32),
                     out_adj=(u.out_adj[e0:e1] - b0).astype(np.int32),
                     in_off=(u.in_off[b0:b1 + 1] - u.in_off[b0]).astype(np.uint32),
                     in_adj=(u.in_adj[f0:f1] - b0).astype(np.int32))


def mixed_graphs(msc, seed=0):
    """Graphs of every shape the batch path must handle, empty and edgeless ones included."""
    rng = np.random.default_rng(seed)
    gs = [msc.gen_rmat(10, 8, seed=seed + 1), msc.create_csr(0, [], []), msc.create_csr(1, [], []),
          msc.create_csr(4, [0, 0, 1], [1, 2, 3]),  # DAG: no pivot
          msc.gen_grid(20, 30, seed=seed + 2), msc.gen_planted(3000, max_size=300, seed=seed + 3)[0],
          msc.create_csr(3, [0, 1], [1, 0])]
    for k in range(12):
        n = int(rng.integers(1, 2500))
        m = int(n * rng.uniform(0.5, 4.0))
        gs.append(msc.create_csr(n, rng.integers(0, n, m), rng.integers(0, n, m)))
    return gs


# ---- CPU: packer and generator --------------------------------------------------------------
def test_pack_roundtrip(msc):
    gs = mixed_graphs(msc)
    u, base = msc.pack_graphs(gs)
    assert base[0] == 0 and base[-1] == u.n == sum(g.n for g in gs)
    assert u.m == sum(g.m for g in gs)
    assert np.all(np.diff(base) == [g.n for g in gs])
    for i, g in enumerate(gs):
        h = unpack(msc, u, base, i)
        for f in ("out_off", "out_adj", "in_off", "in_adj"):
            assert np.array_equal(getattr(h, f), getattr(g, f)), (i, f)


def test_gen_batch_structure(msc, orc):
    u, base, kinds = msc.gen_batch(40, median_n=600, sigma=1.0, min_n=64, max_n=5000, seed=9)
    u2, base2, kinds2 = msc.gen_batch(40, median_n=600, sigma=1.0, min_n=64, max_n=5000, seed=9)
    assert np.array_equal(base, base2) and np.array_equal(kinds, kinds2)
    assert np.array_equal(u.out_adj, u2.out_adj) and np.array_equal(u.in_adj, u2.in_adj)
    sizes = np.diff(base)
    assert sizes.min() >= 64 and sizes.max() <= 5000
    assert set(np.unique(kinds)) <= {0, 1} and 0 < kinds.sum() < 40
    src, dst = u.edges()
    g_of = np.searchsorted(base, np.arange(u.n), side="right") - 1
    assert np.array_equal(g_of[src], g_of[dst]), "an edge leaves its graph"
    assert np.all(src != dst)
    # components of the union are the graphs' components side by side
    a = sp.csr_matrix((np.ones(u.m, dtype=np.int8), u.out_adj, u.out_off.astype(np.int64)), shape=(u.n, u.n))
    _, lab = connected_components(a, directed=True, connection="strong")
    want = np.concatenate([orc.canonical(orc.tarjan(unpack(msc, u, base, i))) + base[i]
                           for i in range(40)]).astype(np.int32)
    assert np.array_equal(orc.canonical(lab.astype(np.int32)), want)


def test_batch_pivots_follow_create_csr(msc, orc):
    """batch_pivots is create_csr's rule per graph: the oracle's pivot and the C++ builder's."""
    gs = mixed_graphs(msc, seed=5)
    u, base = msc.pack_graphs(gs)
    piv = msc.batch_pivots(u, base)
    assert piv.dtype == np.int32 and piv.shape == (len(gs),)
    for i, g in enumerate(gs):
        assert piv[i] == g.max_deg_vert == orc.pivot_scc(g)[0], f"graph {i}"
    assert piv[1] == -1 and piv[2] == -1 and piv[3] == 1  # empty and edgeless graphs have none
    u, base, _ = msc.gen_batch(30, median_n=400, sigma=1.0, min_n=64, max_n=3000, seed=6)
    piv = msc.batch_pivots(u, base)
    for i in range(30):
        assert piv[i] == orc.pivot_scc(unpack(msc, u, base, i))[0], f"generated graph {i}"


def test_gen_batch_other_seed_differs(msc):
    a = msc.gen_batch(8, median_n=300, sigma=0.5, min_n=64, max_n=2000, seed=1)
    b = msc.gen_batch(8, median_n=300, sigma=0.5, min_n=64, max_n=2000, seed=2)
    assert not np.array_equal(a[1], b[1]) or not np.array_equal(a[0].out_adj, b[0].out_adj)


# ---- GPU: every slice equals the per-graph result -------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_scc_batch_matches_per_graph(msc, orc, gpu_ctx, seed):
    _say(f"case: scc batch mixed seed={seed}")
    gs = mixed_graphs(msc, seed)
    u, base = msc.pack_graphs(gs)
    raw = msc.run_scc_batch(u, base, labels=msc.LABELS_RAW, ctx=gpu_ctx, validate=True)
    canon = msc.run_scc_batch(u, base, ctx=gpu_ctx)
    for i, g in enumerate(gs):
        sl = slice(int(base[i]), int(base[i + 1]))
        piv, avg = orc.pivot_scc(g)
        want = orc.scc(g, piv, avg, vert_cutoff=0)
        assert np.array_equal(raw[sl], want), f"graph {i}: raw labels differ"
        assert np.array_equal(canon[sl], orc.canonical(want)), f"graph {i}: canonical labels differ"


@pytest.mark.gpu
def test_cc_batch_matches_per_graph(msc, orc, gpu_ctx):
    _say("case: cc batch mixed")
    gs = mixed_graphs(msc, 5)
    u, base = msc.pack_graphs(gs)
    raw = msc.run_cc_batch(u, base, labels=msc.LABELS_RAW, ctx=gpu_ctx)
    canon = msc.run_cc_batch(u, base, ctx=gpu_ctx)
    for i, g in enumerate(gs):
        sl = slice(int(base[i]), int(base[i + 1]))
        want = orc.cc(g)
        assert np.array_equal(raw[sl], want), f"graph {i}: raw labels differ"
        assert np.array_equal(canon[sl], orc.canonical(want))


@pytest.mark.gpu
def test_batch_given_pivots(msc, orc, gpu_ctx):
    _say("case: scc batch with caller pivots")
    gs = [msc.gen_rmat(9, 8, seed=s) for s in range(6)]
    u, base = msc.pack_graphs(gs)
    rng = np.random.default_rng(3)
    pivs = np.array([int(rng.integers(0, g.n)) for g in gs], dtype=np.int32)
    pivs[2] = -1  # no FW-BW for graph 2
    raw = msc.run_scc_batch(u, base, pivs, labels=msc.LABELS_RAW, ctx=gpu_ctx)
    for i, g in enumerate(gs):
        want = orc.scc(g, int(pivs[i]), g.avg_degree, vert_cutoff=0)
        assert np.array_equal(raw[int(base[i]):int(base[i + 1])], want), f"graph {i}"


@pytest.mark.gpu
def test_generated_batch(msc, orc, gpu_ctx):
    _say("case: generated batch 96 graphs")
    u, base, kinds = msc.gen_batch(96, median_n=1500, sigma=1.2, min_n=256, max_n=40000, seed=4)
    raw = msc.run_scc_batch(u, base, labels=msc.LABELS_RAW, ctx=gpu_ctx)
    wcc = msc.run_cc_batch(u, base, ctx=gpu_ctx)
    # pivots computed once on the host are the ones the device computes per call
    given = msc.run_scc_batch(u, base, msc.batch_pivots(u, base), labels=msc.LABELS_RAW, ctx=gpu_ctx)
    assert np.array_equal(given, raw)
    for i in range(96):
        g = unpack(msc, u, base, i)
        sl = slice(int(base[i]), int(base[i + 1]))
        assert np.array_equal(raw[sl], orc.scc(g, vert_cutoff=0)), f"graph {i} (kind {kinds[i]})"
        assert np.array_equal(wcc[sl], orc.canonical(orc.cc(g))), f"graph {i} wcc"


@pytest.mark.gpu
def test_batch_rejects_bad_input(msc, gpu_ctx):
    gs = [msc.gen_rmat(6, 4, seed=1), msc.gen_rmat(6, 4, seed=2)]
    u, base = msc.pack_graphs(gs)
    with pytest.raises(msc.MscError):  # bases do not end at n
        msc.run_scc_batch(u, np.array([0, 64, 100]), ctx=gpu_ctx)
    with pytest.raises(msc.MscError):  # decreasing bases
        msc.run_scc_batch(u, np.array([0, 100, 64, 128]), ctx=gpu_ctx)
    with pytest.raises(msc.MscError):  # pivot outside its graph
        msc.run_scc_batch(u, base, np.array([0, 64], dtype=np.int32), ctx=gpu_ctx)
    # an edge that crosses from graph 0 into graph 1 is caught by validation
    src, dst = u.edges()
    dst = dst.copy()
    k = int(np.nonzero(src < 64)[0][0])
    dst[k] = 100
    bad = msc.create_csr(u.n, src, dst)
    with pytest.raises(msc.MscError):
        msc.run_scc_batch(bad, base, ctx=gpu_ctx, validate=True)
