"""GPU parity: libmsc on gfx950 against the oracle (CPU restatement of multistep).

Raw labels must equal the oracle's run with vert_cutoff=0 bit for bit (trim -> itself,
pivot's SCC -> pivot, other SCCs -> max id; the GPU colours to empty like the Kokkos loop),
and canonical labels must equal the oracle's canonical partition.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(autouse=True, params=["tail", "rounds"])
def colouring(request, gpu_ctx):
    """Every test runs twice: with the colour tail (one workgroup settles a small colouring)
    and with it off, so that the colouring rounds are checked on the same small graphs."""
    saved = gpu_ctx.tuning()
    gpu_ctx.set_tuning(tail_vertices=65536 if request.param == "tail" else 0)
    yield request.param
    gpu_ctx.set_tuning(**saved)


def _say(msg):
    print(msg, flush=True)


def _check_scc(msc, orc, ctx, g, policy=0):
    piv, avg = orc.pivot_scc(g)
    want = orc.scc(g, piv, avg, vert_cutoff=0)
    got = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=ctx, policy=policy)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"raw SCC labels differ at {bad.size} vertices, first {bad[:5]}: {got[bad[:5]]} vs {want[bad[:5]]}"
    canon = msc.run_scc(g, ctx=ctx, policy=policy)  # pivot computed on device
    assert np.array_equal(canon, orc.canonical(want))
    return want


def _check_cc(msc, orc, ctx, g):
    want = orc.cc(g)
    got = msc.run_cc(g, labels=msc.LABELS_RAW, ctx=ctx)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"raw CC labels differ at {bad.size} vertices, first {bad[:5]}"
    assert np.array_equal(msc.run_cc(g, ctx=ctx), orc.canonical(want))


def test_reference_fixture(msc, orc, gpu_ctx):
    # multistep/test.graph; the reference binary prints labels 0 0 2 (scc) and 0 0 0 (cc)
    g = msc.read_edge(os.path.join(GOLDEN, "test.graph"))
    _say("case: multistep/test.graph")
    assert list(msc.run_scc(g, g.max_deg_vert, g.avg_degree, labels=msc.LABELS_RAW, ctx=gpu_ctx)) == [0, 0, 2]
    assert list(msc.run_cc(g, labels=msc.LABELS_RAW, ctx=gpu_ctx)) == [0, 0, 0]


@pytest.mark.parametrize("rounds", [2, 4, 16])
@pytest.mark.parametrize("kind", ["union", "planted", "rmat"])
def test_fwbw_rounds(msc, orc, gpu_ctx, kind, rounds):
    """msc_opts.fwbw_rounds: more FW-BW rounds before colouring leave every label as the
    reference's single round gives it (raw and canonical); on a union of four R-MAT graphs
    (four giant SCCs) the later rounds commit the other three."""
    _say(f"case: fwbw_rounds {kind} rounds={rounds}")
    if kind == "union":
        g, _ = msc.pack_graphs([msc.gen_rmat(11, 8, seed=40 + i) for i in range(4)])
    elif kind == "planted":
        g = msc.gen_planted(1 << 13, max_size=1024, seed=5)[0]
    else:
        g = msc.gen_rmat(13, 8, seed=6)
    piv, avg = orc.pivot_scc(g)
    want = orc.scc(g, piv, avg, vert_cutoff=0)
    _, one = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx, stats=True)
    got, st = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx, stats=True, fwbw_rounds=rounds)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"raw labels differ at {bad.size} vertices, first {bad[:5]}"
    assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx, fwbw_rounds=rounds), orc.canonical(want))
    assert st.num_fwbw >= one.num_fwbw
    if kind == "union":
        assert st.num_fwbw > one.num_fwbw, "the later rounds should commit the other graphs' giant SCCs"


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n,avgdeg", [(50, 1.5), (300, 2.5), (2000, 4.0)])
def test_random_digraphs(msc, orc, gpu_ctx, n, avgdeg, seed):
    _say(f"case: random n={n} deg={avgdeg} seed={seed}")
    rng = np.random.default_rng(1000 * n + seed)
    m = int(n * avgdeg)
    g = msc.create_csr(n, rng.integers(0, n, m), rng.integers(0, n, m))
    _check_scc(msc, orc, gpu_ctx, g)
    _check_cc(msc, orc, gpu_ctx, g)


@pytest.mark.parametrize("scale", [8, 12, 16])
def test_rmat(msc, orc, gpu_ctx, scale):
    _say(f"case: rmat s{scale}")
    g = msc.gen_rmat(scale, 16, seed=1)
    _check_scc(msc, orc, gpu_ctx, g)
    _check_cc(msc, orc, gpu_ctx, g)


@pytest.mark.parametrize("side", [8, 64, 256])
def test_grid(msc, orc, gpu_ctx, side):
    _say(f"case: grid {side}x{side}")
    g = msc.gen_grid(side, side, seed=2)
    _check_scc(msc, orc, gpu_ctx, g)
    _check_cc(msc, orc, gpu_ctx, g)


def test_planted(msc, orc, gpu_ctx):
    _say("case: planted n=20000")
    g, comp = msc.gen_planted(20000, zipf_s=1.5, max_size=5000, seed=4)
    want = _check_scc(msc, orc, gpu_ctx, g)
    # the planted partition is the SCC partition
    mins = np.full(comp.max() + 1, g.n, dtype=np.int64)
    np.minimum.at(mins, comp, np.arange(g.n))
    assert np.array_equal(orc.canonical(want), mins[comp].astype(np.int32))


def test_edge_cases(msc, orc, gpu_ctx):
    _say("case: edge cases")
    # one vertex, no edges: trimmed, label 0
    g = msc.create_csr(1, [], [])
    assert list(msc.run_scc(g, ctx=gpu_ctx)) == [0]
    assert list(msc.run_cc(g, ctx=gpu_ctx)) == [0]
    # self-loop only: its own SCC
    g = msc.create_csr(2, [0], [0])
    assert list(msc.run_scc(g, ctx=gpu_ctx)) == [0, 1]
    # DAG: every product out*in may be 0 -> no pivot; n singletons
    g = msc.create_csr(4, [0, 0, 1], [1, 2, 3])
    assert list(msc.run_scc(g, ctx=gpu_ctx)) == [0, 1, 2, 3]
    assert list(msc.run_cc(g, ctx=gpu_ctx)) == [0, 0, 0, 0]
    # directed cycle: one SCC
    n = 5000
    g = msc.create_csr(n, np.arange(n), (np.arange(n) + 1) % n)
    assert np.all(msc.run_scc(g, ctx=gpu_ctx) == 0)
    # empty graph
    g = msc.create_csr(0, [], [])
    assert msc.run_scc(g, ctx=gpu_ctx).shape == (0,)


def test_device_resident_graph(msc, orc, gpu_ctx):
    _say("case: device graph s14")
    g = msc.gen_rmat(14, 16, seed=5)
    dg = msc.DeviceGraph(gpu_ctx, g)
    a = msc.run_scc(dg, g.max_deg_vert, g.avg_degree, ctx=gpu_ctx)
    b = msc.run_scc(dg, g.max_deg_vert, g.avg_degree, ctx=gpu_ctx)  # workspace reuse
    want = orc.canonical(orc.scc(g, vert_cutoff=0))
    assert np.array_equal(a, want) and np.array_equal(b, want)
    dg.free()


def test_prepared_call(msc, orc, gpu_ctx):
    """The timing loop's prepared C ABI call writes the same labels as run_scc."""
    _say("case: prepared calls s13 / grid 48")
    for g in (msc.gen_rmat(13, 16, seed=9), msc.gen_grid(48, 48, seed=2)):
        dg = msc.DeviceGraph(gpu_ctx, g)
        out = msc.DeviceLabels(gpu_ctx, g.n)
        for kind, want in (("scc", orc.canonical(orc.scc(g, vert_cutoff=0))), ("cc", orc.canonical(orc.cc(g)))):
            call = msc.Prepared(kind, dg, policy=msc.POLICY_PLAIN, ctx=gpu_ctx, out=out)
            for _ in range(2):
                call()
                assert np.array_equal(out.to_host(), want)
        out.free()
        dg.free()


def test_device_pivot(msc, orc, gpu_ctx):
    for seed in range(3):
        g = msc.gen_rmat(12, 8, seed=seed)
        assert msc.scc_pivot(g, ctx=gpu_ctx) == orc.pivot_scc(g)
        # cc_main.cpp:148-160: largest out + in degree, first such vertex; avg over 2m
        assert msc.cc_pivot(g, ctx=gpu_ctx) == orc.pivot_cc(g)
        piv, avg = msc.cc_pivot(g, ctx=gpu_ctx)
        assert np.array_equal(msc.run_cc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx), orc.cc(g))


def test_stats_and_levels(msc, orc, gpu_ctx):
    _say("case: stats s14")
    g = msc.gen_rmat(14, 16, seed=1)
    lab, st = msc.run_scc(g, g.max_deg_vert, g.avg_degree, ctx=gpu_ctx, stats=True, timing=True, record=True)
    _, lv, summ = orc.scc(g, vert_cutoff=0, trace_cap=1000)
    assert st.num_trim == summ["num_trim"]
    assert st.num_fwbw == summ["num_fwbw"]
    recs = gpu_ctx.levels()
    # top-down levels are exact BFS layers: the first level of each search matches
    fw = [r for r in recs if r.phase == 0]
    ofw = [x for x in lv if x[0] == 0]
    assert fw[0].next_size == ofw[0][3]
    assert st.kernels["level"]["launches"] > 0 and st.kernels["level"]["ms"] > 0
    assert st.bytes > 0


@pytest.mark.parametrize("tuning", [
    dict(blocks_per_cu=1, small_push_edges=0, small_cbw_edges=0, first_batch=1, max_batch=1, concurrent_fwbw=0,
         search_batch=1, local_rounds=0, local_us=1, pull_div=0),
    dict(blocks_per_cu=8, small_push_edges=1 << 20, small_cbw_edges=1 << 20, first_batch=8, max_batch=64,
         concurrent_fwbw=1, search_batch=64, local_rounds=64, local_us=100000, pull_div=1024),
    dict(blocks_per_cu=2, small_push_edges=1 << 12, small_cbw_edges=0, first_batch=2, max_batch=16,
         concurrent_fwbw=0, search_batch=3, local_rounds=1, local_us=40, pull_div=1),
    dict(blocks_per_cu=5, small_push_edges=0, small_cbw_edges=0, first_batch=2, max_batch=32,
         concurrent_fwbw=1, search_batch=2, local_rounds=8, local_us=10, pull_div=2),
])
def test_tuning_keeps_labels(msc, orc, gpu_ctx, tuning):
    """Launch parameters change the schedule only: labels stay the oracle's."""
    _say(f"case: tuning {tuning}")
    defaults = gpu_ctx.tuning()
    try:
        got = gpu_ctx.set_tuning(**tuning)
        assert {k: got[k] for k in tuning} == tuning
        for g in (msc.gen_grid(64, 64, seed=3), msc.gen_rmat(12, 8, seed=2),
                  msc.gen_planted(1 << 13, max_size=512, seed=4)[0]):
            _check_scc(msc, orc, gpu_ctx, g)
            _check_cc(msc, orc, gpu_ctx, g)
    finally:
        gpu_ctx.set_tuning(**defaults)


@pytest.mark.parametrize("kind", ["grid", "planted", "rmat", "random"])
def test_backward_search_within_forward_set(msc, orc, gpu_ctx, kind):
    """concurrent_fwbw 2: beside the forward search, the backward one claims only vertices
    of the forward set once that search has finished.  Raw and canonical labels equal the
    oracle's in every mode, and the restricted search never takes more levels than the
    unrestricted one (mode 1)."""
    if kind == "grid":
        g = msc.gen_grid(512, 512, seed=5)
    elif kind == "planted":
        g = msc.gen_planted(1 << 15, max_size=256, seed=6)[0]
    elif kind == "rmat":
        g = msc.gen_rmat(15, 8, seed=7)
    else:
        rng = np.random.default_rng(8)
        g = msc.create_csr(20000, rng.integers(0, 20000, 30000), rng.integers(0, 20000, 30000))
    piv, avg = orc.pivot_scc(g)
    want = orc.scc(g, piv, avg, vert_cutoff=0)
    defaults = gpu_ctx.tuning()
    levels = {}
    try:
        for mode in (0, 1, 2):
            gpu_ctx.set_tuning(concurrent_fwbw=mode)
            _say(f"case: {kind} concurrent_fwbw={mode} n={g.n} m={g.m}")
            got, st = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx, stats=True)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, f"mode {mode}: raw labels differ at {bad.size} vertices, first {bad[:5]}"
            assert np.array_equal(msc.run_scc(g, piv, avg, ctx=gpu_ctx), orc.canonical(want)), mode
            levels[mode] = (st.fw_levels, st.bw_levels)
    finally:
        gpu_ctx.set_tuning(**defaults)
    assert levels[2][0] == levels[1][0], levels  # the forward search is not touched
    assert levels[2][1] <= levels[1][1], levels


def test_tuning_rejects_bad_values(msc, gpu_ctx):
    defaults = gpu_ctx.tuning()
    for bad in (dict(blocks_per_cu=0), dict(blocks_per_cu=9), dict(small_push_edges=-1),
                dict(first_batch=0), dict(first_batch=8, max_batch=4), dict(max_batch=257),
                dict(concurrent_fwbw=3), dict(search_batch=0), dict(search_batch=257), dict(local_rounds=-1),
                dict(local_rounds=257), dict(local_us=0), dict(pull_div=-1), dict(pull_div=1025),
                dict(tail_vertices=-1), dict(tail_vertices=65537), dict(balance=-1), dict(balance=4)):
        with pytest.raises(msc.MscError):
            gpu_ctx.set_tuning(**bad)
        assert gpu_ctx.tuning() == defaults
    with pytest.raises(msc.MscError):
        gpu_ctx.set_tuning(no_such_field=1)


@pytest.mark.parametrize("kind", ["random", "rmat", "grid", "planted", "chains"])
def test_complete_trim(msc, orc, gpu_ctx, kind):
    """TRIM_LEVEL 2 (scc_trim.cpp:141-297): raw labels equal the oracle's run with the same
    trim level, the pivot kept even when the complete trim removes it."""
    _say(f"case: complete trim {kind}")
    rng = np.random.default_rng(11)
    if kind == "random":
        gs = [msc.create_csr(2000, rng.integers(0, 2000, 2600), rng.integers(0, 2000, 2600)) for _ in range(3)]
    elif kind == "rmat":
        gs = [msc.gen_rmat(s, 8, seed=s) for s in (10, 14)]
    elif kind == "grid":
        gs = [msc.gen_grid(64, 64, seed=2)]
    elif kind == "planted":
        gs = [msc.gen_planted(20000, max_size=400, seed=5)[0]]
    else:  # long chains into a cycle: many trim rounds
        n = 3000
        src = list(range(n - 1)) + [n - 1, 1000]
        dst = list(range(1, n)) + [1000, 999]
        gs = [msc.create_csr(n, src, dst)]
    for g in gs:
        piv, avg = orc.pivot_scc(g)
        want = orc.scc(g, piv, avg, vert_cutoff=0, trim_level=2)
        got, st = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx, trim_level=2, stats=True)
        assert np.array_equal(got, want)
        assert st.num_trim == int((orc.trim_complete(g) >= 0).sum())
        assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx, trim_level=2), orc.canonical(want))
    # a batch with the complete trim: every slice as the graph alone
    u, base = msc.pack_graphs(gs)
    raw = msc.run_scc_batch(u, base, labels=msc.LABELS_RAW, ctx=gpu_ctx, trim_level=2)
    for i, g in enumerate(gs):
        assert np.array_equal(raw[int(base[i]):int(base[i + 1])], orc.scc(g, vert_cutoff=0, trim_level=2))


def test_complete_trim_rejected_for_wcc(msc, gpu_ctx):
    g = msc.gen_rmat(8, 4, seed=1)
    msc.run_cc(g, ctx=gpu_ctx)  # the reference's CC has only the simple trim (cc_trim.cpp:47-97)
    with pytest.raises(msc.MscError):
        msc._run("msc_cc_run", g, None, None, msc.POLICY_PLAIN, msc.LABELS_CANONICAL, False, False, False,
                 gpu_ctx, 15.0, 25, None, None, trim_level=2)



_CHAINS = {}


def _chain_cases(msc, orc):
    """Graphs whose colouring is long chains of small frontiers (the one-wave walk of the
    colour levels) with fan-outs past 64 edges (its hand-over to workgroup rounds), with the
    oracle's raw labels (computed once: the oracle's colouring is quadratic in chain length)."""
    if "cases" in _CHAINS:
        return _CHAINS["cases"]
    rng = np.random.default_rng(21)
    gs = []
    # a path of 3-cycles, ids decreasing along it: the largest colour walks the whole path
    n = 1200
    order = np.arange(n)[::-1]
    src = list(order[:-1]) + [order[i + 2] for i in range(0, n - 2, 3)]
    dst = list(order[1:]) + [order[i] for i in range(0, n - 2, 3)]
    gs.append(msc.create_csr(n, src, dst))
    # the same with permuted ids and a few long-range arcs forward along the path
    perm = rng.permutation(n)
    pa = rng.integers(0, n - 1, 12)
    pb = pa + 1 + (rng.integers(0, n, 12) % (n - 1 - pa))
    s2 = list(perm[np.asarray(src)]) + list(perm[order[pa]])
    d2 = list(perm[np.asarray(dst)]) + list(perm[order[pb]])
    gs.append(msc.create_csr(n, s2, d2))
    # a path with 4-cycles, ids decreasing, where every 300 steps a vertex fans out to 100
    # rungs that rejoin the path at the next vertex
    m = 1200
    src3, dst3 = [], []
    for i in range(m - 1):
        src3.append(i); dst3.append(i + 1)
        if i % 7 == 3:
            src3.append(i + 1); dst3.append(i - 2)
    nxt = m
    for h in range(150, m - 1, 300):
        for k in range(100):
            src3 += [h, nxt]; dst3 += [nxt, h + 1]
            nxt += 1
    rev = np.concatenate([np.arange(m)[::-1], np.arange(m, nxt)])
    gs.append(msc.create_csr(nxt, list(rev[np.asarray(src3)]), list(rev[np.asarray(dst3)])))
    cases = []
    for g in gs:
        piv, avg = orc.pivot_scc(g)
        cases.append((g, piv, avg, orc.scc(g, piv, avg, vert_cutoff=0)))
    _CHAINS["cases"] = cases
    return cases


@pytest.mark.parametrize("tuning", [None, dict(local_us=1), dict(local_rounds=256, local_us=100000)])
def test_colour_chains(msc, orc, gpu_ctx, tuning):
    """Colour levels along long chains (the one-wave walk, its time-budget stop and its
    hand-over at fan-outs): raw and canonical labels equal the oracle's, alone and batched."""
    _say(f"case: colour chains, tuning {tuning}")
    cases = _chain_cases(msc, orc)
    saved = gpu_ctx.tuning()
    try:
        if tuning:
            gpu_ctx.set_tuning(**tuning)
        for g, piv, avg, want in cases:
            got = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx)
            assert np.array_equal(got, want)
            assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx), orc.canonical(want))
        u, base = msc.pack_graphs([c[0] for c in cases])
        raw = msc.run_scc_batch(u, base, labels=msc.LABELS_RAW, ctx=gpu_ctx)
        for i, c in enumerate(cases):
            assert np.array_equal(raw[int(base[i]):int(base[i + 1])], c[3])
    finally:
        gpu_ctx.set_tuning(**saved)
