"""Pins the oracle (CPU restatement of multistep) before it is trusted as the GPU checker.

Each check is chosen so that a plausible slip in the restatement fails it:
* the reference fixture (multistep/test.graph) and its known labels;
* brute force: every digraph on 3 vertices, every loop-free digraph on 4, random ones on
  up to 12, against the transitive-closure partition;
* closed forms: DAG -> singletons, cycles, bidirected grid -> one SCC, disjoint cycles;
* scipy's strong / weak components (Pearce's algorithm), and an independent Tarjan and
  union-find in the oracle itself;
* the reference's own multistep sources compiled into oracle/_ref: raw labels equal for
  every phase, the Tarjan tail included when both run on one thread;
* the direction-switch rule: top-down levels are exact BFS layers, checked against a
  numpy BFS, and the switch decisions against the rule applied to those sizes.
"""
import ctypes
import itertools

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components


def partition_closure(n, src, dst):
    """SCC partition by transitive closure (boolean matrix powers), canonical min ids."""
    R = np.eye(n, dtype=bool)
    for s, d in zip(src, dst):
        R[s, d] = True
    for _ in range(max(1, int(np.ceil(np.log2(max(n, 2)))) + 1)):
        R = R | ((R.astype(np.int32) @ R.astype(np.int32)) > 0)
    both = R & R.T
    return np.array([np.nonzero(both[v])[0].min() for v in range(n)], dtype=np.int32)


def weak_closure(n, src, dst):
    R = np.eye(n, dtype=bool)
    for s, d in zip(src, dst):
        R[s, d] = R[d, s] = True
    for _ in range(max(1, int(np.ceil(np.log2(max(n, 2)))) + 1)):
        R = R | ((R.astype(np.int32) @ R.astype(np.int32)) > 0)
    return np.array([np.nonzero(R[v])[0].min() for v in range(n)], dtype=np.int32)


def scipy_canon(g, conn):
    src, dst = g.edges()
    A = csr_matrix((np.ones(g.m), (src, dst)), shape=(g.n, g.n))
    k, lab = connected_components(A, directed=True, connection=conn)
    mins = np.full(k, g.n, dtype=np.int64)
    np.minimum.at(mins, lab, np.arange(g.n))
    return mins[lab].astype(np.int32)


def test_reference_fixture(msc, orc):
    g = msc.read_edge("tests/golden/test.graph")
    assert g.n == 3 and g.m == 5
    assert list(orc.scc(g, g.max_deg_vert, g.avg_degree)) == [0, 0, 2]
    assert list(orc.cc(g)) == [0, 0, 0]


def test_brute_force_n3_all_with_loops(msc, orc):
    pairs = [(i, j) for i in range(3) for j in range(3)]
    for mask in range(1 << len(pairs)):
        e = [pairs[k] for k in range(len(pairs)) if mask >> k & 1]
        s = [a for a, _ in e]
        d = [b for _, b in e]
        g = msc.create_csr(3, s, d)
        want = partition_closure(3, s, d)
        for cut in (0, 10000):
            assert np.array_equal(orc.canonical(orc.scc(g, vert_cutoff=cut)), want), (e, cut)
        assert np.array_equal(orc.canonical(orc.cc(g)), weak_closure(3, s, d))


def test_brute_force_n4_all(msc, orc):
    pairs = [(i, j) for i in range(4) for j in range(4) if i != j]
    for mask in range(1 << len(pairs)):
        e = [pairs[k] for k in range(len(pairs)) if mask >> k & 1]
        s = [a for a, _ in e]
        d = [b for _, b in e]
        g = msc.create_csr(4, s, d)
        want = partition_closure(4, s, d)
        assert np.array_equal(orc.canonical(orc.scc(g, vert_cutoff=0)), want), e
        assert np.array_equal(orc.canonical(orc.scc(g)), want), e


@pytest.mark.parametrize("seed", range(40))
def test_brute_force_random_small(msc, orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 13))
    m = int(rng.integers(0, 3 * n))
    s, d = rng.integers(0, n, m), rng.integers(0, n, m)
    g = msc.create_csr(n, s, d)
    want = partition_closure(n, s, d)
    for cut in (0, 3, 10000):
        assert np.array_equal(orc.canonical(orc.scc(g, vert_cutoff=cut)), want)
    assert np.array_equal(orc.canonical(orc.cc(g)), weak_closure(n, s, d))


def test_closed_forms(msc, orc):
    n = 1000
    # DAG: every edge goes up -> n singletons
    rng = np.random.default_rng(7)
    a, b = rng.integers(0, n, 5000), rng.integers(0, n, 5000)
    s, d = np.minimum(a, b), np.maximum(a, b)
    keep = s != d
    g = msc.create_csr(n, s[keep], d[keep])
    assert np.array_equal(orc.canonical(orc.scc(g, vert_cutoff=0)), np.arange(n))
    # one directed cycle -> one SCC
    g = msc.create_csr(n, np.arange(n), (np.arange(n) + 1) % n)
    assert np.all(orc.canonical(orc.scc(g, vert_cutoff=0)) == 0)
    # k disjoint cycles of sizes s_i -> k SCCs of those sizes
    sizes = [1, 2, 3, 7, 50, 400]
    src, dst, base = [], [], 0
    for k in sizes:
        for i in range(k):
            src.append(base + i)
            dst.append(base + (i + 1) % k)
        base += k
    g = msc.create_csr(base, src, dst)
    lab = orc.scc(g, vert_cutoff=0)
    counts = orc.verify_counts(lab)
    assert counts == (len(sizes), len(sizes) - 1, 400, 0)
    # grid with every edge in both directions -> one SCC
    g = msc.gen_grid(30, 40, p_one=0.0, p_both=1.0, arc_frac=0.0)
    assert np.all(orc.canonical(orc.scc(g, vert_cutoff=0)) == 0)
    # self-loop-only vertices are singletons
    g = msc.create_csr(4, [0, 1, 2], [0, 1, 2])
    assert list(orc.canonical(orc.scc(g))) == [0, 1, 2, 3]


@pytest.mark.parametrize("scale", [10, 13, 16])
def test_against_scipy_rmat(msc, orc, scale):
    g = msc.gen_rmat(scale, 16, seed=1)
    want = scipy_canon(g, "strong")
    for cut in (0, 10000):
        assert np.array_equal(orc.canonical(orc.scc(g, vert_cutoff=cut)), want)
    assert np.array_equal(orc.tarjan(g), want)
    weak = scipy_canon(g, "weak")
    assert np.array_equal(orc.canonical(orc.cc(g)), weak)
    assert np.array_equal(orc.wcc(g), weak)


def test_against_scipy_grid_and_planted(msc, orc):
    g = msc.gen_grid(128, 128, seed=3)
    assert np.array_equal(orc.canonical(orc.scc(g, vert_cutoff=0)), scipy_canon(g, "strong"))
    assert np.array_equal(orc.canonical(orc.cc(g)), scipy_canon(g, "weak"))
    g, comp = msc.gen_planted(5000, max_size=500, seed=9)
    mins = np.full(comp.max() + 1, g.n, dtype=np.int64)
    np.minimum.at(mins, comp, np.arange(g.n))
    assert np.array_equal(orc.canonical(orc.scc(g, vert_cutoff=0)), mins[comp].astype(np.int32))


def test_pivot_rule_matches_reference_create_csr(msc, orc):
    if not orc.have_ref():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(11)
    for trial in range(20):
        n = int(rng.integers(1, 300))
        m = int(rng.integers(0, 4 * n))
        s, d = rng.integers(0, n, m), rng.integers(0, n, m)
        oo, oa, io, ia, piv, avg = orc.ref_create_csr(n, s, d)
        g = msc.create_csr(n, s, d)
        # the product's CSR build keeps the reference's order exactly
        assert np.array_equal(g.out_off, oo) and np.array_equal(g.out_adj, oa)
        assert np.array_equal(g.in_off, io) and np.array_equal(g.in_adj, ia)
        if m and np.any(np.diff(oo.astype(np.int64)) * np.diff(io.astype(np.int64)) > 0):
            assert (g.max_deg_vert, g.avg_degree) == (piv, avg)
            assert orc.pivot_scc(g) == (piv, avg)


def _one_thread():
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        return True
    except OSError:
        return False


@pytest.mark.parametrize("kind", ["random", "rmat", "grid"])
def test_raw_labels_equal_reference(msc, orc, kind):
    """Oracle raw labels == the reference's own code, every phase, bit for bit."""
    if not orc.have_ref():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    graphs = []
    if kind == "random":
        rng = np.random.default_rng(5)
        for n, deg in [(200, 1.2), (2000, 2.0), (20000, 3.0)]:
            m = int(n * deg)
            graphs.append(msc.create_csr(n, rng.integers(0, n, m), rng.integers(0, n, m)))
    elif kind == "rmat":
        graphs = [msc.gen_rmat(s, 16, seed=2) for s in (10, 14, 16)]
    else:
        graphs = [msc.gen_grid(64, 64, seed=1), msc.gen_grid(200, 100, seed=2)]
    for g in graphs:
        piv, avg = orc.pivot_scc(g)
        for cut in (0, 50):  # cut 0: trim, FW-BW and colouring labels only
            ours = orc.scc(g, piv, avg, vert_cutoff=cut)
            ref = orc.ref_run_scc(g, piv, avg, vert_cutoff=cut)
            if cut == 0:
                assert np.array_equal(ours, ref)
            assert np.array_equal(orc.canonical(ours), orc.canonical(ref))


def test_tarjan_tail_labels_equal_reference_one_thread(msc, orc):
    """With one thread the reference's valid list is in id order, so even the Tarjan tail's
    DFS-root labels are fixed; the oracle's iterative Tarjan must reproduce them."""
    if not orc.have_ref():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import subprocess
    import sys

    code = (
        "import ctypes, numpy as np, sys; sys.path.insert(0, '.');"
        "ctypes.CDLL('libgomp.so.1').omp_set_num_threads(1);"
        "import connectivity_amd as msc, oracle as orc;"
        "bad = 0\n"
        "for g in [msc.gen_rmat(14, 16, seed=3), msc.gen_grid(100, 100, seed=4)]:\n"
        "    piv, avg = orc.pivot_scc(g)\n"
        "    for cut in (10000, 300):\n"
        "        bad += int(not np.array_equal(orc.scc(g, piv, avg, vert_cutoff=cut), orc.ref_run_scc(g, piv, avg, vert_cutoff=cut)))\n"
        "sys.exit(bad)"
    )
    r = subprocess.run([sys.executable, "-c", code], env={**__import__("os").environ, "OMP_NUM_THREADS": "1"},
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_cc_raw_labels_equal_reference(msc, orc):
    if not orc.have_ref():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(8)
    done = 0
    for trial in range(30):
        n = int(rng.integers(10, 3000))
        m = int(rng.integers(0, 2 * n))
        s, d = rng.integers(0, n, m), rng.integers(0, n, m)
        g = msc.create_csr(n, s, d)
        ours = orc.cc(g)
        piv, _ = orc.pivot_cc(g)
        # the reference overflows a list when vertex 0 is valid but outside the pivot's
        # component (multistep/cc_bfs.cpp:231); such inputs are skipped
        deg0 = (g.out_off[1] - g.out_off[0]) + (g.in_off[1] - g.in_off[0])
        if piv < 0 or (deg0 > 0 and ours[0] != piv):
            continue
        ref = orc.ref_run_cc(n, s, d)
        assert np.array_equal(ours, ref)
        done += 1
    assert done >= 5


def bfs_layers(g, root, valid):
    """Exact BFS layer sizes from root over out-edges within `valid` (numpy)."""
    seen = np.zeros(g.n, dtype=bool)
    seen[root] = True
    frontier = np.array([root])
    sizes = []
    deg = np.diff(g.out_off.astype(np.int64))
    while frontier.size:
        starts = g.out_off[frontier].astype(np.int64)
        nbrs = np.concatenate([g.out_adj[s:s + k] for s, k in zip(starts, deg[frontier])]) if frontier.size else []
        nbrs = np.unique(np.asarray(nbrs, dtype=np.int64))
        nbrs = nbrs[valid[nbrs] & ~seen[nbrs]]
        seen[nbrs] = True
        sizes.append(int(nbrs.size))
        frontier = nbrs
    return sizes


def test_switch_rule_and_topdown_layers(msc, orc):
    """Top-down levels are exact BFS layers; each switch follows
    multistep/scc_fwbw.cpp:168-195 applied to those sizes."""
    for seed in (1, 2):
        g = msc.gen_rmat(14, 16, seed=seed)
        piv, avg = orc.pivot_scc(g)
        _, lv, _ = orc.scc(g, piv, avg, vert_cutoff=0, trace_cap=200)
        deg_o = np.diff(g.out_off.astype(np.int64))
        deg_i = np.diff(g.in_off.astype(np.int64))
        valid = (deg_o > 0) & (deg_i > 0)
        layers = bfs_layers(g, piv, valid)
        fw = [x for x in lv if x[0] == 0]
        # levels before the first bottom-up one are exact layers
        found, bu, switched = 0, False, False
        for i, (_, level, bottom_up, nxt) in enumerate(fw):
            assert bottom_up == int(bu)
            if not bu:
                if i < len(layers) and not any(x[2] for x in fw[:i]):
                    assert nxt == layers[i]
            found += nxt
            if not switched and not bu:
                if (g.n - found) * avg / 15.0 < nxt * avg:
                    bu = True
            elif not switched and g.n // 25 > nxt:
                bu, switched = False, True


# ---- complete trim (TRIM_LEVEL 2) -------------------------------------------------------------
def test_trim_complete_closed_forms(msc, orc):
    n = 200
    chain = msc.create_csr(n, np.arange(n - 1), np.arange(1, n))          # a path: all trimmed
    assert np.array_equal(orc.trim_complete(chain), np.arange(n))
    cycle = msc.create_csr(n, np.arange(n), (np.arange(n) + 1) % n)        # a cycle: none
    assert np.all(orc.trim_complete(cycle) == -1)
    # a cycle 0..9 with a tail 9 -> 10 -> ... -> 19 and a spur 20 -> 0: only the cycle stays
    src = list(range(10)) + list(range(9, 19)) + [20]
    dst = [(i + 1) % 10 for i in range(10)] + list(range(10, 20)) + [0]
    g = msc.create_csr(21, src, dst)
    want = np.array([-1] * 10 + list(range(10, 21)), dtype=np.int32)
    assert np.array_equal(orc.trim_complete(g), want)
    # self-loops do not keep a vertex (scc_trim.cpp:183, :192 skip out == vert)
    g = msc.create_csr(3, [0, 1, 1, 2], [0, 1, 2, 1])
    assert list(orc.trim_complete(g)) == [0, -1, -1]
    # duplicated arcs count once per entry but change nothing
    g = msc.create_csr(3, [0, 0, 1, 2, 2], [1, 1, 2, 0, 0])
    assert np.all(orc.trim_complete(g) == -1)


@pytest.mark.parametrize("seed", range(8))
def test_trim_complete_is_the_largest_closed_set(msc, orc, seed):
    """Brute-force fixpoint: repeat 'drop every vertex lacking a valid in- or out-neighbour
    other than itself' until nothing changes; the peeling restatement must agree."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 60))
    m = int(n * rng.uniform(0.8, 2.0))
    s, d = rng.integers(0, n, m), rng.integers(0, n, m)
    g = msc.create_csr(n, s, d)
    valid = np.ones(n, bool)
    while True:
        keep = np.array([valid[v] and any(valid[w] and w != v for w in d[s == v])
                         and any(valid[u] and u != v for u in s[d == v]) for v in range(n)])
        if np.array_equal(keep, valid):
            break
        valid = keep
    assert np.array_equal(orc.trim_complete(g), np.where(valid, -1, np.arange(n)).astype(np.int32))


@pytest.mark.parametrize("kind", ["random", "rmat", "planted"])
def test_trim_complete_equals_reference(msc, orc, kind):
    if not orc.have_ref():
        pytest.skip("reference build (oracle/_ref) absent")
    rng = np.random.default_rng(7)
    if kind == "random":
        gs = [msc.create_csr(3000, rng.integers(0, 3000, 4500), rng.integers(0, 3000, 4500)) for _ in range(4)]
    elif kind == "rmat":
        gs = [msc.gen_rmat(s, 8, seed=s) for s in (10, 14, 16)]
    else:
        gs = [msc.gen_planted(20000, max_size=500, seed=3)[0]]
    for g in gs:
        assert np.array_equal(orc.trim_complete(g), orc.ref_trim_complete(g))


def test_trim_level_two_keeps_the_partition(msc, orc):
    for g in (msc.gen_rmat(13, 8, seed=4), msc.gen_grid(40, 40, seed=4), msc.gen_planted(5000, seed=4)[0]):
        a = orc.scc(g, vert_cutoff=0, trim_level=2)
        b = orc.scc(g, vert_cutoff=0, trim_level=1)
        assert np.array_equal(orc.canonical(a), orc.canonical(b))
        assert np.array_equal(orc.canonical(a), orc.canonical(orc.tarjan(g)))
        # every vertex the complete trim removes keeps its own id as raw label
        t = orc.trim_complete(g)
        assert np.array_equal(a[t >= 0], t[t >= 0])
