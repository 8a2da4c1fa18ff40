"""Seeded random parity sweep: graphs of every generator (R-MAT scales 8-14, grids with random
edge probabilities, planted SCCs, uniform random digraphs with duplicates and self-loops) and
small packed batches, each under a random tuning (searches side by side, in turn or
restricted, every load balance, colour tail on or off, workgroups per CU, pull group width)
and policy.
SCC and WCC labels must equal the oracle's: raw labels under the plain rule (multistep/
scc_main.cpp:210-317, cc_main.cpp:184-245), canonical ones under the learned policy, whose
abandoned searches leave the pivot's component to colouring (labelled with its largest
member instead of the pivot); batches graph by graph against the oracle on that graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _slice(msc, b, base, i):
    b0, b1 = int(base[i]), int(base[i + 1])
    e0, e1 = int(b.out_off[b0]), int(b.out_off[b1])
    f0, f1 = int(b.in_off[b0]), int(b.in_off[b1])
    return msc.Graph(n=b1 - b0, out_off=(b.out_off[b0:b1 + 1] - b.out_off[b0]).astype(np.uint32),
                     out_adj=(b.out_adj[e0:e1] - b0).astype(np.int32),
                     in_off=(b.in_off[b0:b1 + 1] - b.in_off[b0]).astype(np.uint32),
                     in_adj=(b.in_adj[f0:f1] - b0).astype(np.int32))


def _graph(msc, rng, kind, seed):
    if kind == 0:
        return msc.gen_rmat(int(rng.integers(8, 15)), int(rng.choice([4, 8, 16, 32])), seed=seed)
    if kind == 1:
        side = int(rng.integers(16, 160))
        return msc.gen_grid(side, side, float(rng.uniform(0.3, 0.7)), float(rng.uniform(0, 0.3)),
                            float(rng.uniform(0, 0.02)), seed=seed)
    if kind == 2:
        return msc.gen_planted(int(rng.integers(500, 20000)), max_size=int(rng.integers(2, 2000)), seed=seed)[0]
    n = int(rng.integers(10, 20000))
    m = int(rng.integers(n // 2, 4 * n))
    return msc.create_csr(n, rng.integers(0, n, m), rng.integers(0, n, m))


@pytest.mark.parametrize("seed", [11, 12])
def test_random_graphs_and_tunings(msc, orc, gpu_ctx, seed):
    msc.load_policy(gpu_ctx)
    rng = np.random.default_rng(seed)
    defaults = gpu_ctx.tuning()
    try:
        for case in range(150):
            kind = int(rng.integers(0, 5))
            gseed = int(rng.integers(1, 1 << 30))
            tune = dict(concurrent_fwbw=int(rng.integers(0, 3)), balance=int(rng.choice([0, 1, 2, 3])),
                        tail_vertices=int(rng.choice([0, 65536])), blocks_per_cu=int(rng.choice([1, 2, 5, 8])),
                        first_batch=int(rng.choice([1, 2, 16])), group_words=int(rng.choice([0, 1, 8, 64, 256])))
            pol = msc.POLICY_LEARNED if rng.integers(0, 2) else msc.POLICY_PLAIN
            gpu_ctx.set_tuning(**dict(defaults, **tune))
            what = f"case {case} kind {kind} seed {gseed} tuning {tune} policy {pol}"
            print(what, flush=True)
            fix = (lambda x: orc.canonical(x)) if pol == msc.POLICY_LEARNED else (lambda x: x)
            if kind == 4:  # a small packed batch
                b, base, _ = msc.gen_batch(int(rng.integers(2, 24)), median_n=int(rng.choice([256, 1024])),
                                           seed=gseed)
                got = msc.run_scc_batch(b, base, labels=msc.LABELS_RAW, ctx=gpu_ctx, policy=pol)
                gotc = msc.run_cc_batch(b, base, labels=msc.LABELS_RAW, ctx=gpu_ctx, policy=pol)
                for i in range(len(base) - 1):
                    h = _slice(msc, b, base, i)
                    b0, b1 = int(base[i]), int(base[i + 1])
                    assert np.array_equal(fix(got[b0:b1]), fix(orc.scc(h, vert_cutoff=0))), (what, i)
                    assert np.array_equal(fix(gotc[b0:b1]), fix(orc.cc(h))), (what, i)
                continue
            g = _graph(msc, rng, kind, gseed)
            piv, avg = orc.pivot_scc(g)
            got = msc.run_scc(g, piv, avg, labels=msc.LABELS_RAW, ctx=gpu_ctx, policy=pol)
            assert np.array_equal(fix(got), fix(orc.scc(g, piv, avg, vert_cutoff=0))), what
            gotc = msc.run_cc(g, labels=msc.LABELS_RAW, ctx=gpu_ctx, policy=pol)
            assert np.array_equal(fix(gotc), fix(orc.cc(g))), what
    finally:
        gpu_ctx.set_tuning(**defaults)
