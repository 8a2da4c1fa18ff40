"""Learned direction / phase policy.

CPU: the float64 restatement of the MLP (oracle/mlp_ref.py) against central finite
differences.  GPU: the device training kernels' loss / gradient and batched logits against
that restatement; the level closer's in-kernel MLP picks the same actions as the batched
kernel on the features it records; forced schedules and learned runs leave labels
unchanged; training on device reduces the loss.
"""
import numpy as np
import pytest

from oracle import mlp_ref


def random_blob(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, scale, mlp_ref.PARAMS)
    mean = rng.normal(0, 1, mlp_ref.IN)
    std = rng.uniform(0.5, 2.0, mlp_ref.IN)
    return np.concatenate([w, mean, std]).astype(np.float32)


def random_batch(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 2, (n, mlp_ref.IN)).astype(np.float32)
    y = rng.integers(0, 3, n).astype(np.int32)
    w = rng.uniform(0.1, 2.0, n).astype(np.float32)
    return x, y, w


def test_restatement_matches_finite_differences():
    blob = random_blob(1).astype(np.float64)
    x, y, w = random_batch(2, 7)
    loss, grad = mlp_ref.loss_and_grad(blob, x, y, w)
    rng = np.random.default_rng(3)
    for p in rng.choice(mlp_ref.PARAMS, 60, replace=False):
        h = 1e-6
        bp, bm = blob.copy(), blob.copy()
        bp[p] += h
        bm[p] -= h
        fd = (mlp_ref.loss_and_grad(bp, x, y, w)[0] - mlp_ref.loss_and_grad(bm, x, y, w)[0]) / (2 * h)
        assert abs(fd - grad[p]) <= 1e-5 * max(1.0, abs(fd)), p
    # a dropped term would show: bias gradients equal the summed deltas
    assert np.isfinite(loss) and loss > 0


def test_feature_and_blob_layout(msc):
    from connectivity_amd import policy

    assert msc.MLP_PARAMS == mlp_ref.PARAMS == 1571 and msc.MLP_BLOB == mlp_ref.BLOB == 1595
    assert len(policy.FEATURES) == msc.MLP_IN
    b = policy.init_blob(np.zeros(12), np.ones(12), seed=0)
    assert b.shape == (msc.MLP_BLOB,) and np.all(b[-12:] == 1.0)


def test_schedule_family_shapes():
    from connectivity_amd import policy

    c = policy.phase_schedules(6)
    names = [n for n, _ in c]
    assert "all_td" in names and "bu1-None" in names and "stop1" in names
    for _, acts in c:
        assert all(a in (-1, 0, 1, 2) for a in acts) and len(acts) <= 64


@pytest.mark.gpu
def test_device_gradient_matches_restatement(msc, gpu_ctx):
    blob = random_blob(4)
    for n in (1, 63, 64, 65, 500):
        x, y, w = random_batch(10 + n, n)
        loss, grad = msc.mlp_grad(gpu_ctx, blob, x, y, w)
        rl, rg = mlp_ref.loss_and_grad(blob, x, y, w)
        assert abs(loss - rl) <= 1e-4 * max(1.0, abs(rl))
        np.testing.assert_allclose(grad, rg, rtol=2e-4, atol=2e-4 * np.abs(rg).max())
    # unweighted == weights of one
    x, y, _ = random_batch(99, 100)
    l1, g1 = msc.mlp_grad(gpu_ctx, blob, x, y)
    l2, g2 = msc.mlp_grad(gpu_ctx, blob, x, y, np.ones(100, np.float32))
    assert abs(l1 - l2) < 1e-4 * abs(l1) and np.allclose(g1, g2, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_device_forward_matches_restatement(msc, gpu_ctx):
    blob = random_blob(5)
    x, _, _ = random_batch(6, 300)
    z = msc.mlp_forward(gpu_ctx, blob, x)
    np.testing.assert_allclose(z, mlp_ref.forward(blob, x)[-1], rtol=1e-4, atol=1e-4)


@pytest.mark.gpu
def test_training_reduces_loss(msc, gpu_ctx):
    from connectivity_amd import policy

    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (4000, 12)).astype(np.float32)
    y = (x[:, 0] > 0).astype(np.int32) + (x[:, 1] > 1).astype(np.int32)  # learnable labels 0..2
    blob0 = policy.init_blob(x.mean(0), x.std(0), seed=1)
    blob, curve = msc.mlp_train(gpu_ctx, blob0, x, y, epochs=20, batch=256, lr=3e-3)
    assert curve[-1] < 0.5 * curve[0]
    acc = np.mean(np.argmax(msc.mlp_forward(gpu_ctx, blob, x), 1) == y)
    assert acc > 0.9
    # the trained loss equals the restatement's loss of the returned weights
    l_dev, _ = msc.mlp_grad(gpu_ctx, blob, x, y)
    l_ref, _ = mlp_ref.loss_and_grad(blob, x, y)
    assert abs(l_dev - l_ref) < 1e-3 * l_ref


@pytest.mark.gpu
@pytest.mark.parametrize("graph", ["rmat", "grid"])
def test_forced_schedules_keep_labels(msc, orc, gpu_ctx, graph):
    g = msc.gen_rmat(14, 16, seed=3) if graph == "rmat" else msc.gen_grid(96, 96, seed=3)
    print(f"case: forced schedules {graph}", flush=True)
    want = orc.canonical(orc.scc(g, vert_cutoff=0))
    want_cc = orc.canonical(orc.cc(g))
    for fw, bw in [([0] * 64, [0] * 64), ([1] * 64, [1] * 64), ([0, 1, 1, 0], [1, 0, 1]), ([2], None),
                   (None, [2]), ([0, 0, 2], [1, 2])]:
        sched = msc.schedule_array(fw=fw, bw=bw)
        lab, st = msc.run_scc(g, ctx=gpu_ctx, schedule=sched, stats=True)
        assert np.array_equal(lab, want), (fw, bw)
    for cc in ([1] * 64, [2], [0, 1]):
        assert np.array_equal(msc.run_cc(g, ctx=gpu_ctx, schedule=msc.schedule_array(cc=cc)), want_cc)


@pytest.mark.gpu
@pytest.mark.parametrize("concurrent", [0, 1, 2])
def test_abandoned_searches_every_mode(msc, orc, gpu_ctx, concurrent):
    """An abandoned search (the policy's third action) leaves its vertices to colouring with
    the searches side by side and in turn, for one graph and for a batch (which runs its
    searches in turn under mode 2): canonical labels stay the oracle's.  Abandoning the
    forward search in turn leaves no backward search to run."""
    g = msc.gen_rmat(13, 16, seed=5)
    b, base, _ = msc.gen_batch(24, seed=6)
    want = orc.canonical(orc.scc(g, vert_cutoff=0))
    defaults = gpu_ctx.tuning()
    try:
        gpu_ctx.set_tuning(concurrent_fwbw=concurrent)
        for fw, bw in (([2], None), (None, [2]), ([0, 2], [2]), ([1, 1, 2], None)):
            print(f"case: abandon fw={fw} bw={bw} concurrent_fwbw={concurrent}", flush=True)
            sched = msc.schedule_array(fw=fw, bw=bw)
            assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx, schedule=sched), want), (fw, bw)
            got = msc.run_scc_batch(b, base, ctx=gpu_ctx, schedule=sched)
            for i in range(len(base) - 1):
                b0, b1 = int(base[i]), int(base[i + 1])
                e0, e1 = int(b.out_off[b0]), int(b.out_off[b1])
                f0, f1 = int(b.in_off[b0]), int(b.in_off[b1])
                h = msc.Graph(n=b1 - b0, out_off=(b.out_off[b0:b1 + 1] - b.out_off[b0]).astype(np.uint32),
                              out_adj=(b.out_adj[e0:e1] - b0).astype(np.int32),
                              in_off=(b.in_off[b0:b1 + 1] - b.in_off[b0]).astype(np.uint32),
                              in_adj=(b.in_adj[f0:f1] - b0).astype(np.int32))
                assert np.array_equal(got[b0:b1], orc.canonical(orc.scc(h, vert_cutoff=0))), (fw, bw, i)
    finally:
        gpu_ctx.set_tuning(**defaults)


@pytest.mark.gpu
def test_forced_actions_are_recorded(msc, gpu_ctx):
    g = msc.gen_rmat(14, 16, seed=3)
    sched = msc.schedule_array(fw=[1, 1, 0, 0, 0, 0, 0, 0])
    msc.run_scc(g, ctx=gpu_ctx, schedule=sched, record=True)
    fw = [r for r in gpu_ctx.levels() if r.phase == 0]
    assert [r.action for r in fw[:2]] == [1, 1]
    assert [r.bottom_up for r in fw[1:3]] == [1, 1]
    assert all(r.level_us > 0 for r in fw)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_learned_runs_keep_labels_and_match_batched_mlp(msc, orc, gpu_ctx, seed):
    blob = random_blob(seed, scale=0.5)
    blob[-24:-12] = 0.0   # mean
    blob[-12:] = 3.0      # std
    msc.set_policy(gpu_ctx, blob)
    for g in (msc.gen_rmat(13, 8, seed=seed), msc.gen_grid(64, 64, seed=seed)):
        want = orc.canonical(orc.scc(g, vert_cutoff=0))
        lab = msc.run_scc(g, ctx=gpu_ctx, policy=msc.POLICY_LEARNED, record=True)
        assert np.array_equal(lab, want)
        recs = gpu_ctx.levels()
        feats = np.array([r.feat for r in recs if r.phase in (0, 1)], np.float32)
        acts = np.array([r.action for r in recs if r.phase in (0, 1)])
        if len(feats):
            z = msc.mlp_forward(gpu_ctx, blob, feats)
            assert np.array_equal(np.argmax(z, 1), acts)
        assert np.array_equal(msc.run_cc(g, ctx=gpu_ctx, policy=msc.POLICY_LEARNED), orc.canonical(orc.cc(g)))


@pytest.mark.gpu
def test_cc_network_decides_the_cc_search(msc, orc, gpu_ctx):
    """With two networks loaded (msc_policy_set_weights, 2 x MLP_BLOB floats) the SCC
    searches act on the first one's logits and the CC search on the second one's; a wrong
    float count is refused."""
    nets = []
    for seed, act in ((21, 0), (22, 1)):  # the first always top-down, the second bottom-up
        b = random_blob(seed, scale=0.5)
        b[-24:-12] = 0.0
        b[-12:] = 3.0
        b[mlp_ref.PARAMS - 3 + act] = 50.0  # output bias b3
        nets.append(b)
    msc.set_policy(gpu_ctx, nets[0], nets[1])
    g = msc.gen_rmat(13, 8, seed=5)
    assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx, policy=msc.POLICY_LEARNED, record=True),
                          orc.canonical(orc.scc(g, vert_cutoff=0)))
    recs = [r for r in gpu_ctx.levels() if r.phase in (0, 1)]
    feats = np.array([r.feat for r in recs], np.float32)
    assert np.array_equal(np.argmax(msc.mlp_forward(gpu_ctx, nets[0], feats), 1), [r.action for r in recs])
    assert np.array_equal(msc.run_cc(g, ctx=gpu_ctx, policy=msc.POLICY_LEARNED, record=True),
                          orc.canonical(orc.cc(g)))
    assert recs and all(r.action == 0 for r in recs)
    recs = [r for r in gpu_ctx.levels() if r.phase == 2]
    feats = np.array([r.feat for r in recs], np.float32)
    acts = [r.action for r in recs]
    assert np.array_equal(np.argmax(msc.mlp_forward(gpu_ctx, nets[1], feats), 1), acts)
    assert acts and all(a == 1 for a in acts)
    import ctypes as C

    three = np.concatenate(nets + nets[:1]).astype(np.float32)
    rc = msc._lib().msc_policy_set_weights(gpu_ctx.handle, three.ctypes.data_as(C.c_void_p), three.shape[0])
    assert rc != 0


@pytest.mark.gpu
def test_shipped_weights_keep_labels(msc, orc, gpu_ctx):
    import os

    if not os.path.exists(msc.POLICY_WEIGHTS):
        pytest.skip("no trained weights shipped yet")
    msc.load_policy(gpu_ctx)
    for g in (msc.gen_rmat(15, 16, seed=77), msc.gen_grid(128, 128, seed=77), msc.gen_planted(20000, seed=77)[0]):
        want = orc.canonical(orc.scc(g, vert_cutoff=0))
        assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx, policy=msc.POLICY_LEARNED), want)
        assert np.array_equal(msc.run_cc(g, ctx=gpu_ctx, policy=msc.POLICY_LEARNED), orc.canonical(orc.cc(g)))


@pytest.mark.gpu
def test_recorded_features_follow_their_definition(msc, gpu_ctx):
    """Features the level closer computes, recomputed on the host from the same record.
    Counts enter as log fractions of the graph, so they do not grow with the scale."""
    for g in (msc.gen_rmat(14, 16, seed=5), msc.gen_rmat(17, 16, seed=5)):
        msc.run_scc(g, g.max_deg_vert, g.avg_degree, ctx=gpu_ctx, record=True)
        recs = [r for r in gpu_ctx.levels() if r.phase in (0, 1)]
        assert recs
        n, m = np.float32(g.n), np.float32(g.avg_degree) * np.float32(g.n)
        for r in recs:
            f = np.array(r.feat, np.float32)
            want = {
                0: np.log(max(r.next_size, 1) / n),
                1: np.log(max(r.next_edges, 1) / m),
                2: np.log(max(r.num_valid - r.found, 1) / n),
                3: np.log1p(r.level),
                4: r.next_size / n,
                5: r.next_edges / m,
                6: np.log1p(r.next_edges / r.next_size) if r.next_size else 0.0,
                7: np.log1p(g.avg_degree),
                8: float(r.bottom_up),
                10: r.found / r.num_valid,
                11: float(r.phase),
            }
            for i, w in want.items():
                assert abs(f[i] - w) <= 1e-4 * max(1.0, abs(w)), (r.phase, r.level, i, f[i], w)


@pytest.mark.gpu
@pytest.mark.parametrize("scale,seed", [(20, 424243), (22, 424244)])
def test_learned_policy_gain(msc, orc, gpu_ctx, scale, seed):
    """What the trained network gains over the plain ALPHA/BETA rule on held-out R-MAT
    graphs (seeds outside the training set): the FW-BW searches' summed level durations
    must drop by at least 1.3x, with labels identical to the plain run and the oracle."""
    import os

    if not os.path.exists(msc.POLICY_WEIGHTS):
        pytest.skip("no trained weights shipped yet")
    msc.load_policy(gpu_ctx)
    g = msc.gen_rmat(scale, 16, seed=seed)
    dg = msc.DeviceGraph(gpu_ctx, g)
    want = orc.canonical(orc.scc(g, g.max_deg_vert, g.avg_degree, vert_cutoff=0))
    cost = {}
    try:
        for name, pol in (("plain", msc.POLICY_PLAIN), ("learned", msc.POLICY_LEARNED)):
            best = None
            for _ in range(3):
                lab = msc.run_scc(dg, g.max_deg_vert, g.avg_degree, ctx=gpu_ctx, policy=pol, timing=True,
                                  record=True)
                assert np.array_equal(lab, want), name
                recs = [r for r in gpu_ctx.levels() if r.phase in (0, 1)]
                us = sum(r.level_us for r in recs)
                edges = sum((r.inspected if r.bottom_up else r.q_edges) for r in recs)
                if best is None or us < best[0]:
                    best = (us, edges, len(recs))
            cost[name] = best
    finally:
        dg.free()
    gain = cost["plain"][0] / cost["learned"][0]
    print(f"R-MAT s{scale} seed {seed}: plain {cost['plain'][0]:.0f} us over {cost['plain'][2]} levels "
          f"({cost['plain'][1]} edges), learned {cost['learned'][0]:.0f} us over {cost['learned'][2]} levels "
          f"({cost['learned'][1]} edges): {gain:.2f}x", flush=True)
    assert gain >= 1.3, cost
