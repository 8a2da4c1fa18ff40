"""msc_opts.replay: a call captured into a HIP graph and replayed on later identical calls
gives exactly the labels (and statistics) of an ordinary call, on every replay, with
ordinary calls on other graphs and host synchronisations in between, after a tuning change
(which captures anew) and in both label modes, for SCC and WCC."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs(msc):
    return [("rmat16", msc.gen_rmat(16, 16, seed=3)), ("grid256", msc.gen_grid(256, 256, seed=4)),
            ("planted", msc.gen_planted(20000, max_size=2000, seed=6)[0])]


@pytest.mark.parametrize("kind", ["scc", "cc"])
def test_replay_matches_ordinary_calls(msc, orc, gpu_ctx, kind):
    other = msc.gen_rmat(12, 8, seed=9)  # ordinary work between the replays
    other_want = orc.canonical(orc.scc(other, vert_cutoff=0) if kind == "scc" else orc.cc(other))
    run = msc.run_scc if kind == "scc" else msc.run_cc
    defaults = gpu_ctx.tuning()
    try:
        for name, g in _graphs(msc):
            piv, avg = (orc.pivot_scc(g) if kind == "scc" else orc.pivot_cc(g))
            raw = orc.scc(g, piv, avg, vert_cutoff=0) if kind == "scc" else orc.cc(g, piv, avg)
            dg = msc.DeviceGraph(gpu_ctx, g)
            out = msc.DeviceLabels(gpu_ctx, g.n)
            try:
                for labels, want in ((msc.LABELS_RAW, raw), (msc.LABELS_CANONICAL, orc.canonical(raw))):
                    _, ref = run(dg, piv, avg, ctx=gpu_ctx, out=out, labels=labels, stats=True)
                    assert np.array_equal(out.to_host(), want), (name, "ordinary call")
                    for it in range(4):
                        print(f"case: replay {kind} {name} labels={labels} call {it}", flush=True)
                        _, st = run(dg, piv, avg, ctx=gpu_ctx, out=out, labels=labels, stats=True, replay=True)
                        got = out.to_host()
                        assert np.array_equal(got, want), (name, labels, it, int((got != want).sum()))
                        assert (st.num_trim, st.num_fwbw, st.color_rounds) == \
                            (ref.num_trim, ref.num_fwbw, ref.color_rounds), (name, it)
                        # ordinary work and a synchronisation between replays
                        assert np.array_equal(run(other, ctx=gpu_ctx), other_want)
                        msc._lib().msc_ctx_sync(gpu_ctx.handle)
                    # a tuning change is part of the captured call: it captures anew
                    gpu_ctx.set_tuning(local_us=120 if gpu_ctx.tuning()["local_us"] != 120 else 200)
                    run(dg, piv, avg, ctx=gpu_ctx, out=out, labels=labels, replay=True)
                    assert np.array_equal(out.to_host(), want), (name, "after a tuning change")
                    run(dg, piv, avg, ctx=gpu_ctx, out=out, labels=labels, replay=True)
                    assert np.array_equal(out.to_host(), want), (name, "replay after the tuning change")
                    gpu_ctx.set_tuning(**defaults)
            finally:
                out.free()
                dg.free()
    finally:
        gpu_ctx.set_tuning(**defaults)


def test_replay_ignored_where_it_does_not_apply(msc, orc, gpu_ctx):
    """Host graphs, host labels and a device-computed pivot run as ordinary calls."""
    g = msc.gen_rmat(13, 8, seed=2)
    want = orc.canonical(orc.scc(g, vert_cutoff=0))
    for _ in range(2):
        assert np.array_equal(msc.run_scc(g, ctx=gpu_ctx, replay=True), want)
        assert np.array_equal(msc.run_scc(g, g.max_deg_vert, g.avg_degree, ctx=gpu_ctx, replay=True), want)
